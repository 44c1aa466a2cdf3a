// FP64 operator application with no device allocation (see noalloc.hip).
#pragma once

#include "fft_fast.h"
#include "runtime.h"

namespace tdda {

// fixed device scratch: the job descriptor, then the pass maps (8 bytes per tile)
constexpr size_t kNoAllocScratchBytes = size_t(256) << 10;
constexpr size_t kNoAllocJobBytes = 1024;

// y = (diag(al) + A) x with FP64 transforms for one system; x, y: device, 3 ndip
// complex.  Enqueued on st and not waited for; allocates nothing.  Throws
// (ERR_MEMORY) when the system's pass maps exceed the scratch.
void apply_noalloc(System& s, const double2* x, double2* y, hipStream_t st);

}  // namespace tdda

// y = (diag(al) + A) x in FP64 for one problem with no device allocation:
// the job descriptor and the (job, tile) maps of the five passes, which the
// allocating apply keeps in library-allocated buffers, live in a fixed
// device array of this file instead.  For callers that cannot allocate
// (HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
#include "noalloc.h"

#include <mutex>

namespace tdda {

namespace {

// Every call on a GPU shares that GPU's copy: a call rewrites all of it (the
// descriptor and every map entry it launches over), so nothing has to start at
// zero, and two calls that overlap on one GPU race.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device (the array's name is
// a host-side stand-in and must not reach a kernel or a copy).
void* scratch_address() {
    int dev = 0;
    TDDA_HIP(hipGetDevice(&dev));
    TDDA_REQUIRE(dev >= 0 && dev < kMaxDevices, ERR_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[dev] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            throw Error(ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

// Region starts of the maps in scratch (after the descriptor), in launch order.
struct NaLayout {
    int start[NPASS + 1];   // PASS_X, PASS_Y, PASS_Z, PASS_XF, end (map entries)
};

// job -> scratch[0]; maps: region k holds (0, t) for its tiles t
__global__ void k_noalloc_setup(GridJob J, unsigned char* base, NaLayout lay) {
    GridJob* jobs = reinterpret_cast<GridJob*>(base);
    int2* maps = reinterpret_cast<int2*>(base + kNoAllocJobBytes);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) jobs[0] = J;
    if (i >= lay.start[NPASS]) return;
    int k = 0;
    while (i >= lay.start[k + 1]) ++k;
    maps[i] = make_int2(0, i - lay.start[k]);
}

}  // namespace

void apply_noalloc(System& s, const double2* x, double2* y, hipStream_t st) {
    static_assert(sizeof(GridJob) <= kNoAllocJobBytes, "descriptor region too small");
    TDDA_REQUIRE(x != nullptr && y != nullptr, ERR_INVALID, "null vector");
    GridJob J = s.job;
    J.xin = x;
    J.yout = y;
    J.ks = nullptr;
    const int tiles[NPASS] = {J.tx64, J.ty64, J.tz64, J.txf64};
    NaLayout lay{};
    for (int k = 0; k < NPASS; ++k) lay.start[k + 1] = lay.start[k] + tiles[k];
    const size_t need = kNoAllocJobBytes + (size_t)lay.start[NPASS] * sizeof(int2);
    TDDA_REQUIRE(need <= kNoAllocScratchBytes, ERR_MEMORY,
                 "problem needs " + std::to_string(need) + " bytes of pass maps; the no-allocation apply holds " +
                     std::to_string(kNoAllocScratchBytes) + " (use tdda_problems_apply)");
    unsigned char* base = static_cast<unsigned char*>(scratch_address());
    const int n = lay.start[NPASS];
    hipLaunchKernelGGL(k_noalloc_setup, dim3((n + 255) / 256), dim3(256), 0, st, J, base, lay);
    TDDA_HIP(hipGetLastError());

    LaunchSet ls;
    ls.jobs = reinterpret_cast<const GridJob*>(base);
    ls.njobs = 1;
    ls.refined = s.refined;   // FP64 transforms; the plain mode reads and writes x, y in FP64
    const int2* maps = reinterpret_cast<const int2*>(base + kNoAllocJobBytes);
    const int fast[NPASS] = {J.fx, J.fy, J.fz, J.fx};
    for (int k = 0; k < NPASS; ++k) {
        PassMap& pm = ls.pm[1][k][fast[k]];
        pm.map = maps + lay.start[k];
        pm.n = tiles[k];
        pm.lds = lds_bytes(J, k, false);
        if (k == PASS_Z && J.fz) {
            const int zc = z_class(J.lz, J.var[1][PASS_Z]);
            pm.cls = 1 << zc;
            pm.lds_cls[zc] = pm.lds;
        }
    }
    launch_matvec(ls, PlanRegistry::instance().pools(), false, XF_PLAIN, XI_PLAIN, GATE_ALWAYS, st);
}

}  // namespace tdda

// Green's-tensor setup kernels (FP64).
#pragma once

#include "common.h"

namespace tdda {

struct GreenParams {
    int nx, ny, nz;
    double d;        // dipole spacing (nm)
    double K;        // 2 pi nback / lambda
    int method;      // 0 FCD, 1 LDR
    int maxm, maxn;  // image range
    double lm, ln;   // image displacement (nm)
    long nsici;      // Si/Ci table length
    double delta;    // Si/Ci table spacing assumed by the lookup
};

// octant tensor [6][nz][ny][nx]
void launch_green_octant(const GreenParams& P, const double* si, const double* ci, double2* out, hipStream_t st);
// one parity-reduced DFT axis: in [6][outer][no][inner] -> out [6][outer][h][inner]
void launch_green_dft(const double2* in, double2* out, int outer, int no, int h, int inner, int L, int axis,
                      const double2* tw, hipStream_t st);
// [6][rows][hx] component-major spectrum (rows = hz * hy) -> the layouts the
// Z passes read, where a run of consecutive kx is one contiguous range per
// load: FP64 [rows][6][hx] (double2), c64 [rows][3][hx] of float4 holding
// the component pairs (00,01), (02,11), (12,22).
void launch_green_interleave(const double2* in, long rows, int hx, double2* out64, float2* out32, hipStream_t st);

}  // namespace tdda

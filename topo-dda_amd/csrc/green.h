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

// Batched set-up: the same steps for several operators, one launch per step.
// A job's blocks are [block0, block0 + its block count); jobs are listed in
// increasing block0 and a block finds its job by bisection.
struct GreenSingleJob {   // octant tensor without images
    GreenParams P;
    const double *si, *ci;
    double2* out;
    int block0;
};
struct GreenDftJob {      // one DFT axis (as launch_green_dft)
    const double2* in;
    double2* out;
    const double2* tw;
    int outer, no, h, inner, L, axis, ng;
    int block0;
};
struct GreenInterleaveJob {
    const double2* in;
    double2* out64;
    float2* out32;
    long rows;
    int hx;
    int block0;
};
// outputs per thread (8 or 16) and blocks of one DFT axis
int green_dft_kb(int h);
int green_dft_blocks(int outer, int h, int inner);
inline int green_single_blocks(const GreenParams& P) { return (int)(((long)P.nx * P.ny * P.nz + 255) / 256); }
inline int green_interleave_blocks(long rows, int hx) { return (int)((6 * rows * hx + 255) / 256); }
// jobs: device tables; blocks: total over the jobs
void launch_green_single_batch(const GreenSingleJob* jobs, int nj, int blocks, hipStream_t st);
void launch_green_dft_batch(const GreenDftJob* jobs, int nj, int blocks, int kb, hipStream_t st);
void launch_green_interleave_batch(const GreenInterleaveJob* jobs, int nj, int blocks, hipStream_t st);

}  // namespace tdda

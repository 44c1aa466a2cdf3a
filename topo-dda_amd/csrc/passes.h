// Launch interface of the FFT-convolution passes and the Krylov kernels.
#pragma once

#include "dda_types.h"

namespace tdda {

enum Gate : int { GATE_ALWAYS = 0, GATE_ACTIVE = 1, GATE_RU = 2 };
// X forward pass: what feeds the transform
enum XfMode : int {
    XF_PLAIN = 0,   // u = xin
    XF_P = 1,       // p = r + beta (p - zeta v); u = p   (DDAModel.cpp:976)
    XF_T = 2,       // t = r - alpha v;           u = t   (DDAModel.cpp:979)
    XF_SOL = 3,     // u = P (FP64 solution; reliable update / warm start)
};
// X inverse pass: what is done with y = A u + al u
enum XiMode : int {
    XI_PLAIN = 0,   // yout = y
    XI_V = 1,       // v = y, alpha = rho / <r0, v>
    XI_S = 2,       // s = y, zeta = <s,t>/<s,s>
    XI_RESID = 3,   // r = rhs - y (true residual), reliable-update bookkeeping
};

// Per-precision caps on pencils*length per tile.
template <typename Real> struct Caps;
// (the X pass holds three components per row, so 3*Lx <= X is required)
template <> struct Caps<float> { static constexpr int X = 2048, Y = 4096, Z = 2048; };
template <> struct Caps<double> { static constexpr int X = 1024, Y = 2048, Z = 1024; };

// A launch over a batch: device arrays of jobs and of (job, tile) per block.
struct LaunchSet {
    const GridJob* jobs = nullptr;   // device
    const int2* map_x32 = nullptr; int n_x32 = 0;
    const int2* map_y32 = nullptr; int n_y32 = 0;
    const int2* map_z32 = nullptr; int n_z32 = 0;
    const int2* map_x64 = nullptr; int n_x64 = 0;
    const int2* map_y64 = nullptr; int n_y64 = 0;
    const int2* map_z64 = nullptr; int n_z64 = 0;
    const int2* map_upd = nullptr; int n_upd = 0;
    size_t lds_x32 = 0, lds_y32 = 0, lds_z32 = 0, lds_x64 = 0, lds_y64 = 0, lds_z64 = 0;
    bool refined = false;            // Krylov vectors are c64
};

// Twiddle pools (device), indexed by FftPlan::tw_off
struct TwiddlePools {
    const float2* tw32 = nullptr;
    const double2* tw64 = nullptr;
};

// y = A u + al u over the batch (gate decides which jobs run).
// Real: transform precision; krylov_f32: Krylov vectors are c64.
void launch_matvec(const LaunchSet& ls, const TwiddlePools& tw, bool c64, int xf_mode, int xi_mode,
                   int gate, hipStream_t st);
// Same launches as launch_matvec (plain mode), each bracketed by HIP events on
// `st`; adds the elapsed ms of X fwd, Y fwd, Z, Y inv, X inv to ms5.
void launch_matvec_timed(const LaunchSet& ls, const TwiddlePools& tw, bool c64, hipStream_t st, int reps,
                         double* ms5);
// Elementwise BiCGSTAB update and convergence logic.
void launch_update(const LaunchSet& ls, int gate, hipStream_t st);
// Initial residual, r0, p = v = 0 and scalars; warm: r = rhs - yout.
void launch_init(const LaunchSet& ls, bool warm, hipStream_t st);

// Host-side tile geometry for a job (fills the b*/t* fields).
void plan_tiles(GridJob& J);
size_t lds_bytes_x(const GridJob& J, bool c64);
size_t lds_bytes_y(const GridJob& J, bool c64);
size_t lds_bytes_z(const GridJob& J, bool c64);
constexpr int UPD_EPT = 4;   // elements per thread in the update kernel

}  // namespace tdda

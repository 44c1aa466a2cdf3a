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

// One launch of one pass: device map of (job, tile) per block, its length,
// and the dynamic LDS bytes (the largest tile of any job in the map).
struct PassMap {
    const int2* map = nullptr;
    int n = 0;
    size_t lds = 0;
    const int* limit = nullptr;   // device count of valid map entries (blocks beyond it return)
    // specialised Z pass: bit c is set when a job of length class c is in the
    // map (each class is its own kernel), lds_cls[c] the LDS of that launch
    int cls = 0;
    size_t lds_cls[2] = {0, 0};
};
// PASS_X: X inverse tiles (its dot partials are per tile); PASS_XF: X forward
// tiles, sized apart from the inverse.
enum PassKind : int { PASS_X = 0, PASS_Y = 1, PASS_Z = 2, PASS_XF = 3 };
static_assert(PASS_X == 0 && PASS_Y == 1 && PASS_Z == 2 && PASS_XF == 3, "with_split_pass numbers the passes so");
constexpr int NPASS = 4;
constexpr int RU_REGIONS = 2 * NPASS;

// Length classes of the fused y/z pass (c64): Ly = 16 LY1 with LY1 in
// {2, 3, 4, 6, 8}, Lz in {32, 48, 64}.
constexpr int NYZ = 15;
// Classes of the plane passes (fft_plane.h: L in {32, 48, 64, 96, 128})
constexpr int NXY_CLASSES = 5;

// A launch over a batch: device array of jobs and, per transform precision
// and pass, the blocks of the generic kernels and of the length-specialised
// kernels (a job uses the specialised kernel of a pass when that pass's
// length has a compile-time split).
struct LaunchSet {
    const GridJob* jobs = nullptr;   // device
    int njobs = 0;
    PassMap pm[2][NPASS][2];
    // device count of jobs (of any batch) that wait for a reliable update
    int* ru_jobs = nullptr;
    // reliable-update launches run over compacted copies of the c128 maps
    // (only tiles of waiting jobs): ru_map holds the RU_REGIONS regions in
    // the order of pm[1] (pass-major, generic then specialised) at ru_off[k],
    // ru_count[k] their filled lengths
    int2* ru_map = nullptr;
    int* ru_count = nullptr;
    const int2* ru_src = nullptr;    // the c128 maps, contiguous, same regions
    const int* ru_ranges = nullptr;  // [RU_REGIONS][njobs + 1]: start of job j's tiles in region k
    int ru_off[RU_REGIONS + 1] = {};
    const int2* map_upd = nullptr;
    int n_upd = 0;
    bool refined = false;            // Krylov vectors are c64
    // fused y/z pass of the c64 matvec, one map per length class; its jobs
    // are in no c64 Y or Z map
    PassMap yz[NYZ];
    // plane passes of the c64 matvec, one (job, z) map per class; their jobs
    // are in no c64 X, X-forward or Y map
    PassMap xy[NXY_CLASSES];
    int n_xy = 0;   // blocks over all classes
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
                         double* ms5, int xf_mode = XF_PLAIN, int xi_mode = XI_PLAIN);
// The end-of-step vector update of a BiCGSTAB step is applied by the X
// forward pass of the next matvec; jobs that stop at a step's decision still
// owe it: call after every batch of iterations (no-op for the others).
void launch_finalize(const LaunchSet& ls, hipStream_t st);
// Initial residual, r0, p = v = 0 and scalars; warm: r = rhs - yout.
void launch_init(const LaunchSet& ls, bool warm, hipStream_t st);
// Every job's scalars (GridJob::ks) set to t, in one launch.
void launch_ks_template(const LaunchSet& ls, const KScalars& t, hipStream_t st);

// The device counter behind LaunchSet::ru_jobs (one per process, zeroed).
int* ru_job_counter();

// The tuned schedule (runtime.cpp): variant of a pass at length L for
// precision pr (0 c64, 1 c128; 0 when none is recorded), and whether the
// fused y/z pass is enabled for (Ly, Lz) (off when none is recorded).
int schedule_variant(int pr, int pass, int L);
bool schedule_fused(int ly, int lz);
bool schedule_plane(int L);
// Host-side tile geometry for a job (fills the b*/t*/f*/var fields).
void plan_tiles(GridJob& J);
// Dynamic LDS bytes of the job's tile in a pass (generic or specialised).
size_t lds_bytes(const GridJob& J, int pass, bool c64);

// Length-specialised kernels (passes_fast.hip).
void launch_fast_xfwd(const PassMap& pm, const GridJob* jobs, const TwiddlePools& tw, bool c64, bool krylov32,
                      int mode, int gate, hipStream_t st);
void launch_fast_xinv(const PassMap& pm, const GridJob* jobs, const TwiddlePools& tw, bool c64, bool krylov32,
                      int mode, int gate, hipStream_t st);
void launch_fast_y(const PassMap& pm, const GridJob* jobs, const TwiddlePools& tw, bool c64, int dir, int gate,
                   hipStream_t st);
void launch_fast_z(const PassMap& pm, const GridJob* jobs, const TwiddlePools& tw, bool c64, int gate,
                   hipStream_t st);
// Fused y/z pass (passes_yz.hip): class of a job's lengths (-1: none), LDS
// bytes of a class, and the launches of every non-empty class map.
int yz_class(int ly, int lz, int ny, int nz);
size_t yz_lds_bytes(int cls);
void launch_fast_yz(const PassMap* yz, const GridJob* jobs, const TwiddlePools& tw, int gate, hipStream_t st);
// Which eligible c64 jobs batches built from now on run through the fused
// y/z pass: FUSED_OFF none, FUSED_SCHEDULE those whose (Ly, Lz) the schedule
// enables (the default), FUSED_ALL all.
enum FusedMode : int { FUSED_OFF = 0, FUSED_SCHEDULE = 1, FUSED_ALL = 2 };
void set_fused_yz(int mode);
int fused_yz();
// Plane passes (passes_xy.hip): X forward + Y forward and Y inverse + X
// inverse of c64 matvecs, one workgroup per (job, z-plane) of every
// non-empty class map; modes XF_PLAIN/XF_P/XF_T and XI_PLAIN/XI_V/XI_S.
void launch_plane_fwd(const PassMap* xy, const GridJob* jobs, const TwiddlePools& tw, int mode, int gate,
                      hipStream_t st);
void launch_plane_inv(const PassMap* xy, const GridJob* jobs, const TwiddlePools& tw, int mode, int gate,
                      hipStream_t st);
size_t plane_lds_bytes(int cls, bool inverse);
int plane_class_of(int lx, int ly, int nx, int ny);
// Whether batches built from now on run eligible c64 jobs through the plane
// passes (on by default).
void set_plane_passes(bool on);
bool plane_passes();
// Rows (X) or pencils (Y, Z) per tile and LDS bytes of the specialised path
// for length L; false if L has no specialised split.
bool fast_tiles(int L, int pass, bool c64, int v, int* per_tile, size_t* lds_bytes);
constexpr int UPD_EPT = 4;   // elements per thread in the update kernel

}  // namespace tdda

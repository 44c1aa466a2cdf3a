// Device-visible descriptors of one linear system ("job") and its Krylov state.
#pragma once

#include "common.h"
#include "fft_core.h"

namespace tdda {

// Solver status codes (device and host)
enum JobStatus : int {
    JOB_ACTIVE = 0,
    JOB_CONVERGED = 1,
    JOB_BREAKDOWN = 2,   // <r0, r_prev> == 0 (DDAModel.cpp:952-970)
    JOB_MAXIT = 3,       // ran out of iterations (DDAModel.cpp:1005-1008)
    JOB_IDLE = 4,        // not started / retired
};

// BiCGSTAB scalars of one job, all FP64 (DDAModel.cpp:920-1004 names:
// alpha, beta, zeta; rho = <r0, r>).
struct KScalars {
    double2 alpha, beta, zeta;
    double2 rho;        // <r0, r_k>, numerator of alpha in the current step
    double2 rho_prev;   // <r0, r_{k-1}>  (the reference's r0.dot(rl))
    double enorm2;      // ||rhs||^2
    double rnorm2;      // ||r||^2 of the latest residual (recursive or true)
    double maxr2;       // reliable-update reference level
    double tol;         // relative tolerance
    double ru_delta;    // reliable-update trigger ratio (refined mode)
    double err;         // final ||r||/||rhs||
    int iter;           // loop index of the step in flight (-1: initial step)
    int maxit;
    int status;         // JobStatus
    int need_ru;        // refined mode: true-residual recompute pending
    int refined;        // 1: FP32 Krylov with FP64 reliable updates
    int iterations;     // reported iteration count (reference ITERATION)
    int n_ru;           // number of reliable updates done
    int pending;        // 1: P += alpha p + zeta t, r = t - zeta s of the last step not yet applied
    int dirty;          // refined mode: dP may hold steps not yet added to P
    // Refined mode, field-error control.  The relative error of the local
    // field E = f P (f = al 3/(eps + 2) per dipole) behaves as amp * ||r||/||rhs||
    // over a solve, with amp a property of the system; amp is measured from
    // the change of E between two reliable updates whose true residuals are
    // well apart: ||dE||/||E|| ~ amp (r_prev - r_now).  A solve converges when
    // the true residual meets tol and safety * amp * residual meets tol_e.
    double tol_e;       // relative field-error target (0: residual only)
    double res_target;  // relative residual that triggers the next reliable update at convergence
    double amp;         // current amplification estimate
    double prev_true;   // relative true residual of the last reliable update (0: none yet)
    double err_est;     // estimated relative field error of the returned solution
};

// prior amplification before any measurement, and the safety factor on the
// estimate (measured amplifications of the DDA systems are 1.5 to 3; over 96
// served requests the estimate was 0.69 to 2.3 times the actual field error)
constexpr double kAmpPrior = 3.0;
constexpr double kErrSafety = 2.0;
// a measurement needs the true residual to have fallen by this factor since
// the previous reliable update
constexpr double kAmpSeparation = 0.5;

// doubles per tile in the per-job partial-sum buffer
constexpr int PARTIAL_STRIDE = 8;
// doubles per X forward tile of a reliable update: ||f dP||^2, ||f P||^2
constexpr int XF_NACC = 2;

// One system (diag(al) + A) P = rhs on an Nx x Ny x Nz box.
struct GridJob {
    int nx, ny, nz, ndip;
    int lx, ly, lz;
    int hx, hy, hz;            // L/2 + 1: extents of the octant spectrum
    FftPlan px, py, pz;
    int bx32, by32, bz32;      // pencils (x: rows) per tile, c64 launches
    int bx64, by64, bz64;      // same for c128 launches
    int tx32, ty32, tz32;      // tiles per pass, c64
    int tx64, ty64, tz64;      // tiles per pass, c128
    int bxf32, bxf64, txf32, txf64;   // rows per tile / tiles of the X forward pass
    int tupd;                  // tiles of the elementwise update kernel
    int fx, fy, fz;            // 1: the pass uses the length-specialised kernels
    int var[2][4];             // schedule variant per [c64, c128][PassKind] (with_split_pass)
    // fused y/z pass of the c64 matvec: its length class (-1: none); tr32 = 1
    // in a batch that runs this job through it (the X passes then keep the
    // work grid as [kx][3][nz][ny])
    int yzc, tr32;
    // plane passes of the c64 matvec (X and Y transforms of a z-plane in one
    // kernel per direction): the class of (Lx, Ly, Nx, Ny) (-1: none); xy32 = 1
    // in a batch that runs this job through them (its X inverse partials are
    // then per plane: tx32 = nz in that batch's copy)
    int xyc, xy32;
    unsigned long long mlx, mny;   // ceil(2^32 / lx), ceil(2^32 / ny): fast quotients
    const int* box_map;        // [nz][ny][nx] -> dipole or -1; null: identity
    // operator data
    const float2* green32;     // octant spectrum, [hz][hy][3][hx] of component pairs
                               // (FP64: [hz][hy][6][hx])
    const double2* green64;
    const float2* al32;        // [ndip] 1/alpha
    const double2* al64;
    // Right Jacobi preconditioning (jacobi != 0): BiCGSTAB runs on
    // (A diag(al)^-1) Y = rhs with P = Y / al.  The residual is that of the
    // original system, so the stopping rule is unchanged; the transform input
    // is the Krylov vector times 1/al and the X inverse epilogue adds the
    // vector itself (al (x / al) = x).  P holds Y during a solve and is
    // scaled back at its end.
    int jacobi;
    const float2* ial32;       // [ndip] alpha (= 1 / (1/alpha))
    const double2* ial64;
    // work grid [3][nz][ly][lx]; the c64 view aliases the c128 buffer
    void* grid;
    // Krylov vectors, 3*ndip entries each, Krylov precision (float2 or double2)
    void *r, *r0, *p, *v, *t, *s;
    double2* P;                // solution, FP64
    // refined mode: c64 sum of the steps since the last reliable update (the
    // solution is P + dP; reliable updates and the end of a solve fold it
    // into P); null in FP64 mode
    void* dP;
    const double2* rhs;        // right-hand side, FP64
    // plain operator application y = A x + al x (tests, gradients)
    const void* xin;
    void* yout;
    KScalars* ks;
    double* partial;           // [max tiles][PARTIAL_STRIDE]
    // refined mode: partial sums of the reliable update's X forward pass
    // (||f dP||^2, ||f P||^2 per tile), apart from the X inverse partials
    double* partial_xf;        // [X forward tiles][2]
    const float2* efac;        // [ndip] local-field factor f = al 3/(eps + 2); null: f = 1
    // r0_inc = 1: the shadow residual r0 of BiCGSTAB is the incident field
    // (a solve from P = 0 whose right-hand side is that field, full box): the
    // X inverse epilogues evaluate it (incident_phase / incident_wave with
    // K d, n_k, e0 n_e0) instead of loading r0
    int r0_inc;
    double inc_kd;
    double inc_nk[3];
    double inc_en[3];
};

}  // namespace tdda

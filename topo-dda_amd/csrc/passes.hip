// FFT-convolution matvec passes and fused BiCGSTAB kernels (gfx950).
//
// The operator y = A u + al*u (AProductCore::Aproduct + DDAModel::Aproductwithalb,
// src/AProductCore.cpp:461-515, src/DDAModel.cpp:1194-1203) is evaluated on an
// Lx*Ly*Lz grid (L >= 2N-1 per axis, 2/3/5-smooth) in five passes that keep only
// the non-zero part of the zero-padded grid resident:
//
//   X fwd : dipoles (x < Nx) -> rows (z < Nz, y < Ny, all x)        [fused vector op]
//   Y fwd : rows y < Ny -> all y
//   Z     : z < Nz -> FFT -> 3x3 symmetric Green's multiply -> IFFT -> z < Nz
//   Y inv : all y -> y < Ny
//   X inv : all x -> dipoles (x < Nx), /M, + al*u                    [fused epilogue + dots]
//
// The work grid is [3][Nz][Ly][Lx] (component, z, y, x; x fastest).  Every
// launch covers a batch of jobs; block b handles (job, tile) = map[b].  Tiles
// depend only on a job's own shape, and reductions are per job in a fixed
// order, so a job's result does not depend on what else is in the batch.
#include "fft_fast.h"
#include "pass_util.h"

namespace tdda {

// ----------------------------------------------------------------------------
// X forward pass
// ----------------------------------------------------------------------------
template <typename Real, typename KReal, int MODE>
__global__ void __launch_bounds__(FFT_THREADS)
k_xfwd(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap, const cx_t<Real>* __restrict__ tw, int gate, const int* __restrict__ limit) {
    using C = cx_t<Real>;
    using KC = cx_t<KReal>;
    constexpr int CAP = Caps<Real>::X;
    extern __shared__ __align__(16) unsigned char smem_x[];
    C* lds = reinterpret_cast<C*>(smem_x);
    if (beyond_limit(limit)) return;
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    if (!gate_open(J, gate)) return;
    const int bx = Tile<Real>::bxf(J);
    const int nx = J.nx, ny = J.ny, lx = J.lx;
    const int nrows = ny * J.nz;
    const int q0 = jt.y * bx;
    const int nb = min(bx, nrows - q0);
    const int np = 3 * nb;
    const int bp = np | 1;
    const int nel = 3 * nx * nb;
    const XfScalars sc = xf_scalars<MODE>(J);
    double xacc[XF_NACC] = {};
    for (int e = threadIdx.x; e < nel; e += FFT_THREADS) {
        const int c = e % 3;
        const int rem = e / 3;
        const int x = rem % nx;
        const int b = rem / nx;
        const int vi = vec_index(J, q0 + b, x, c);
        lds[x * bp + 3 * b + c] = xf_element<C, KC, MODE>(J, vi, sc, xacc);
    }
    xf_finish<MODE>(J, jt.y, xacc);
    __syncthreads();
    lds_fft<-1, CAP, 1>(J.px, lds, bp, np, nx, tw);
    C* g = static_cast<C*>(J.grid);
    const long S = (long)J.nz * J.ly * lx;
    const int nout = np * lx;
    for (int e = threadIdx.x; e < nout; e += FFT_THREADS) {
        const int x = e % lx;
        const int rem = e / lx;
        const int b = rem % nb;
        const int c = rem / nb;
        const int q = q0 + b;
        const int z = q / ny;
        const int y = q - z * ny;
        g[c * S + ((long)z * J.ly + y) * lx + x] = lds[x * bp + 3 * b + c];
    }
}

// ----------------------------------------------------------------------------
// X inverse pass with the operator epilogue
// ----------------------------------------------------------------------------
template <typename Real, typename KReal, int MODE>
__global__ void __launch_bounds__(FFT_THREADS)
k_xinv(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap, const cx_t<Real>* __restrict__ tw, int gate, const int* __restrict__ limit) {
    using C = cx_t<Real>;
    using KC = cx_t<KReal>;
    constexpr int CAP = Caps<Real>::X;
    extern __shared__ __align__(16) unsigned char smem_xi[];
    C* lds = reinterpret_cast<C*>(smem_xi);
    if (beyond_limit(limit)) return;
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    if (!gate_open(J, gate)) return;
    const int bx = Tile<Real>::bx(J);
    const int nx = J.nx, ny = J.ny, lx = J.lx;
    const int nrows = ny * J.nz;
    const int q0 = jt.y * bx;
    const int nb = min(bx, nrows - q0);
    const int np = 3 * nb;
    const int bp = np | 1;
    const C* g = static_cast<const C*>(J.grid);
    const long S = (long)J.nz * J.ly * lx;
    const int nin = np * lx;
    for (int e = threadIdx.x; e < nin; e += FFT_THREADS) {
        const int x = e % lx;
        const int rem = e / lx;
        const int b = rem % nb;
        const int c = rem / nb;
        const int q = q0 + b;
        const int z = q / ny;
        const int y = q - z * ny;
        lds[x * bp + 3 * b + c] = g[c * S + ((long)z * J.ly + y) * lx + x];
    }
    __syncthreads();
    lds_fft<1, CAP, 1>(J.px, lds, bp, np, lx, tw);

    const Real inv_m = (Real)(1.0 / ((double)J.lx * (double)J.ly * (double)J.lz));
    const C* al = Tile<Real>::al(J);
    double acc[XI_NACC] = {};
    const int nel = 3 * nx * nb;
    for (int e = threadIdx.x; e < nel; e += FFT_THREADS) {
        const int c = e % 3;
        const int rem = e / 3;
        const int x = rem % nx;
        const int b = rem / nx;
        const int vi = vec_index(J, q0 + b, x, c);
        if (vi < 0) continue;
        xi_element<C, KC, MODE>(J, vi, cscale(lds[x * bp + 3 * b + c], inv_m), al, acc);
    }
    xi_finish<MODE>(J, jt.y, acc);
}

// ----------------------------------------------------------------------------
// Y pass (strided pencils, global -> global through LDS)
// ----------------------------------------------------------------------------
template <typename Real, int DIR>
__global__ void __launch_bounds__(FFT_THREADS)
k_ypass(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap, const cx_t<Real>* __restrict__ twpool, int gate, const int* __restrict__ limit) {
    using C = cx_t<Real>;
    constexpr int CAP = Caps<Real>::Y;
    extern __shared__ __align__(16) unsigned char smem_y[];
    C* lds = reinterpret_cast<C*>(smem_y);
    if (beyond_limit(limit)) return;
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    if (!gate_open(J, gate)) return;
    const int by = Tile<Real>::by(J);
    const int lx = J.lx, ly = J.ly;
    const int npen = 3 * J.nz * lx;
    const int pl0 = jt.y * by;
    const int np = min(by, npen - pl0);
    const int bp = np | 1;
    const int n_in = DIR < 0 ? J.ny : ly;
    const int n_out = DIR < 0 ? ly : J.ny;
    C* g = static_cast<C*>(J.grid);
    const FftPlan& pl = J.py;
    const int L = pl.L;
    const cx_t<Real>* tw = twpool + pl.tw_off;
    auto gaddr = [&](int b, int pos) -> long {
        const int p = pl0 + b;
        const int cz = p / lx;
        const int x = p - cz * lx;
        return ((long)cz * ly + pos) * lx + x;
    };
    auto gload = [&](int, int b, int pos) -> C { return pos < n_in ? g[gaddr(b, pos)] : mk<C>(0.0, 0.0); };
    auto gstore = [&](int, int b, int pos, C v) { if (pos < n_out) g[gaddr(b, pos)] = v; };
    auto lload = [&](int, int b, int pos) -> C { return lds[pos * bp + b]; };
    auto lstore = [&](int, int b, int pos, C v) { lds[pos * bp + b] = v; };
    const int nst = pl.nst;
    if (nst == 1) {
        io_stage_rt<DIR, CAP, 1, C>(pl.radix[0], np, L, 1, tw, gload, gstore, false);
        return;
    }
    int ns = 1;
    int R = stage_radix<DIR>(pl, 0);
    io_stage_rt<DIR, CAP, 1, C>(R, np, L, ns, tw, gload, lstore, false);
    __syncthreads();
    ns *= R;
    for (int s = 1; s < nst - 1; ++s) {
        R = stage_radix<DIR>(pl, s);
        lds_stage_rt<DIR, CAP, 1>(R, lds, bp, np, L, ns, tw);
        ns *= R;
    }
    R = stage_radix<DIR>(pl, nst - 1);
    io_stage_rt<DIR, CAP, 1, C>(R, np, L, ns, tw, lload, gstore, false);
}

// ----------------------------------------------------------------------------
// Z pass: forward FFT, Green's-tensor multiply, inverse FFT, all in one kernel
// ----------------------------------------------------------------------------
// Last forward stage (radix R), multiply in registers, first inverse stage
// (same radix, so every thread already holds its inputs).
template <int R, int CAP, typename C, typename TW, typename LoadF, typename StoreF, typename MulF>
__device__ inline void z_bridge(int np, int L, const TW* __restrict__ tw, LoadF load, StoreF store, MulF mul,
                                bool from_lds) {
    constexpr int U = Umax<R, CAP>::value;
    StageRegs<R, U, 3, C> s;
    const int ns = L / R;
    stage_compute<R, -1, U, 3>(s, np, L, ns, tw, load);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (s.pos[u] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) mul(s.b[u], s.pos[u] + r * ns, s.v[0][u][r], s.v[1][u][r], s.v[2][u][r]);
#pragma unroll
            for (int c = 0; c < 3; ++c) Dft<R, 1, C>::run(s.v[c][u]);
            s.pos[u] *= R;   // inverse stage 0: outputs at j*R + r
        }
    }
    if (from_lds) __syncthreads();
    stage_store<R, U, 3>(s, 1, store);
}

template <int CAP, typename C, typename TW, typename LoadF, typename StoreF, typename MulF>
__device__ inline void z_bridge_rt(int R, int np, int L, const TW* __restrict__ tw, LoadF load, StoreF store,
                                   MulF mul, bool from_lds) {
    switch (R) {
        case 2: z_bridge<2, CAP, C>(np, L, tw, load, store, mul, from_lds); break;
        case 3: z_bridge<3, CAP, C>(np, L, tw, load, store, mul, from_lds); break;
        case 4: z_bridge<4, CAP, C>(np, L, tw, load, store, mul, from_lds); break;
        case 5: z_bridge<5, CAP, C>(np, L, tw, load, store, mul, from_lds); break;
        case 8: z_bridge<8, CAP, C>(np, L, tw, load, store, mul, from_lds); break;
        default: break;
    }
}

template <typename Real>
__global__ void __launch_bounds__(FFT_THREADS)
k_zpass(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap, const cx_t<Real>* __restrict__ twpool, int gate, const int* __restrict__ limit) {
    using C = cx_t<Real>;
    constexpr int CAP = Caps<Real>::Z;
    extern __shared__ __align__(16) unsigned char smem_z[];
    C* lds = reinterpret_cast<C*>(smem_z);
    if (beyond_limit(limit)) return;
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    if (!gate_open(J, gate)) return;
    const int bz = Tile<Real>::bz(J);
    const int lx = J.lx, nz = J.nz;
    const long plane = (long)J.ly * lx;
    const int npen = J.ly * lx;
    const int pl0 = jt.y * bz;
    const int np = min(bz, npen - pl0);
    const int bp = np | 1;
    const long S = (long)nz * plane;
    C* g = static_cast<C*>(J.grid);
    const C* G = Tile<Real>::green(J);
    const FftPlan& pl = J.pz;
    const int L = pl.L;
    const cx_t<Real>* tw = twpool + pl.tw_off;
    auto gload = [&](int c, int b, int pos) -> C {
        return pos < nz ? g[c * S + pos * plane + pl0 + b] : mk<C>(0.0, 0.0);
    };
    auto gstore = [&](int c, int b, int pos, C v) { if (pos < nz) g[c * S + pos * plane + pl0 + b] = v; };
    auto lload = [&](int c, int b, int pos) -> C { return lds[(c * L + pos) * bp + b]; };
    auto lstore = [&](int c, int b, int pos, C v) { lds[(c * L + pos) * bp + b] = v; };
    auto mul = [&](int b, int kz, C& v0, C& v1, C& v2) {
        const int p = pl0 + b;
        const int ky = p / lx;
        const int kx = p - ky * lx;
        green_apply(G, J, kx, ky, kz, v0, v1, v2);
    };
    const int nst = pl.nst;
    if (nst == 1) {
        z_bridge_rt<CAP, C>(pl.radix[0], np, L, tw, gload, gstore, mul, false);
        return;
    }
    // forward stage 0: global -> LDS
    int R = pl.radix[0];
    io_stage_rt<-1, CAP, 3, C>(R, np, L, 1, tw, gload, lstore, false);
    __syncthreads();
    int ns = R;
    for (int s = 1; s < nst - 1; ++s) {
        R = pl.radix[s];
        lds_stage_rt<-1, CAP, 3>(R, lds, bp, np, L, ns, tw);
        ns *= R;
    }
    // last forward stage + multiply + first inverse stage: LDS -> LDS
    R = pl.radix[nst - 1];
    z_bridge_rt<CAP, C>(R, np, L, tw, lload, lstore, mul, true);
    __syncthreads();
    ns = R;
    for (int s = nst - 2; s >= 1; --s) {
        R = pl.radix[s];
        lds_stage_rt<1, CAP, 3>(R, lds, bp, np, L, ns, tw);
        ns *= R;
    }
    // last inverse stage: LDS -> global (z < Nz only)
    R = pl.radix[0];
    io_stage_rt<1, CAP, 3, C>(R, np, L, ns, tw, lload, gstore, false);
}

// ----------------------------------------------------------------------------
// Pending end-of-step update of jobs that stopped at the XI_S decision
// (converged, iteration limit or breakdown with the step's P and r update
// still owed): P += alpha p + zeta t, r = t - zeta s; in refined mode the
// c64 increments dP are folded into P as well.  k_finalize_clear, launched
// after it, clears the flags, so a second pair changes nothing.
// ----------------------------------------------------------------------------
template <typename KReal>
__global__ void __launch_bounds__(FFT_THREADS)
k_finalize(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap) {
    using KC = cx_t<KReal>;
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    const KScalars* ks = J.ks;
    if (ks == nullptr || ks->status == JOB_ACTIVE) return;
    const bool fold = kSplitSolution<KC> && ks->dirty != 0 && J.dP != nullptr;
    if (ks->pending == 0 && !fold) return;
    XfScalars sc{ks->alpha, ks->beta, ks->zeta, 1};
    const long n3 = 3L * J.ndip;
    const long base = (long)jt.y * FFT_THREADS * UPD_EPT;
    const KC* p = static_cast<const KC*>(J.p);
#pragma unroll
    for (int k = 0; k < UPD_EPT; ++k) {
        const long i = base + (long)k * FFT_THREADS + threadIdx.x;
        if (i >= n3) continue;
        if (ks->pending != 0) apply_pending<KC>(J, (unsigned)i, sc, to_d(p[i]), true);
        else fold_solution<KC>(J, (unsigned)i, make_double2(0.0, 0.0));
    }
}

// one thread per job
__global__ void k_finalize_clear(const GridJob* __restrict__ jobs, int njobs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    KScalars* ks = jobs[j].ks;
    if (ks != nullptr && ks->status != JOB_ACTIVE) {
        ks->pending = 0;
        ks->dirty = 0;
    }
}

// ----------------------------------------------------------------------------
// compaction of the c128 pass maps to the tiles of jobs waiting for a
// reliable update (jobs in any order, each job's tiles contiguous; counts
// zeroed by the caller)
// ----------------------------------------------------------------------------
struct RuRegions {
    int off[RU_REGIONS + 1];
};
__global__ void __launch_bounds__(FFT_THREADS)
k_ru_compact(const GridJob* __restrict__ jobs, int njobs, const int2* __restrict__ src, RuRegions rg,
             const int* __restrict__ ranges, int2* __restrict__ dst, int* __restrict__ count,
             const int* __restrict__ ru_jobs) {
    // one block per job; its tiles are contiguous in every region of src
    if (__hip_atomic_load(ru_jobs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
    const int j = blockIdx.x;
    const KScalars* ks = jobs[j].ks;
    if (ks->status != JOB_ACTIVE || ks->need_ru == 0) return;
    __shared__ int base[RU_REGIONS];
    if (threadIdx.x < RU_REGIONS) {
        const int* r = ranges + threadIdx.x * (njobs + 1);
        const int n = r[j + 1] - r[j];
        base[threadIdx.x] = n > 0 ? atomicAdd(&count[threadIdx.x], n) : 0;
    }
    __syncthreads();
    for (int k = 0; k < RU_REGIONS; ++k) {
        const int* r = ranges + k * (njobs + 1);
        const int s0 = r[j], n = r[j + 1] - s0;
        for (int t = threadIdx.x; t < n; t += FFT_THREADS) dst[rg.off[k] + base[k] + t] = src[rg.off[k] + s0 + t];
    }
}

__global__ void k_clear_ints(int* __restrict__ p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// ----------------------------------------------------------------------------
// per-job sums of the X-inverse dot partials and the scalar step (one block
// per job, after the X inverse launches of a matvec)
// ----------------------------------------------------------------------------
template <typename Real, int MODE>
__global__ void __launch_bounds__(FFT_THREADS)
k_xi_reduce(const GridJob* __restrict__ jobs, int gate, int* __restrict__ ru_jobs) {
    if (ru_idle(gate, ru_jobs)) return;
    const GridJob& J = jobs[blockIdx.x];
    if (!gate_open(J, gate)) return;
    double sum[XI_NACC];
    job_sum<XI_NACC>(J, Tile<Real>::tx(J), sum);
    // a reliable update's X forward sums (its tiles are the X forward tiles)
    double xs[XF_NACC] = {0.0, 0.0};
    if (MODE == XI_RESID && J.partial_xf != nullptr) job_sum<XF_NACC>(J.partial_xf, XF_NACC, Tile<Real>::txf(J), xs);
    if (threadIdx.x == 0) xi_decide<MODE>(J.ks, sum, ru_jobs, xs);
}

// ----------------------------------------------------------------------------
// initial state: r = rhs (- A P0 if warm), r0 = r, p = v = 0
// ----------------------------------------------------------------------------
template <typename KReal>
__global__ void __launch_bounds__(FFT_THREADS)
k_init(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap, int warm) {
    using KC = cx_t<KReal>;
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    const long n3 = 3L * J.ndip;
    const long base = (long)jt.y * FFT_THREADS * UPD_EPT;
    double acc[2] = {0.0, 0.0};
    const KC zero = mk<KC>(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < UPD_EPT; ++k) {
        const long i = base + (long)k * FFT_THREADS + threadIdx.x;
        if (i < n3) {
            const double2 b = J.rhs[i];
            double2 rr = b;
            if (warm) rr = csub(b, static_cast<const double2*>(J.yout)[i]);
            const KC rk = from_d<KC>(rr);
            static_cast<KC*>(J.r)[i] = rk;
            static_cast<KC*>(J.r0)[i] = rk;
            static_cast<KC*>(J.p)[i] = zero;
            static_cast<KC*>(J.v)[i] = zero;
            if (J.dP != nullptr) static_cast<KC*>(J.dP)[i] = zero;
            const double2 rd = to_d(rk);
            acc[0] += rd.x * rd.x + rd.y * rd.y;
            acc[1] += b.x * b.x + b.y * b.y;
        }
    }
    block_sum<2>(acc);
    write_partials<2>(J, jt.y, acc);
}

template <typename KReal>
__global__ void __launch_bounds__(FFT_THREADS)
k_init_reduce(const GridJob* __restrict__ jobs, int warm) {
    const GridJob& J = jobs[blockIdx.x];
    double sum[2];
    job_sum<2>(J, J.tupd, sum);
    if (threadIdx.x != 0) return;
    KScalars* ks = J.ks;
    ks->alpha = make_double2(0, 0);
    ks->beta = make_double2(0, 0);
    ks->zeta = make_double2(0, 0);
    ks->rho = make_double2(sum[0], 0.0);
    ks->rho_prev = make_double2(0, 0);
    ks->enorm2 = sum[1];
    ks->rnorm2 = sum[0];
    ks->maxr2 = sum[0];
    ks->iter = -1;
    ks->need_ru = 0;
    ks->pending = 0;
    ks->dirty = 0;
    ks->n_ru = 0;
    ks->iterations = 0;
    ks->err = 0.0;
    ks->amp = kAmpPrior;
    ks->prev_true = 0.0;   // the first reliable update only sets the reference level
    ks->err_est = 0.0;
    ks->res_target = ks->tol_e > 0.0 ? fmin(ks->tol, ks->tol_e / (kErrSafety * kAmpPrior)) : ks->tol;
    // a zero right-hand side has the exact solution P = 0
    ks->status = (sum[1] == 0.0 && !warm) ? JOB_CONVERGED : JOB_ACTIVE;
}

// ----------------------------------------------------------------------------
// host launchers
// ----------------------------------------------------------------------------
template <typename Real, typename KReal>
static void xfwd_generic(const PassMap& pm, const GridJob* jobs, const cx_t<Real>* tw, int xf, int gate,
                         hipStream_t st) {
    if (pm.n == 0) return;
    const dim3 grid(pm.n), blk(FFT_THREADS);
    switch (xf) {
        case XF_PLAIN: hipLaunchKernelGGL((k_xfwd<Real, KReal, XF_PLAIN>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        case XF_P: hipLaunchKernelGGL((k_xfwd<Real, KReal, XF_P>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        case XF_T: hipLaunchKernelGGL((k_xfwd<Real, KReal, XF_T>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        case XF_SOL: hipLaunchKernelGGL((k_xfwd<Real, KReal, XF_SOL>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        default: throw Error(ERR_INVALID, "bad X-forward mode");
    }
}

template <typename Real, typename KReal>
static void xinv_generic(const PassMap& pm, const GridJob* jobs, const cx_t<Real>* tw, int xi, int gate,
                         hipStream_t st) {
    if (pm.n == 0) return;
    const dim3 grid(pm.n), blk(FFT_THREADS);
    switch (xi) {
        case XI_PLAIN: hipLaunchKernelGGL((k_xinv<Real, KReal, XI_PLAIN>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        case XI_V: hipLaunchKernelGGL((k_xinv<Real, KReal, XI_V>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        case XI_S: hipLaunchKernelGGL((k_xinv<Real, KReal, XI_S>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        case XI_RESID: hipLaunchKernelGGL((k_xinv<Real, KReal, XI_RESID>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        default: throw Error(ERR_INVALID, "bad X-inverse mode");
    }
}

template <typename Real>
static void ypass_generic(const PassMap& pm, const GridJob* jobs, const cx_t<Real>* tw, int dir, int gate,
                          hipStream_t st) {
    if (pm.n == 0) return;
    const dim3 grid(pm.n), blk(FFT_THREADS);
    if (dir < 0) hipLaunchKernelGGL((k_ypass<Real, -1>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit);
    else hipLaunchKernelGGL((k_ypass<Real, 1>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit);
}

template <typename Real>
static void zpass_generic(const PassMap& pm, const GridJob* jobs, const cx_t<Real>* tw, int gate,
                          hipStream_t st) {
    if (pm.n == 0) return;
    hipLaunchKernelGGL((k_zpass<Real>), dim3(pm.n), dim3(FFT_THREADS), pm.lds, st, jobs, pm.map, tw, gate, pm.limit);
}

// One pass of the matvec (0 X fwd, 1 Y fwd, 2 Z, 3 Y inv, 4 X inv): the
// generic launch and the specialised launch of the batch.
template <typename Real, typename KReal>
static void pass_impl(int k, const LaunchSet& ls, const TwiddlePools& tw, int xf, int xi, int gate, hipStream_t st) {
    constexpr bool c64 = sizeof(Real) == 4;
    constexpr bool k32 = sizeof(KReal) == 4;
    const int pr = c64 ? 0 : 1;
    const cx_t<Real>* t = c64 ? reinterpret_cast<const cx_t<Real>*>(tw.tw32) : reinterpret_cast<const cx_t<Real>*>(tw.tw64);
    switch (k) {
        case 0:
            xfwd_generic<Real, KReal>(ls.pm[pr][PASS_XF][0], ls.jobs, t, xf, gate, st);
            launch_fast_xfwd(ls.pm[pr][PASS_XF][1], ls.jobs, tw, c64, k32, xf, gate, st);
            if (c64) launch_plane_fwd(ls.xy, ls.jobs, tw, xf, gate, st);   // x and y of the plane jobs
            break;
        case 1:
        case 3:
            ypass_generic<Real>(ls.pm[pr][PASS_Y][0], ls.jobs, t, k == 1 ? -1 : 1, gate, st);
            launch_fast_y(ls.pm[pr][PASS_Y][1], ls.jobs, tw, c64, k == 1 ? -1 : 1, gate, st);
            break;
        case 2:
            zpass_generic<Real>(ls.pm[pr][PASS_Z][0], ls.jobs, t, gate, st);
            launch_fast_z(ls.pm[pr][PASS_Z][1], ls.jobs, tw, c64, gate, st);
            if (c64) launch_fast_yz(ls.yz, ls.jobs, tw, gate, st);   // y, z and inverse y of the fused jobs
            break;
        default:
            xinv_generic<Real, KReal>(ls.pm[pr][PASS_X][0], ls.jobs, t, xi, gate, st);
            launch_fast_xinv(ls.pm[pr][PASS_X][1], ls.jobs, tw, c64, k32, xi, gate, st);
            if (c64) launch_plane_inv(ls.xy, ls.jobs, tw, xi, gate, st);   // y and x of the plane jobs
            if (xi != XI_PLAIN && ls.njobs > 0) {
                const dim3 grid(ls.njobs), blk(FFT_THREADS);
                switch (xi) {
                    case XI_V: hipLaunchKernelGGL((k_xi_reduce<Real, XI_V>), grid, blk, 0, st, ls.jobs, gate, ls.ru_jobs); break;
                    case XI_S: hipLaunchKernelGGL((k_xi_reduce<Real, XI_S>), grid, blk, 0, st, ls.jobs, gate, ls.ru_jobs); break;
                    case XI_RESID: hipLaunchKernelGGL((k_xi_reduce<Real, XI_RESID>), grid, blk, 0, st, ls.jobs, gate, ls.ru_jobs); break;
                    default: throw Error(ERR_INVALID, "bad X-inverse mode");
                }
            }
            break;
    }
}

template <typename Real, typename KReal>
static void matvec_impl(const LaunchSet& ls, const TwiddlePools& tw, int xf, int xi, int gate, hipStream_t st) {
    const int pr = sizeof(Real) == 4 ? 0 : 1;
    if (ls.pm[pr][PASS_X][0].n + ls.pm[pr][PASS_X][1].n + (pr == 0 ? ls.n_xy : 0) == 0) return;
    if (gate == GATE_RU && pr == 1 && ls.ru_map != nullptr && ls.ru_ranges != nullptr && ls.ru_off[RU_REGIONS] > 0) {
        // only the tiles of jobs that wait for their reliable update (the
        // counts are cleared by a kernel, not a memset: this sequence is
        // captured into graphs, where memset nodes are avoided)
        hipLaunchKernelGGL(k_clear_ints, dim3(1), dim3(64), 0, st, ls.ru_count, RU_REGIONS);
        RuRegions rg;
        for (int k = 0; k <= RU_REGIONS; ++k) rg.off[k] = ls.ru_off[k];
        hipLaunchKernelGGL(k_ru_compact, dim3(ls.njobs), dim3(FFT_THREADS), 0, st, ls.jobs, ls.njobs, ls.ru_src, rg,
                           ls.ru_ranges, ls.ru_map, ls.ru_count, ls.ru_jobs);
        LaunchSet lr = ls;
        for (int p = 0; p < NPASS; ++p)
            for (int k = 0; k < 2; ++k) {
                PassMap& pm = lr.pm[1][p][k];
                pm.map = ls.ru_map + ls.ru_off[2 * p + k];
                pm.limit = ls.ru_count + 2 * p + k;
            }
        for (int k = 0; k < 5; ++k) pass_impl<Real, KReal>(k, lr, tw, xf, xi, gate, st);
    } else {
        for (int k = 0; k < 5; ++k) pass_impl<Real, KReal>(k, ls, tw, xf, xi, gate, st);
    }
    TDDA_HIP(hipGetLastError());
}

void launch_matvec(const LaunchSet& ls, const TwiddlePools& tw, bool c64, int xf_mode, int xi_mode, int gate,
                   hipStream_t st) {
    if (c64) {
        matvec_impl<float, float>(ls, tw, xf_mode, xi_mode, gate, st);
    } else if (ls.refined) {
        // FP64 transform feeding c64 Krylov vectors: reliable updates only
        if (!(xf_mode == XF_SOL && xi_mode == XI_RESID) && !(xf_mode == XF_PLAIN && xi_mode == XI_PLAIN))
            throw Error(ERR_INVALID, "FP64 transform in refined mode supports residual/plain only");
        matvec_impl<double, float>(ls, tw, xf_mode, xi_mode, gate, st);
    } else {
        matvec_impl<double, double>(ls, tw, xf_mode, xi_mode, gate, st);
    }
}

// true when pass k of the matvec launches nothing for this batch (the Y
// passes of a batch whose jobs all run through the plane or fused y/z passes)
static bool pass_empty(int k, const LaunchSet& ls, int pr) {
    if (k != 1 && k != 3) return false;
    return ls.pm[pr][PASS_Y][0].n + ls.pm[pr][PASS_Y][1].n == 0;
}

template <typename Real, typename KReal>
static void matvec_timed_impl(const LaunchSet& ls, const TwiddlePools& tw, hipStream_t st, int reps, double* ms5,
                              int xf, int xi) {
    const int pr = sizeof(Real) == 4 ? 0 : 1;
    hipEvent_t ev[6];
    for (auto& e : ev) TDDA_HIP(hipEventCreate(&e));
    for (int r = 0; r < reps; ++r) {
        // empty passes get no event pair (an event record costs a few us)
        int last = 0;
        int from[5];
        TDDA_HIP(hipEventRecord(ev[0], st));
        for (int k = 0; k < 5; ++k) {
            from[k] = -1;
            if (pass_empty(k, ls, pr)) continue;
            pass_impl<Real, KReal>(k, ls, tw, xf, xi, GATE_ALWAYS, st);
            TDDA_HIP(hipEventRecord(ev[k + 1], st));
            from[k] = last;
            last = k + 1;
        }
        TDDA_HIP(hipEventSynchronize(ev[last]));
        for (int k = 0; k < 5; ++k) {
            if (from[k] < 0) continue;
            float ms = 0.f;
            TDDA_HIP(hipEventElapsedTime(&ms, ev[from[k]], ev[k + 1]));
            ms5[k] += ms;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    TDDA_HIP(hipGetLastError());
}

void launch_matvec_timed(const LaunchSet& ls, const TwiddlePools& tw, bool c64, hipStream_t st, int reps,
                         double* ms5, int xf_mode, int xi_mode) {
    if (c64)
        matvec_timed_impl<float, float>(ls, tw, st, reps, ms5, xf_mode, xi_mode);
    else if (ls.refined)
        matvec_timed_impl<double, float>(ls, tw, st, reps, ms5, xf_mode, xi_mode);
    else
        matvec_timed_impl<double, double>(ls, tw, st, reps, ms5, xf_mode, xi_mode);
}

void launch_finalize(const LaunchSet& ls, hipStream_t st) {
    if (ls.n_upd == 0) return;
    if (ls.refined)
        hipLaunchKernelGGL((k_finalize<float>), dim3(ls.n_upd), dim3(FFT_THREADS), 0, st, ls.jobs, ls.map_upd);
    else
        hipLaunchKernelGGL((k_finalize<double>), dim3(ls.n_upd), dim3(FFT_THREADS), 0, st, ls.jobs, ls.map_upd);
    hipLaunchKernelGGL(k_finalize_clear, dim3((ls.njobs + 255) / 256), dim3(256), 0, st, ls.jobs, ls.njobs);
    TDDA_HIP(hipGetLastError());
}

__global__ void k_ks_template(const GridJob* __restrict__ jobs, int n, KScalars t) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) *jobs[j].ks = t;
}

void launch_ks_template(const LaunchSet& ls, const KScalars& t, hipStream_t st) {
    if (ls.njobs == 0) return;
    hipLaunchKernelGGL(k_ks_template, dim3((ls.njobs + 63) / 64), dim3(64), 0, st, ls.jobs, ls.njobs, t);
    TDDA_HIP(hipGetLastError());
}

void launch_init(const LaunchSet& ls, bool warm, hipStream_t st) {
    if (ls.n_upd == 0) return;
    if (ls.refined)
        hipLaunchKernelGGL((k_init<float>), dim3(ls.n_upd), dim3(FFT_THREADS), 0, st, ls.jobs, ls.map_upd, (int)warm);
    else
        hipLaunchKernelGGL((k_init<double>), dim3(ls.n_upd), dim3(FFT_THREADS), 0, st, ls.jobs, ls.map_upd, (int)warm);
    if (ls.refined)
        hipLaunchKernelGGL((k_init_reduce<float>), dim3(ls.njobs), dim3(FFT_THREADS), 0, st, ls.jobs, (int)warm);
    else
        hipLaunchKernelGGL((k_init_reduce<double>), dim3(ls.njobs), dim3(FFT_THREADS), 0, st, ls.jobs, (int)warm);
    TDDA_HIP(hipGetLastError());
}

// ----------------------------------------------------------------------------
// tile geometry
// ----------------------------------------------------------------------------
static unsigned long long quotient_magic(int d) {
    return ((1ull << 32) + (unsigned long long)d - 1ull) / (unsigned long long)d;
}

void plan_tiles(GridJob& J) {
    auto cdiv = [](long a, long b) { return (int)((a + b - 1) / b); };
    size_t unused = 0;
    J.mlx = quotient_magic(J.lx);
    J.mny = quotient_magic(J.ny);
    // schedule variants (both variants of a length are specialised, so the
    // fast / generic choice below does not depend on them)
    for (int pr = 0; pr < 2; ++pr)
        for (int pass = 0; pass < NPASS; ++pass) {
            const int L = (pass == PASS_X || pass == PASS_XF) ? J.lx : pass == PASS_Y ? J.ly : J.lz;
            J.var[pr][pass] = split_variants(pass, L) > 1 ? schedule_variant(pr, pass, L) : 0;
        }
    J.fx = fast_tiles(J.lx, PASS_X, true, J.var[0][PASS_X], &J.bx32, &unused) ? 1 : 0;
    if (J.fx) {
        fast_tiles(J.lx, PASS_X, false, J.var[1][PASS_X], &J.bx64, &unused);
        fast_tiles(J.lx, PASS_XF, true, J.var[0][PASS_XF], &J.bxf32, &unused);
        fast_tiles(J.lx, PASS_XF, false, J.var[1][PASS_XF], &J.bxf64, &unused);
    } else {
        J.bx32 = std::max(1, Caps<float>::X / (3 * J.lx));
        J.bx64 = std::max(1, Caps<double>::X / (3 * J.lx));
        J.bxf32 = J.bx32;
        J.bxf64 = J.bx64;
    }
    J.fy = fast_tiles(J.ly, PASS_Y, true, J.var[0][PASS_Y], &J.by32, &unused) ? 1 : 0;
    if (J.fy) {
        fast_tiles(J.ly, PASS_Y, false, J.var[1][PASS_Y], &J.by64, &unused);
    } else {
        J.by32 = std::max(1, Caps<float>::Y / J.ly);
        J.by64 = std::max(1, Caps<double>::Y / J.ly);
    }
    J.fz = fast_tiles(J.lz, PASS_Z, true, J.var[0][PASS_Z], &J.bz32, &unused) ? 1 : 0;
    if (J.fz) {
        fast_tiles(J.lz, PASS_Z, false, J.var[1][PASS_Z], &J.bz64, &unused);
    } else {
        J.bz32 = std::max(1, Caps<float>::Z / J.lz);
        J.bz64 = std::max(1, Caps<double>::Z / J.lz);
    }
    const long nrows = (long)J.ny * J.nz;
    J.tx32 = cdiv(nrows, J.bx32);
    J.txf32 = cdiv(nrows, J.bxf32);
    J.txf64 = cdiv(nrows, J.bxf64);
    J.ty32 = cdiv(3L * J.nz * J.lx, J.by32);
    J.tz32 = cdiv((long)J.ly * J.lx, J.bz32);
    J.tx64 = cdiv(nrows, J.bx64);
    J.ty64 = cdiv(3L * J.nz * J.lx, J.by64);
    J.tz64 = cdiv((long)J.ly * J.lx, J.bz64);
    J.tupd = cdiv(3L * J.ndip, (long)FFT_THREADS * UPD_EPT);
    J.yzc = J.fx ? yz_class(J.ly, J.lz, J.ny, J.nz) : -1;
    J.tr32 = 0;   // set per batch
    J.xyc = plane_class_of(J.lx, J.ly, J.nx, J.ny);
    J.xy32 = 0;   // set per batch
}

size_t lds_bytes(const GridJob& J, int pass, bool c64) {
    const int L = (pass == PASS_X || pass == PASS_XF) ? J.lx : pass == PASS_Y ? J.ly : J.lz;
    const bool fast = (pass == PASS_X || pass == PASS_XF) ? J.fx : pass == PASS_Y ? J.fy : J.fz;
    const size_t cs = c64 ? sizeof(float2) : sizeof(double2);
    if (fast) {
        int per = 0;
        size_t bytes = 0;
        fast_tiles(L, pass, c64, J.var[c64 ? 0 : 1][pass], &per, &bytes);
        return bytes;
    }
    if (pass == PASS_X) return (size_t)J.lx * ((3 * (c64 ? J.bx32 : J.bx64)) | 1) * cs;
    if (pass == PASS_XF) return (size_t)J.lx * ((3 * (c64 ? J.bxf32 : J.bxf64)) | 1) * cs;
    if (pass == PASS_Y) return (size_t)J.ly * ((c64 ? J.by32 : J.by64) | 1) * cs;
    return (size_t)3 * J.lz * ((c64 ? J.bz32 : J.bz64) | 1) * cs;
}

}  // namespace tdda

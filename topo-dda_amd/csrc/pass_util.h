// Device helpers shared by the generic and the length-specialised passes:
// job gates, deterministic reductions, BiCGSTAB scalar logic, the vector
// operations fused into the X passes, and the Green's-tensor multiply.
#pragma once

#include "passes.h"

namespace tdda {

// Global-memory accesses through pointers read from job descriptors: without
// an address-space cast the compiler cannot tell global from LDS and emits
// FLAT accesses, which also occupy the LDS counter and serialise waits.  The
// HIP vector types cannot be copied out of address space 1 directly, so the
// accessors move the same bytes as native clang vectors.  Offsets are 32-bit
// (every array of one job is far below 2^31 elements).
#define TDDA_GLOBAL __attribute__((address_space(1)))
template <typename T> struct NativeVec { using type = T; };
template <> struct NativeVec<float2> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct NativeVec<double2> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct NativeVec<float4> { typedef float type __attribute__((ext_vector_type(4))); };

template <typename T>
__device__ __forceinline__ T gload(const void* base, unsigned i) {
    using V = typename NativeVec<T>::type;
    const V v = ((const TDDA_GLOBAL V*)base)[i];
    return __builtin_bit_cast(T, v);
}
template <typename T>
__device__ __forceinline__ void gstore(void* base, unsigned i, T x) {
    using V = typename NativeVec<T>::type;
    ((TDDA_GLOBAL V*)base)[i] = __builtin_bit_cast(V, x);
}

template <typename Real> struct Tile;
template <> struct Tile<float> {
    __device__ static int bx(const GridJob& J) { return J.bx32; }
    __device__ static int bxf(const GridJob& J) { return J.bxf32; }
    __device__ static int by(const GridJob& J) { return J.by32; }
    __device__ static int bz(const GridJob& J) { return J.bz32; }
    __device__ static int tx(const GridJob& J) { return J.tx32; }
    __device__ static int txf(const GridJob& J) { return J.txf32; }
    __device__ static const float2* green(const GridJob& J) { return J.green32; }
    __device__ static const float2* al(const GridJob& J) { return J.al32; }
    __device__ static const float2* ial(const GridJob& J) { return J.ial32; }
};
template <> struct Tile<double> {
    __device__ static int bx(const GridJob& J) { return J.bx64; }
    __device__ static int bxf(const GridJob& J) { return J.bxf64; }
    __device__ static int by(const GridJob& J) { return J.by64; }
    __device__ static int bz(const GridJob& J) { return J.bz64; }
    __device__ static int tx(const GridJob& J) { return J.tx64; }
    __device__ static int txf(const GridJob& J) { return J.txf64; }
    __device__ static const double2* green(const GridJob& J) { return J.green64; }
    __device__ static const double2* al(const GridJob& J) { return J.al64; }
    __device__ static const double2* ial(const GridJob& J) { return J.ial64; }
};

// GATE_ACTIVE: a running job not waiting for its reliable update;
// GATE_RU: a running job waiting for it.
__device__ inline bool gate_open(const GridJob& J, int gate) {
    if (gate == GATE_ALWAYS) return true;
    const KScalars* ks = J.ks;
    if (ks->status != JOB_ACTIVE) return false;
    if (gate == GATE_RU) return ks->need_ru != 0;
    return ks->need_ru == 0;
}

// Blocks of a compacted launch beyond the filled length return at once.
__device__ inline bool beyond_limit(const int* limit) {
    return limit != nullptr && (int)blockIdx.x >= __hip_atomic_load(limit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// True when a reliable-update launch has nothing to do: no job of any batch
// waits for one (ru_jobs counts them; one load instead of the per-block
// map -> job -> scalars chain).  The per-job gate still decides otherwise.
__device__ inline bool ru_idle(int gate, const int* ru_jobs) {
    return gate == GATE_RU && ru_jobs != nullptr &&
           __hip_atomic_load(ru_jobs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
}

// Deterministic block reduction of NV doubles; every thread gets the sums.
// Butterfly over the 64 lanes of each wave (a fixed pattern, so the result
// is reproducible), then the NT/64 wave sums in wave order (NT: threads of
// the block).
template <int NV, int NT = FFT_THREADS>
__device__ inline void block_sum(double (&v)[NV]) {
    static_assert(NT % 64 == 0, "whole waves");
    constexpr int NW = NT / 64;
    __shared__ double red[NV][NW];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) red[k][threadIdx.x >> 6] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = red[k][0];
#pragma unroll
        for (int w = 1; w < NW; ++w) s += red[k][w];
        v[k] = s;
    }
    __syncthreads();   // red may be reused by the next call
}

// Cross-block sums go through a per-tile partial buffer: every block of a
// pass writes its block sums (write_partials), and a one-block-per-job kernel
// launched after the pass adds them in tile order (job_sum) and advances the
// job's scalars.  The kernel boundary makes the partials visible; no block
// waits on another, and no device-scope fence is issued per block (on a
// multi-XCD part that fence writes back the XCD's whole L2).
template <int NV>
__device__ inline void write_partials(const GridJob& J, int tile, const double (&v)[NV]) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) J.partial[(long)tile * PARTIAL_STRIDE + k] = v[k];
    }
}

// Called by every thread of a one-block-per-job kernel: thread t sums tiles
// t, t+256, ... and the block reduces, a fixed order for a given job shape.
template <int NV>
__device__ inline void job_sum(const double* part, int stride, int ntiles, double (&out)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = 0.0;
    for (int t = threadIdx.x; t < ntiles; t += FFT_THREADS) {
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] += part[(long)t * stride + k];
    }
    block_sum<NV>(out);
}
template <int NV>
__device__ inline void job_sum(const GridJob& J, int ntiles, double (&out)[NV]) {
    job_sum<NV>(J.partial, PARTIAL_STRIDE, ntiles, out);
}

// ----------------------------------------------------------------------------
// scalar logic (one thread of the per-job reduction), DDAModel.cpp:945-1004
// ----------------------------------------------------------------------------
// conv_ok: the field-error test of a refined solve passed (always true for
// the recursive residual of an FP64 solve, which stops as the reference does)
__device__ inline void bicg_decide(KScalars* ks, double2 rho_new, double rn2, bool conv_ok = true) {
    ks->rnorm2 = rn2;
    const double rel = sqrt(rn2 / ks->enorm2);
    const int it = ks->iter;
    if (it >= 0 && rel <= ks->tol && conv_ok) {      // ||r||/||E|| <= tol after a loop step
        ks->status = JOB_CONVERGED;
        ks->iterations = it + 1;
        ks->err = rel;
        return;
    }
    const int next = it + 1;
    if (next >= ks->maxit) {
        ks->status = JOB_MAXIT;
        ks->iterations = ks->maxit;
        ks->err = rel;
        return;
    }
    if (ks->rho.x == 0.0 && ks->rho.y == 0.0) {   // r0.dot(rl) == 0 at the top of the loop
        ks->status = JOB_BREAKDOWN;
        ks->iterations = next + 1;
        ks->err = rel;
        return;
    }
    // beta = (alpha/zeta) * <r0,r> / <r0,rl>
    ks->beta = ddiv(dcmul(ddiv(ks->alpha, ks->zeta), rho_new), ks->rho);
    ks->rho_prev = ks->rho;
    ks->rho = rho_new;
    ks->iter = next;
}

// One BiCGSTAB step's end (DDAModel.cpp:981-1004) in refined mode first:
// a residual small against the last reliable update, or one that meets the
// tolerance or the iteration limit, asks for the true residual instead.
__device__ inline void step_decide(KScalars* ks, double2 rho_new, double rn2) {
    if (ks->refined) {
        const double rel2 = rn2 / ks->enorm2;
        const bool trigger = rn2 <= ks->ru_delta * ks->ru_delta * ks->maxr2 ||
                             rel2 <= ks->res_target * ks->res_target || ks->iter + 1 >= ks->maxit;
        if (trigger && ks->iter >= 0) {
            ks->need_ru = 1;     // true residual decides (k_xinv XI_RESID)
            ks->rnorm2 = rn2;
            return;
        }
        if (rn2 > ks->maxr2) ks->maxr2 = rn2;
        // only a reliable update (true residual and field-error estimate)
        // ends a refined solve; the recursive residual may sit between the
        // trigger target and tol
        bicg_decide(ks, rho_new, rn2, false);
        return;
    }
    bicg_decide(ks, rho_new, rn2);
}

// ----------------------------------------------------------------------------
// X forward input: the vector fed to the transform, with the BiCGSTAB vector
// operations fused in.  The end-of-step update of the previous step
// (DDAModel.cpp:981-986: P += alpha p + zeta t, which equals the reference's
// alpha p + zeta r - alpha zeta v since t = r - alpha v; r = t - zeta s) is
// applied here, element by element, when ks->pending is set; then
// p = r + beta (p - zeta v) (DDAModel.cpp:976) or t = r - alpha v (:979).
// vi < 0: outside the geometry.
// ----------------------------------------------------------------------------
struct XfScalars {
    double2 alpha, beta, zeta;
    int pending;
};
template <int MODE>
__device__ inline XfScalars xf_scalars(const GridJob& J) {
    XfScalars s{make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), 0};
    const KScalars* ks = J.ks;
    if (MODE == XF_P || MODE == XF_SOL) {
        s.alpha = ks->alpha;
        s.beta = ks->beta;
        s.zeta = ks->zeta;
        s.pending = ks->pending;
    }
    if (MODE == XF_T) s.alpha = ks->alpha;
    return s;
}

// c64 Krylov vectors: the solution increments go to dP between reliable updates
template <typename KC>
constexpr bool kSplitSolution = sizeof(KC) == sizeof(float2);

// P + dP into P (dP cleared); returns the new P and the change in *delta
template <typename KC>
__device__ inline double2 fold_solution(const GridJob& J, unsigned vi, double2 add, double2* delta = nullptr) {
    double2 Pn = cadd(gload<double2>(J.P, vi), add);
    double2 d = add;
    if constexpr (kSplitSolution<KC>) {
        const double2 dp = to_d(gload<KC>(J.dP, vi));
        Pn = cadd(Pn, dp);
        d = cadd(d, dp);
        gstore<KC>(J.dP, vi, mk<KC>(0.0, 0.0));
    }
    gstore<double2>(J.P, vi, Pn);
    if (delta) *delta = d;
    return Pn;
}

// the pending step update of one element, with dP folded into P; returns
// the new P (and the change in *delta)
template <typename KC>
__device__ inline double2 apply_pending(const GridJob& J, unsigned vi, const XfScalars& sc, double2 pv, bool write_r,
                                        double2* delta = nullptr) {
    const double2 td = to_d(gload<KC>(J.t, vi));
    const double2 Pn = fold_solution<KC>(J, vi, cadd(dcmul(sc.alpha, pv), dcmul(sc.zeta, td)), delta);
    if (write_r) {
        const double2 sd = to_d(gload<KC>(J.s, vi));
        gstore<KC>(J.r, vi, from_d<KC>(csub(td, dcmul(sc.zeta, sd))));
    }
    return Pn;
}

// The vector a mode feeds to the transform (before the Jacobi scaling), with
// the mode's vector updates applied.
template <typename C, typename KC, int MODE>
__device__ inline C xf_vector(const GridJob& J, unsigned i, const XfScalars& sc, double (&xacc)[XF_NACC]) {
    if (MODE == XF_SOL) {
        double2 delta = make_double2(0.0, 0.0);
        double2 Pn;
        if (sc.pending) {
            const double2 pv = to_d(gload<KC>(J.p, i));
            Pn = apply_pending<KC>(J, i, sc, pv, false, &delta);
        } else if constexpr (kSplitSolution<KC>) {
            Pn = fold_solution<KC>(J, i, make_double2(0.0, 0.0), &delta);
        } else {
            Pn = gload<double2>(J.P, i);
        }
        if (J.jacobi) {   // P holds Y = al P: the polarisation and its change are Y / al
            const double2 ia = gload<double2>(J.ial64, i / 3u);
            Pn = dcmul(Pn, ia);
            delta = dcmul(delta, ia);
        }
        if (J.partial_xf != nullptr) {
            const double2 f = J.efac != nullptr ? to_d(gload<float2>(J.efac, i / 3u)) : make_double2(1.0, 0.0);
            const double2 fd = dcmul(f, delta), fp = dcmul(f, Pn);
            xacc[0] += fd.x * fd.x + fd.y * fd.y;
            xacc[1] += fp.x * fp.x + fp.y * fp.y;
        }
        return from_d<C>(Pn);
    } else if (MODE == XF_P) {
        const double2 pv = to_d(gload<KC>(J.p, i));
        double2 rv;
        if (sc.pending) {
            const double2 td = to_d(gload<KC>(J.t, i));
            const double2 sd = to_d(gload<KC>(J.s, i));
            const double2 step = cadd(dcmul(sc.alpha, pv), dcmul(sc.zeta, td));
            if constexpr (kSplitSolution<KC>)
                gstore<KC>(J.dP, i, from_d<KC>(cadd(to_d(gload<KC>(J.dP, i)), step)));
            else
                gstore<double2>(J.P, i, cadd(gload<double2>(J.P, i), step));
            const KC rk = from_d<KC>(csub(td, dcmul(sc.zeta, sd)));
            gstore<KC>(J.r, i, rk);
            rv = to_d(rk);
        } else {
            rv = to_d(gload<KC>(J.r, i));
        }
        const double2 w = csub(pv, dcmul(sc.zeta, to_d(gload<KC>(J.v, i))));   // p - u, u = zeta v
        const KC pk = from_d<KC>(cadd(rv, dcmul(sc.beta, w)));
        gstore<KC>(J.p, i, pk);
        return from_d<C>(to_d(pk));
    } else {   // XF_T
        const KC tk = from_d<KC>(csub(to_d(gload<KC>(J.r, i)), dcmul(sc.alpha, to_d(gload<KC>(J.v, i)))));
        gstore<KC>(J.t, i, tk);
        return from_d<C>(to_d(tk));
    }
}

// X forward input of element vi (< 0: outside the geometry): the mode's
// vector, times 1/al under Jacobi preconditioning.
template <typename C, typename KC, int MODE>
__device__ inline C xf_element(const GridJob& J, int vi, const XfScalars& sc, double (&xacc)[XF_NACC]) {
    if (vi < 0) return mk<C>(0.0, 0.0);
    const unsigned i = (unsigned)vi;
    if (MODE == XF_PLAIN) return gload<C>(J.xin, i);
    const C u = xf_vector<C, KC, MODE>(J, i, sc, xacc);
    if (MODE != XF_SOL && J.jacobi) {   // the transform sees x / al (XF_SOL scales inside)
        using Real = decltype(u.x);
        return cmul(u, gload<C>(Tile<Real>::ial(J), i / 3u));
    }
    return u;
}

// dipole vector index of box cell (row q = z*ny + y, x) component c, or -1
__device__ inline int vec_index(const GridJob& J, int q, int x, int c) {
    const int cell = q * J.nx + x;
    if (J.box_map == nullptr) return 3 * cell + c;
    const int dip = ((const TDDA_GLOBAL int*)J.box_map)[(unsigned)cell];
    return dip >= 0 ? 3 * dip + c : -1;
}

// ----------------------------------------------------------------------------
// X inverse epilogue: y = conv + al u and what the mode does with it.
// conv is already divided by M.  acc gathers the per-thread dot partials:
//   XI_V     : <r0, v>
//   XI_S     : <s, t>, <s, s>, <r0, t>, <r0, s>, <t, t>
//   XI_RESID : <r0, r>, <r, r>
// ----------------------------------------------------------------------------
constexpr int XI_NACC = 8;

// BiCGSTAB's shadow residual r0 at vector index i = 3 cell + c: loaded, or
// the incident field when the job says so (full box)
__device__ inline void cell_of(const GridJob& J, unsigned cell, int& gx, int& gy, int& gz) {
    const unsigned row = cell / (unsigned)J.nx;
    gx = (int)(cell - row * (unsigned)J.nx);
    gy = (int)(row % (unsigned)J.ny);
    gz = (int)(row / (unsigned)J.ny);
}
template <typename KC>
__device__ inline KC r0_at(const GridJob& J, unsigned i) {
    if (!J.r0_inc) return gload<KC>(J.r0, i);
    const unsigned cell = i / 3u;
    int gx, gy, gz;
    cell_of(J, cell, gx, gy, gz);
    return from_d<KC>(incident_wave(incident_phase(J.inc_kd, J.inc_nk, gx, gy, gz), J.inc_en[i - 3u * cell]));
}
// the three components at cell i / 3 (i a multiple of 3)
template <typename KC>
__device__ inline void r0_cell(const GridJob& J, unsigned i, KC (&rz)[3]) {
    if (!J.r0_inc) {
        static_for<3>([&](auto ci) { rz[decltype(ci)::value] = gload<KC>(J.r0, i + decltype(ci)::value); });
        return;
    }
    int gx, gy, gz;
    cell_of(J, i / 3u, gx, gy, gz);
    const double2 ph = incident_phase(J.inc_kd, J.inc_nk, gx, gy, gz);
    static_for<3>([&](auto ci) {
        constexpr int c = decltype(ci)::value;
        rz[c] = from_d<KC>(incident_wave(ph, J.inc_en[c]));
    });
}

template <typename C, typename KC, int MODE>
__device__ inline void xi_element(const GridJob& J, int vi, C conv, const C* __restrict__ al,
                                  double (&acc)[XI_NACC]) {
    const unsigned i = (unsigned)vi;
    // Jacobi: the operator is A diag(al)^-1, whose diagonal part al (x / al) is x
    const C a = (MODE != XI_PLAIN && J.jacobi) ? mk<C>(1.0, 0.0) : gload<C>(al, i / 3u);
    if (MODE == XI_PLAIN) {
        const C u = gload<C>(J.xin, i);
        gstore<C>(J.yout, i, cadd(conv, cmul(a, u)));
    } else if (MODE == XI_V) {
        const C y = cadd(conv, cmul(a, from_d<C>(to_d(gload<KC>(J.p, i)))));
        const KC vk = from_d<KC>(to_d(y));
        gstore<KC>(J.v, i, vk);
        const double2 d = cmulc(to_d(vk), to_d(r0_at<KC>(J, i)));   // conj(r0) v
        acc[0] += d.x;
        acc[1] += d.y;
    } else if (MODE == XI_S) {
        const KC tk = gload<KC>(J.t, i);
        const C y = cadd(conv, cmul(a, from_d<C>(to_d(tk))));
        const KC sk = from_d<KC>(to_d(y));
        gstore<KC>(J.s, i, sk);
        const double2 sd = to_d(sk), td = to_d(tk);
        const double2 rz = to_d(r0_at<KC>(J, i));
        const double2 d = cmulc(td, sd);    // conj(s) t
        const double2 e = cmulc(td, rz);    // conj(r0) t
        const double2 f = cmulc(sd, rz);    // conj(r0) s
        acc[0] += d.x;
        acc[1] += d.y;
        acc[2] += sd.x * sd.x + sd.y * sd.y;
        acc[3] += e.x;
        acc[4] += e.y;
        acc[5] += f.x;
        acc[6] += f.y;
        acc[7] += td.x * td.x + td.y * td.y;
    } else if (MODE == XI_RESID) {
        const double2 y = cadd(to_d(conv), dcmul(to_d(a), gload<double2>(J.P, i)));
        const double2 rr = csub(gload<double2>(J.rhs, i), y);
        const KC rk = from_d<KC>(rr);
        gstore<KC>(J.r, i, rk);
        const double2 d = cmulc(to_d(rk), to_d(r0_at<KC>(J, i)));
        acc[0] += d.x;
        acc[1] += d.y;
        acc[2] += rr.x * rr.x + rr.y * rr.y;
    }
}

// Block sums of the dot partials of an X inverse tile (all threads call it).
template <int MODE>
__device__ inline void xi_finish(const GridJob& J, int tile, double (&acc)[XI_NACC]) {
    if (MODE == XI_PLAIN) return;
    block_sum<XI_NACC>(acc);
    write_partials<XI_NACC>(J, tile, acc);
}

// Advance a job's scalars from the summed dot partials (one thread):
//   XI_V     : alpha = <r0,r>/<r0,Ap>; the pending update was applied by the
//              X forward pass of this matvec
//   XI_S     : zeta = <At,t>/<At,At>; with r = t - zeta s,
//              <r0, r> = <r0,t> - zeta <r0,s> and
//              |r|^2 = |t|^2 - |<s,t>|^2/|s|^2 (zeta minimises |t - zeta s|);
//              the vector update becomes pending and the step is decided
//   XI_RESID : true residual of the updated P (the X forward pass applied the
//              pending update and folded dP into P): reliable-update
//              bookkeeping and decision
// Block sums of the X forward partials of a reliable update (all threads call it).
template <int MODE>
__device__ inline void xf_finish(const GridJob& J, int tile, double (&xacc)[XF_NACC]) {
    if (MODE != XF_SOL || J.partial_xf == nullptr) return;   // uniform over the block (one job)
    block_sum<XF_NACC>(xacc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < XF_NACC; ++k) J.partial_xf[(long)tile * XF_NACC + k] = xacc[k];
    }
}

// Field-error bookkeeping of a reliable update (refined mode): xs holds
// ||f dP||^2 and ||f P||^2 of the update (dP = the change since the previous
// one), rel its true relative residual.  Returns whether the field-error
// target is met.
__device__ inline bool error_decide(KScalars* ks, const double (&xs)[XF_NACC], double rel) {
    if (xs[1] > 0.0) {
        const double de = sqrt(xs[0] / xs[1]);
        if (ks->prev_true > 0.0 && rel <= kAmpSeparation * ks->prev_true) ks->amp = de / (ks->prev_true - rel);
        ks->prev_true = rel;
    }
    ks->err_est = ks->amp * rel;
    if (ks->tol_e <= 0.0) {
        ks->res_target = ks->tol;
        return true;
    }
    ks->res_target = fmin(ks->tol, ks->tol_e / (kErrSafety * ks->amp));
    return kErrSafety * ks->err_est <= ks->tol_e;
}

// Advance a job's scalars from the summed dot partials (one thread):
//   XI_V     : alpha = <r0,r>/<r0,Ap>; the pending update was applied by the
//              X forward pass of this matvec
//   XI_S     : zeta = <At,t>/<At,At>; with r = t - zeta s,
//              <r0, r> = <r0,t> - zeta <r0,s> and
//              |r|^2 = |t|^2 - |<s,t>|^2/|s|^2 (zeta minimises |t - zeta s|);
//              the vector update becomes pending and the step is decided
//   XI_RESID : true residual of the updated P (the X forward pass applied the
//              pending update and folded dP into P): reliable-update
//              bookkeeping, field-error estimate and decision
template <int MODE>
__device__ inline void xi_decide(KScalars* ks, const double (&sum)[XI_NACC], int* ru_jobs,
                                 const double (&xs)[XF_NACC]) {
    const double2 dot = make_double2(sum[0], sum[1]);
    if (MODE == XI_V) {
        ks->alpha = ddiv(ks->rho, dot);
        ks->pending = 0;
    } else if (MODE == XI_S) {
        const double ss = sum[2];
        const double2 zeta = make_double2(dot.x / ss, dot.y / ss);
        ks->zeta = zeta;
        const double2 r0t = make_double2(sum[3], sum[4]);
        const double2 r0s = make_double2(sum[5], sum[6]);
        const double2 rho_new = csub(r0t, dcmul(zeta, r0s));
        const double rn2 = fmax(sum[7] - (dot.x * dot.x + dot.y * dot.y) / ss, 0.0);
        ks->pending = 1;
        if (ks->refined) ks->dirty = 1;   // the X forward pass adds the step to dP
        step_decide(ks, rho_new, rn2);
        if (ks->need_ru != 0 && ru_jobs != nullptr) atomicAdd(ru_jobs, 1);
    } else if (MODE == XI_RESID) {
        if (ks->need_ru != 0 && ru_jobs != nullptr) atomicSub(ru_jobs, 1);
        ks->pending = 0;
        ks->dirty = 0;                    // the X forward pass folded dP into P
        ks->need_ru = 0;
        ks->n_ru += 1;
        ks->maxr2 = sum[2];
        const bool err_ok = error_decide(ks, xs, sqrt(sum[2] / ks->enorm2));
        if (ks->status == JOB_ACTIVE) bicg_decide(ks, dot, sum[2], err_ok);
    }
}

// ----------------------------------------------------------------------------
// Green's-tensor multiply at spectral point (kx, ky, kz).
// Component (00,01,02,11,12,22) is odd in the axes {}, {x,y}, {x,z}, {},
// {y,z}, {} respectively, so only k <= L/2 per axis is stored and
// Ghat(L-k) = +-Ghat(k).
// ----------------------------------------------------------------------------
// Spectrum rows (kz, ky) hold each component (FP64) or component pair (c64,
// pairs (00,01), (02,11), (12,22) in one float4) as a run over kx, so the
// lanes of a wave, which take consecutive kx, load contiguous bytes.
__device__ inline void green_load6(const float2* G, unsigned row, unsigned hx, unsigned kx, float2 (&g)[6]) {
    const unsigned o = 3u * row * hx + kx;
    const float4 a = gload<float4>(G, o), b = gload<float4>(G, o + hx), c = gload<float4>(G, o + 2u * hx);
    g[0] = make_float2(a.x, a.y); g[1] = make_float2(a.z, a.w);
    g[2] = make_float2(b.x, b.y); g[3] = make_float2(b.z, b.w);
    g[4] = make_float2(c.x, c.y); g[5] = make_float2(c.z, c.w);
}
__device__ inline void green_load6(const double2* G, unsigned row, unsigned hx, unsigned kx, double2 (&g)[6]) {
    const unsigned o = 6u * row * hx + kx;
#pragma unroll
    for (int k = 0; k < 6; ++k) g[k] = gload<double2>(G, o + (unsigned)k * hx);
}

// The record of spectral point (kx, ky, kz) from the octant, with the
// parity signs of the mirrored axes applied.
template <typename C>
__device__ inline void green_fetch(const C* G, const GridJob& J, int kx, int ky, int kz, C (&g)[6]) {
    int sx = 1, sy = 1, sz = 1;
    if (2 * kx > J.lx) { kx = J.lx - kx; sx = -1; }
    if (2 * ky > J.ly) { ky = J.ly - ky; sy = -1; }
    if (2 * kz > J.lz) { kz = J.lz - kz; sz = -1; }
    green_load6(G, (unsigned)kz * J.hy + ky, J.hx, kx, g);
    if (sx * sy < 0) g[1] = cscale(g[1], -1);
    if (sx * sz < 0) g[2] = cscale(g[2], -1);
    if (sy * sz < 0) g[4] = cscale(g[4], -1);
}

// (v0, v1, v2) <- G (v0, v1, v2), G symmetric
template <typename C>
__device__ inline void green_mul(const C (&g)[6], C& v0, C& v1, C& v2) {
    const C a0 = v0, a1 = v1, a2 = v2;
    v0 = cadd(cadd(cmul(g[0], a0), cmul(g[1], a1)), cmul(g[2], a2));
    v1 = cadd(cadd(cmul(g[1], a0), cmul(g[3], a1)), cmul(g[4], a2));
    v2 = cadd(cadd(cmul(g[2], a0), cmul(g[4], a1)), cmul(g[5], a2));
}

template <typename C>
__device__ inline void green_apply(const C* G, const GridJob& J, int kx, int ky, int kz, C& v0, C& v1, C& v2) {
    C g[6];
    green_fetch(G, J, kx, ky, kz, g);
    green_mul(g, v0, v1, v2);
}

}  // namespace tdda

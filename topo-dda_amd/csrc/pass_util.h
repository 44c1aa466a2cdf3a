// Device helpers shared by the generic and the length-specialised passes:
// job gates, deterministic reductions, BiCGSTAB scalar logic, the vector
// operations fused into the X passes, and the Green's-tensor multiply.
#pragma once

#include "passes.h"

namespace tdda {

template <typename Real> struct Tile;
template <> struct Tile<float> {
    __device__ static int bx(const GridJob& J) { return J.bx32; }
    __device__ static int by(const GridJob& J) { return J.by32; }
    __device__ static int bz(const GridJob& J) { return J.bz32; }
    __device__ static int tx(const GridJob& J) { return J.tx32; }
    __device__ static const float2* green(const GridJob& J) { return J.green32; }
    __device__ static const float2* al(const GridJob& J) { return J.al32; }
};
template <> struct Tile<double> {
    __device__ static int bx(const GridJob& J) { return J.bx64; }
    __device__ static int by(const GridJob& J) { return J.by64; }
    __device__ static int bz(const GridJob& J) { return J.bz64; }
    __device__ static int tx(const GridJob& J) { return J.tx64; }
    __device__ static const double2* green(const GridJob& J) { return J.green64; }
    __device__ static const double2* al(const GridJob& J) { return J.al64; }
};

__device__ inline bool gate_open(const GridJob& J, int gate) {
    if (gate == GATE_ALWAYS) return true;
    const KScalars* ks = J.ks;
    if (ks->status != JOB_ACTIVE) return false;
    if (gate == GATE_RU) return ks->need_ru != 0;
    return true;
}

// Deterministic block reduction of NV doubles; every thread gets the sums.
template <int NV>
__device__ inline void block_sum(double (&v)[NV]) {
    __shared__ double red[NV][FFT_THREADS];
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = FFT_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < NV; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = red[k][0];
}

// Publish this block's partial sums; returns true (in every thread) for the
// block that completes the job.  That block then sums all partials in tile
// order (fixed, hence deterministic) through sum_partials().
template <int NV>
__device__ inline bool publish_partials(const GridJob& J, int tile, int ntiles, const double (&v)[NV]) {
    __shared__ int s_last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) J.partial[(long)tile * 6 + k] = v[k];
        __threadfence();
        const unsigned prev = atomicAdd(J.counter, 1u);
        s_last = (prev + 1 == (unsigned)ntiles) ? 1 : 0;
    }
    __syncthreads();
    return s_last != 0;
}

// Called by every thread of the completing block: thread t sums tiles t,
// t+256, ... and the block tree-reduces, a fixed order for a given job shape.
template <int NV>
__device__ inline void sum_partials(const GridJob& J, int ntiles, double (&out)[NV]) {
    __threadfence();
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = 0.0;
    for (int t = threadIdx.x; t < ntiles; t += FFT_THREADS) {
#pragma unroll
        for (int k = 0; k < NV; ++k)
            out[k] += __hip_atomic_load(&J.partial[(long)t * 6 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum<NV>(out);
    if (threadIdx.x == 0) __hip_atomic_store(J.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ----------------------------------------------------------------------------
// scalar logic (runs in one thread of the last block), DDAModel.cpp:945-1004
// ----------------------------------------------------------------------------
__device__ inline void bicg_decide(KScalars* ks, double2 rho_new, double rn2) {
    ks->rnorm2 = rn2;
    const double rel = sqrt(rn2 / ks->enorm2);
    const int it = ks->iter;
    if (it >= 0 && rel <= ks->tol) {      // ||r||/||E|| <= tol after a loop step
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

// ----------------------------------------------------------------------------
// X forward input: the vector fed to the transform, with the BiCGSTAB vector
// update fused in (DDAModel.cpp:976, 979).  vi < 0: outside the geometry.
// ----------------------------------------------------------------------------
struct XfScalars {
    double2 alpha, beta, zeta;
};
template <int MODE>
__device__ inline XfScalars xf_scalars(const GridJob& J) {
    XfScalars s{make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    if (MODE == XF_P) { s.beta = J.ks->beta; s.zeta = J.ks->zeta; }
    if (MODE == XF_T) { s.alpha = J.ks->alpha; }
    return s;
}

template <typename C, typename KC, int MODE>
__device__ inline C xf_element(const GridJob& J, long vi, const XfScalars& sc) {
    if (vi < 0) return mk<C>(0.0, 0.0);
    if (MODE == XF_PLAIN) {
        return static_cast<const C*>(J.xin)[vi];
    } else if (MODE == XF_SOL) {
        return from_d<C>(J.P[vi]);
    } else if (MODE == XF_P) {
        KC* r = static_cast<KC*>(J.r);
        KC* p = static_cast<KC*>(J.p);
        KC* v = static_cast<KC*>(J.v);
        const double2 w = csub(to_d(p[vi]), dcmul(sc.zeta, to_d(v[vi])));   // p - u, u = zeta v
        const KC pk = from_d<KC>(cadd(to_d(r[vi]), dcmul(sc.beta, w)));
        p[vi] = pk;
        return from_d<C>(to_d(pk));
    } else {   // XF_T
        KC* r = static_cast<KC*>(J.r);
        KC* v = static_cast<KC*>(J.v);
        KC* t = static_cast<KC*>(J.t);
        const KC tk = from_d<KC>(csub(to_d(r[vi]), dcmul(sc.alpha, to_d(v[vi]))));
        t[vi] = tk;
        return from_d<C>(to_d(tk));
    }
}

// dipole vector index of box cell (row q = z*ny + y, x) component c, or -1
__device__ inline long vec_index(const GridJob& J, int q, int x, int c) {
    if (J.box_map == nullptr) return 3L * ((long)q * J.nx + x) + c;
    const int dip = J.box_map[(long)q * J.nx + x];
    return dip >= 0 ? 3L * dip + c : -1;
}

// ----------------------------------------------------------------------------
// X inverse epilogue: y = conv + al u and what the mode does with it.
// conv is already divided by M.  acc gathers the per-thread dot partials.
// ----------------------------------------------------------------------------
template <typename C, typename KC, int MODE>
__device__ inline void xi_element(const GridJob& J, long vi, C conv, const C* __restrict__ al, double (&acc)[3]) {
    const C a = al[vi / 3];
    if (MODE == XI_PLAIN) {
        const C u = static_cast<const C*>(J.xin)[vi];
        static_cast<C*>(J.yout)[vi] = cadd(conv, cmul(a, u));
    } else if (MODE == XI_V) {
        const KC* p = static_cast<const KC*>(J.p);
        const KC* r0 = static_cast<const KC*>(J.r0);
        const C y = cadd(conv, cmul(a, from_d<C>(to_d(p[vi]))));
        const KC vk = from_d<KC>(to_d(y));
        static_cast<KC*>(J.v)[vi] = vk;
        const double2 d = cmulc(to_d(vk), to_d(r0[vi]));   // conj(r0) v
        acc[0] += d.x;
        acc[1] += d.y;
    } else if (MODE == XI_S) {
        const KC* t = static_cast<const KC*>(J.t);
        const KC tk = t[vi];
        const C y = cadd(conv, cmul(a, from_d<C>(to_d(tk))));
        const KC sk = from_d<KC>(to_d(y));
        static_cast<KC*>(J.s)[vi] = sk;
        const double2 sd = to_d(sk), td = to_d(tk);
        const double2 d = cmulc(td, sd);   // conj(s) t
        acc[0] += d.x;
        acc[1] += d.y;
        acc[2] += sd.x * sd.x + sd.y * sd.y;
    } else if (MODE == XI_RESID) {
        const double2 y = cadd(to_d(conv), dcmul(to_d(a), J.P[vi]));
        const double2 rr = csub(J.rhs[vi], y);
        const KC rk = from_d<KC>(rr);
        static_cast<KC*>(J.r)[vi] = rk;
        const double2 d = cmulc(to_d(rk), to_d(static_cast<const KC*>(J.r0)[vi]));
        acc[0] += d.x;
        acc[1] += d.y;
        acc[2] += rr.x * rr.x + rr.y * rr.y;
    }
}

// Reduce the dot partials of a job over its X tiles and set alpha / zeta /
// the reliable-update state in the completing block (all threads call it).
template <int MODE>
__device__ inline void xi_finish(const GridJob& J, int tile, int ntiles, double (&acc)[3]) {
    if (MODE == XI_PLAIN) return;
    block_sum<3>(acc);
    if (!publish_partials<3>(J, tile, ntiles, acc)) return;
    double sum[3];
    sum_partials<3>(J, ntiles, sum);
    if (threadIdx.x != 0) return;
    KScalars* ks = J.ks;
    const double2 dot = make_double2(sum[0], sum[1]);
    if (MODE == XI_V) {
        ks->alpha = ddiv(ks->rho, dot);                          // alpha = <r0,r>/<r0,Ap>
    } else if (MODE == XI_S) {
        ks->zeta = make_double2(dot.x / sum[2], dot.y / sum[2]);  // zeta = <At,t>/<At,At>
    } else if (MODE == XI_RESID) {
        ks->need_ru = 0;
        ks->n_ru += 1;
        ks->maxr2 = sum[2];
        if (ks->status == JOB_ACTIVE) bicg_decide(ks, dot, sum[2]);
    }
}

// ----------------------------------------------------------------------------
// Green's-tensor multiply at spectral point (kx, ky, kz).
// Component (00,01,02,11,12,22) is odd in the axes {}, {x,y}, {x,z}, {},
// {y,z}, {} respectively, so only k <= L/2 per axis is stored and
// Ghat(L-k) = +-Ghat(k).
// ----------------------------------------------------------------------------
template <typename C>
__device__ inline void green_apply(const C* __restrict__ G, const GridJob& J, int kx, int ky, int kz,
                                   C& v0, C& v1, C& v2) {
    int sx = 1, sy = 1, sz = 1;
    if (2 * kx > J.lx) { kx = J.lx - kx; sx = -1; }
    if (2 * ky > J.ly) { ky = J.ly - ky; sy = -1; }
    if (2 * kz > J.lz) { kz = J.lz - kz; sz = -1; }
    const long cs = (long)J.hz * J.hy * J.hx;
    const long off = ((long)kz * J.hy + ky) * J.hx + kx;
    const C g0 = G[off];
    C g1 = G[cs + off];
    C g2 = G[2 * cs + off];
    const C g3 = G[3 * cs + off];
    C g4 = G[4 * cs + off];
    const C g5 = G[5 * cs + off];
    if (sx * sy < 0) g1 = cscale(g1, -1);
    if (sx * sz < 0) g2 = cscale(g2, -1);
    if (sy * sz < 0) g4 = cscale(g4, -1);
    const C a0 = v0, a1 = v1, a2 = v2;
    v0 = cadd(cadd(cmul(g0, a0), cmul(g1, a1)), cmul(g2, a2));
    v1 = cadd(cadd(cmul(g1, a0), cmul(g3, a1)), cmul(g4, a2));
    v2 = cadd(cadd(cmul(g2, a0), cmul(g4, a1)), cmul(g5, a2));
}

}  // namespace tdda

// Length-specialised FFT passes (gfx950): the same five-pass matvec as the
// generic kernels, for jobs whose transform length has a compile-time split
// L = L1 * L2 (fft_small.h).  Every pencil goes through one LDS exchange per
// 1-D transform (two for the fused z forward / multiply / inverse), with
// butterflies, twiddles and indices resolved at compile time.
#include "fft_fast.h"
#include "pass_util.h"

namespace tdda {

namespace {

template <typename C>
__device__ inline void load_twiddles(C* twl, const C* __restrict__ twg, int L) {
    for (int i = threadIdx.x; i < L; i += FFT_THREADS) twl[i] = gload<C>(twg, (unsigned)i);
}

// ----------------------------------------------------------------------------
// X forward (vector op fused) and X inverse (epilogue fused)
// ----------------------------------------------------------------------------
template <typename Real, typename KReal, int MODE>
__global__ void __launch_bounds__(FFT_THREADS)
k_xfwd_fast(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap, const cx_t<Real>* __restrict__ twpool,
            int gate, const int* __restrict__ limit) {
    using C = cx_t<Real>;
    using KC = cx_t<KReal>;
    extern __shared__ __align__(16) unsigned char smem_xff[];
    if (beyond_limit(limit)) return;
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    if (!gate_open(J, gate)) return;
    with_split_pass<PASS_XF>(J.lx, J.var[sizeof(Real) == 4 ? 0 : 1][PASS_XF], [&](auto s1, auto s2) {
        constexpr int L1 = decltype(s1)::value, L2 = decltype(s2)::value, L = L1 * L2;
        using X = XFast<L1, L2, fast_nr_x<Real>(L1, L2), C>;
        C* lds = reinterpret_cast<C*>(smem_xff);
        C* twl = lds + X::LDS;
        load_twiddles(twl, twpool + J.px.tw_off, L);
        const int q0 = jt.y * fast_nr_x<Real>(L1, L2);
        const int nr = min(fast_nr_x<Real>(L1, L2), J.ny * J.nz - q0);
        const XfScalars sc = xf_scalars<MODE>(J);
        double xacc[XF_NACC] = {};
        __syncthreads();
        for (int t = threadIdx.x; t < X::TA; t += FFT_THREADS)
            X::fwd_a(t, lds, twl, nr, J.nx, [&](int r, int c, int x) {
                return xf_element<C, KC, MODE>(J, vec_index(J, q0 + r, x, c), sc, xacc);
            });
        xf_finish<MODE>(J, jt.y, xacc);
        __syncthreads();
        C* g = static_cast<C*>(J.grid);
        const unsigned ny = J.ny, ly = J.ly;
        const unsigned S = (unsigned)J.nz * ly * L;
        const unsigned long long mny = J.mny;
        if (sizeof(Real) == 4 && J.tr32) {   // [kx][3][nz][ny] for the fused y/z pass
            const unsigned nzy = (unsigned)J.nz * ny, slab = 3u * nzy;
            for (int t = threadIdx.x; t < X::TB; t += FFT_THREADS)
                X::fwd_b(t, lds, nr, [&](int r, int c, int x, C v) {
                    gstore<C>(g, x * slab + c * nzy + (unsigned)(q0 + r), v);
                });
            return;
        }
        for (int t = threadIdx.x; t < X::TB; t += FFT_THREADS)
            X::fwd_b(t, lds, nr, [&](int r, int c, int x, C v) {
                const unsigned q = q0 + r;
                const unsigned z = fdiv(q, mny);
                const unsigned y = q - z * ny;
                gstore<C>(g, c * S + (z * ly + y) * L + x, v);
            });
    });
}

template <typename Real, typename KReal, int MODE>
__global__ void __launch_bounds__(FFT_THREADS)
k_xinv_fast(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap, const cx_t<Real>* __restrict__ twpool,
            int gate, const int* __restrict__ limit) {
    using C = cx_t<Real>;
    using KC = cx_t<KReal>;
    extern __shared__ __align__(16) unsigned char smem_xif[];
    if (beyond_limit(limit)) return;
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    if (!gate_open(J, gate)) return;
    double acc[XI_NACC] = {};
    with_split_pass<PASS_X>(J.lx, J.var[sizeof(Real) == 4 ? 0 : 1][PASS_X], [&](auto s1, auto s2) {
        constexpr int L1 = decltype(s1)::value, L2 = decltype(s2)::value, L = L1 * L2;
        using X = XFast<L1, L2, fast_nr_x<Real>(L1, L2), C>;
        C* lds = reinterpret_cast<C*>(smem_xif);
        C* twl = lds + X::LDS;
        load_twiddles(twl, twpool + J.px.tw_off, L);
        const int q0 = jt.y * fast_nr_x<Real>(L1, L2);
        const int nr = min(fast_nr_x<Real>(L1, L2), J.ny * J.nz - q0);
        const C* g = static_cast<const C*>(J.grid);
        const unsigned ny = J.ny, ly = J.ly;
        const unsigned S = (unsigned)J.nz * ly * L;
        const unsigned long long mny = J.mny;
        __syncthreads();
        if (sizeof(Real) == 4 && J.tr32) {   // [kx][3][nz][ny] after the fused y/z pass
            const unsigned nzy = (unsigned)J.nz * ny, slab = 3u * nzy;
            for (int t = threadIdx.x; t < X::TA; t += FFT_THREADS)
                X::inv_a(t, lds, twl, nr, [&](int r, int c, int x) -> C {
                    return gload<C>(g, x * slab + c * nzy + (unsigned)(q0 + r));
                });
        } else {
            for (int t = threadIdx.x; t < X::TA; t += FFT_THREADS)
                X::inv_a(t, lds, twl, nr, [&](int r, int c, int x) -> C {
                    const unsigned q = q0 + r;
                    const unsigned z = fdiv(q, mny);
                    const unsigned y = q - z * ny;
                    return gload<C>(g, c * S + (z * ly + y) * L + x);
                });
        }
        __syncthreads();
        const Real inv_m = (Real)(1.0 / ((double)J.lx * (double)J.ly * (double)J.lz));
        const C* al = Tile<Real>::al(J);
        for (int t = threadIdx.x; t < X::TB; t += FFT_THREADS)
            X::inv_b(t, lds, nr, J.nx, [&](int r, int c, int x, C v) {
                const int vi = vec_index(J, q0 + r, x, c);
                if (vi >= 0) xi_element<C, KC, MODE>(J, vi, cscale(v, inv_m), al, acc);
            });
    });
    xi_finish<MODE>(J, jt.y, acc);
}

// ----------------------------------------------------------------------------
// Y pass (forward: y < Ny -> all y; inverse: all y -> y < Ny)
// ----------------------------------------------------------------------------
template <typename Real, int DIR>
__global__ void __launch_bounds__(FFT_THREADS)
k_ypass_fast(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap, const cx_t<Real>* __restrict__ twpool,
             int gate, const int* __restrict__ limit) {
    using C = cx_t<Real>;
    extern __shared__ __align__(16) unsigned char smem_yf[];
    if (beyond_limit(limit)) return;
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    if (!gate_open(J, gate)) return;
    with_split_pass<PASS_Y>(J.ly, J.var[sizeof(Real) == 4 ? 0 : 1][PASS_Y], [&](auto s1, auto s2) {
        constexpr int L1 = decltype(s1)::value, L2 = decltype(s2)::value, L = L1 * L2;
        constexpr int NP = fast_np_y<Real>(L1, L2);
        using Y = YFast<L1, L2, NP, DIR, C>;
        C* lds = reinterpret_cast<C*>(smem_yf);
        C* twl = lds + Y::LDS;
        load_twiddles(twl, twpool + J.py.tw_off, L);
        const unsigned lx = J.lx;
        const int pl0 = jt.y * NP;
        const int nb = min(NP, 3 * J.nz * (int)lx - pl0);
        const unsigned long long mlx = J.mlx;
        C* g = static_cast<C*>(J.grid);
        auto base = [&](int b) -> unsigned {
            const unsigned p = pl0 + b;
            const unsigned cz = fdiv(p, mlx);
            return cz * L * lx + (p - cz * lx);
        };
        const int n_in = DIR < 0 ? J.ny : L;
        const int n_out = DIR < 0 ? L : J.ny;
        __syncthreads();
        for (int t = threadIdx.x; t < Y::TA; t += FFT_THREADS)
            Y::step_a(t, lds, twl, nb, n_in, [&](int b, int pos) -> C { return gload<C>(g, base(b) + (unsigned)pos * lx); });
        __syncthreads();
        for (int t = threadIdx.x; t < Y::TB; t += FFT_THREADS)
            Y::step_b(t, lds, nb, n_out, [&](int b, int pos, C v) { gstore<C>(g, base(b) + (unsigned)pos * lx, v); });
    });
}

// ----------------------------------------------------------------------------
// Z pass: forward z FFT, Green's multiply, inverse z FFT (z < Nz in and out)
// ----------------------------------------------------------------------------
// z-mirror pairs in the multiply step (one octant record per pair of
// spectral points) measured slower than one task per k1: fewer threads then
// carry the step's loads.  The c64 kernel of the L2 = 4 splits shares the
// records between mirror lanes of a wave instead (ZFast XCHG).
constexpr bool kZPairs = false;

template <typename Real, int CLS>
__global__ void __launch_bounds__(FFT_THREADS)
k_zpass_fast(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap, const cx_t<Real>* __restrict__ twpool,
             int gate, const int* __restrict__ limit) {
    using C = cx_t<Real>;
    extern __shared__ __align__(16) unsigned char smem_zf[];
    if (beyond_limit(limit)) return;
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    if (!gate_open(J, gate)) return;
    with_split_pass<PASS_Z>(J.lz, J.var[sizeof(Real) == 4 ? 0 : 1][PASS_Z], [&](auto s1, auto s2) {
        if constexpr (z_class_of<decltype(s2)>() != CLS) return;   // the other class's launch
        constexpr int L1 = decltype(s1)::value, L2 = decltype(s2)::value, L = L1 * L2;
        constexpr int NP = fast_np_z<Real>(L1, L2);
        using Z = ZFast<L1, L2, NP, C, kZPairs, sizeof(Real) == 4 && L2 == 4>;
        C* lds = reinterpret_cast<C*>(smem_zf);
        C* twl = lds + Z::LDS;
        load_twiddles(twl, twpool + J.pz.tw_off, L);
        const unsigned lx = J.lx;
        const int nz = J.nz;
        const unsigned plane = (unsigned)J.ly * lx;
        const unsigned S = (unsigned)nz * plane;
        const int pl0 = jt.y * NP;
        const int nb = min(NP, (int)plane - pl0);
        C* g = static_cast<C*>(J.grid) + pl0;
        const C* G = Tile<Real>::green(J);
        const unsigned long long mlx = J.mlx;
        __syncthreads();
        for (int t = threadIdx.x; t < Z::TA; t += FFT_THREADS)
            Z::fwd_a(t, lds, twl, nb, nz, [&](int c, int b, int pos) -> C { return gload<C>(g, c * S + pos * plane + b); });
        __syncthreads();
        for (int t = threadIdx.x; t < Z::TB; t += FFT_THREADS)
            Z::mid(
                t, lds, twl, nb,
                [&](int b, int kz, C (&gg)[6]) {
                    const unsigned p = pl0 + b;
                    const unsigned ky = fdiv(p, mlx);
                    green_fetch(G, J, (int)(p - ky * lx), (int)ky, kz, gg);
                },
                [&](int, int, const C (&mine)[6], C (&out)[6]) {
                    static_for<6>([&](auto q) {
                        constexpr int k = decltype(q)::value;
                        out[k].x = __shfl_xor(mine[k].x, 32);
                        out[k].y = __shfl_xor(mine[k].y, 32);
                    });
                });
        __syncthreads();
        for (int t = threadIdx.x; t < Z::TA; t += FFT_THREADS)
            Z::inv_a(t, lds, nb, nz, [&](int c, int b, int pos, C v) { gstore<C>(g, c * S + pos * plane + b, v); });
    });
}

template <typename Real, typename KReal>
void xfwd_fast(const PassMap& pm, const GridJob* jobs, const cx_t<Real>* tw, int mode, int gate,
               hipStream_t st) {
    const dim3 grid(pm.n), blk(FFT_THREADS);
    switch (mode) {
        case XF_PLAIN: hipLaunchKernelGGL((k_xfwd_fast<Real, KReal, XF_PLAIN>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        case XF_P: hipLaunchKernelGGL((k_xfwd_fast<Real, KReal, XF_P>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        case XF_T: hipLaunchKernelGGL((k_xfwd_fast<Real, KReal, XF_T>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        case XF_SOL: hipLaunchKernelGGL((k_xfwd_fast<Real, KReal, XF_SOL>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        default: throw Error(ERR_INVALID, "bad X-forward mode");
    }
}

template <typename Real, typename KReal>
void xinv_fast(const PassMap& pm, const GridJob* jobs, const cx_t<Real>* tw, int mode, int gate,
               hipStream_t st) {
    const dim3 grid(pm.n), blk(FFT_THREADS);
    switch (mode) {
        case XI_PLAIN: hipLaunchKernelGGL((k_xinv_fast<Real, KReal, XI_PLAIN>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        case XI_V: hipLaunchKernelGGL((k_xinv_fast<Real, KReal, XI_V>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        case XI_S: hipLaunchKernelGGL((k_xinv_fast<Real, KReal, XI_S>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        case XI_RESID: hipLaunchKernelGGL((k_xinv_fast<Real, KReal, XI_RESID>), grid, blk, pm.lds, st, jobs, pm.map, tw, gate, pm.limit); break;
        default: throw Error(ERR_INVALID, "bad X-inverse mode");
    }
}

}  // namespace

// Krylov precision follows the transform precision except for FP64 transforms
// feeding c64 Krylov vectors (reliable updates in refined mode).
void launch_fast_xfwd(const PassMap& pm, const GridJob* jobs, const TwiddlePools& tw, bool c64, bool krylov32,
                      int mode, int gate, hipStream_t st) {
    if (pm.n == 0) return;
    if (c64) xfwd_fast<float, float>(pm, jobs, tw.tw32, mode, gate, st);
    else if (krylov32) xfwd_fast<double, float>(pm, jobs, tw.tw64, mode, gate, st);
    else xfwd_fast<double, double>(pm, jobs, tw.tw64, mode, gate, st);
}

void launch_fast_xinv(const PassMap& pm, const GridJob* jobs, const TwiddlePools& tw, bool c64, bool krylov32,
                      int mode, int gate, hipStream_t st) {
    if (pm.n == 0) return;
    if (c64) xinv_fast<float, float>(pm, jobs, tw.tw32, mode, gate, st);
    else if (krylov32) xinv_fast<double, float>(pm, jobs, tw.tw64, mode, gate, st);
    else xinv_fast<double, double>(pm, jobs, tw.tw64, mode, gate, st);
}

void launch_fast_y(const PassMap& pm, const GridJob* jobs, const TwiddlePools& tw, bool c64, int dir, int gate,
                   hipStream_t st) {
    if (pm.n == 0) return;
    const dim3 grid(pm.n), blk(FFT_THREADS);
    if (c64) {
        if (dir < 0) hipLaunchKernelGGL((k_ypass_fast<float, -1>), grid, blk, pm.lds, st, jobs, pm.map, tw.tw32, gate, pm.limit);
        else hipLaunchKernelGGL((k_ypass_fast<float, 1>), grid, blk, pm.lds, st, jobs, pm.map, tw.tw32, gate, pm.limit);
    } else {
        if (dir < 0) hipLaunchKernelGGL((k_ypass_fast<double, -1>), grid, blk, pm.lds, st, jobs, pm.map, tw.tw64, gate, pm.limit);
        else hipLaunchKernelGGL((k_ypass_fast<double, 1>), grid, blk, pm.lds, st, jobs, pm.map, tw.tw64, gate, pm.limit);
    }
}

void launch_fast_z(const PassMap& pm, const GridJob* jobs, const TwiddlePools& tw, bool c64, int gate,
                   hipStream_t st) {
    if (pm.n == 0) return;
    const dim3 grid(pm.n), blk(FFT_THREADS);
    // one launch per length class present; blocks of the other class return
    if (pm.cls & 1) {
        if (c64) hipLaunchKernelGGL((k_zpass_fast<float, 0>), grid, blk, pm.lds_cls[0], st, jobs, pm.map, tw.tw32, gate, pm.limit);
        else hipLaunchKernelGGL((k_zpass_fast<double, 0>), grid, blk, pm.lds_cls[0], st, jobs, pm.map, tw.tw64, gate, pm.limit);
    }
    if (pm.cls & 2) {
        if (c64) hipLaunchKernelGGL((k_zpass_fast<float, 1>), grid, blk, pm.lds_cls[1], st, jobs, pm.map, tw.tw32, gate, pm.limit);
        else hipLaunchKernelGGL((k_zpass_fast<double, 1>), grid, blk, pm.lds_cls[1], st, jobs, pm.map, tw.tw64, gate, pm.limit);
    }
}

// tile geometry of the specialised path (host)
bool fast_tiles(int L, int pass, bool c64, int v, int* per_tile, size_t* lds_bytes) {
    auto tiles = [&](auto s1, auto s2) {
        constexpr int L1 = decltype(s1)::value, L2 = decltype(s2)::value, LL = L1 * L2;
        const size_t cs = c64 ? sizeof(float2) : sizeof(double2);
        if (pass == PASS_X || pass == PASS_XF) {
            *per_tile = c64 ? fast_nr_x<float>(L1, L2) : fast_nr_x<double>(L1, L2);
            *lds_bytes = ((size_t)3 * *per_tile * fast_ps_x(L1, L2) + LL) * cs;
        } else if (pass == PASS_Y) {
            *per_tile = c64 ? fast_np_y<float>(L1, L2) : fast_np_y<double>(L1, L2);
            *lds_bytes = ((size_t)*per_tile * LL + LL) * cs;
        } else {
            *per_tile = c64 ? fast_np_z<float>(L1, L2) : fast_np_z<double>(L1, L2);
            *lds_bytes = ((size_t)3 * *per_tile * LL + LL) * cs;
        }
    };
    if (pass == PASS_Z) return with_split_pass<PASS_Z>(L, v, tiles);
    if (pass == PASS_XF) return with_split_pass<PASS_XF>(L, v, tiles);
    if (pass == PASS_Y) return with_split_pass<PASS_Y>(L, v, tiles);
    return with_split_pass<PASS_X>(L, v, tiles);
}

}  // namespace tdda

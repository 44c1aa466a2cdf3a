// Mixed-radix Stockham FFT building blocks for LDS-resident pencil tiles.
//
// A workgroup of FFT_THREADS threads owns a tile of `np` pencils of length L.
// The tile lives in LDS as [pos][bp] (pencil index fastest, bp >= np), so that
// threads with consecutive butterfly ids touch consecutive LDS words and, for
// strided pencils, consecutive global addresses.
//
// Stage s of radix R (Stockham autosort, Van Loan 1992, sec. 1.7): with Ns the
// product of the radices already applied, butterfly j (0 <= j < L/R) of a
// pencil reads positions j + r*L/R, applies twiddles w^(r*k), k = j mod Ns,
// w = exp(DIR*2*pi*i/(Ns*R)), does an R-point DFT and writes positions
// (j/Ns)*Ns*R + k + r*Ns.  After the last stage the pencil is in natural order.
#pragma once

#include "common.h"

namespace tdda {

constexpr int FFT_THREADS = 256;
constexpr int FFT_MAX_STAGES = 8;

// host+device description of one transform length
struct FftPlan {
    int L;                        // length
    int nst;                      // number of stages
    int radix[FFT_MAX_STAGES];    // radices in forward order
    int tw_off;                   // offset of the L-entry twiddle table in the pool
};

// ---------------------------------------------------------------------------
// R-point DFTs in registers.  DIR = -1 forward, +1 inverse (unnormalised).
// ---------------------------------------------------------------------------
template <int R, int DIR, typename C> struct Dft;

template <int DIR, typename C> struct Dft<2, DIR, C> {
    __host__ __device__ static inline void run(C* v) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int DIR, typename C> struct Dft<3, DIR, C> {
    __host__ __device__ static inline void run(C* v) {
        using R = decltype(v[0].x);
        const R h = (R)0.8660254037844386467637231707529362;  // sqrt(3)/2
        C a = v[0];
        C s = cadd(v[1], v[2]);
        C d = csub(v[1], v[2]);
        C m = csub(a, cscale(s, (R)0.5));
        C e = mul_i<DIR>(cscale(d, h));
        v[0] = cadd(a, s);
        v[1] = cadd(m, e);
        v[2] = csub(m, e);
    }
};

template <int DIR, typename C> struct Dft<4, DIR, C> {
    __host__ __device__ static inline void run(C* v) {
        C t0 = cadd(v[0], v[2]);
        C t1 = csub(v[0], v[2]);
        C t2 = cadd(v[1], v[3]);
        C t3 = mul_i<DIR>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <int DIR, typename C> struct Dft<5, DIR, C> {
    __host__ __device__ static inline void run(C* v) {
        using R = decltype(v[0].x);
        const R c1 = (R)0.3090169943749474241022934171828191;   // cos(2pi/5)
        const R c2 = (R)-0.8090169943749474241022934171828191;  // cos(4pi/5)
        const R s1 = (R)0.9510565162951535721164393333793821;   // sin(2pi/5)
        const R s2 = (R)0.5877852522924731291687059546390728;   // sin(4pi/5)
        C x0 = v[0];
        C a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        C a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        C t1 = cadd(x0, cadd(cscale(a1, c1), cscale(a2, c2)));
        C t2 = cadd(x0, cadd(cscale(a1, c2), cscale(a2, c1)));
        C u1 = mul_i<DIR>(cadd(cscale(b1, s1), cscale(b2, s2)));
        C u2 = mul_i<DIR>(csub(cscale(b1, s2), cscale(b2, s1)));
        v[0] = cadd(x0, cadd(a1, a2));
        v[1] = cadd(t1, u1);
        v[4] = csub(t1, u1);
        v[2] = cadd(t2, u2);
        v[3] = csub(t2, u2);
    }
};

template <int DIR, typename C> struct Dft<8, DIR, C> {
    __host__ __device__ static inline void run(C* v) {
        using R = decltype(v[0].x);
        const R h = (R)0.7071067811865475244008443621048490;  // sqrt(2)/2
        C e[4] = {v[0], v[2], v[4], v[6]};
        C o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, DIR, C>::run(e);
        Dft<4, DIR, C>::run(o);
        // w^1 = h(1 + DIR i), w^2 = DIR i, w^3 = h(-1 + DIR i)
        C w1o = cscale(cadd(o[1], mul_i<DIR>(o[1])), h);
        C w2o = mul_i<DIR>(o[2]);
        C w3o = cscale(csub(mul_i<DIR>(o[3]), o[3]), h);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], w1o);
        v[5] = csub(e[1], w1o);
        v[2] = cadd(e[2], w2o);
        v[6] = csub(e[2], w2o);
        v[3] = cadd(e[3], w3o);
        v[7] = csub(e[3], w3o);
    }
};

template <int R, int CAP> struct Umax {
    static constexpr int value = (CAP + R * FFT_THREADS - 1) / (R * FFT_THREADS);
};

// Registers of one stage for up to U butterflies of radix R over NC
// components (the same butterfly ids for every component).
template <int R, int U, int NC, typename C> struct StageRegs {
    C v[NC][U][R];
    int pos[U];  // output base position (first of R, stride ns)
    int b[U];    // pencil
};

// Load the inputs of every butterfly this thread owns, twiddle, DFT.
// load(c, b, pos) must return zero for positions the caller treats as empty.
template <int R, int DIR, int U, int NC, typename C, typename TW, typename LoadF>
__device__ inline void stage_compute(StageRegs<R, U, NC, C>& s, int np, int L, int ns,
                                     const TW* __restrict__ tw, LoadF load) {
    const int lr = L / R;
    const int nbf = np * lr;
    const int twstep = L / (ns * R);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int bf = threadIdx.x + u * FFT_THREADS;
        if (bf < nbf) {
            const int b = bf % np;
            const int j = bf / np;
            const int k = j % ns;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
#pragma unroll
                for (int r = 0; r < R; ++r) s.v[c][u][r] = load(c, b, j + r * lr);
            }
            if (ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    TW w = tw[r * k * twstep];
                    C wc;
                    wc.x = w.x;
                    wc.y = DIR < 0 ? w.y : -w.y;
#pragma unroll
                    for (int c = 0; c < NC; ++c) s.v[c][u][r] = cmul(s.v[c][u][r], wc);
                }
            }
#pragma unroll
            for (int c = 0; c < NC; ++c) Dft<R, DIR, C>::run(s.v[c][u]);
            s.pos[u] = (j / ns) * ns * R + k;
            s.b[u] = b;
        } else {
            s.pos[u] = -1;
            s.b[u] = 0;
        }
    }
}

// Write the stage outputs: store(c, b, pos, value).
template <int R, int U, int NC, typename C, typename StoreF>
__device__ inline void stage_store(const StageRegs<R, U, NC, C>& s, int ns, StoreF store) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (s.pos[u] >= 0) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
#pragma unroll
                for (int r = 0; r < R; ++r) store(c, s.b[u], s.pos[u] + r * ns, s.v[c][u][r]);
            }
        }
    }
}

// A full LDS-resident stage: every thread reads, barrier, every thread writes,
// barrier.  lds layout: comp c, position pos, pencil b at (c*L + pos)*bp + b.
template <int R, int DIR, int CAP, int NC, typename C, typename TW>
__device__ inline void lds_stage(C* lds, int bp, int np, int L, int ns, const TW* __restrict__ tw) {
    constexpr int U = Umax<R, CAP>::value;
    StageRegs<R, U, NC, C> s;
    stage_compute<R, DIR, U, NC>(s, np, L, ns, tw,
                                 [&](int c, int b, int pos) { return lds[(c * L + pos) * bp + b]; });
    __syncthreads();
    stage_store<R, U, NC>(s, ns, [&](int c, int b, int pos, C v) { lds[(c * L + pos) * bp + b] = v; });
    __syncthreads();
}

// Dispatch a runtime radix to the compile-time stage.
template <int DIR, int CAP, int NC, typename C, typename TW>
__device__ inline void lds_stage_rt(int R, C* lds, int bp, int np, int L, int ns, const TW* __restrict__ tw) {
    switch (R) {
        case 2: lds_stage<2, DIR, CAP, NC>(lds, bp, np, L, ns, tw); break;
        case 3: lds_stage<3, DIR, CAP, NC>(lds, bp, np, L, ns, tw); break;
        case 4: lds_stage<4, DIR, CAP, NC>(lds, bp, np, L, ns, tw); break;
        case 5: lds_stage<5, DIR, CAP, NC>(lds, bp, np, L, ns, tw); break;
        case 8: lds_stage<8, DIR, CAP, NC>(lds, bp, np, L, ns, tw); break;
        default: break;
    }
}

// One stage with caller-supplied load/store (global or LDS); `mid_sync`
// inserts the barrier needed when load and store alias the same LDS tile.
template <int R, int DIR, int CAP, int NC, typename C, typename TW, typename LoadF, typename StoreF>
__device__ inline void io_stage(int np, int L, int ns, const TW* __restrict__ tw, LoadF load, StoreF store,
                                bool mid_sync) {
    constexpr int U = Umax<R, CAP>::value;
    StageRegs<R, U, NC, C> s;
    stage_compute<R, DIR, U, NC>(s, np, L, ns, tw, load);
    if (mid_sync) __syncthreads();
    stage_store<R, U, NC>(s, ns, store);
}

template <int DIR, int CAP, int NC, typename C, typename TW, typename LoadF, typename StoreF>
__device__ inline void io_stage_rt(int R, int np, int L, int ns, const TW* __restrict__ tw, LoadF load,
                                   StoreF store, bool mid_sync) {
    switch (R) {
        case 2: io_stage<2, DIR, CAP, NC, C>(np, L, ns, tw, load, store, mid_sync); break;
        case 3: io_stage<3, DIR, CAP, NC, C>(np, L, ns, tw, load, store, mid_sync); break;
        case 4: io_stage<4, DIR, CAP, NC, C>(np, L, ns, tw, load, store, mid_sync); break;
        case 5: io_stage<5, DIR, CAP, NC, C>(np, L, ns, tw, load, store, mid_sync); break;
        case 8: io_stage<8, DIR, CAP, NC, C>(np, L, ns, tw, load, store, mid_sync); break;
        default: break;
    }
}

// Radix of stage s in direction DIR (inverse runs the plan backwards).
template <int DIR> __device__ inline int stage_radix(const FftPlan& pl, int s) {
    return DIR < 0 ? pl.radix[s] : pl.radix[pl.nst - 1 - s];
}

// Whole transform, LDS in -> LDS out.  The first stage treats positions
// >= n_in as zero (the pruned, zero-padded half of the pencil).
template <int DIR, int CAP, int NC, typename C, typename TW>
__device__ inline void lds_fft(const FftPlan& pl, C* lds, int bp, int np, int n_in, const TW* __restrict__ twpool) {
    const TW* tw = twpool + pl.tw_off;
    const int L = pl.L;
    int ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = stage_radix<DIR>(pl, s);
        if (s == 0 && n_in < L) {
            io_stage_rt<DIR, CAP, NC, C>(
                R, np, L, ns, tw,
                [&](int c, int b, int pos) {
                    return pos < n_in ? lds[(c * L + pos) * bp + b] : mk<C>(0.0, 0.0);
                },
                [&](int c, int b, int pos, C v) { lds[(c * L + pos) * bp + b] = v; }, true);
            __syncthreads();
        } else {
            lds_stage_rt<DIR, CAP, NC>(R, lds, bp, np, L, ns, tw);
        }
        ns *= R;
    }
}

}  // namespace tdda

// Register-resident small DFTs with compile-time twiddles, and the two-step
// (L = L1 * L2) pencil FFT used by the specialised passes.
//
// A length-L transform of a pencil is split Cooley-Tukey style: with input
// index n = n2 + L2*n1 and output index k = k1 + L1*k2,
//
//   X[k1 + L1 k2] = sum_n2 w_L2^(n2 k2) [ w_L^(n2 k1) sum_n1 w_L1^(n1 k1) x[n2 + L2 n1] ]
//
// Step A (one thread per (pencil, n2)) does the inner L1-point DFT in
// registers and applies w_L^(n2 k1); step B (one thread per (pencil, k1)) does
// the L2-point DFT.  One LDS exchange separates them.  Every index below is a
// compile-time constant once L1 and L2 are fixed.
#pragma once

#include <type_traits>

#include "fft_core.h"

#ifndef TDDA_HD
#define TDDA_HD __host__ __device__ __forceinline__
#endif

namespace tdda {

// ---------------------------------------------------------------------------
// compile-time unit roots: (cos, sin)(2 pi k / N), exact for multiples of pi/2
// ---------------------------------------------------------------------------
struct CS {
    double c, s;
};
constexpr double kPiC = 3.14159265358979323846264338327950288;

constexpr double sin_series(double x) {   // |x| <= pi/4
    double x2 = x * x, term = x, sum = x;
    for (int k = 1; k < 14; ++k) {
        term *= -x2 / (double)((2 * k) * (2 * k + 1));
        sum += term;
    }
    return sum;
}
constexpr double cos_series(double x) {
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int k = 1; k < 14; ++k) {
        term *= -x2 / (double)((2 * k - 1) * (2 * k));
        sum += term;
    }
    return sum;
}
constexpr CS unit_root(long k, long N) {
    k = ((k % N) + N) % N;
    const long q = (4 * k) / N;          // quarter turns
    const long r = 4 * k - q * N;        // remainder angle = (pi/2) r / N
    double c = 1.0, s = 0.0;
    if (r != 0) {
        if (2 * r <= N) {
            const double phi = (kPiC / 2) * (double)r / (double)N;
            c = cos_series(phi);
            s = sin_series(phi);
        } else {
            const double phi = (kPiC / 2) * (double)(N - r) / (double)N;
            c = sin_series(phi);
            s = cos_series(phi);
        }
    }
    for (long i = 0; i < q; ++i) {
        const double t = c;
        c = -s;
        s = t;
    }
    return CS{c, s};
}

// ---------------------------------------------------------------------------
// compile-time loops
// ---------------------------------------------------------------------------
template <int I> using IC = std::integral_constant<int, I>;
template <int I, int N, typename F>
TDDA_HD void sfor_impl(F& f) {
    if constexpr (I < N) {
        f(IC<I>{});
        sfor_impl<I + 1, N>(f);
    }
}
template <int N, typename F>
TDDA_HD void static_for(F&& f) {
    sfor_impl<0, N>(f);
}

// v * exp(DIR * 2 pi i E / N) with E, N compile-time; trivial factors cost nothing.
template <int E, int N, int DIR, typename C>
TDDA_HD C tw_const(C v) {
    constexpr int e = ((E % N) + N) % N;
    using R = decltype(v.x);
    if constexpr (e == 0) {
        return v;
    } else if constexpr ((4 * e) % N == 0) {
        constexpr int q = (4 * e) / N;   // 1, 2, 3 quarter turns of exp(DIR i pi/2)
        if constexpr (q == 2) return mk<C>(-v.x, -v.y);
        else if constexpr ((q == 1) == (DIR > 0)) return mul_i<1>(v);   // * i
        else return mul_i<-1>(v);                                        // * -i
    } else {
        constexpr CS w = unit_root((long)DIR * e, N);
        C r;
        r.x = v.x * (R)w.c - v.y * (R)w.s;
        r.y = v.x * (R)w.s + v.y * (R)w.c;
        return r;
    }
}

// ---------------------------------------------------------------------------
// In-place natural-order DFT of N points held in registers
// ---------------------------------------------------------------------------
template <int N, int DIR, typename C> struct SDft;

template <int DIR, typename C> struct SDft<1, DIR, C> { TDDA_HD static void run(C*) {} };
template <int DIR, typename C> struct SDft<2, DIR, C> { TDDA_HD static void run(C* v) { Dft<2, DIR, C>::run(v); } };
template <int DIR, typename C> struct SDft<3, DIR, C> { TDDA_HD static void run(C* v) { Dft<3, DIR, C>::run(v); } };
template <int DIR, typename C> struct SDft<4, DIR, C> { TDDA_HD static void run(C* v) { Dft<4, DIR, C>::run(v); } };
template <int DIR, typename C> struct SDft<5, DIR, C> { TDDA_HD static void run(C* v) { Dft<5, DIR, C>::run(v); } };
template <int DIR, typename C> struct SDft<8, DIR, C> { TDDA_HD static void run(C* v) { Dft<8, DIR, C>::run(v); } };

// N = A*B in registers: n = b + B a, k = ka + A kb.
template <int A, int B, int DIR, typename C>
TDDA_HD void ct_dft(C* v) {
    C t[A][B];
    static_for<B>([&](auto bi) {
        constexpr int b = decltype(bi)::value;
        C col[A];
        static_for<A>([&](auto ai) { col[decltype(ai)::value] = v[b + B * decltype(ai)::value]; });
        SDft<A, DIR, C>::run(col);
        static_for<A>([&](auto ki) {
            constexpr int ka = decltype(ki)::value;
            t[ka][b] = tw_const<ka * b, A * B, DIR>(col[ka]);
        });
    });
    static_for<A>([&](auto ki) {
        constexpr int ka = decltype(ki)::value;
        SDft<B, DIR, C>::run(t[ka]);
        static_for<B>([&](auto kbi) {
            constexpr int kb = decltype(kbi)::value;
            v[ka + A * kb] = t[ka][kb];
        });
    });
}
template <int DIR, typename C> struct SDft<6, DIR, C> { TDDA_HD static void run(C* v) { ct_dft<2, 3, DIR>(v); } };
template <int DIR, typename C> struct SDft<10, DIR, C> { TDDA_HD static void run(C* v) { ct_dft<2, 5, DIR>(v); } };
template <int DIR, typename C> struct SDft<12, DIR, C> { TDDA_HD static void run(C* v) { ct_dft<4, 3, DIR>(v); } };
template <int DIR, typename C> struct SDft<16, DIR, C> { TDDA_HD static void run(C* v) { ct_dft<4, 4, DIR>(v); } };

// N-point DFT whose inputs n >= N/2 are zero (N even): even outputs are the
// N/2-point DFT of x, odd outputs that of x[n] w_N^n.
template <int N, int DIR, typename C>
TDDA_HD void dft_half_in(C* v) {
    static_assert(N % 2 == 0, "half-input DFT needs even N");
    constexpr int H = N / 2;
    C e[H], o[H];
    static_for<H>([&](auto ni) {
        constexpr int n = decltype(ni)::value;
        e[n] = v[n];
        o[n] = tw_const<n, N, DIR>(v[n]);
    });
    SDft<H, DIR, C>::run(e);
    SDft<H, DIR, C>::run(o);
    static_for<H>([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        v[2 * m] = e[m];
        v[2 * m + 1] = o[m];
    });
}

// N-point DFT of which only outputs k < N/2 are wanted (N even): split the
// inputs into even and odd n; out[k] = E[k] + w_N^k O[k], k < N/2.
template <int N, int DIR, typename C>
TDDA_HD void dft_half_out(C* v) {
    static_assert(N % 2 == 0, "half-output DFT needs even N");
    constexpr int H = N / 2;
    C e[H], o[H];
    static_for<H>([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        e[m] = v[2 * m];
        o[m] = v[2 * m + 1];
    });
    SDft<H, DIR, C>::run(e);
    SDft<H, DIR, C>::run(o);
    static_for<H>([&](auto ki) {
        constexpr int k = decltype(ki)::value;
        v[k] = cadd(e[k], tw_const<k, N, DIR>(o[k]));
    });
}

// ---------------------------------------------------------------------------
// Lengths with a specialised two-step plan: L = L1 * L2 (L1 >= L2)
// ---------------------------------------------------------------------------
struct Split {
    int L, L1, L2;
};
constexpr Split kSplits[] = {
    {16, 4, 4},  {20, 5, 4},  {24, 6, 4},   {32, 8, 4},   {40, 8, 5},
    {48, 8, 6},  {64, 8, 8},  {80, 10, 8},  {96, 12, 8},  {128, 16, 8},
};
constexpr int kNumSplits = sizeof(kSplits) / sizeof(kSplits[0]);

// host+device lookup: index into kSplits or -1
__host__ __device__ inline int split_index(int L) {
    for (int i = 0; i < kNumSplits; ++i)
        if (kSplits[i].L == L) return i;
    return -1;
}

}  // namespace tdda

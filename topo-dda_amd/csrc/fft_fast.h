// Length-specialised pencil passes (L = L1 * L2 with both factors fixed at
// compile time).  Each pass is split into per-task functions: the kernels
// call them in block-strided loops separated by barriers, and a CPU harness
// can call them in plain loops, so indexing and arithmetic are checkable
// without a GPU.
//
// Layout of the LDS exchange between step A (task = (pencil, n2), inner
// L1-point DFT over n1, twiddle w_L^(n2 k1)) and step B (task = (pencil, k1),
// L2-point DFT over n2):
//   Y, Z : [component][k1*L2 + n2][pencil]   (pencil fastest; pencils are
//          consecutive x columns, so global accesses are coalesced across
//          the pencil index)
//   X    : [pencil][k1*(L2+1) + n2] with an odd pencil stride (rows are
//          contiguous in memory, so global accesses are coalesced across
//          positions instead).
#pragma once

#include "fft_small.h"

#ifndef TDDA_HD
#define TDDA_HD __host__ __device__ __forceinline__
#endif

namespace tdda {

// twiddle w_L^(j) from the table (forward sign), conjugated for DIR > 0
template <int DIR, typename C>
TDDA_HD C tw_at(const C* __restrict__ tw, int j) {
    C w = tw[j];
    if (DIR > 0) w.y = -w.y;
    return w;
}

// Twiddles w[k] = w_L^(n2 k) (conjugated for DIR > 0), k = 1..K-1, read
// before the task's first LDS store: the table shares the LDS allocation with
// the tile, so a table read issued after a tile store waits for that store,
// and reads interleaved with stores would each pay the full LDS latency.
template <int K, int DIR, typename C>
TDDA_HD void load_tw(const C* __restrict__ tw, int n2, C (&w)[K]) {
    static_for<K>([&](auto i) {
        constexpr int k = decltype(i)::value;
        if constexpr (k > 0) w[k] = tw_at<DIR>(tw, n2 * k);
    });
}

// Pencils / rows per tile: step-A task counts stay at or below one pass of
// a 256-thread block (FP64 tiles hold half as many).
template <typename Real> TDDA_HD constexpr int fast_np_y(int L1, int L2) {
    return (256 / L2) / (sizeof(Real) == 8 ? 2 : 1) + 0 * L1;
}
// (c64, L1 = 12: 16 pencils rather than 21, so that the Z multiply step can
// pair z-mirror tasks 32 lanes apart; 64 of the 256 threads idle)
template <typename Real> TDDA_HD constexpr int fast_np_z(int L1, int L2) {
    return (L1 == 12 && sizeof(Real) == 4 ? 16 : (256 / L1) / (sizeof(Real) == 8 ? 2 : 1)) + 0 * L2;
}
template <typename Real> TDDA_HD constexpr int fast_nr_x(int L1, int L2) {
    return (256 / (3 * L2)) / (sizeof(Real) == 8 ? 2 : 1) + 0 * L1;
}
TDDA_HD constexpr int fast_ps_x(int L1, int L2) { return (L1 * (L2 + 1)) | 1; }

// 32-bit quotient a / d by a precomputed m = ceil(2^32 / d) (exact for a*d < 2^32)
TDDA_HD unsigned fdiv(unsigned a, unsigned long long m) {
    return (unsigned)(((unsigned long long)a * m) >> 32);
}

// ---------------------------------------------------------------------------
// Y pass: np pencils along y, positions pos in [0, L), global address of
// (pencil b, pos) supplied by the caller.
// ---------------------------------------------------------------------------
template <int L1, int L2, int NP, int DIR, typename C>
struct YFast {
    static constexpr int L = L1 * L2;
    static constexpr int TA = NP * L2;
    static constexpr int TB = NP * L1;
    static constexpr int LDS = NP * L;   // complex entries (+ L for twiddles)

    template <typename Load>
    TDDA_HD static void step_a(int tau, C* __restrict__ lds, const C* __restrict__ tw, int nb, int n_in, Load load) {
        const int b = tau % NP;
        const int n2 = tau / NP;
        if (b >= nb) return;
        C w[L1];
        load_tw<L1, DIR>(tw, n2, w);
        C v[L1];
        bool done = false;
        if constexpr (L1 % 2 == 0) {
            if (2 * n_in <= L) {
                static_for<L1 / 2>([&](auto i) {
                    constexpr int n1 = decltype(i)::value;
                    const int pos = n2 + L2 * n1;
                    v[n1] = pos < n_in ? load(b, pos) : mk<C>(0.0, 0.0);
                });
                dft_half_in<L1, DIR>(v);
                done = true;
            }
        }
        if (!done) {
            static_for<L1>([&](auto i) {
                constexpr int n1 = decltype(i)::value;
                const int pos = n2 + L2 * n1;
                v[n1] = pos < n_in ? load(b, pos) : mk<C>(0.0, 0.0);
            });
            SDft<L1, DIR, C>::run(v);
        }
        static_for<L1>([&](auto i) {
            constexpr int k1 = decltype(i)::value;
            if constexpr (k1 > 0) v[k1] = cmul(v[k1], w[k1]);
            lds[(k1 * L2 + n2) * NP + b] = v[k1];
        });
    }

    template <typename Store>
    TDDA_HD static void step_b(int tau, const C* __restrict__ lds, int nb, int n_out, Store store) {
        const int b = tau % NP;
        const int k1 = tau / NP;
        if (b >= nb) return;
        C v[L2];
        static_for<L2>([&](auto i) {
            constexpr int n2 = decltype(i)::value;
            v[n2] = lds[(k1 * L2 + n2) * NP + b];
        });
        if constexpr (L2 % 2 == 0) {
            if (2 * n_out <= L) {
                dft_half_out<L2, DIR>(v);
                static_for<L2 / 2>([&](auto i) {
                    constexpr int k2 = decltype(i)::value;
                    const int pos = k1 + L1 * k2;
                    if (pos < n_out) store(b, pos, v[k2]);
                });
                return;
            }
        }
        SDft<L2, DIR, C>::run(v);
        static_for<L2>([&](auto i) {
            constexpr int k2 = decltype(i)::value;
            const int pos = k1 + L1 * k2;
            if (pos < n_out) store(b, pos, v[k2]);
        });
    }
};

// ---------------------------------------------------------------------------
// Z pass: forward transform of three components, 3x3 multiply per spectral
// point, inverse transform; input and output positions < n (pruned).
// ---------------------------------------------------------------------------
template <int L1, int L2, int NP, typename C, bool PAIR = true, bool XCHG = false>
struct ZFast {
    static constexpr int L = L1 * L2;
    static constexpr int TA = 3 * NP * L2;
    static constexpr int LDS = 3 * NP * L;

    TDDA_HD static int at(int c, int k1, int n2, int b) { return ((c * L + k1 * L2 + n2) * NP) + b; }

    // forward, step A: load(c, b, pos)
    template <typename Load>
    TDDA_HD static void fwd_a(int tau, C* __restrict__ lds, const C* __restrict__ tw, int nb, int n, Load load) {
        const int b = tau % NP;
        const int rest = tau / NP;
        const int n2 = rest % L2;
        const int c = rest / L2;
        if (b >= nb) return;
        C w[L1];
        load_tw<L1, -1>(tw, n2, w);
        C v[L1];
        bool done = false;
        if constexpr (L1 % 2 == 0) {
            if (2 * n <= L) {
                static_for<L1 / 2>([&](auto i) {
                    constexpr int n1 = decltype(i)::value;
                    const int pos = n2 + L2 * n1;
                    v[n1] = pos < n ? load(c, b, pos) : mk<C>(0.0, 0.0);
                });
                dft_half_in<L1, -1>(v);
                done = true;
            }
        }
        if (!done) {
            static_for<L1>([&](auto i) {
                constexpr int n1 = decltype(i)::value;
                const int pos = n2 + L2 * n1;
                v[n1] = pos < n ? load(c, b, pos) : mk<C>(0.0, 0.0);
            });
            SDft<L1, -1, C>::run(v);
        }
        static_for<L1>([&](auto i) {
            constexpr int k1 = decltype(i)::value;
            if constexpr (k1 > 0) v[k1] = cmul(v[k1], w[k1]);
            lds[at(c, k1, n2, b)] = v[k1];
        });
    }

    // Step B: forward L2-point DFTs, 3x3 multiply per spectral point,
    // inverse L2-point DFTs, in place (a task reads and writes the same LDS
    // entries).  Tasks are (pencil, unit j): unit j pairs k1 = j with
    // k1 = L1 - j, whose points kz = j + L1 k2 and L - kz (at k2' = L2-1-k2)
    // are z-mirrors, so one octant record serves both (the z-odd components
    // 02 and 12 change sign); j = 0 and j = L1/2 stand alone.  Without PAIR
    // (fewer registers) every k1 is its own task.
    // fetch(b, kz, C (&g)[6]) gives the components (00,01,02,11,12,22) at
    // (pencil b, kz) with their parity signs.
    static constexpr int NU = PAIR ? L1 / 2 + 1 : L1;
    static constexpr int TB = NP * NU;

    // Lane exchange (XCHG, one task per k1): the tasks of a 64-wide wave are
    // ordered so that lane l + 32 holds the z-mirror task k1' = L1 - k1 of
    // lane l's pencil; each lane loads the octant records of its points
    // k2 < L2/2 and takes those of k2 >= L2/2 from its mirror lane, whose
    // own points k2' = L2-1-k2 are their z-mirrors (the records differ only
    // in the sign of the z-odd components 02 and 12).  k1 = 0 and k1 = L1/2
    // are their own mirrors (they share one pair of wave halves): they reuse
    // their own records, and k1 = 0 also loads the one at kz = L/2, issued
    // with the others.
    static constexpr bool kX = XCHG && !PAIR && L1 % 2 == 0 && L2 % 2 == 0 && 32 % NP == 0 && TB % 64 == 0;
    // k1 of task slot u = tau / NP under the exchange order
    TDDA_HD static int xchg_k1(int u) {
        constexpr int h = 32 / NP;                  // mirror pairs per wave
        const int gp = (u / (2 * h)) * h + u % h;   // pair index; the last is (0, L1/2)
        const bool hi = (u % (2 * h)) >= h;
        if (gp == L1 / 2 - 1) return hi ? L1 / 2 : 0;
        return hi ? L1 - (gp + 1) : gp + 1;
    }

    TDDA_HD static void sym_mul(const C (&g)[6], C& v0, C& v1, C& v2) {
        const C a0 = v0, a1 = v1, a2 = v2;
        v0 = cadd(cadd(cmul(g[0], a0), cmul(g[1], a1)), cmul(g[2], a2));
        v1 = cadd(cadd(cmul(g[1], a0), cmul(g[3], a1)), cmul(g[4], a2));
        v2 = cadd(cadd(cmul(g[2], a0), cmul(g[4], a1)), cmul(g[5], a2));
    }

    TDDA_HD static void load_fwd(const C* __restrict__ lds, int k1, int b, C (&v)[3][L2]) {
        static_for<3>([&](auto ci) {
            constexpr int c = decltype(ci)::value;
            static_for<L2>([&](auto i) {
                constexpr int n2 = decltype(i)::value;
                v[c][n2] = lds[at(c, k1, n2, b)];
            });
            SDft<L2, -1, C>::run(v[c]);
        });
    }

    // w: load_tw<L2, 1>(tw, k1, w), read before the task's first store
    TDDA_HD static void inv_store(C* __restrict__ lds, const C (&w)[L2], int k1, int b, C (&v)[3][L2]) {
        static_for<3>([&](auto ci) {
            constexpr int c = decltype(ci)::value;
            SDft<L2, 1, C>::run(v[c]);
            static_for<L2>([&](auto i) {
                constexpr int n2 = decltype(i)::value;
                C x = v[c][n2];
                if constexpr (n2 > 0) x = cmul(x, w[n2]);
                lds[at(c, k1, n2, b)] = x;
            });
        });
    }

    // exch(b, kz_m, mine, out): out = the record the mirror lane loaded for
    // its point kz_m = L - kz, given this lane's record `mine` for the same
    // instruction (called by every lane of the wave alike; only when kX)
    TDDA_HD static int uniform_wave(int tau) {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_readfirstlane(tau >> 6);
#else
        return tau >> 6;
#endif
    }

    template <typename Fetch, typename Exch>
    TDDA_HD static void mid(int tau, C* __restrict__ lds, const C* __restrict__ tw, int nb, Fetch fetch, Exch exch) {
        if constexpr (kX) {
            const int b = tau % NP;
            const int j = xchg_k1(tau / NP);
            if (b >= nb) return;
            const bool self = j == 0 || 2 * j == L1;
            C w[L2];
            load_tw<L2, 1>(tw, j, w);
            C own[L2 / 2][6], half[6];
            static_for<L2 / 2>([&](auto i) {
                constexpr int k2 = decltype(i)::value;
                fetch(b, j + L1 * k2, own[k2]);
            });
            // kz = L/2 is its own mirror: k1 = 0 loads that record too; the
            // self pair sits in the last wave, which loads it for all its lanes
            // (a wave-uniform branch)
            if (uniform_wave(tau) == TB / 64 - 1) fetch(b, L / 2, half);
            C v[3][L2];
            load_fwd(lds, j, b, v);
            static_for<L2>([&](auto i) {
                constexpr int k2 = decltype(i)::value;
                C g[6];
                if constexpr (k2 < L2 / 2) {
                    static_for<6>([&](auto q) { g[decltype(q)::value] = own[k2][decltype(q)::value]; });
                } else {
                    C ex[6];
                    exch(b, L - (j + L1 * k2), own[L2 - 1 - k2], ex);
                    // the z-mirror record: the mirror lane's, or for k1 = L1/2
                    // (kz' at k2' = L2-1-k2) and k1 = 0 (kz' at k2' = L2-k2)
                    // this lane's own
                    static_for<6>([&](auto q) {
                        constexpr int c = decltype(q)::value;
                        C m;
                        if constexpr (2 * k2 == L2)
                            m = self ? own[L2 - 1 - k2][c] : ex[c];
                        else
                            m = !self ? ex[c] : (j != 0 ? own[L2 - 1 - k2][c] : own[L2 - k2][c]);
                        if constexpr (c == 2 || c == 4) m = mk<C>(-m.x, -m.y);
                        if constexpr (2 * k2 == L2) g[c] = j == 0 ? half[c] : m;
                        else g[c] = m;
                    });
                }
                sym_mul(g, v[0][k2], v[1][k2], v[2][k2]);
            });
            inv_store(lds, w, j, b, v);
            return;
        }
        const int b = tau % NP;
        const int j = tau / NP;
        if (b >= nb) return;
        if (!PAIR || j == 0 || 2 * j == L1) {
            C w[L2];
            load_tw<L2, 1>(tw, j, w);
            C v[3][L2];
            load_fwd(lds, j, b, v);
            static_for<L2>([&](auto i) {
                constexpr int k2 = decltype(i)::value;
                C g[6];
                fetch(b, j + L1 * k2, g);
                sym_mul(g, v[0][k2], v[1][k2], v[2][k2]);
            });
            inv_store(lds, w, j, b, v);
            return;
        }
        C wa[L2], wb[L2];
        load_tw<L2, 1>(tw, j, wa);
        load_tw<L2, 1>(tw, L1 - j, wb);
        C va[3][L2], vb[3][L2];
        load_fwd(lds, j, b, va);
        load_fwd(lds, L1 - j, b, vb);
        static_for<L2>([&](auto i) {
            constexpr int k2 = decltype(i)::value;
            constexpr int m2 = L2 - 1 - k2;
            C g[6];
            fetch(b, j + L1 * k2, g);
            sym_mul(g, va[0][k2], va[1][k2], va[2][k2]);
            g[2] = mk<C>(-g[2].x, -g[2].y);
            g[4] = mk<C>(-g[4].x, -g[4].y);
            sym_mul(g, vb[0][m2], vb[1][m2], vb[2][m2]);
        });
        inv_store(lds, wa, j, b, va);
        inv_store(lds, wb, L1 - j, b, vb);
    }

    // inverse, step A: store(c, b, pos, value) for pos < n
    template <typename Store>
    TDDA_HD static void inv_a(int tau, const C* __restrict__ lds, int nb, int n, Store store) {
        const int b = tau % NP;
        const int rest = tau / NP;
        const int n2 = rest % L2;
        const int c = rest / L2;
        if (b >= nb) return;
        C v[L1];
        static_for<L1>([&](auto i) {
            constexpr int k1 = decltype(i)::value;
            v[k1] = lds[at(c, k1, n2, b)];
        });
        if constexpr (L1 % 2 == 0) {
            if (2 * n <= L) {
                dft_half_out<L1, 1>(v);
                static_for<L1 / 2>([&](auto i) {
                    constexpr int n1 = decltype(i)::value;
                    const int pos = n2 + L2 * n1;
                    if (pos < n) store(c, b, pos, v[n1]);
                });
                return;
            }
        }
        SDft<L1, 1, C>::run(v);
        static_for<L1>([&](auto i) {
            constexpr int n1 = decltype(i)::value;
            const int pos = n2 + L2 * n1;
            if (pos < n) store(c, b, pos, v[n1]);
        });
    }
};

// ---------------------------------------------------------------------------
// X pass: NR rows of three components (pencil b = 3*row + c) along x.
// ---------------------------------------------------------------------------
template <int L1, int L2, int NR, typename C>
struct XFast {
    static constexpr int L = L1 * L2;
    static constexpr int PS = fast_ps_x(L1, L2);
    static constexpr int TA = 3 * NR * L2;
    static constexpr int TB = 3 * NR * L1;
    static constexpr int LDS = 3 * NR * PS;

    TDDA_HD static int at(int b, int k1, int n2) { return b * PS + k1 * (L2 + 1) + n2; }

    // forward A: task (c fastest, n2, row); load(row, c, x) for x < n_in
    template <typename Load>
    TDDA_HD static void fwd_a(int tau, C* __restrict__ lds, const C* __restrict__ tw, int nr, int n_in, Load load) {
        const int c = tau % 3;
        const int rest = tau / 3;
        const int n2 = rest % L2;
        const int r = rest / L2;
        if (r >= nr) return;
        C w[L1];
        load_tw<L1, -1>(tw, n2, w);
        C v[L1];
        bool done = false;
        if constexpr (L1 % 2 == 0) {
            if (2 * n_in <= L) {
                static_for<L1 / 2>([&](auto i) {
                    constexpr int n1 = decltype(i)::value;
                    const int x = n2 + L2 * n1;
                    v[n1] = x < n_in ? load(r, c, x) : mk<C>(0.0, 0.0);
                });
                dft_half_in<L1, -1>(v);
                done = true;
            }
        }
        if (!done) {
            static_for<L1>([&](auto i) {
                constexpr int n1 = decltype(i)::value;
                const int x = n2 + L2 * n1;
                v[n1] = x < n_in ? load(r, c, x) : mk<C>(0.0, 0.0);
            });
            SDft<L1, -1, C>::run(v);
        }
        const int b = 3 * r + c;
        static_for<L1>([&](auto i) {
            constexpr int k1 = decltype(i)::value;
            if constexpr (k1 > 0) v[k1] = cmul(v[k1], w[k1]);
            lds[at(b, k1, n2)] = v[k1];
        });
    }

    // forward B: task (k1 fastest, pencil); store(row, c, x, value), all x < L
    template <typename Store>
    TDDA_HD static void fwd_b(int tau, const C* __restrict__ lds, int nr, Store store) {
        const int k1 = tau % L1;
        const int b = tau / L1;
        if (b >= 3 * nr) return;
        C v[L2];
        static_for<L2>([&](auto i) {
            constexpr int n2 = decltype(i)::value;
            v[n2] = lds[at(b, k1, n2)];
        });
        SDft<L2, -1, C>::run(v);
        const int r = b / 3;
        const int c = b - 3 * r;
        static_for<L2>([&](auto i) {
            constexpr int k2 = decltype(i)::value;
            store(r, c, k1 + L1 * k2, v[k2]);
        });
    }

    // inverse A: task (n2 fastest, pencil); load(row, c, x) for all x < L
    template <typename Load>
    TDDA_HD static void inv_a(int tau, C* __restrict__ lds, const C* __restrict__ tw, int nr, Load load) {
        const int n2 = tau % L2;
        const int b = tau / L2;
        if (b >= 3 * nr) return;
        const int r = b / 3;
        const int c = b - 3 * r;
        C w[L1];
        load_tw<L1, 1>(tw, n2, w);
        C v[L1];
        static_for<L1>([&](auto i) {
            constexpr int n1 = decltype(i)::value;
            v[n1] = load(r, c, n2 + L2 * n1);
        });
        SDft<L1, 1, C>::run(v);
        static_for<L1>([&](auto i) {
            constexpr int k1 = decltype(i)::value;
            if constexpr (k1 > 0) v[k1] = cmul(v[k1], w[k1]);
            lds[at(b, k1, n2)] = v[k1];
        });
    }

    // inverse B: task (c fastest, k1, row); epi(row, c, x, value) for x < n_out
    template <typename Epi>
    TDDA_HD static void inv_b(int tau, const C* __restrict__ lds, int nr, int n_out, Epi epi) {
        const int c = tau % 3;
        const int rest = tau / 3;
        const int k1 = rest % L1;
        const int r = rest / L1;
        if (r >= nr) return;
        const int b = 3 * r + c;
        C v[L2];
        static_for<L2>([&](auto i) {
            constexpr int n2 = decltype(i)::value;
            v[n2] = lds[at(b, k1, n2)];
        });
        if constexpr (L2 % 2 == 0) {
            if (2 * n_out <= L) {
                dft_half_out<L2, 1>(v);
                static_for<L2 / 2>([&](auto i) {
                    constexpr int k2 = decltype(i)::value;
                    const int x = k1 + L1 * k2;
                    if (x < n_out) epi(r, c, x, v[k2]);
                });
                return;
            }
        }
        SDft<L2, 1, C>::run(v);
        static_for<L2>([&](auto i) {
            constexpr int k2 = decltype(i)::value;
            const int x = k1 + L1 * k2;
            if (x < n_out) epi(r, c, x, v[k2]);
        });
    }
};

// Call f(IC<L1>, IC<L2>) for a length with a specialised split; false if none.
template <typename F>
TDDA_HD bool with_split(int L, F&& f) {
    switch (L) {
        case 16: f(IC<4>{}, IC<4>{}); return true;
        case 20: f(IC<5>{}, IC<4>{}); return true;
        case 24: f(IC<6>{}, IC<4>{}); return true;
        case 32: f(IC<8>{}, IC<4>{}); return true;
        case 40: f(IC<8>{}, IC<5>{}); return true;
        case 48: f(IC<8>{}, IC<6>{}); return true;
        case 64: f(IC<8>{}, IC<8>{}); return true;
        case 80: f(IC<10>{}, IC<8>{}); return true;
        case 96: f(IC<12>{}, IC<8>{}); return true;
        case 128: f(IC<16>{}, IC<8>{}); return true;
        case 256: f(IC<16>{}, IC<16>{}); return true;
        default: return false;
    }
}

// Splits of the X and Y passes: a tile holds about L1 values per thread in
// LDS, and every block of a launch reserves the largest tile of the launch,
// so L1 = 8 for the long lengths (more resident waves per CU).
template <typename F>
TDDA_HD bool with_split_xy(int L, F&& f) {
    switch (L) {
        case 80: f(IC<8>{}, IC<10>{}); return true;
        case 96: f(IC<8>{}, IC<12>{}); return true;
        case 128: f(IC<8>{}, IC<16>{}); return true;
        default: return with_split(L, f);
    }
}

// Splits of the Z pass: the multiply step holds 3 x L2 values per thread and
// the tile holds 3 x L2 entries per thread in LDS, so L2 = 4 wherever L1 stays
// within a 16-point register DFT (more resident waves per CU).
template <typename F>
TDDA_HD bool with_split_z(int L, F&& f) {
    switch (L) {
        case 40: f(IC<10>{}, IC<4>{}); return true;
        case 48: f(IC<12>{}, IC<4>{}); return true;
        case 64: f(IC<16>{}, IC<4>{}); return true;
        // 256 keeps the generic Z pass under either variant: a 16 x 16 tile
        // of three components needs about 100 KB of LDS per block, and its
        // registers would set those of every other split of its kernel class
        case 256: return false;
        default: return with_split(L, f);
    }
}

// Schedule variants.  Variant 0 of a pass uses its table above, variant 1 the
// other table (X forward: with_split_xy / with_split; X inverse and Y:
// with_split / with_split_xy; Z: with_split_z / with_split).  A length has a
// variant 1 only where both tables hold a split and the splits differ, so a
// length is specialised under either variant.  The autotuner times both and
// records the faster per length, pass and precision.
// PASS: 0 X inverse, 1 Y, 2 Z, 3 X forward (the PassKind values).
template <int PASS, typename F>
TDDA_HD bool with_split_pass(int L, int v, F&& f) {
    if constexpr (PASS == 3) return v ? with_split(L, f) : with_split_xy(L, f);
    else if constexpr (PASS == 2) return v ? (L != 256 && with_split(L, f)) : with_split_z(L, f);
    else return v ? with_split_xy(L, f) : with_split(L, f);
}

// 2 when the pass has a variant 1 at length L, else 1
inline int split_variants(int pass, int L) {
    int a1 = 0, a2 = 0, b1 = 0, b2 = 0;
    auto ga = [&](auto s1, auto s2) { a1 = decltype(s1)::value; a2 = decltype(s2)::value; };
    auto gb = [&](auto s1, auto s2) { b1 = decltype(s1)::value; b2 = decltype(s2)::value; };
    bool ha = false, hb = false;
    switch (pass) {
        case 3: ha = with_split_xy(L, ga); hb = with_split(L, gb); break;
        case 2: ha = with_split_z(L, ga); hb = L != 256 && with_split(L, gb); break;
        default: ha = with_split(L, ga); hb = with_split_xy(L, gb); break;
    }
    return (ha && hb && (a1 != b1 || a2 != b2)) ? 2 : 1;
}

// Z lengths fall in two kernel classes: 0 for the splits with L2 = 4, 1 for
// the others (their multiply step holds 3 x L2 >= 15 values per thread).
// Each class is its own kernel, so the register budget of the short splits
// is not set by the long ones.
template <typename S2>
constexpr int z_class_of() { return S2::value == 4 ? 0 : 1; }
// -1: no specialised split (v: the schedule variant of the Z pass)
inline int z_class(int L, int v = 0) {
    int c = -1;
    with_split_pass<2>(L, v, [&](auto, auto s2) { c = z_class_of<decltype(s2)>(); });
    return c;
}

}  // namespace tdda

// Plane transforms of the c64 matvec: the x and y FFTs of one z-plane of
// one component, with the plane resident in LDS between them.
//
// The forward pair (x < Nx, y < Ny inputs -> all kx, ky outputs) replaces
// the separate X forward and Y forward passes, and the inverse pair (all
// kx, ky -> x < Nx, y < Ny) the Y inverse and X inverse passes: the
// intermediate Lx x Ny grid never goes through HBM.  Lx = Ly = L with
// Nx, Ny <= L/2 (L even), so the zero half of each padded axis is pruned:
// half-input DFTs forward, half-output DFTs inverse.
//
// Both transforms use the two-step split of fft_fast.h (L = L1 L2, input
// n2 + L2 n1, twiddle w_L^(n2 k1) between the steps, output k1 + L1 k2).
// The x split is A1 x A2, the y split B1 x B2.  LDS is one region of L x L
// complex entries, reused by the phases of a component:
//   E  (x exchange)  [y][k1 (A2+1) + n2], row stride SE
//   P  (the plane)   [y][kx], row stride SP
//   E2 (y exchange)  [k1' B2 + n2'][kx]
// A phase whose output overlaps its input reads everything first (the
// *_read functions, into registers), then the block synchronises and the
// *_write functions store.  Strides are padded so that the lanes of a wave
// (task index fastest as listed) hit distinct LDS banks.
//
// Every function is one task of one phase; kernels call them in
// block-strided loops between barriers, and a CPU harness calls them in
// plain loops.
#pragma once

#include "fft_fast.h"

namespace tdda {

// smallest s >= base with s = want (mod 32) (8-byte LDS units)
TDDA_HD constexpr int pad_to(int base, int want) { return base + (((want - base) % 32) + 32) % 32; }

// Forward: A1 x A2 (x), B1 x B2 (y).
template <int L_, int A1, int A2, int B1, int B2, typename C>
struct PlaneFwd {
    static constexpr int L = L_;
    static_assert(A1 * A2 == L && B1 * B2 == L, "splits must multiply to L");
    static_assert(A1 % 2 == 0 && B1 % 2 == 0, "half-input steps need even first factors");
    static constexpr int NYM = L / 2;                        // rows y < Ny <= L/2
    static constexpr int SE = pad_to(A1 * (A2 + 1), A2);     // XA lanes: n2 fastest, then y
    static constexpr int SP = pad_to(L, A1);                 // XB lanes: k1 fastest, then y
    static constexpr int REGION = (L * L > NYM * SE ? (L * L > NYM * SP ? L * L : NYM * SP)
                                                    : (NYM * SE > NYM * SP ? NYM * SE : NYM * SP));
    // task counts
    static constexpr int TXA = NYM * A2;   // (y, n2)
    static constexpr int TXB = NYM * A1;   // (y, k1)
    static constexpr int TYA = L * B2;     // (kx, n2')
    static constexpr int TYB = L * B1;     // (kx, k1')
    static constexpr int NIN = A1 / 2;     // x inputs per XA task

    TDDA_HD static int e_at(int y, int k1, int n2) { return y * SE + k1 * (A2 + 1) + n2; }
    TDDA_HD static int p_at(int y, int kx) { return y * SP + kx; }
    TDDA_HD static int e2_at(int k1, int n2, int kx) { return (k1 * B2 + n2) * L + kx; }

    // XA task (y, n2): its inputs are x = n2 + A2 n1, n1 < NIN
    TDDA_HD static int xa_y(int tau) { return tau / A2; }
    TDDA_HD static int xa_x(int tau, int n1) { return tau % A2 + A2 * n1; }
    // u[n1]: the inputs (zero at x >= Nx); half-input A1-point DFT, twiddle, E
    TDDA_HD static void xa_store(int tau, C* __restrict__ lds, const C* __restrict__ tw, const C (&u)[NIN]) {
        const int n2 = tau % A2, y = tau / A2;
        C w[A1];
        load_tw<A1, -1>(tw, n2, w);
        C v[A1];
        static_for<NIN>([&](auto i) { v[decltype(i)::value] = u[decltype(i)::value]; });
        dft_half_in<A1, -1>(v);
        static_for<A1>([&](auto i) {
            constexpr int k1 = decltype(i)::value;
            C x = v[k1];
            if constexpr (k1 > 0) x = cmul(x, w[k1]);
            lds[e_at(y, k1, n2)] = x;
        });
    }

    // XB task (y, k1): A2-point DFT over n2 -> kx = k1 + A1 k2 of row y
    TDDA_HD static int xb_y(int tau) { return tau / A1; }
    TDDA_HD static void xb_read(int tau, const C* __restrict__ lds, C (&v)[A2]) {
        const int k1 = tau % A1, y = tau / A1;
        static_for<A2>([&](auto i) { v[decltype(i)::value] = lds[e_at(y, k1, decltype(i)::value)]; });
    }
    TDDA_HD static void xb_write(int tau, C* __restrict__ lds, C (&v)[A2]) {
        const int k1 = tau % A1, y = tau / A1;
        SDft<A2, -1, C>::run(v);
        static_for<A2>([&](auto i) {
            constexpr int k2 = decltype(i)::value;
            lds[p_at(y, k1 + A1 * k2)] = v[k2];
        });
    }

    // YA task (kx, n2'): inputs y = n2' + B2 n1' < Ny (n1' < B1/2)
    TDDA_HD static void ya_read(int tau, const C* __restrict__ lds, int ny, C (&v)[B1 / 2]) {
        const int kx = tau % L, n2 = tau / L;
        static_for<B1 / 2>([&](auto i) {
            constexpr int n1 = decltype(i)::value;
            const int y = n2 + B2 * n1;
            v[n1] = y < ny ? lds[p_at(y, kx)] : mk<C>(0.0, 0.0);
        });
    }
    TDDA_HD static void ya_write(int tau, C* __restrict__ lds, const C* __restrict__ tw, const C (&u)[B1 / 2]) {
        const int kx = tau % L, n2 = tau / L;
        C w[B1];
        load_tw<B1, -1>(tw, n2, w);
        C v[B1];
        static_for<B1 / 2>([&](auto i) { v[decltype(i)::value] = u[decltype(i)::value]; });
        dft_half_in<B1, -1>(v);
        static_for<B1>([&](auto i) {
            constexpr int k1 = decltype(i)::value;
            C x = v[k1];
            if constexpr (k1 > 0) x = cmul(x, w[k1]);
            lds[e2_at(k1, n2, kx)] = x;
        });
    }

    // YB task (kx, k1'): B2-point DFT -> ky = k1' + B1 k2'; store(ky, kx, value)
    template <typename Store>
    TDDA_HD static void yb(int tau, const C* __restrict__ lds, Store store) {
        const int kx = tau % L, k1 = tau / L;
        C v[B2];
        static_for<B2>([&](auto i) { v[decltype(i)::value] = lds[e2_at(k1, decltype(i)::value, kx)]; });
        SDft<B2, -1, C>::run(v);
        static_for<B2>([&](auto i) {
            constexpr int k2 = decltype(i)::value;
            store(k1 + B1 * k2, kx, v[k2]);
        });
    }
};

// Inverse: B1 x B2 (y), then A1 x A2 (x); outputs y < Ny, x < Nx.
template <int L_, int A1, int A2, int B1, int B2, typename C>
struct PlaneInv {
    static constexpr int L = L_;
    static_assert(A1 * A2 == L && B1 * B2 == L, "splits must multiply to L");
    static_assert(A2 % 2 == 0 && B2 % 2 == 0, "half-output steps need even second factors");
    static constexpr int NYM = L / 2;
    static constexpr int SE = pad_to(A1 * (A2 + 1), A2);     // XIA lanes: n2 fastest, then y
    static constexpr int SP = pad_to(L, A2);                 // XIA reads: n2 fastest, then y
    static constexpr int REGION = (L * L > NYM * SE ? (L * L > NYM * SP ? L * L : NYM * SP)
                                                    : (NYM * SE > NYM * SP ? NYM * SE : NYM * SP));
    static constexpr int TYA = L * B2;     // (kx, n2')
    static constexpr int TYB = L * B1;     // (kx, k1')
    static constexpr int TXA = NYM * A2;   // (y, n2)
    static constexpr int TXB = NYM * A1;   // (y, k1)
    static constexpr int NOUT = A2 / 2;    // x outputs per XB task

    TDDA_HD static int e_at(int y, int k1, int n2) { return y * SE + k1 * (A2 + 1) + n2; }
    TDDA_HD static int p_at(int y, int kx) { return y * SP + kx; }
    TDDA_HD static int e2_at(int k1, int n2, int kx) { return (k1 * B2 + n2) * L + kx; }

    // YA task (kx, n2'): load(ky, kx) for ky = n2' + B2 n1', all n1' < B1
    template <typename Load>
    TDDA_HD static void ya(int tau, C* __restrict__ lds, const C* __restrict__ tw, Load load) {
        const int kx = tau % L, n2 = tau / L;
        C w[B1];
        load_tw<B1, 1>(tw, n2, w);
        C v[B1];
        static_for<B1>([&](auto i) {
            constexpr int n1 = decltype(i)::value;
            v[n1] = load(n2 + B2 * n1, kx);
        });
        SDft<B1, 1, C>::run(v);
        static_for<B1>([&](auto i) {
            constexpr int k1 = decltype(i)::value;
            C x = v[k1];
            if constexpr (k1 > 0) x = cmul(x, w[k1]);
            lds[e2_at(k1, n2, kx)] = x;
        });
    }

    // YB task (kx, k1'): half-output B2-point DFT -> y = k1' + B1 k2' < Ny
    TDDA_HD static void yb_read(int tau, const C* __restrict__ lds, C (&v)[B2]) {
        const int kx = tau % L, k1 = tau / L;
        static_for<B2>([&](auto i) { v[decltype(i)::value] = lds[e2_at(k1, decltype(i)::value, kx)]; });
    }
    TDDA_HD static void yb_write(int tau, C* __restrict__ lds, int ny, C (&v)[B2]) {
        const int kx = tau % L, k1 = tau / L;
        dft_half_out<B2, 1>(v);
        static_for<B2 / 2>([&](auto i) {
            constexpr int k2 = decltype(i)::value;
            const int y = k1 + B1 * k2;
            if (y < ny) lds[p_at(y, kx)] = v[k2];
        });
    }

    // XA task (y, n2): A1-point DFT over kx = n2 + A2 n1 of row y
    TDDA_HD static int xa_y(int tau) { return tau / A2; }
    TDDA_HD static void xa_read(int tau, const C* __restrict__ lds, C (&v)[A1]) {
        const int n2 = tau % A2, y = tau / A2;
        static_for<A1>([&](auto i) {
            constexpr int n1 = decltype(i)::value;
            v[n1] = lds[p_at(y, n2 + A2 * n1)];
        });
    }
    TDDA_HD static void xa_write(int tau, C* __restrict__ lds, const C* __restrict__ tw, C (&v)[A1]) {
        const int n2 = tau % A2, y = tau / A2;
        C w[A1];
        load_tw<A1, 1>(tw, n2, w);
        SDft<A1, 1, C>::run(v);
        static_for<A1>([&](auto i) {
            constexpr int k1 = decltype(i)::value;
            C x = v[k1];
            if constexpr (k1 > 0) x = cmul(x, w[k1]);
            lds[e_at(y, k1, n2)] = x;
        });
    }

    // XB task (y, k1): half-output A2-point DFT; out[k2] is x = k1 + A1 k2
    TDDA_HD static int xb_y(int tau) { return tau / A1; }
    TDDA_HD static int xb_x(int tau, int k2) { return tau % A1 + A1 * k2; }
    TDDA_HD static void xb(int tau, const C* __restrict__ lds, C (&out)[NOUT]) {
        const int k1 = tau % A1, y = tau / A1;
        C v[A2];
        static_for<A2>([&](auto i) { v[decltype(i)::value] = lds[e_at(y, k1, decltype(i)::value)]; });
        dft_half_out<A2, 1>(v);
        static_for<NOUT>([&](auto i) { out[decltype(i)::value] = v[decltype(i)::value]; });
    }
};

// Classes of the plane passes: L, x split forward (A1 x A2), x split
// inverse, y split (B1 x B2, both directions), threads per block.  Each
// direction has exactly one task per thread in the phase that touches the
// Krylov vectors (XA forward, XB inverse): NT = (L/2) A2 forward and
// (L/2) A1 inverse.
constexpr int NXY = 5;
template <typename F>
TDDA_HD bool with_plane(int cls, F&& f) {
    switch (cls) {
        case 0: f(IC<32>{}, IC<4>{}, IC<8>{}, IC<8>{}, IC<4>{}, IC<8>{}, IC<4>{}); return true;
        case 1: f(IC<48>{}, IC<6>{}, IC<8>{}, IC<8>{}, IC<6>{}, IC<8>{}, IC<6>{}); return true;
        case 2: f(IC<64>{}, IC<8>{}, IC<8>{}, IC<8>{}, IC<8>{}, IC<8>{}, IC<8>{}); return true;
        case 3: f(IC<96>{}, IC<8>{}, IC<12>{}, IC<12>{}, IC<8>{}, IC<12>{}, IC<8>{}); return true;
        case 4: f(IC<128>{}, IC<8>{}, IC<16>{}, IC<16>{}, IC<8>{}, IC<16>{}, IC<8>{}); return true;
        default: return false;
    }
}
// class of a box (-1: the separate X and Y passes)
TDDA_HD int plane_class(int lx, int ly, int nx, int ny) {
    if (lx != ly || 2 * nx > lx || 2 * ny > ly) return -1;
    switch (lx) {
        case 32: return 0;
        case 48: return 1;
        case 64: return 2;
        case 96: return 3;
        case 128: return 4;
        default: return -1;
    }
}

}  // namespace tdda

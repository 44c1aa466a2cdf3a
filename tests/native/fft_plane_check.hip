// CPU check of the plane transforms (fft_plane.h) against naive long-double
// 2-D DFTs: every class, both directions, both precisions, pruned inputs and
// outputs.  The phases run in the kernels' order; a phase whose output
// overlaps its input reads all its tasks first, as the kernels do across a
// barrier.  LDS is an exact-size vector filled with NaN, so a read of an
// entry no earlier phase wrote shows up as a NaN result, and (with
// AddressSanitizer on the host) any index outside the region fails.
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>
#include "fft_plane.h"
using namespace tdda;
typedef std::complex<long double> cld;
static std::mt19937_64 rng(7);
static double urand() { return std::uniform_real_distribution<double>(-1, 1)(rng); }
static double worst = 0;
static int fails = 0;

template <typename C> std::vector<C> table(int L) {
    std::vector<C> t(L);
    for (int m = 0; m < L; ++m) {
        long double a = -2.0L * M_PIl * m / L;
        t[m] = mk<C>((double)cosl(a), (double)sinl(a));
    }
    return t;
}
static cld ex(int dir, long num, int L) {
    long double a = dir * 2.0L * M_PIl * (long double)(num % L) / L;
    return cld(cosl(a), sinl(a));
}
static void rep(const char* what, int L, int nx, int ny, double err, double tol) {
    if (!(err == err) || err > tol) {
        printf("FAIL %s L=%d nx=%d ny=%d err=%.3e\n", what, L, nx, ny, err);
        ++fails;
    }
    if (err > worst || !(err == err)) worst = err == err ? err : 1.0;
}

template <typename PF, typename C> void check_fwd(int nx, int ny, double tol) {
    constexpr int L = PF::L;
    const C nan = mk<C>(NAN, NAN);
    std::vector<C> lds(PF::REGION, nan);
    auto tw = table<C>(L);
    std::vector<cld> u(L * L, cld(0));
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) u[y * L + x] = cld(urand(), urand());
    auto cv = [](cld z) { return mk<C>((double)z.real(), (double)z.imag()); };
    for (int t = 0; t < PF::TXA; ++t) {
        if (PF::xa_y(t) >= ny) continue;
        C in[PF::NIN];
        for (int n1 = 0; n1 < PF::NIN; ++n1) {
            const int x = PF::xa_x(t, n1);
            in[n1] = x < nx ? cv(u[PF::xa_y(t) * L + x]) : mk<C>(0.0, 0.0);
        }
        PF::xa_store(t, lds.data(), tw.data(), in);
    }
    constexpr int A2 = L / (2 * PF::NIN);
    std::vector<std::array<C, A2>> sb(PF::TXB);
    for (int t = 0; t < PF::TXB; ++t)
        if (PF::xb_y(t) < ny) { C v[A2]; PF::xb_read(t, lds.data(), v); for (int i = 0; i < A2; ++i) sb[t][i] = v[i]; }
    for (int t = 0; t < PF::TXB; ++t)
        if (PF::xb_y(t) < ny) { C v[A2]; for (int i = 0; i < A2; ++i) v[i] = sb[t][i]; PF::xb_write(t, lds.data(), v); }
    constexpr int HB1 = (PF::TYA / L) == 0 ? 1 : L / (PF::TYA / L) / 2;   // B1 / 2
    std::vector<std::array<C, HB1>> sa(PF::TYA);
    for (int t = 0; t < PF::TYA; ++t) { C v[HB1]; PF::ya_read(t, lds.data(), ny, v); for (int i = 0; i < HB1; ++i) sa[t][i] = v[i]; }
    for (int t = 0; t < PF::TYA; ++t) { C v[HB1]; for (int i = 0; i < HB1; ++i) v[i] = sa[t][i]; PF::ya_write(t, lds.data(), tw.data(), v); }
    std::vector<C> out(L * L, nan);
    std::vector<int> cnt(L * L, 0);
    for (int t = 0; t < PF::TYB; ++t)
        PF::yb(t, lds.data(), [&](int ky, int kx, C v) {
            if (ky < 0 || ky >= L || kx < 0 || kx >= L) { printf("FAIL fwd store index\n"); ++fails; return; }
            out[ky * L + kx] = v;
            cnt[ky * L + kx]++;
        });
    // reference: direct DFT along x of every row, then along y of every column
    std::vector<cld> rowt(L * L, cld(0)), ref(L * L, cld(0));
    for (int y = 0; y < ny; ++y)
        for (int kx = 0; kx < L; ++kx) {
            cld s = 0;
            for (int x = 0; x < nx; ++x) s += u[y * L + x] * ex(-1, (long)kx * x, L);
            rowt[y * L + kx] = s;
        }
    for (int ky = 0; ky < L; ++ky)
        for (int kx = 0; kx < L; ++kx) {
            cld s = 0;
            for (int y = 0; y < ny; ++y) s += rowt[y * L + kx] * ex(-1, (long)ky * y, L);
            ref[ky * L + kx] = s;
        }
    double e = 0, nrm = 0;
    for (int ky = 0; ky < L; ++ky)
        for (int kx = 0; kx < L; ++kx) {
            if (cnt[ky * L + kx] != 1) { printf("FAIL fwd (%d,%d) written %d times\n", ky, kx, cnt[ky * L + kx]); ++fails; }
            const cld s = ref[ky * L + kx];
            const cld g(out[ky * L + kx].x, out[ky * L + kx].y);
            e = std::max(e, (double)std::abs(g - s));
            nrm = std::max(nrm, (double)std::abs(s));
        }
    rep(sizeof(C) == 8 ? "fwd c64" : "fwd c128", L, nx, ny, e / nrm, tol);
}

template <typename PI, typename C> void check_inv(int nx, int ny, double tol) {
    constexpr int L = PI::L;
    const C nan = mk<C>(NAN, NAN);
    std::vector<C> lds(PI::REGION, nan);
    auto tw = table<C>(L);
    std::vector<cld> G(L * L);
    for (auto& g : G) g = cld(urand(), urand());
    auto cv = [](cld z) { return mk<C>((double)z.real(), (double)z.imag()); };
    for (int t = 0; t < PI::TYA; ++t) PI::ya(t, lds.data(), tw.data(), [&](int ky, int kx) { return cv(G[ky * L + kx]); });
    constexpr int B2 = PI::TYA / L;
    std::vector<std::array<C, B2>> sb(PI::TYB);
    for (int t = 0; t < PI::TYB; ++t) { C v[B2]; PI::yb_read(t, lds.data(), v); for (int i = 0; i < B2; ++i) sb[t][i] = v[i]; }
    for (int t = 0; t < PI::TYB; ++t) { C v[B2]; for (int i = 0; i < B2; ++i) v[i] = sb[t][i]; PI::yb_write(t, lds.data(), ny, v); }
    constexpr int A1 = PI::TXB / PI::NYM;
    std::vector<std::array<C, A1>> sa(PI::TXA);
    for (int t = 0; t < PI::TXA; ++t)
        if (PI::xa_y(t) < ny) { C v[A1]; PI::xa_read(t, lds.data(), v); for (int i = 0; i < A1; ++i) sa[t][i] = v[i]; }
    for (int t = 0; t < PI::TXA; ++t)
        if (PI::xa_y(t) < ny) { C v[A1]; for (int i = 0; i < A1; ++i) v[i] = sa[t][i]; PI::xa_write(t, lds.data(), tw.data(), v); }
    std::vector<C> res(L * L, nan);
    std::vector<int> cnt(L * L, 0);
    for (int t = 0; t < PI::TXB; ++t) {
        const int y = PI::xb_y(t);
        if (y >= ny) continue;
        C o[PI::NOUT];
        PI::xb(t, lds.data(), o);
        for (int k2 = 0; k2 < PI::NOUT; ++k2) {
            const int x = PI::xb_x(t, k2);
            if (x < nx) { res[y * L + x] = o[k2]; cnt[y * L + x]++; }
        }
    }
    // reference: direct inverse DFT along y of every column (rows y < Ny), then along x
    std::vector<cld> colt(L * L, cld(0));
    for (int y = 0; y < ny; ++y)
        for (int kx = 0; kx < L; ++kx) {
            cld s = 0;
            for (int ky = 0; ky < L; ++ky) s += G[ky * L + kx] * ex(1, (long)ky * y, L);
            colt[y * L + kx] = s;
        }
    double e = 0, nrm = 0;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            if (cnt[y * L + x] != 1) { printf("FAIL inv (%d,%d) written %d times\n", y, x, cnt[y * L + x]); ++fails; }
            cld s = 0;
            for (int kx = 0; kx < L; ++kx) s += colt[y * L + kx] * ex(1, (long)kx * x, L);
            const cld g(res[y * L + x].x, res[y * L + x].y);
            e = std::max(e, (double)std::abs(g - s));
            nrm = std::max(nrm, (double)std::abs(s));
        }
    rep(sizeof(C) == 8 ? "inv c64" : "inv c128", L, nx, ny, e / nrm, tol);
}

template <int L, int FA1, int FA2, int IA1, int IA2, int B1, int B2> void check_class() {
    using F32 = PlaneFwd<L, FA1, FA2, B1, B2, float2>;
    using F64 = PlaneFwd<L, FA1, FA2, B1, B2, double2>;
    using I32 = PlaneInv<L, IA1, IA2, B1, B2, float2>;
    using I64 = PlaneInv<L, IA1, IA2, B1, B2, double2>;
    static_assert(F32::TXA == I32::TXB, "forward and inverse use the same threads per block");
    const int sizes[][2] = {{L / 2, L / 2}, {L / 2 - 1, L / 2 - 3}, {3, 1}, {1, L / 2}};
    for (auto& s : sizes) {
        check_fwd<F64, double2>(s[0], s[1], 1e-13);
        check_fwd<F32, float2>(s[0], s[1], 3e-6);
        check_inv<I64, double2>(s[0], s[1], 1e-13);
        check_inv<I32, float2>(s[0], s[1], 3e-6);
    }
    printf("class L=%d: threads %d, LDS %d + %d entries (fwd), %d + %d (inv)\n", L, F32::TXA, F32::REGION, L,
           I32::REGION, L);
}

int main() {
    for (int cls = 0; cls < NXY; ++cls)
        with_plane(cls, [&](auto l, auto fa1, auto fa2, auto ia1, auto ia2, auto b1, auto b2) {
            check_class<decltype(l)::value, decltype(fa1)::value, decltype(fa2)::value, decltype(ia1)::value,
                        decltype(ia2)::value, decltype(b1)::value, decltype(b2)::value>();
        });
    // class selection
    if (plane_class(128, 128, 64, 64) != 4 || plane_class(128, 128, 65, 64) != -1 || plane_class(96, 64, 48, 32) != -1 ||
        plane_class(64, 64, 31, 32) != 2 || plane_class(40, 40, 20, 20) != -1) {
        printf("FAIL plane_class\n");
        ++fails;
    }
    printf("fails %d worst %.3e\n", fails, worst);
    return fails ? 1 : 0;
}

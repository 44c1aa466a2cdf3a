// CPU check of the length-specialised pencil passes (fft_fast.h) against
// naive long-double DFTs: every split, both precisions, pruned inputs and
// outputs, partial tiles, and no writes outside the requested range.
#include <cstdio>
#include <cmath>
#include <vector>
#include <random>
#include <complex>
#include "fft_fast.h"
using namespace tdda;
typedef std::complex<long double> cld;
static std::mt19937_64 rng(1);
static double urand() { return std::uniform_real_distribution<double>(-1, 1)(rng); }

template <typename C> std::vector<C> table(int L) {
    std::vector<C> t(L);
    for (int m = 0; m < L; ++m) { long double a = -2.0L * M_PIl * m / L; t[m] = mk<C>((double)cosl(a), (double)sinl(a)); }
    return t;
}
static std::vector<cld> dft(const std::vector<cld>& x, int dir) {
    int L = x.size(); std::vector<cld> X(L);
    for (int k = 0; k < L; ++k) { cld s = 0; for (int n = 0; n < L; ++n) { long double a = dir * 2.0L * M_PIl * (long double)((long)n * k % L) / L; s += x[n] * cld(cosl(a), sinl(a)); } X[k] = s; }
    return X;
}
double worst = 0;
template <typename C> void rep(const char* what, int L, double err, double tol) {
    if (err > worst) worst = err;
    if (err > tol || !(err == err)) printf("FAIL %s L=%d err=%.3e\n", what, L, err);
}

template <int L1, int L2, typename Real> void check_y(int n_in, int n_out, int dir) {
    using C = cx_t<Real>;
    constexpr int L = L1 * L2;
    constexpr int NP = fast_np_y<Real>(L1, L2);
    auto tw = table<C>(L);
    int nb = NP - 1;   // partial tile
    std::vector<cld> X(NP * L);
    std::vector<C> g(NP * L), out(NP * L, mk<C>(7, 7));
    for (auto& v : X) v = cld(urand(), urand());
    for (int b = 0; b < NP; ++b) for (int p = 0; p < L; ++p) g[b * L + p] = mk<C>((double)X[b * L + p].real(), (double)X[b * L + p].imag());
    std::vector<C> lds(YFast<L1, L2, NP, -1, C>::LDS);
    auto run = [&](auto Y) {
        using YY = decltype(Y);
        for (int t = 0; t < YY::TA; ++t) YY::step_a(t, lds.data(), tw.data(), nb, n_in, [&](int b, int pos) { return g[b * L + pos]; });
        for (int t = 0; t < YY::TB; ++t) YY::step_b(t, lds.data(), nb, n_out, [&](int b, int pos, C v) { out[b * L + pos] = v; });
    };
    if (dir < 0) run(YFast<L1, L2, NP, -1, C>{}); else run(YFast<L1, L2, NP, 1, C>{});
    double e = 0, nrm = 0;
    for (int b = 0; b < nb; ++b) {
        std::vector<cld> x(L);
        for (int p = 0; p < L; ++p) x[p] = p < n_in ? X[b * L + p] : cld(0);
        auto R = dft(x, dir);
        for (int p = 0; p < n_out; ++p) { cld d = cld(out[b*L+p].x, out[b*L+p].y) - R[p]; e = std::max(e, (double)std::abs(d)); nrm = std::max(nrm, (double)std::abs(R[p])); }
        for (int p = n_out; p < L; ++p) if (out[b*L+p].x != 7) { printf("Y wrote beyond n_out\n"); worst = 1; }
    }
    if (out[(NP-1)*L].x != 7) { printf("Y wrote partial pencil\n"); worst = 1; }
    rep<C>("Y", L, e / nrm, sizeof(Real) == 4 ? 3e-6 : 1e-14);
}

template <int L1, int L2, typename Real, bool PAIR, bool XCHG = false> void check_z_impl(int n) {
    using C = cx_t<Real>;
    constexpr int L = L1 * L2;
    constexpr int NP = fast_np_z<Real>(L1, L2);
    using Z = ZFast<L1, L2, NP, C, PAIR, XCHG>;
    if constexpr (Z::kX) {
        // lane l + 32 of every wave holds the same pencil and the z-mirror k1
        std::vector<int> seen(L1, 0);
        for (int tau = 0; tau < Z::TB; ++tau) {
            const int b = tau % NP, k1 = Z::xchg_k1(tau / NP);
            if (b == 0) seen[k1]++;
            const int pt = tau ^ 32, pk1 = Z::xchg_k1(pt / NP);
            const bool self = k1 == 0 || 2 * k1 == L1;
            const bool ok = pt % NP == b && (self ? (pk1 == 0 || 2 * pk1 == L1) && pk1 != k1 : pk1 == L1 - k1);
            if (!ok) { printf("FAIL Z exchange order L=%d tau=%d\n", L, tau); worst = 1; }
        }
        for (int k = 0; k < L1; ++k)
            if (seen[k] != 1) { printf("FAIL Z exchange k1 cover L=%d k1=%d\n", L, k); worst = 1; }
    }
    auto tw = table<C>(L);
    int nb = NP - 3 > 0 ? NP - 3 : NP;
    std::vector<cld> X(3 * NP * L);
    for (auto& v : X) v = cld(urand(), urand());
    // test spectrum with the z parity of the Green's tensor: components 02
    // and 12 are odd in z, the others even
    std::vector<cld> G(6 * NP * L);
    for (int q = 0; q < 6; ++q)
        for (int b = 0; b < NP; ++b)
            for (int k = 0; k < L; ++k) {
                const int m = (L - k) % L;
                G[(q * NP + b) * L + k] = k <= m ? cld(urand(), urand())
                                                 : (q == 2 || q == 4 ? -1.0L : 1.0L) * G[(q * NP + b) * L + m];
            }
    std::vector<C> g(3 * NP * L), out(3 * NP * L, mk<C>(7, 7));
    for (size_t i = 0; i < g.size(); ++i) g[i] = mk<C>((double)X[i].real(), (double)X[i].imag());
    std::vector<C> lds(Z::LDS);
    auto idx = [&](int c, int b, int pos) { return (c * NP + b) * L + pos; };
    for (int t = 0; t < Z::TA; ++t) Z::fwd_a(t, lds.data(), tw.data(), nb, n, [&](int c, int b, int pos) { return g[idx(c, b, pos)]; });
    auto fetch = [&](int b, int kz, C (&gg)[6]) {
        for (int k = 0; k < 6; ++k) { cld z = G[(k * NP + b) * L + kz]; gg[k] = mk<C>((double)z.real(), (double)z.imag()); }
    };
    // the mirror lane's record for its point kz_m (what the wave exchange delivers)
    auto exch = [&](int b, int kz_m, const C (&)[6], C (&out)[6]) { fetch(b, kz_m, out); };
    for (int t = 0; t < Z::TB; ++t) Z::mid(t, lds.data(), tw.data(), nb, fetch, exch);
    for (int t = 0; t < Z::TA; ++t) Z::inv_a(t, lds.data(), nb, n, [&](int c, int b, int pos, C v) { out[idx(c, b, pos)] = v; });
    double e = 0, nrm = 0;
    for (int b = 0; b < nb; ++b) {
        std::vector<cld> F[3];
        for (int c = 0; c < 3; ++c) { std::vector<cld> x(L); for (int p = 0; p < L; ++p) x[p] = p < n ? X[idx(c, b, p)] : cld(0); F[c] = dft(x, -1); }
        std::vector<cld> Y[3]; for (int c = 0; c < 3; ++c) Y[c].resize(L);
        for (int k = 0; k < L; ++k) {
            cld gk[6]; for (int q = 0; q < 6; ++q) gk[q] = G[(q * NP + b) * L + k];
            Y[0][k] = gk[0]*F[0][k] + gk[1]*F[1][k] + gk[2]*F[2][k];
            Y[1][k] = gk[1]*F[0][k] + gk[3]*F[1][k] + gk[4]*F[2][k];
            Y[2][k] = gk[2]*F[0][k] + gk[4]*F[1][k] + gk[5]*F[2][k];
        }
        for (int c = 0; c < 3; ++c) {
            auto y = dft(Y[c], 1);
            for (int p = 0; p < n; ++p) { cld d = cld(out[idx(c,b,p)].x, out[idx(c,b,p)].y) - y[p]; e = std::max(e, (double)std::abs(d)); nrm = std::max(nrm, (double)std::abs(y[p])); }
            for (int p = n; p < L; ++p) if (out[idx(c,b,p)].x != 7) { printf("Z wrote beyond n\n"); worst = 1; }
        }
    }
    for (int b = nb; b < NP; ++b) for (int c = 0; c < 3; ++c) if (out[idx(c, b, 0)].x != 7) { printf("Z wrote partial\n"); worst = 1; }
    rep<C>(XCHG ? "Z(exchange)" : PAIR ? "Z(pairs)" : "Z", L, e / nrm, sizeof(Real) == 4 ? 5e-6 : 1e-14);
}
// the multiply-step task layouts: z-mirror pairs, one task per k1, and one
// task per k1 in the lane-exchange order
template <int L1, int L2, typename Real> void check_z(int n) {
    check_z_impl<L1, L2, Real, true>(n);
    check_z_impl<L1, L2, Real, false>(n);
    check_z_impl<L1, L2, Real, false, true>(n);
}

template <int L1, int L2, typename Real> void check_x(int nx) {
    using C = cx_t<Real>;
    constexpr int L = L1 * L2;
    constexpr int NR = fast_nr_x<Real>(L1, L2);
    using X = XFast<L1, L2, NR, C>;
    auto tw = table<C>(L);
    int nr = NR > 1 ? NR - 1 : NR;
    // forward: rows (r, c) inputs x < nx -> all L outputs
    std::vector<cld> In(3 * NR * L);
    for (auto& v : In) v = cld(urand(), urand());
    std::vector<C> fo(3 * NR * L, mk<C>(7, 7));
    std::vector<C> lds(X::LDS);
    for (int t = 0; t < X::TA; ++t) X::fwd_a(t, lds.data(), tw.data(), nr, nx, [&](int r, int c, int x) { cld z = In[(r * 3 + c) * L + x]; return mk<C>((double)z.real(), (double)z.imag()); });
    for (int t = 0; t < X::TB; ++t) X::fwd_b(t, lds.data(), nr, [&](int r, int c, int x, C v) { fo[(r * 3 + c) * L + x] = v; });
    double e = 0, nrm = 0;
    for (int b = 0; b < 3 * nr; ++b) {
        std::vector<cld> x(L); for (int p = 0; p < L; ++p) x[p] = p < nx ? In[b * L + p] : cld(0);
        auto R = dft(x, -1);
        for (int p = 0; p < L; ++p) { cld d = cld(fo[b*L+p].x, fo[b*L+p].y) - R[p]; e = std::max(e, (double)std::abs(d)); nrm = std::max(nrm, (double)std::abs(R[p])); }
    }
    for (int c = 0; c < 3; ++c) if (nr < NR && fo[((NR - 1) * 3 + c) * L].x != 7) { printf("X wrote partial row\n"); worst = 1; }
    rep<C>("Xf", L, e / nrm, sizeof(Real) == 4 ? 3e-6 : 1e-14);
    // inverse: all L inputs -> outputs x < nx
    std::vector<C> io(3 * NR * L, mk<C>(7, 7));
    for (int t = 0; t < X::TA; ++t) X::inv_a(t, lds.data(), tw.data(), nr, [&](int r, int c, int x) { cld z = In[(r * 3 + c) * L + x]; return mk<C>((double)z.real(), (double)z.imag()); });
    for (int t = 0; t < X::TB; ++t) X::inv_b(t, lds.data(), nr, nx, [&](int r, int c, int x, C v) { io[(r * 3 + c) * L + x] = v; });
    e = 0; nrm = 0;
    for (int b = 0; b < 3 * nr; ++b) {
        std::vector<cld> x(In.begin() + b * L, In.begin() + (b + 1) * L);
        auto R = dft(x, 1);
        for (int p = 0; p < nx; ++p) { cld d = cld(io[b*L+p].x, io[b*L+p].y) - R[p]; e = std::max(e, (double)std::abs(d)); nrm = std::max(nrm, (double)std::abs(R[p])); }
        for (int p = nx; p < L; ++p) if (io[b*L+p].x != 7) { printf("X wrote beyond nx\n"); worst = 1; }
    }
    rep<C>("Xi", L, e / nrm, sizeof(Real) == 4 ? 3e-6 : 1e-14);
}

template <int L1, int L2> void check_len() {
    constexpr int L = L1 * L2;
    for (int n : {L / 2, L / 2 - 1, (L + 1) / 2 + 1, L, 1, 3}) {
        if (n < 1 || n > L) continue;
        check_y<L1, L2, float>(n, L, -1); check_y<L1, L2, double>(n, L, -1);
        check_y<L1, L2, float>(L, n, 1); check_y<L1, L2, double>(L, n, 1);
        check_z<L1, L2, float>(n); check_z<L1, L2, double>(n);
        check_x<L1, L2, float>(n); check_x<L1, L2, double>(n);
    }
    printf("L=%d done (worst so far %.3e)\n", L, worst);
}

int main() {
    // small DFTs
    check_len<4, 4>(); check_len<5, 4>(); check_len<6, 4>(); check_len<8, 4>(); check_len<8, 5>();
    check_len<8, 6>(); check_len<8, 8>(); check_len<10, 8>(); check_len<12, 8>(); check_len<16, 8>();
    // Z-pass splits with L2 = 4
    for (int n : {20, 19, 22, 40, 1}) { check_z<10, 4, float>(n); check_z<10, 4, double>(n); }
    for (int n : {24, 23, 26, 48, 1}) { check_z<12, 4, float>(n); check_z<12, 4, double>(n); }
    for (int n : {32, 31, 34, 64, 1}) { check_z<16, 4, float>(n); check_z<16, 4, double>(n); }
    printf("Z splits with L2 = 4 done (worst so far %.3e)\n", worst);
    // X / Y splits with L1 = 8
    check_len<8, 10>(); check_len<8, 12>(); check_len<8, 16>(); check_len<16, 16>();
    printf("worst rel err %.3e\n", worst);
    return worst > 1e-5 ? 1 : 0;
}

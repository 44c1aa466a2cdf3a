/*
 * green_oracle.c -- FP64 CPU restatement of the reference's interaction-tensor
 * construction, used ONLY as test infrastructure (tests/, the smoke check and
 * the CPU-baseline leg of bench.py).  No product code links or loads this.
 *
 * What it restates (reference file:line):
 *   - SiCi::get_Si / get_Ci table lookup ........ src/SiCi.cpp:12-44
 *   - LDR_inter (point-dipole interaction) ....... src/AProductCore.cpp:351-386
 *   - FCD_inter (filtered coupled dipole) ........ src/AProductCore.cpp:388-458
 *   - A_dic_generator(x,y,z,m,n): image shift and
 *     exp(-(0.01 K r)^4) damping ................. src/AProductCore.cpp:335-349
 *   - lattice sum over m in [-MAXm,MAXm] (outer),
 *     n in [-MAXn,MAXn] (inner), components
 *     (00,01,02,11,12,22) ........................ src/AProductCore.cpp:63-244
 *
 * Output layout: a centred offset grid out[ix][iy][iz][6][2] with
 * ix = ox + (Nx-1) in [0, 2Nx-2] (and likewise y, z), i.e. offset ox in
 * [-(Nx-1), Nx-1].  Callers place it on whatever FFT grid they use.
 *
 * With octant_only != 0 the sum is evaluated for ox,oy,oz >= 0 only and the
 * other seven octants are filled from the per-axis parity of each component
 * (xy odd in x and y, xz odd in x and z, yz odd in y and z, diagonal even).
 * tests/ check that both modes agree on small grids, which is what lets the
 * large golden case use the cheaper mode.
 *
 * Parity: pinned by the golden objective of python_tests/test_DDAModelWrapper.py:40,
 * by the textbook free-space dyadic Green's function (LDR), by the FCD -> LDR
 * limit with exact Si/Ci, and by brute-force image sums on tiny grids.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    const double* si;
    const double* ci;
    long n;
    double delta;
} sici_table;

/* SiCi.cpp:12-44: index p1 = floor(x/delta)-1, p2 = ceil(x/delta)-1; past the
 * end returns the last entry; otherwise linear interpolation between the
 * abscissae (p+1)*delta.  An exact multiple of delta gives 0/0 (NaN) exactly
 * as the reference does.  p1 < 0 would read before the table in the
 * reference; it never happens for r >= d (x >= pi - K d), and is clamped here. */
static double table_lookup(const double* t, long n, double delta, double x) {
    long p1 = (long)(floor(x / delta) - 1.0);
    long p2 = (long)(ceil(x / delta) - 1.0);
    if (p2 >= n) return t[n - 1];
    if (p1 < 0) p1 = 0;
    if (p2 < 0) p2 = 0;
    double v1 = t[p1], v2 = t[p2];
    double a1 = (double)(p1 + 1) * delta, a2 = (double)(p2 + 1) * delta;
    return v1 + ((v2 - v1) / (a2 - a1)) * (x - a1);
}

/* six symmetric components (00,01,02,11,12,22), complex as re/im pairs */
typedef struct { double re[6], im[6]; } sym6;

/* AProductCore.cpp:351-386.  A = exp(iKr)/r^3 * [c2 (r^2 - 3 x_i x_j) ... ]
 * with c2 = (1 - iKr)/r^2 and c3 = K^2 - 3 c2 for the off-diagonals. */
static sym6 ldr_tensor(double K, double x, double y, double z) {
    sym6 s; memset(&s, 0, sizeof s);
    double xx = x * x, yy = y * y, zz = z * z;
    double r = sqrt(xx + yy + zz);
    if (r == 0.0) return s;
    double r2 = r * r, r3 = r2 * r;
    /* pre = exp(iKr)/r^3 */
    double pre_re = cos(K * r) / r3, pre_im = sin(K * r) / r3;
    /* c2 = (1 - iKr)/r^2 */
    double c2_re = 1.0 / r2, c2_im = -K * r / r2;
    /* c3 = K^2 - 3 c2 */
    double c3_re = K * K - 3.0 * c2_re, c3_im = -3.0 * c2_im;
    double k2 = K * K;
    double diag_re[3], diag_im[3];
    diag_re[0] = c2_re * (r2 - 3 * xx) - k2 * (yy + zz); diag_im[0] = c2_im * (r2 - 3 * xx);
    diag_re[1] = c2_re * (r2 - 3 * yy) - k2 * (zz + xx); diag_im[1] = c2_im * (r2 - 3 * yy);
    diag_re[2] = c2_re * (r2 - 3 * zz) - k2 * (xx + yy); diag_im[2] = c2_im * (r2 - 3 * zz);
    double off[3] = {x * y, z * x, y * z}; /* 01, 02, 12 */
    double in_re[6] = {diag_re[0], off[0] * c3_re, off[1] * c3_re, diag_re[1], off[2] * c3_re, diag_re[2]};
    double in_im[6] = {diag_im[0], off[0] * c3_im, off[1] * c3_im, diag_im[1], off[2] * c3_im, diag_im[2]};
    for (int c = 0; c < 6; ++c) {
        s.re[c] = pre_re * in_re[c] - pre_im * in_im[c];
        s.im[c] = pre_re * in_im[c] + pre_im * in_re[c];
    }
    return s;
}

/* AProductCore.cpp:388-458.  Filtered coupled dipole interaction with
 * Kf = pi/d; Si/Ci come from the tables at (Kf +- K) r. */
static sym6 fcd_tensor(double K, double d, const sici_table* tb, double x, double y, double z) {
    sym6 s; memset(&s, 0, sizeof s);
    double xx = x * x, yy = y * y, zz = z * z;
    double r = sqrt(xx + yy + zz);
    if (r == 0.0) return s;
    double r2 = r * r, r3 = r2 * r;
    double kf = M_PI / d;
    double ap = (kf + K) * r, am = (kf - K) * r;
    double ci_p = table_lookup(tb->ci, tb->n, tb->delta, ap);
    double ci_m = table_lookup(tb->ci, tb->n, tb->delta, am);
    double si_p = table_lookup(tb->si, tb->n, tb->delta, ap);
    double si_m = table_lookup(tb->si, tb->n, tb->delta, am);
    double cos_p = cos(ap), sin_p = sin(ap), cos_m = cos(am), sin_m = sin(am);
    /* first and second radial derivatives of cos(a r)/.. and sin(a r)/.. terms */
    double c1_p = cos_p / r, c1_m = cos_m / r, s1_p = sin_p / r, s1_m = sin_m / r;
    double c2_p = (-ap * sin_p - cos_p) / r2, c2_m = (-am * sin_m - cos_m) / r2;
    double s2_p = (ap * cos_p - sin_p) / r2, s2_m = (am * cos_m - sin_m) / r2;
    double kr = K * r;
    double skr = sin(kr), ckr = cos(kr);
    double skf = sin(kf * r), ckf = cos(kf * r);
    /* CC = i*pi + Ci(am) - Ci(ap) */
    double cc_re = ci_m - ci_p, cc_im = M_PI;
    double dc1 = c1_m - c1_p, dc2 = c2_m - c2_p;
    double ss = si_p + si_m, ss1 = s1_p + s1_m, ss2 = s2_p + s2_m;
    /* gf = (sin(kr) CC + cos(kr) SS) / (pi r) */
    double gf_re = (skr * cc_re + ckr * ss) / (M_PI * r);
    double gf_im = (skr * cc_im) / (M_PI * r);
    /* q1 = (kr cos - sin) CC - (kr sin + cos) SS + r sin dc1 + r cos ss1 */
    double q1_re = (kr * ckr - skr) * cc_re - (kr * skr + ckr) * ss + r * skr * dc1 + r * ckr * ss1;
    double q1_im = (kr * ckr - skr) * cc_im;
    /* q2 = (-K kr sin) CC - (K kr cos) SS + 2 kr cos dc1 - 2 kr sin ss1 + r sin dc2 + r cos ss2 */
    double q2_re = (-K * kr * skr) * cc_re - (K * kr * ckr) * ss + (2 * kr * ckr) * dc1 - (2 * kr * skr) * ss1 + r * skr * dc2 + r * ckr * ss2;
    double q2_im = (-K * kr * skr) * cc_im;
    double g1_re = q1_re / (M_PI * r2), g1_im = q1_im / (M_PI * r2);
    double g2_re = (r * q2_re - 2.0 * q1_re) / (M_PI * r3), g2_im = (r * q2_im - 2.0 * q1_im) / (M_PI * r3);
    double hr = (skf - kf * r * ckf) / (2.0 * M_PI * M_PI * r3);
    double t1_re = K * K * gf_re + g1_re / r + (4.0 * M_PI / 3.0) * hr;
    double t1_im = K * K * gf_im + g1_im / r;
    double t2_re = (g2_re - g1_re / r) / r2, t2_im = (g2_im - g1_im / r) / r2;
    double w[6] = {xx, x * y, z * x, yy, y * z, zz};
    int diag[6] = {1, 0, 0, 1, 0, 1};
    for (int c = 0; c < 6; ++c) {
        s.re[c] = -(diag[c] * t1_re + w[c] * t2_re);
        s.im[c] = -(diag[c] * t1_im + w[c] * t2_im);
    }
    return s;
}

/* AProductCore.cpp:335-349 plus the (m,n) image loop of :63-244. */
static sym6 lattice_sum(int method, double K, double d, const sici_table* tb,
                        double x, double y, double z, int maxm, int maxn, double lm, double ln) {
    sym6 acc; memset(&acc, 0, sizeof acc);
    for (int m = -maxm; m <= maxm; ++m) {
        for (int n = -maxn; n <= maxn; ++n) {
            double xs = x + m * lm, ys = y + n * ln;
            sym6 t = method == 0 ? fcd_tensor(K, d, tb, xs, ys, z) : ldr_tensor(K, xs, ys, z);
            double r = sqrt(xs * xs + ys * ys + z * z);
            double damp = exp(-pow(0.01 * K * r, 4));
            for (int c = 0; c < 6; ++c) {
                acc.re[c] += t.re[c] * damp;
                acc.im[c] += t.im[c] * damp;
            }
        }
    }
    return acc;
}

/* parity of component c under x -> -x, y -> -y, z -> -z */
static const int kParity[6][3] = {
    {1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}, {1, -1, -1}, {1, 1, 1}};

/*
 * method: 0 = FCD, 1 = LDR.  lm, ln are the image displacements in nm
 * (cells * d, DDAModel.cpp:281).  out has (2Nx-1)(2Ny-1)(2Nz-1)*12 doubles.
 * nthreads > 1 splits the ox planes over pthreads-free OpenMP if available.
 */
int oracle_green_grid(int nx, int ny, int nz, double d, double K, int method,
                      int maxm, int maxn, double lm, double ln,
                      const double* si, const double* ci, long nsici, double delta,
                      int octant_only, double* out) {
    sici_table tb = {si, ci, nsici, delta};
    const int gx = 2 * nx - 1, gy = 2 * ny - 1, gz = 2 * nz - 1;
    const int x0 = octant_only ? 0 : -(nx - 1);
    const int y0 = octant_only ? 0 : -(ny - 1);
    const int z0 = octant_only ? 0 : -(nz - 1);
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int ox = x0; ox <= nx - 1; ++ox) {
        for (int oy = y0; oy <= ny - 1; ++oy) {
            for (int oz = z0; oz <= nz - 1; ++oz) {
                sym6 a = lattice_sum(method, K, d, &tb, d * ox, d * oy, d * oz, maxm, maxn, lm, ln);
                int sx_n = octant_only ? 2 : 1;
                for (int fx = 0; fx < sx_n; ++fx)
                for (int fy = 0; fy < (octant_only ? 2 : 1); ++fy)
                for (int fz = 0; fz < (octant_only ? 2 : 1); ++fz) {
                    if ((fx && ox == 0) || (fy && oy == 0) || (fz && oz == 0)) continue;
                    int px = fx ? -ox : ox, py = fy ? -oy : oy, pz = fz ? -oz : oz;
                    long cell = ((long)(px + nx - 1) * gy + (py + ny - 1)) * gz + (pz + nz - 1);
                    for (int c = 0; c < 6; ++c) {
                        int sgn = (fx ? kParity[c][0] : 1) * (fy ? kParity[c][1] : 1) * (fz ? kParity[c][2] : 1);
                        out[cell * 12 + 2 * c] = sgn * a.re[c];
                        out[cell * 12 + 2 * c + 1] = sgn * a.im[c];
                    }
                }
            }
        }
    }
    (void)gx;
    return 0;
}

/* Single-point evaluation (no damping, no images) for closed-form tests. */
int oracle_green_point(double K, double d, int method, const double* si, const double* ci,
                       long nsici, double delta, double x, double y, double z, double* out12) {
    sici_table tb = {si, ci, nsici, delta};
    sym6 a = method == 0 ? fcd_tensor(K, d, &tb, x, y, z) : ldr_tensor(K, x, y, z);
    for (int c = 0; c < 6; ++c) { out12[2 * c] = a.re[c]; out12[2 * c + 1] = a.im[c]; }
    return 0;
}

/* Table lookup exposed for tests of the Si/Ci semantics. */
double oracle_sici_lookup(const double* t, long n, double delta, double x) {
    return table_lookup(t, n, delta, x);
}

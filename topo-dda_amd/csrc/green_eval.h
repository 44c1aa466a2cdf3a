// Device evaluation of the interaction tensor A(r) = -G(r) of one dipole pair
// (AProductCore::LDR_inter / FCD_inter, src/AProductCore.cpp:351-458) and the
// Si/Ci table lookup it uses (src/SiCi.cpp:12-44).  Components (00,01,02,11,12,22).
#pragma once

#include "common.h"

namespace tdda {

// SiCi.cpp:12-44 table lookup (abscissae (p+1)*delta; past the end -> last entry).
__device__ inline double sici_lookup(const double* __restrict__ t, long n, double delta, double x) {
    long p1 = (long)(floor(x / delta) - 1.0);
    long p2 = (long)(ceil(x / delta) - 1.0);
    if (p2 >= n) return t[n - 1];
    if (p1 < 0) p1 = 0;   // r >= d keeps x >= pi - K d > delta; clamp only guards memory
    if (p2 < 0) p2 = 0;
    const double v1 = t[p1], v2 = t[p2];
    const double a1 = (double)(p1 + 1) * delta, a2 = (double)(p2 + 1) * delta;
    return v1 + ((v2 - v1) / (a2 - a1)) * (x - a1);
}

struct Sym6 {
    double re[6], im[6];
};

// AProductCore.cpp:351-386 (point dipole)
__device__ inline void ldr_tensor(double K, double x, double y, double z, Sym6& s) {
    const double xx = x * x, yy = y * y, zz = z * z;
    const double r = sqrt(xx + yy + zz);
    for (int c = 0; c < 6; ++c) { s.re[c] = 0.0; s.im[c] = 0.0; }
    if (r == 0.0) return;
    const double r2 = r * r, r3 = r2 * r;
    double sk, ck;
    sincos(K * r, &sk, &ck);
    const double pre_re = ck / r3, pre_im = sk / r3;
    const double c2_re = 1.0 / r2, c2_im = -K * r / r2;
    const double c3_re = K * K - 3.0 * c2_re, c3_im = -3.0 * c2_im;
    const double k2 = K * K;
    const double in_re[6] = {c2_re * (r2 - 3 * xx) - k2 * (yy + zz), x * y * c3_re, z * x * c3_re,
                             c2_re * (r2 - 3 * yy) - k2 * (zz + xx), y * z * c3_re,
                             c2_re * (r2 - 3 * zz) - k2 * (xx + yy)};
    const double in_im[6] = {c2_im * (r2 - 3 * xx), x * y * c3_im, z * x * c3_im, c2_im * (r2 - 3 * yy),
                             y * z * c3_im, c2_im * (r2 - 3 * zz)};
    for (int c = 0; c < 6; ++c) {
        s.re[c] = pre_re * in_re[c] - pre_im * in_im[c];
        s.im[c] = pre_re * in_im[c] + pre_im * in_re[c];
    }
}

// AProductCore.cpp:388-458 (filtered coupled dipole, Kf = pi/d)
__device__ inline void fcd_tensor(double K, double d, const double* __restrict__ si, const double* __restrict__ ci,
                                  long nt, double delta, double x, double y, double z, Sym6& s) {
    const double xx = x * x, yy = y * y, zz = z * z;
    const double r = sqrt(xx + yy + zz);
    for (int c = 0; c < 6; ++c) { s.re[c] = 0.0; s.im[c] = 0.0; }
    if (r == 0.0) return;
    const double r2 = r * r, r3 = r2 * r;
    const double kf = M_PI / d;
    const double ap = (kf + K) * r, am = (kf - K) * r;
    const double ci_p = sici_lookup(ci, nt, delta, ap);
    const double ci_m = sici_lookup(ci, nt, delta, am);
    const double si_p = sici_lookup(si, nt, delta, ap);
    const double si_m = sici_lookup(si, nt, delta, am);
    double sin_p, cos_p, sin_m, cos_m, skr, ckr, skf, ckf;
    sincos(ap, &sin_p, &cos_p);
    sincos(am, &sin_m, &cos_m);
    const double kr = K * r;
    sincos(kr, &skr, &ckr);
    sincos(kf * r, &skf, &ckf);
    const double c1_p = cos_p / r, c1_m = cos_m / r, s1_p = sin_p / r, s1_m = sin_m / r;
    const double c2_p = (-ap * sin_p - cos_p) / r2, c2_m = (-am * sin_m - cos_m) / r2;
    const double s2_p = (ap * cos_p - sin_p) / r2, s2_m = (am * cos_m - sin_m) / r2;
    const double cc_re = ci_m - ci_p, cc_im = M_PI;
    const double dc1 = c1_m - c1_p, dc2 = c2_m - c2_p;
    const double ss = si_p + si_m, ss1 = s1_p + s1_m, ss2 = s2_p + s2_m;
    const double gf_re = (skr * cc_re + ckr * ss) / (M_PI * r);
    const double gf_im = (skr * cc_im) / (M_PI * r);
    const double q1_re = (kr * ckr - skr) * cc_re - (kr * skr + ckr) * ss + r * skr * dc1 + r * ckr * ss1;
    const double q1_im = (kr * ckr - skr) * cc_im;
    const double q2_re = (-K * kr * skr) * cc_re - (K * kr * ckr) * ss + (2 * kr * ckr) * dc1 -
                         (2 * kr * skr) * ss1 + r * skr * dc2 + r * ckr * ss2;
    const double q2_im = (-K * kr * skr) * cc_im;
    const double g1_re = q1_re / (M_PI * r2), g1_im = q1_im / (M_PI * r2);
    const double g2_re = (r * q2_re - 2.0 * q1_re) / (M_PI * r3), g2_im = (r * q2_im - 2.0 * q1_im) / (M_PI * r3);
    const double hr = (skf - kf * r * ckf) / (2.0 * M_PI * M_PI * r3);
    const double t1_re = K * K * gf_re + g1_re / r + (4.0 * M_PI / 3.0) * hr;
    const double t1_im = K * K * gf_im + g1_im / r;
    const double t2_re = (g2_re - g1_re / r) / r2, t2_im = (g2_im - g1_im / r) / r2;
    const double w[6] = {xx, x * y, z * x, yy, y * z, zz};
    const double dg[6] = {1.0, 0.0, 0.0, 1.0, 0.0, 1.0};
    for (int c = 0; c < 6; ++c) {
        s.re[c] = -(dg[c] * t1_re + w[c] * t2_re);
        s.im[c] = -(dg[c] * t1_im + w[c] * t2_im);
    }
}

}  // namespace tdda

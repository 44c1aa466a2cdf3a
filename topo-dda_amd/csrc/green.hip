// Green's-tensor setup in FP64 (AProductCore::AProductCore, src/AProductCore.cpp:12-298).
//
// 1. k_green_*: the damped, lattice-summed interaction tensor A(o) for the
//    octant offsets o = (ox,oy,oz) >= 0 (AProductCore.cpp:63-244, 335-458).
//    Every component is even or odd in each axis, and the image sum is
//    symmetric in m -> -m, n -> -n, so the octant determines the full grid.
// 2. k_green_dft (x3): the spectrum on the padded grid restricted to
//    0 <= k <= L/2 per axis: Ghat(k) = G(0) + sum_{o>0} G(o) 2cos(2 pi k o/L)
//    for an even axis, sum_{o>0} G(o) (-2i) sin(2 pi k o/L) for an odd one.
//    This is the DFT of the zero-padded, wrapped grid the reference builds
//    (AProductCore.cpp:91-244 + 262-279) evaluated directly; k > L/2 follows
//    from Ghat(L-k) = +-Ghat(k).
#include "green.h"
#include "green_eval.h"

namespace tdda {

// parity under x->-x, y->-y, z->-z of (00,01,02,11,12,22)
__constant__ int c_odd[6][3] = {{0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 0, 0}, {0, 1, 1}, {0, 0, 0}};

// A_dic_generator(x,y,z,m,n): shifted image times exp(-(0.01 K r)^4)
__device__ inline void image_term(const GreenParams& P, const double* si, const double* ci, double x, double y,
                                  double z, int m, int n, Sym6& t) {
    const double xs = x + m * P.lm, ys = y + n * P.ln;
    if (P.method == 0)
        fcd_tensor(P.K, P.d, si, ci, P.nsici, P.delta, xs, ys, z, t);
    else
        ldr_tensor(P.K, xs, ys, z, t);
    const double rr = sqrt(xs * xs + ys * ys + z * z);
    const double q = 0.01 * P.K * rr;
    const double q2 = q * q;
    const double damp = exp(-(q2 * q2));
    for (int c = 0; c < 6; ++c) { t.re[c] *= damp; t.im[c] *= damp; }
}

// One thread per octant offset (no images).
__global__ void __launch_bounds__(256) k_green_single(GreenParams P, const double* __restrict__ si,
                                                     const double* __restrict__ ci, double2* __restrict__ out) {
    const long n = (long)P.nx * P.ny * P.nz;
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= n) return;
    const int ox = cell % P.nx, oy = (cell / P.nx) % P.ny, oz = cell / ((long)P.nx * P.ny);
    Sym6 t;
    image_term(P, si, ci, P.d * ox, P.d * oy, P.d * oz, 0, 0, t);
    for (int c = 0; c < 6; ++c) out[c * n + cell] = make_double2(t.re[c], t.im[c]);
}

// One workgroup per octant offset; threads split the (2MAXm+1)(2MAXn+1) images
// (m outer, n inner order within a thread) and reduce in a fixed tree.
__global__ void __launch_bounds__(256) k_green_lattice(GreenParams P, const double* __restrict__ si,
                                                      const double* __restrict__ ci, double2* __restrict__ out) {
    const long n = (long)P.nx * P.ny * P.nz;
    const long cell = blockIdx.x;
    const int ox = cell % P.nx, oy = (cell / P.nx) % P.ny, oz = cell / ((long)P.nx * P.ny);
    const double x = P.d * ox, y = P.d * oy, z = P.d * oz;
    const int nn = 2 * P.maxn + 1;
    const int nimg = (2 * P.maxm + 1) * nn;
    double acc[12];
    for (int c = 0; c < 12; ++c) acc[c] = 0.0;
    for (int i = threadIdx.x; i < nimg; i += blockDim.x) {
        const int m = i / nn - P.maxm, nimg_n = i % nn - P.maxn;
        Sym6 t;
        image_term(P, si, ci, x, y, z, m, nimg_n, t);
        for (int c = 0; c < 6; ++c) { acc[2 * c] += t.re[c]; acc[2 * c + 1] += t.im[c]; }
    }
    __shared__ double red[12][256];
    for (int c = 0; c < 12; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 12; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int c = 0; c < 6; ++c) {
            // odd components vanish on their odd axes' zero plane by symmetry
            const bool zero = (c_odd[c][0] && ox == 0) || (c_odd[c][1] && oy == 0) || (c_odd[c][2] && oz == 0);
            out[c * n + cell] = zero ? make_double2(0.0, 0.0) : make_double2(red[2 * c][0], red[2 * c + 1][0]);
        }
    }
}

// out[c][a][k][i] = sum_o W_c(k,o) in[c][a][o][i], 0 <= k < h, 0 <= o < no.
// A thread computes GREEN_KB outputs k = kb + j nk of one column (j < KB,
// nk = ceil(h / KB)), so every input it loads feeds KB products; the twiddle
// table sits in LDS.  Consecutive threads take consecutive columns i (then
// consecutive kb), so the input loads are coalesced along the contiguous axis
// or shared by a wave when inner = 1.
constexpr int GREEN_KB = 4;
constexpr int GREEN_MAX_L = 256;
__global__ void __launch_bounds__(256) k_green_dft(const double2* __restrict__ in, double2* __restrict__ out,
                                                   int outer, int no, int h, int inner, int L, int axis,
                                                   const double2* __restrict__ tw) {
    __shared__ double2 twl[GREEN_MAX_L];
    for (int t = threadIdx.x; t < L; t += blockDim.x) twl[t] = tw[t];
    __syncthreads();
    const int nk = (h + GREEN_KB - 1) / GREEN_KB;
    const long total = 6L * outer * nk * inner;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int i = (int)(idx % inner);
    const int kb = (int)((idx / inner) % nk);
    const long a = (idx / ((long)inner * nk)) % outer;
    const int c = (int)(idx / ((long)inner * nk * outer));
    const bool odd = c_odd[c][axis] != 0;
    const double2* src = in + ((long)c * outer + a) * no * inner + i;
    int k[GREEN_KB], m[GREEN_KB];
    double2 acc[GREEN_KB];
#pragma unroll
    for (int j = 0; j < GREEN_KB; ++j) {
        k[j] = min(kb + j * nk, h - 1);   // clamped lanes compute a duplicate that is not stored
        m[j] = 0;                         // phase index k o mod L, advanced by k per input
        acc[j] = make_double2(0.0, 0.0);
    }
    for (int o = 0; o < no; ++o) {
        const double2 v = src[(long)o * inner];
#pragma unroll
        for (int j = 0; j < GREEN_KB; ++j) {
            const double2 w = twl[m[j]];   // exp(-2 pi i m / L)
            if (!odd) {
                const double f = o == 0 ? 1.0 : 2.0 * w.x;   // 2 cos
                acc[j].x += f * v.x;
                acc[j].y += f * v.y;
            } else if (o > 0) {
                const double s2 = -2.0 * w.y;   // 2 sin; (-i s2) v
                acc[j].x += s2 * v.y;
                acc[j].y -= s2 * v.x;
            }
            m[j] = (m[j] + k[j] >= L) ? m[j] + k[j] - L : m[j] + k[j];
        }
    }
    double2* dst = out + ((long)c * outer + a) * h * inner + i;
#pragma unroll
    for (int j = 0; j < GREEN_KB; ++j)
        if (kb + j * nk < h) dst[(long)(kb + j * nk) * inner] = acc[j];
}

__global__ void k_green_interleave(const double2* __restrict__ in, long rows, int hx, double2* __restrict__ out64,
                                   float2* __restrict__ out32) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;   // over [rows][6][hx]
    if (i >= 6 * rows * hx) return;
    const int x = (int)(i % hx);
    const long rc = i / hx;
    const long row = rc / 6;
    const int c = (int)(rc - 6 * row);
    const double2 v = in[(c * rows + row) * hx + x];
    out64[i] = v;
    out32[((row * 3 + (c >> 1)) * hx + x) * 2 + (c & 1)] = make_float2((float)v.x, (float)v.y);
}

void launch_green_octant(const GreenParams& P, const double* si, const double* ci, double2* out, hipStream_t st) {
    const long n = (long)P.nx * P.ny * P.nz;
    if (P.maxm == 0 && P.maxn == 0)
        hipLaunchKernelGGL(k_green_single, dim3((n + 255) / 256), dim3(256), 0, st, P, si, ci, out);
    else
        hipLaunchKernelGGL(k_green_lattice, dim3(n), dim3(256), 0, st, P, si, ci, out);
    TDDA_HIP(hipGetLastError());
}

void launch_green_dft(const double2* in, double2* out, int outer, int no, int h, int inner, int L, int axis,
                      const double2* tw, hipStream_t st) {
    TDDA_REQUIRE(L <= GREEN_MAX_L, ERR_INVALID, "transform length beyond the Green's DFT table");
    const long total = 6L * outer * ((h + GREEN_KB - 1) / GREEN_KB) * inner;
    hipLaunchKernelGGL(k_green_dft, dim3((total + 255) / 256), dim3(256), 0, st, in, out, outer, no, h, inner, L,
                       axis, tw);
    TDDA_HIP(hipGetLastError());
}

void launch_green_interleave(const double2* in, long rows, int hx, double2* out64, float2* out32, hipStream_t st) {
    const long n = 6 * rows * hx;
    hipLaunchKernelGGL(k_green_interleave, dim3((n + 255) / 256), dim3(256), 0, st, in, rows, hx, out64, out32);
    TDDA_HIP(hipGetLastError());
}

}  // namespace tdda

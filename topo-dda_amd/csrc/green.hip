// Green's-tensor setup in FP64 (AProductCore::AProductCore, src/AProductCore.cpp:12-298).
//
// 1. k_green_*: the damped, lattice-summed interaction tensor A(o) for the
//    octant offsets o = (ox,oy,oz) >= 0 (AProductCore.cpp:63-244, 335-458).
//    Every component is even or odd in each axis, and the image sum is
//    symmetric in m -> -m, n -> -n, so the octant determines the full grid.
// 2. k_green_dft (x3, one axis each): the spectrum on the padded grid restricted to
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
__device__ inline void green_single_cell(const GreenParams& P, const double* __restrict__ si,
                                         const double* __restrict__ ci, double2* __restrict__ out, long cell) {
    const long n = (long)P.nx * P.ny * P.nz;
    if (cell >= n) return;
    const int ox = cell % P.nx, oy = (cell / P.nx) % P.ny, oz = cell / ((long)P.nx * P.ny);
    Sym6 t;
    image_term(P, si, ci, P.d * ox, P.d * oy, P.d * oz, 0, 0, t);
    for (int c = 0; c < 6; ++c) out[c * n + cell] = make_double2(t.re[c], t.im[c]);
}

__global__ void __launch_bounds__(256) k_green_single(GreenParams P, const double* __restrict__ si,
                                                     const double* __restrict__ ci, double2* __restrict__ out) {
    green_single_cell(P, si, ci, out, (long)blockIdx.x * blockDim.x + threadIdx.x);
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

// out[c][a][k][i] = sum_o W_c(k,o) in[c][a][o][i], 0 <= k < h, 0 <= o < no,
// with W_c(k,0) = 1, W_c(k,o) = 2 cos(2 pi k o/L) for an even component and
// W_c(k,0) = 0, W_c(k,o) = -i 2 sin(2 pi k o/L) for an odd one.
// A thread owns one column (a, i) and KB consecutive outputs k0 .. k0+KB-1;
// c and k0 are uniform over the workgroup, so the phase indices k o mod L and
// the twiddles exp(-2 pi i m/L) are scalar values shared by the whole group,
// and every input the thread loads feeds KB multiply-adds.  The sum over o
// runs in order, o = 0 first.
// The sum over o for one parity.  Half the sum is accumulated and doubled at
// the end: scaling by 2 is exact, so this equals summing 2 cos / 2 sin terms
// in the same order.  The extra outputs of the last group repeat k = h - 1
// (so every phase stays in the table) and are not stored; lanes past the
// last column read column 0 and store nothing.
template <int KB, bool ODD>
__device__ inline void green_dft_column(const double2* __restrict__ src, long stride, int no, int k0, int h, int L,
                                        const double2* __restrict__ tw, double2 (&acc)[KB]) {
    // byte offsets of the phases k o mod L in the table (of the sine parts
    // for an odd component)
    const unsigned period = 16u * (unsigned)L;
    unsigned mb[KB];
    const double2 v0 = src[0];
#pragma unroll
    for (int j = 0; j < KB; ++j) {
        mb[j] = 16u * (unsigned)min(k0 + j, h - 1);   // phase at o = 1
        acc[j] = ODD ? make_double2(0.0, 0.0) : make_double2(0.5 * v0.x, 0.5 * v0.y);   // o = 0, weight 1
    }
    const char* twb = reinterpret_cast<const char*>(tw) + (ODD ? 8 : 0);
    auto step = [&](const double2 v) {
        double w[KB];
#pragma unroll
        for (int j = 0; j < KB; ++j) w[j] = *reinterpret_cast<const double*>(twb + mb[j]);   // exp(-2 pi i m / L)
#pragma unroll
        for (int j = 0; j < KB; ++j) {
            if (!ODD) {   // cos v
                acc[j].x += w[j] * v.x;
                acc[j].y += w[j] * v.y;
            } else {      // sin (-i v) = sin (v.y, -v.x), from w = -sin
                acc[j].x -= w[j] * v.y;
                acc[j].y += w[j] * v.x;
            }
            mb[j] += 16u * (unsigned)min(k0 + j, h - 1);
            mb[j] -= mb[j] >= period ? period : 0u;
        }
    };
    // one input per iteration: the phases then stay in scalar registers and
    // each twiddle is a scalar load (42-74 VGPRs, 6-8 waves per SIMD); with
    // four inputs in flight the compiler kept per-lane copies of them and
    // vector loads (214-248 VGPRs, 1-2 waves per SIMD, 30-45 % slower)
#pragma unroll 1
    for (int o = 1; o < no; ++o) step(src[(long)o * stride]);
#pragma unroll
    for (int j = 0; j < KB; ++j) acc[j] = make_double2(2.0 * acc[j].x, 2.0 * acc[j].y);
}

// block b of one DFT axis
template <int KB>
__device__ inline void green_dft_block(const double2* __restrict__ in, double2* __restrict__ out, int outer, int no,
                                       int h, int inner, int L, int axis, int ng, const double2* __restrict__ tw,
                                       long b) {
    const long ncols = (long)outer * inner;
    const long nbc = (ncols + 255) / 256;
    const long cb = b % nbc;
    const int g = (int)((b / nbc) % ng);
    const int c = (int)(b / (nbc * ng));
    const int k0 = g * KB;
    const int nk = min(KB, h - k0);
    const long col = cb * 256 + threadIdx.x;
    const bool valid = col < ncols;
    const long a = valid ? col / inner : 0;
    const int i = valid ? (int)(col - a * inner) : 0;
    const double2* src = in + ((long)c * outer + a) * no * inner + i;
    double2 acc[KB];
    if (c_odd[c][axis])
        green_dft_column<KB, true>(src, inner, no, k0, h, L, tw, acc);
    else
        green_dft_column<KB, false>(src, inner, no, k0, h, L, tw, acc);
    if (!valid) return;
    double2* dst = out + ((long)c * outer + a) * h * inner + i;
#pragma unroll
    for (int j = 0; j < KB; ++j)
        if (j < nk) dst[(long)(k0 + j) * inner] = acc[j];
}

template <int KB>
__global__ void __launch_bounds__(256) k_green_dft(const double2* __restrict__ in, double2* __restrict__ out,
                                                   int outer, int no, int h, int inner, int L, int axis, int ng,
                                                   const double2* __restrict__ tw) {
    green_dft_block<KB>(in, out, outer, no, h, inner, L, axis, ng, tw, blockIdx.x);
}

__device__ inline void green_interleave_elem(const double2* __restrict__ in, long rows, int hx,
                                             double2* __restrict__ out64, float2* __restrict__ out32, long i) {
    if (i >= 6 * rows * hx) return;   // i over [rows][6][hx]
    const int x = (int)(i % hx);
    const long rc = i / hx;
    const long row = rc / 6;
    const int c = (int)(rc - 6 * row);
    const double2 v = in[(c * rows + row) * hx + x];
    out64[i] = v;
    out32[((row * 3 + (c >> 1)) * hx + x) * 2 + (c & 1)] = make_float2((float)v.x, (float)v.y);
}

__global__ void k_green_interleave(const double2* __restrict__ in, long rows, int hx, double2* __restrict__ out64,
                                   float2* __restrict__ out32) {
    green_interleave_elem(in, rows, hx, out64, out32, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// the job of block b: the last with block0 <= b (jobs in increasing block0)
template <typename Job> __device__ inline const Job& job_of_block(const Job* __restrict__ jobs, int nj, int b) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (jobs[mid].block0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return jobs[lo];
}

__global__ void __launch_bounds__(256) k_green_single_batch(const GreenSingleJob* __restrict__ jobs, int nj) {
    const GreenSingleJob& J = job_of_block(jobs, nj, blockIdx.x);
    green_single_cell(J.P, J.si, J.ci, J.out, (long)(blockIdx.x - J.block0) * 256 + threadIdx.x);
}

template <int KB>
__global__ void __launch_bounds__(256) k_green_dft_batch(const GreenDftJob* __restrict__ jobs, int nj) {
    const GreenDftJob& J = job_of_block(jobs, nj, blockIdx.x);
    green_dft_block<KB>(J.in, J.out, J.outer, J.no, J.h, J.inner, J.L, J.axis, J.ng, J.tw, blockIdx.x - J.block0);
}

__global__ void __launch_bounds__(256) k_green_interleave_batch(const GreenInterleaveJob* __restrict__ jobs, int nj) {
    const GreenInterleaveJob& J = job_of_block(jobs, nj, blockIdx.x);
    green_interleave_elem(J.in, J.rows, J.hx, J.out64, J.out32, (long)(blockIdx.x - J.block0) * 256 + threadIdx.x);
}

void launch_green_octant(const GreenParams& P, const double* si, const double* ci, double2* out, hipStream_t st) {
    const long n = (long)P.nx * P.ny * P.nz;
    if (P.maxm == 0 && P.maxn == 0)
        hipLaunchKernelGGL(k_green_single, dim3((n + 255) / 256), dim3(256), 0, st, P, si, ci, out);
    else
        hipLaunchKernelGGL(k_green_lattice, dim3(n), dim3(256), 0, st, P, si, ci, out);
    TDDA_HIP(hipGetLastError());
}

// KB = 8 or 16 outputs per thread, whichever wastes fewer on the last group
int green_dft_kb(int h) {
    const int g8 = (h + 7) / 8, g16 = (h + 15) / 16;
    return 8 * g8 < 16 * g16 ? 8 : 16;
}

int green_dft_blocks(int outer, int h, int inner) {
    const int kb = green_dft_kb(h);
    const long nbc = ((long)outer * inner + 255) / 256;
    return (int)(6L * ((h + kb - 1) / kb) * nbc);
}

void launch_green_dft(const double2* in, double2* out, int outer, int no, int h, int inner, int L, int axis,
                      const double2* tw, hipStream_t st) {
    const int kb = green_dft_kb(h);
    const int ng = (h + kb - 1) / kb;
    const dim3 grid((unsigned)green_dft_blocks(outer, h, inner));
    if (kb == 8)
        hipLaunchKernelGGL(k_green_dft<8>, grid, dim3(256), 0, st, in, out, outer, no, h, inner, L, axis, ng, tw);
    else
        hipLaunchKernelGGL(k_green_dft<16>, grid, dim3(256), 0, st, in, out, outer, no, h, inner, L, axis, ng, tw);
    TDDA_HIP(hipGetLastError());
}

void launch_green_single_batch(const GreenSingleJob* jobs, int nj, int blocks, hipStream_t st) {
    if (nj == 0) return;
    hipLaunchKernelGGL(k_green_single_batch, dim3(blocks), dim3(256), 0, st, jobs, nj);
    TDDA_HIP(hipGetLastError());
}

void launch_green_dft_batch(const GreenDftJob* jobs, int nj, int blocks, int kb, hipStream_t st) {
    if (nj == 0) return;
    if (kb == 8)
        hipLaunchKernelGGL(k_green_dft_batch<8>, dim3(blocks), dim3(256), 0, st, jobs, nj);
    else
        hipLaunchKernelGGL(k_green_dft_batch<16>, dim3(blocks), dim3(256), 0, st, jobs, nj);
    TDDA_HIP(hipGetLastError());
}

void launch_green_interleave_batch(const GreenInterleaveJob* jobs, int nj, int blocks, hipStream_t st) {
    if (nj == 0) return;
    hipLaunchKernelGGL(k_green_interleave_batch, dim3(blocks), dim3(256), 0, st, jobs, nj);
    TDDA_HIP(hipGetLastError());
}

void launch_green_interleave(const double2* in, long rows, int hx, double2* out64, float2* out32, hipStream_t st) {
    const long n = 6 * rows * hx;
    hipLaunchKernelGGL(k_green_interleave, dim3((n + 255) / 256), dim3(256), 0, st, in, rows, hx, out64, out32);
    TDDA_HIP(hipGetLastError());
}

}  // namespace tdda

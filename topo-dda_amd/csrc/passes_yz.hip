// Fused y forward / z / y inverse pass of the c64 matvec (gfx950).
//
// The three middle passes of the matvec (y FFT, z FFT with the 3x3 Green's
// multiply and inverse z FFT, inverse y FFT) only couple values of one x
// frequency kx, so one workgroup takes the whole slab of a kx: 3 components
// x Nz x Ny values, read once from and written once to the work grid, which
// the X passes keep in the [kx][3][nz][ny] layout for these jobs.  The three
// separate passes move the padded grid (3 Nz Ly per kx) through HBM four times.
//
// The y transform (length Ly = 16 LY1) is split as y = n2 + 16 n1,
// ky = k1 + LY1 k2 and streamed over k1:
//   stage 1: T[c][z][n2] = sum_n1 x[c][z][n2 + 16 n1] w_Ly^(k1 (n2 + 16 n1))
//            (the inputs of a thread's tasks stay in registers);
//   stage 2: 16-point DFT over n2 of every row of T -> ky = k1 + LY1 k2;
//   z:       the 16 pencils k2 of T (three components each) through the
//            specialised z transform, multiply and inverse (ZFast);
//   inverse: 16-point inverse DFT over k2 of every row, then
//            x[c][z][y] += w_Ly^(-k1 y) T[c][z][n2] accumulated in registers.
// LDS holds one k1 slice (3 x Nz x 16) and the z tile, about 40 KB at
// Nz = 32, Lz = 64.
#include <atomic>

#include "fft_fast.h"
#include "pass_util.h"

namespace tdda {

namespace {

constexpr int YZ_THREADS = 256;
constexpr int YZ_P2 = 16;          // second-stage length of the y transform = z pencils per step
constexpr int YZ_RS = YZ_P2 + 1;   // LDS row stride of T (rows read whole by one thread)

// class -> (LY1, LZ1, LZ2); LY1 in {2, 3, 4, 6, 8}, Lz in {32, 48, 64}
template <typename F>
TDDA_HD bool with_yz(int cls, F&& f) {
    switch (cls) {
        case 0: f(IC<2>{}, IC<8>{}, IC<4>{}); return true;
        case 1: f(IC<2>{}, IC<12>{}, IC<4>{}); return true;
        case 2: f(IC<2>{}, IC<16>{}, IC<4>{}); return true;
        case 3: f(IC<3>{}, IC<8>{}, IC<4>{}); return true;
        case 4: f(IC<3>{}, IC<12>{}, IC<4>{}); return true;
        case 5: f(IC<3>{}, IC<16>{}, IC<4>{}); return true;
        case 6: f(IC<4>{}, IC<8>{}, IC<4>{}); return true;
        case 7: f(IC<4>{}, IC<12>{}, IC<4>{}); return true;
        case 8: f(IC<4>{}, IC<16>{}, IC<4>{}); return true;
        case 9: f(IC<6>{}, IC<8>{}, IC<4>{}); return true;
        case 10: f(IC<6>{}, IC<12>{}, IC<4>{}); return true;
        case 11: f(IC<6>{}, IC<16>{}, IC<4>{}); return true;
        case 12: f(IC<8>{}, IC<8>{}, IC<4>{}); return true;
        case 13: f(IC<8>{}, IC<12>{}, IC<4>{}); return true;
        case 14: f(IC<8>{}, IC<16>{}, IC<4>{}); return true;
        default: return false;
    }
}

template <int LY1, int LZ1, int LZ2>
struct YzShape {
    static constexpr int LY = YZ_P2 * LY1, LZ = LZ1 * LZ2;
    static constexpr int NIN = (LY1 + 1) / 2;   // rows y = n2 + 16 n1 < Ny <= Ly / 2 per task
    static constexpr int NZMAX = LZ / 2;        // Nz <= Lz / 2
    static constexpr int TPT = (3 * NZMAX * YZ_P2 + YZ_THREADS - 1) / YZ_THREADS;   // stage-1 tasks per thread
    using Z = ZFast<LZ1, LZ2, YZ_P2, float2, false, LZ2 == 4>;
    static constexpr int T_ELEMS = 3 * NZMAX * YZ_RS;
    static constexpr size_t lds_bytes() { return (size_t)(Z::LDS + T_ELEMS + LY + LZ) * sizeof(float2); }
};

template <int LY1, int LZ1, int LZ2>
__global__ void __launch_bounds__(YZ_THREADS)
k_yz_fast(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap, const float2* __restrict__ twpool,
          int gate, const int* __restrict__ limit) {
    using C = float2;
    using S = YzShape<LY1, LZ1, LZ2>;
    using Z = typename S::Z;
    constexpr int LY = S::LY, LZ = S::LZ, NIN = S::NIN, TPT = S::TPT;
    extern __shared__ __align__(16) unsigned char smem_yz[];
    if (beyond_limit(limit)) return;
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    if (!gate_open(J, gate)) return;
    const int kx = jt.y;
    const int nz = J.nz, ny = J.ny;
    C* zl = reinterpret_cast<C*>(smem_yz);   // z tile
    C* T = zl + Z::LDS;                      // one k1 slice, rows (c, z) of YZ_RS
    C* twy = T + S::T_ELEMS;                 // exp(-2 pi i m / Ly)
    C* twz = twy + LY;                       // exp(-2 pi i m / Lz)
    for (int i = threadIdx.x; i < LY; i += YZ_THREADS) twy[i] = gload<C>(twpool + J.py.tw_off, (unsigned)i);
    for (int i = threadIdx.x; i < LZ; i += YZ_THREADS) twz[i] = gload<C>(twpool + J.pz.tw_off, (unsigned)i);
    C* g = static_cast<C*>(J.grid) + (size_t)kx * (3u * nz * ny);
    const int ntask = 3 * nz * YZ_P2;   // task = row (c, z) * 16 + n2
    const int nrow = 3 * nz;
    const C zero = mk<C>(0.0, 0.0);
    C in[TPT][NIN], acc[TPT][NIN];
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
        const int tau = threadIdx.x + u * YZ_THREADS;
        const int row = tau / YZ_P2, n2 = tau - row * YZ_P2;
#pragma unroll
        for (int n1 = 0; n1 < NIN; ++n1) {
            const int y = n2 + YZ_P2 * n1;
            in[u][n1] = (tau < ntask && y < ny) ? gload<C>(g, (unsigned)(row * ny + y)) : zero;
            acc[u][n1] = zero;
        }
    }
    const C* G = J.green32;
    __syncthreads();
    for (int k1 = 0; k1 < LY1; ++k1) {
        // stage 1: the k1 slice (inner LY1-point DFT and the w_Ly^(n2 k1) twiddle in one sum)
#pragma unroll
        for (int u = 0; u < TPT; ++u) {
            const int tau = threadIdx.x + u * YZ_THREADS;
            if (tau < ntask) {
                const int row = tau / YZ_P2, n2 = tau - row * YZ_P2;
                C a = zero;
#pragma unroll
                for (int n1 = 0; n1 < NIN; ++n1) {
                    const int y = n2 + YZ_P2 * n1;
                    if (y < ny) a = cadd(a, cmul(in[u][n1], twy[(k1 * y) % LY]));
                }
                T[row * YZ_RS + n2] = a;
            }
        }
        __syncthreads();
        // stage 2: 16-point DFTs along the rows (natural ky order k2)
        for (int row = threadIdx.x; row < nrow; row += YZ_THREADS) {
            C v[YZ_P2];
#pragma unroll
            for (int i = 0; i < YZ_P2; ++i) v[i] = T[row * YZ_RS + i];
            SDft<YZ_P2, -1, C>::run(v);
#pragma unroll
            for (int i = 0; i < YZ_P2; ++i) T[row * YZ_RS + i] = v[i];
        }
        __syncthreads();
        // z transform, Green's multiply and inverse over the 16 pencils ky = k1 + LY1 k2
        for (int t = threadIdx.x; t < Z::TA; t += YZ_THREADS)
            Z::fwd_a(t, zl, twz, YZ_P2, nz, [&](int c, int b, int pos) -> C { return T[(c * nz + pos) * YZ_RS + b]; });
        __syncthreads();
        for (int t = threadIdx.x; t < Z::TB; t += YZ_THREADS)
            Z::mid(
                t, zl, twz, YZ_P2,
                [&](int b, int kz, C (&gg)[6]) { green_fetch(G, J, kx, k1 + LY1 * b, kz, gg); },
                [&](int, int, const C (&mine)[6], C (&out)[6]) {
                    static_for<6>([&](auto q) {
                        constexpr int k = decltype(q)::value;
                        out[k].x = __shfl_xor(mine[k].x, 32);
                        out[k].y = __shfl_xor(mine[k].y, 32);
                    });
                });
        __syncthreads();
        for (int t = threadIdx.x; t < Z::TA; t += YZ_THREADS)
            Z::inv_a(t, zl, YZ_P2, nz, [&](int c, int b, int pos, C v) { T[(c * nz + pos) * YZ_RS + b] = v; });
        __syncthreads();
        // inverse stage 2 (natural n2 order)
        for (int row = threadIdx.x; row < nrow; row += YZ_THREADS) {
            C v[YZ_P2];
#pragma unroll
            for (int i = 0; i < YZ_P2; ++i) v[i] = T[row * YZ_RS + i];
            SDft<YZ_P2, 1, C>::run(v);
#pragma unroll
            for (int i = 0; i < YZ_P2; ++i) T[row * YZ_RS + i] = v[i];
        }
        __syncthreads();
        // inverse stage 1, accumulated: x[y] += w_Ly^(-k1 y) T[row][n2]
#pragma unroll
        for (int u = 0; u < TPT; ++u) {
            const int tau = threadIdx.x + u * YZ_THREADS;
            if (tau < ntask) {
                const int row = tau / YZ_P2, n2 = tau - row * YZ_P2;
                const C b = T[row * YZ_RS + n2];
#pragma unroll
                for (int n1 = 0; n1 < NIN; ++n1) {
                    const int y = n2 + YZ_P2 * n1;
                    const C w = twy[(k1 * y) % LY];
                    acc[u][n1] = cadd(acc[u][n1], cmul(b, mk<C>(w.x, -w.y)));
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
        const int tau = threadIdx.x + u * YZ_THREADS;
        if (tau < ntask) {
            const int row = tau / YZ_P2, n2 = tau - row * YZ_P2;
#pragma unroll
            for (int n1 = 0; n1 < NIN; ++n1) {
                const int y = n2 + YZ_P2 * n1;
                if (y < ny) gstore<C>(g, (unsigned)(row * ny + y), acc[u][n1]);
            }
        }
    }
}

std::atomic<int> g_fused_yz{FUSED_SCHEDULE};

}  // namespace

void set_fused_yz(int mode) {
    TDDA_REQUIRE(mode >= FUSED_OFF && mode <= FUSED_ALL, ERR_INVALID, "fused y/z mode must be 0, 1 or 2");
    g_fused_yz.store(mode);
}
int fused_yz() { return g_fused_yz.load(); }

int yz_class(int ly, int lz, int ny, int nz) {
    if (ly % YZ_P2 != 0 || 2 * ny > ly || 2 * nz > lz) return -1;
    const int ly1 = ly / YZ_P2;
    int row = -1, col = -1;
    switch (ly1) {
        case 2: row = 0; break;
        case 3: row = 1; break;
        case 4: row = 2; break;
        case 6: row = 3; break;
        case 8: row = 4; break;
        default: return -1;
    }
    switch (lz) {
        case 32: col = 0; break;
        case 48: col = 1; break;
        case 64: col = 2; break;
        default: return -1;
    }
    return 3 * row + col;
}

size_t yz_lds_bytes(int cls) {
    size_t b = 0;
    with_yz(cls, [&](auto a, auto z1, auto z2) {
        b = YzShape<decltype(a)::value, decltype(z1)::value, decltype(z2)::value>::lds_bytes();
    });
    return b;
}

void launch_fast_yz(const PassMap* yz, const GridJob* jobs, const TwiddlePools& tw, int gate, hipStream_t st) {
    for (int cls = 0; cls < NYZ; ++cls) {
        const PassMap& pm = yz[cls];
        if (pm.n == 0) continue;
        with_yz(cls, [&](auto a, auto z1, auto z2) {
            constexpr int LY1 = decltype(a)::value, LZ1 = decltype(z1)::value, LZ2 = decltype(z2)::value;
            constexpr size_t lds = YzShape<LY1, LZ1, LZ2>::lds_bytes();
            hipLaunchKernelGGL((k_yz_fast<LY1, LZ1, LZ2>), dim3(pm.n), dim3(YZ_THREADS), lds, st, jobs, pm.map,
                               tw.tw32, gate, pm.limit);
        });
    }
    TDDA_HIP(hipGetLastError());
}

}  // namespace tdda

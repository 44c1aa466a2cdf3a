// Plane passes of the c64 matvec (gfx950): X forward + Y forward as one
// kernel, Y inverse + X inverse as another, one workgroup per (job, z-plane).
//
// A workgroup takes the three components of its plane in turn through one
// LDS region (fft_plane.h).  Forward, each thread owns one x-step-A task
// (row y, n2) and reads the Krylov vectors of its x positions for all three
// components at once (24 contiguous bytes per dipole, with the BiCGSTAB
// vector operations of the mode applied), feeds component 0 to the transform
// and keeps components 1 and 2 in registers for the next two rounds.  The
// y-step-B output goes straight to the work grid [3][nz][ly][lx] that the Z
// pass reads.  Inverse, each thread owns one x-step-B task (row y, k1) and
// keeps the convolution of components 0 and 1 in registers until component
// 2 is done, then runs the epilogue of the mode (y = conv / M + al x and its
// dot products) on all three components of its dipoles.  The dot partials
// are per plane (tile = z), summed per job by k_xi_reduce.
//
// Against the separate passes this saves the Lx x Ny x Nz x 3 intermediate
// grid's write and read in each direction and two launches per matvec.
#include <atomic>

#include "fft_plane.h"
#include "pass_util.h"

namespace tdda {

static_assert(NXY == NXY_CLASSES, "plane class tables agree");

namespace {

// block_sum for the large blocks of the plane passes: the same butterflies
// and wave order, but only thread 0 adds the wave sums (the only one that
// writes the partials); every thread reading all NV x NT/64 sums would not
// fit in the registers of a 1024-thread block.
template <int NV, int NT>
__device__ inline void block_sum_first(double (&v)[NV]) {
    static_assert(NT % 64 == 0, "whole waves");
    constexpr int NW = NT / 64;
    __shared__ double red[NV][NW];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) red[k][threadIdx.x >> 6] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; ++k) {
            double s = red[k][0];
            for (int w = 1; w < NW; ++w) s += red[k][w];
            v[k] = s;
        }
    }
}

// X forward input of the three components of one dipole (vector index
// i0 = 3 dip, < 0 outside the geometry): the operations of xf_element (same
// helpers, same order, so the same values), with every load of the dipole
// issued before its stores.  Element by element, each element's loads wait
// behind the previous element's stores, which the compiler must assume may
// alias them: with one or two workgroups per CU that serialises the memory
// round trips of the X step A.
template <int MODE>
__device__ inline void xf_dipole(const GridJob& J, int i0, const XfScalars& sc, float2 (&u)[3]) {
    using C = float2;
    if (i0 < 0) {
        static_for<3>([&](auto ci) { u[decltype(ci)::value] = mk<C>(0.0, 0.0); });
        return;
    }
    const unsigned i = (unsigned)i0;
    if constexpr (MODE == XF_PLAIN) {
        static_for<3>([&](auto ci) { u[decltype(ci)::value] = gload<C>(J.xin, i + decltype(ci)::value); });
        return;
    } else {
        const bool jac = J.jacobi != 0;
        const C ia = jac ? gload<C>(J.ial32, i / 3u) : mk<C>(1.0, 0.0);
        if constexpr (MODE == XF_P) {
            C pv[3], vv[3], a[3], b[3], dp[3];
            static_for<3>([&](auto ci) {
                constexpr int c = decltype(ci)::value;
                pv[c] = gload<C>(J.p, i + c);
                vv[c] = gload<C>(J.v, i + c);
            });
            if (sc.pending) {
                static_for<3>([&](auto ci) {
                    constexpr int c = decltype(ci)::value;
                    a[c] = gload<C>(J.t, i + c);
                    b[c] = gload<C>(J.s, i + c);
                    dp[c] = gload<C>(J.dP, i + c);
                });
                static_for<3>([&](auto ci) {
                    constexpr int c = decltype(ci)::value;
                    const double2 p_ = to_d(pv[c]), td = to_d(a[c]), sd = to_d(b[c]);
                    const double2 step = cadd(dcmul(sc.alpha, p_), dcmul(sc.zeta, td));
                    dp[c] = from_d<C>(cadd(to_d(dp[c]), step));
                    a[c] = from_d<C>(csub(td, dcmul(sc.zeta, sd)));   // r
                });
                static_for<3>([&](auto ci) {
                    constexpr int c = decltype(ci)::value;
                    gstore<C>(J.dP, i + c, dp[c]);
                    gstore<C>(J.r, i + c, a[c]);
                });
            } else {
                static_for<3>([&](auto ci) { a[decltype(ci)::value] = gload<C>(J.r, i + decltype(ci)::value); });
            }
            static_for<3>([&](auto ci) {
                constexpr int c = decltype(ci)::value;
                const double2 w = csub(to_d(pv[c]), dcmul(sc.zeta, to_d(vv[c])));   // p - zeta v
                pv[c] = from_d<C>(cadd(to_d(a[c]), dcmul(sc.beta, w)));
                gstore<C>(J.p, i + c, pv[c]);
                u[c] = jac ? cmul(pv[c], ia) : pv[c];
            });
        } else {   // XF_T
            C rv[3], vv[3];
            static_for<3>([&](auto ci) {
                constexpr int c = decltype(ci)::value;
                rv[c] = gload<C>(J.r, i + c);
                vv[c] = gload<C>(J.v, i + c);
            });
            static_for<3>([&](auto ci) {
                constexpr int c = decltype(ci)::value;
                const C tk = from_d<C>(csub(to_d(rv[c]), dcmul(sc.alpha, to_d(vv[c]))));
                gstore<C>(J.t, i + c, tk);
                u[c] = jac ? cmul(tk, ia) : tk;
            });
        }
    }
}

// X inverse epilogue of the three components of one dipole (conv already
// divided by M): xi_element's operations with the dipole's loads first.
template <int MODE>
__device__ inline void xi_dipole(const GridJob& J, unsigned i, const float2 (&conv)[3], double (&acc)[XI_NACC]) {
    using C = float2;
    const C a = (MODE != XI_PLAIN && J.jacobi) ? mk<C>(1.0, 0.0) : gload<C>(J.al32, i / 3u);
    if constexpr (MODE == XI_PLAIN) {
        C xv[3];
        static_for<3>([&](auto ci) { xv[decltype(ci)::value] = gload<C>(J.xin, i + decltype(ci)::value); });
        static_for<3>([&](auto ci) {
            constexpr int c = decltype(ci)::value;
            gstore<C>(J.yout, i + c, cadd(conv[c], cmul(a, xv[c])));
        });
    } else if constexpr (MODE == XI_V) {
        C pv[3], rz[3];
        static_for<3>([&](auto ci) { pv[decltype(ci)::value] = gload<C>(J.p, i + decltype(ci)::value); });
        r0_cell(J, i, rz);
        static_for<3>([&](auto ci) {
            constexpr int c = decltype(ci)::value;
            const C vk = cadd(conv[c], cmul(a, pv[c]));
            gstore<C>(J.v, i + c, vk);
            const double2 d = cmulc(to_d(vk), to_d(rz[c]));   // conj(r0) v
            acc[0] += d.x;
            acc[1] += d.y;
        });
    } else {   // XI_S
        C tv[3], rz[3];
        static_for<3>([&](auto ci) { tv[decltype(ci)::value] = gload<C>(J.t, i + decltype(ci)::value); });
        r0_cell(J, i, rz);
        static_for<3>([&](auto ci) {
            constexpr int c = decltype(ci)::value;
            const C sk = cadd(conv[c], cmul(a, tv[c]));
            gstore<C>(J.s, i + c, sk);
            const double2 sd = to_d(sk), td = to_d(tv[c]), r_ = to_d(rz[c]);
            const double2 d = cmulc(td, sd);   // conj(s) t
            const double2 e = cmulc(td, r_);   // conj(r0) t
            const double2 f = cmulc(sd, r_);   // conj(r0) s
            acc[0] += d.x;
            acc[1] += d.y;
            acc[2] += sd.x * sd.x + sd.y * sd.y;
            acc[3] += e.x;
            acc[4] += e.y;
            acc[5] += f.x;
            acc[6] += f.y;
            acc[7] += td.x * td.x + td.y * td.y;
        });
    }
}

template <int L, int A1, int A2, int B1, int B2, int MODE>
__global__ void __launch_bounds__((L / 2) * A2)
k_plane_fwd(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap, const float2* __restrict__ twpool,
            int gate) {
    using C = float2;
    using P = PlaneFwd<L, A1, A2, B1, B2, C>;
    constexpr int NT = P::TXA;
    constexpr int NIN = P::NIN;
    constexpr int PXB = (P::TXB + NT - 1) / NT;
    constexpr int PYA = (P::TYA + NT - 1) / NT;
    constexpr int PYB = (P::TYB + NT - 1) / NT;
    extern __shared__ __align__(16) unsigned char smem_pf[];
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    if (!gate_open(J, gate)) return;
    C* lds = reinterpret_cast<C*>(smem_pf);
    C* twl = lds + P::REGION;
    for (int i = threadIdx.x; i < L; i += NT) twl[i] = gload<C>(twpool + J.px.tw_off, (unsigned)i);
    const int z = jt.y;
    const int nx = J.nx, ny = J.ny;
    const int tau = threadIdx.x;   // this thread's x-step-A task
    const int y0 = P::xa_y(tau);
    const int q = z * ny + y0;
    const XfScalars sc = xf_scalars<MODE>(J);
    C hold[2][NIN];
    C* g = static_cast<C*>(J.grid);
    const unsigned S = (unsigned)J.nz * L * L;
    __syncthreads();
    static_for<3>([&](auto ci) {
        constexpr int c = decltype(ci)::value;
        // x step A: vectors (component 0 now, 1 and 2 kept) -> E
        if (y0 < ny) {
            C u[NIN];
            static_for<NIN>([&](auto i) {
                constexpr int n1 = decltype(i)::value;
                if constexpr (c == 0) {
                    const int x = P::xa_x(tau, n1);
                    C w3[3];
                    xf_dipole<MODE>(J, x < nx ? vec_index(J, q, x, 0) : -1, sc, w3);
                    u[n1] = w3[0];
                    hold[0][n1] = w3[1];
                    hold[1][n1] = w3[2];
                } else {
                    u[n1] = hold[c - 1][n1];
                }
            });
            P::xa_store(tau, lds, twl, u);
        }
        __syncthreads();
        // x step B: E -> plane (in place: read all, then write)
        {
            C sb[PXB][A2];
            static_for<PXB>([&](auto ki) {
                constexpr int k = decltype(ki)::value;
                const int t = tau + k * NT;
                if (t < P::TXB && P::xb_y(t) < ny) P::xb_read(t, lds, sb[k]);
            });
            __syncthreads();
            static_for<PXB>([&](auto ki) {
                constexpr int k = decltype(ki)::value;
                const int t = tau + k * NT;
                if (t < P::TXB && P::xb_y(t) < ny) P::xb_write(t, lds, sb[k]);
            });
        }
        __syncthreads();
        // y step A: plane -> E2 (in place)
        {
            C sa[PYA][B1 / 2];
            static_for<PYA>([&](auto ki) {
                constexpr int k = decltype(ki)::value;
                const int t = tau + k * NT;
                if (t < P::TYA) P::ya_read(t, lds, ny, sa[k]);
            });
            __syncthreads();
            static_for<PYA>([&](auto ki) {
                constexpr int k = decltype(ki)::value;
                const int t = tau + k * NT;
                if (t < P::TYA) P::ya_write(t, lds, twl, sa[k]);
            });
        }
        __syncthreads();
        // y step B: E2 -> work grid (lanes over kx: coalesced stores)
        static_for<PYB>([&](auto ki) {
            constexpr int k = decltype(ki)::value;
            const int t = tau + k * NT;
            if (t < P::TYB)
                P::yb(t, lds, [&](int ky, int kx, C v) {
                    gstore<C>(g, c * S + ((unsigned)z * L + (unsigned)ky) * L + (unsigned)kx, v);
                });
        });
        __syncthreads();   // the next component's step A overwrites E2
    });
}

template <int L, int A1, int A2, int B1, int B2, int MODE>
__global__ void __launch_bounds__((L / 2) * A1)
k_plane_inv(const GridJob* __restrict__ jobs, const int2* __restrict__ bmap, const float2* __restrict__ twpool,
            int gate) {
    using C = float2;
    using P = PlaneInv<L, A1, A2, B1, B2, C>;
    constexpr int NT = P::TXB;
    constexpr int NOUT = P::NOUT;
    constexpr int PYA = (P::TYA + NT - 1) / NT;
    constexpr int PYB = (P::TYB + NT - 1) / NT;
    constexpr int PXA = (P::TXA + NT - 1) / NT;
    extern __shared__ __align__(16) unsigned char smem_pi[];
    const int2 jt = bmap[blockIdx.x];
    const GridJob& J = jobs[jt.x];
    if (!gate_open(J, gate)) return;
    C* lds = reinterpret_cast<C*>(smem_pi);
    C* twl = lds + P::REGION;
    for (int i = threadIdx.x; i < L; i += NT) twl[i] = gload<C>(twpool + J.px.tw_off, (unsigned)i);
    const int z = jt.y;
    const int nx = J.nx, ny = J.ny;
    const int tau = threadIdx.x;   // this thread's x-step-B task
    const int y0 = P::xb_y(tau);
    const int q = z * ny + y0;
    double acc[XI_NACC] = {};
    C hold[2][NOUT];
    const C* g = static_cast<const C*>(J.grid);
    const unsigned S = (unsigned)J.nz * L * L;
    const float inv_m = (float)(1.0 / ((double)L * (double)L * (double)J.lz));
    __syncthreads();
    static_for<3>([&](auto ci) {
        constexpr int c = decltype(ci)::value;
        // y step A: work grid -> E2 (lanes over kx: coalesced loads)
        static_for<PYA>([&](auto ki) {
            constexpr int k = decltype(ki)::value;
            const int t = tau + k * NT;
            if (t < P::TYA)
                P::ya(t, lds, twl, [&](int ky, int kx) -> C {
                    return gload<C>(g, c * S + ((unsigned)z * L + (unsigned)ky) * L + (unsigned)kx);
                });
        });
        __syncthreads();
        // y step B: E2 -> plane rows y < Ny (in place)
        {
            C sb[PYB][B2];
            static_for<PYB>([&](auto ki) {
                constexpr int k = decltype(ki)::value;
                const int t = tau + k * NT;
                if (t < P::TYB) P::yb_read(t, lds, sb[k]);
            });
            __syncthreads();
            static_for<PYB>([&](auto ki) {
                constexpr int k = decltype(ki)::value;
                const int t = tau + k * NT;
                if (t < P::TYB) P::yb_write(t, lds, ny, sb[k]);
            });
        }
        __syncthreads();
        // x step A: plane -> E (in place)
        {
            C sa[PXA][A1];
            static_for<PXA>([&](auto ki) {
                constexpr int k = decltype(ki)::value;
                const int t = tau + k * NT;
                if (t < P::TXA && P::xa_y(t) < ny) P::xa_read(t, lds, sa[k]);
            });
            __syncthreads();
            static_for<PXA>([&](auto ki) {
                constexpr int k = decltype(ki)::value;
                const int t = tau + k * NT;
                if (t < P::TXA && P::xa_y(t) < ny) P::xa_write(t, lds, twl, sa[k]);
            });
        }
        __syncthreads();
        // x step B; the epilogue once all three components are in registers
        if (y0 < ny) {
            C o[NOUT];
            P::xb(tau, lds, o);
            if constexpr (c < 2) {
                static_for<NOUT>([&](auto i) { hold[c][decltype(i)::value] = o[decltype(i)::value]; });
            } else {
                // one dipole at a time: a scheduling barrier between them keeps
                // the loads of later dipoles from being hoisted (their registers
                // would otherwise spill in the 1024-thread blocks)
                static_for<NOUT>([&](auto i) {
                    constexpr int k2 = decltype(i)::value;
                    const int x = P::xb_x(tau, k2);
                    if (x < nx) {
                        const int v0 = vec_index(J, q, x, 0);
                        if (v0 >= 0) {
                            const C cv[3] = {cscale(hold[0][k2], inv_m), cscale(hold[1][k2], inv_m),
                                             cscale(o[k2], inv_m)};
                            xi_dipole<MODE>(J, (unsigned)v0, cv, acc);
                        }
                    }
                    __builtin_amdgcn_sched_barrier(0);
                });
            }
        }
        __syncthreads();   // the next component's step A overwrites E
    });
    if constexpr (MODE != XI_PLAIN) {
        block_sum_first<XI_NACC, NT>(acc);
        write_partials<XI_NACC>(J, z, acc);
    }
}

template <int L, int A1, int A2, int B1, int B2, int MODE>
void launch_fwd_one(const PassMap& pm, const GridJob* jobs, const float2* tw, int gate, hipStream_t st) {
    using P = PlaneFwd<L, A1, A2, B1, B2, float2>;
    constexpr size_t lds = (size_t)(P::REGION + L) * sizeof(float2);
    auto kern = k_plane_fwd<L, A1, A2, B1, B2, MODE>;
    static std::atomic<bool> attr{false};
    if (!attr.exchange(true))
        (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kern, dim3(pm.n), dim3(P::TXA), lds, st, jobs, pm.map, tw, gate);
}

template <int L, int A1, int A2, int B1, int B2, int MODE>
void launch_inv_one(const PassMap& pm, const GridJob* jobs, const float2* tw, int gate, hipStream_t st) {
    using P = PlaneInv<L, A1, A2, B1, B2, float2>;
    constexpr size_t lds = (size_t)(P::REGION + L) * sizeof(float2);
    auto kern = k_plane_inv<L, A1, A2, B1, B2, MODE>;
    static std::atomic<bool> attr{false};
    if (!attr.exchange(true))
        (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kern, dim3(pm.n), dim3(P::TXB), lds, st, jobs, pm.map, tw, gate);
}

std::atomic<int> g_plane_mode{1};

}  // namespace

size_t plane_lds_bytes(int cls, bool inverse) {
    size_t b = 0;
    with_plane(cls, [&](auto l, auto fa1, auto fa2, auto ia1, auto ia2, auto b1, auto b2) {
        constexpr int L = decltype(l)::value, B1 = decltype(b1)::value, B2 = decltype(b2)::value;
        if (inverse)
            b = (PlaneInv<L, decltype(ia1)::value, decltype(ia2)::value, B1, B2, float2>::REGION + L) * sizeof(float2);
        else
            b = (PlaneFwd<L, decltype(fa1)::value, decltype(fa2)::value, B1, B2, float2>::REGION + L) * sizeof(float2);
    });
    return b;
}

void launch_plane_fwd(const PassMap* xy, const GridJob* jobs, const TwiddlePools& tw, int mode, int gate,
                      hipStream_t st) {
    for (int cls = 0; cls < NXY; ++cls) {
        const PassMap& pm = xy[cls];
        if (pm.n == 0) continue;
        with_plane(cls, [&](auto l, auto fa1, auto fa2, auto, auto, auto b1, auto b2) {
            constexpr int L = decltype(l)::value, A1 = decltype(fa1)::value, A2 = decltype(fa2)::value;
            constexpr int B1 = decltype(b1)::value, B2 = decltype(b2)::value;
            switch (mode) {
                case XF_PLAIN: launch_fwd_one<L, A1, A2, B1, B2, XF_PLAIN>(pm, jobs, tw.tw32, gate, st); break;
                case XF_P: launch_fwd_one<L, A1, A2, B1, B2, XF_P>(pm, jobs, tw.tw32, gate, st); break;
                case XF_T: launch_fwd_one<L, A1, A2, B1, B2, XF_T>(pm, jobs, tw.tw32, gate, st); break;
                default: throw Error(ERR_INVALID, "plane pass: X-forward mode not supported in c64");
            }
        });
    }
}

void launch_plane_inv(const PassMap* xy, const GridJob* jobs, const TwiddlePools& tw, int mode, int gate,
                      hipStream_t st) {
    for (int cls = 0; cls < NXY; ++cls) {
        const PassMap& pm = xy[cls];
        if (pm.n == 0) continue;
        with_plane(cls, [&](auto l, auto, auto, auto ia1, auto ia2, auto b1, auto b2) {
            constexpr int L = decltype(l)::value, A1 = decltype(ia1)::value, A2 = decltype(ia2)::value;
            constexpr int B1 = decltype(b1)::value, B2 = decltype(b2)::value;
            switch (mode) {
                case XI_PLAIN: launch_inv_one<L, A1, A2, B1, B2, XI_PLAIN>(pm, jobs, tw.tw32, gate, st); break;
                case XI_V: launch_inv_one<L, A1, A2, B1, B2, XI_V>(pm, jobs, tw.tw32, gate, st); break;
                case XI_S: launch_inv_one<L, A1, A2, B1, B2, XI_S>(pm, jobs, tw.tw32, gate, st); break;
                default: throw Error(ERR_INVALID, "plane pass: X-inverse mode not supported in c64");
            }
        });
    }
}

int plane_class_of(int lx, int ly, int nx, int ny) { return plane_class(lx, ly, nx, ny); }

void set_plane_passes(bool on) { g_plane_mode = on ? 1 : 0; }
bool plane_passes() { return g_plane_mode.load() != 0; }

}  // namespace tdda

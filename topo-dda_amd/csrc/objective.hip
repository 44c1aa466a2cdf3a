// Objectives of a solved problem and their adjoint gradient w.r.t. the voxel
// densities, all on the device.
//
// IntegratedE (ObjDDAModel.cpp:130-212): f = log sum_i sum_c |al_i P_ic|^p.
// Target field: f = (1/T) sum_t |E(r_t)|^2 with the field at external points
//   E(r_t) = E_inc(r_t) - sum_j A(r_t - r_j) P_j    (DDAModel::solve_E,
//   src/DDAModel.cpp:1153-1183, whose RResult branch the reference never enables;
//   the incident field here is E0 e exp(i K k.r) for every component).
// Adjoint (DDAModel.cpp:398-469 without finite differences):
//   devp = 2 df/dP (Wirtinger), lambda = S^-1 devp with S = diag(al) + A
//   (complex symmetric), df/drho_i = explicit_i - Re sum_c lambda_ic dal_i P_ic,
//   dal_i = d(1/alpha)/d eps * d eps/d rho.
#include "green_eval.h"
#include "objective.h"

namespace tdda {

namespace {

struct ObjParams {
    int nx, ny, nz, n;
    double d, K;
    double nk[3], ne0[3], e0;
    double2 m0, m1;
    double S;
    int power;
    int method;
    long nsici;
    double delta;
};

__device__ inline double2 eps_of(const ObjParams& P, double x) {
    x = fmin(fmax(x, 0.0), 1.0);
    if (x >= 1.0) return P.m1;
    return make_double2(P.m0.x + x * (P.m1.x - P.m0.x), P.m0.y + x * (P.m1.y - P.m0.y));
}

// d(1/alpha)/d rho = [-(4 pi/d^3)/(eps-1)^2 + K^2 (b2 + b3 S)/d] (m1 - m0)
// inside [0, 1]; the density is clipped to [0, 1] (eps_of), so the objective
// is flat outside and the derivative there is zero (at the ends 0 and 1 the
// one-sided derivative from inside is returned)
__device__ inline double2 dal_drho(const ObjParams& P, double x) {
    if (x < 0.0 || x > 1.0) return make_double2(0.0, 0.0);
    const double b2 = 0.1648469, b3 = -1.7700004;
    const double2 e = eps_of(P, x);
    const double2 em1 = make_double2(e.x - 1.0, e.y);
    const double2 sq = cmul(em1, em1);
    const double c = 3.0 * P.d * P.d * P.d / (4.0 * M_PI);
    const double2 t1 = ddiv(make_double2(-3.0 / c, 0.0), sq);
    const double2 de = make_double2(t1.x + P.K * P.K * (b2 + b3 * P.S) / P.d, t1.y);
    return cmul(de, make_double2(P.m1.x - P.m0.x, P.m1.y - P.m0.y));
}

template <int NV>
__device__ inline void block_sum_d(double (&v)[NV]) {
    __shared__ double red[NV][256];
    for (int k = 0; k < NV; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < NV; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    for (int k = 0; k < NV; ++k) v[k] = red[k][0];
}

__device__ inline double pow_int(double x, int p) {
    double r = 1.0;
    for (int k = 0; k < p; ++k) r *= x;
    return r;
}

// partial[b] = sum over this block's dipoles of sum_c |al P|^p (fixed order)
__global__ void __launch_bounds__(256) k_obj_integrated(ObjParams P, const double2* __restrict__ al,
                                                        const double2* __restrict__ Pv, double* __restrict__ partial) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v[1] = {0.0};
    if (i < P.n) {
        for (int c = 0; c < 3; ++c) {
            const double2 w = cmul(al[i], Pv[3 * i + c]);
            v[0] += pow_int(sqrt(w.x * w.x + w.y * w.y), P.power);
        }
    }
    block_sum_d<1>(v);
    if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

// devp = p |w|^(p-2) conj(w) al / E,  w = al P
__global__ void k_devp_integrated(ObjParams P, double eint, const double2* __restrict__ al,
                                  const double2* __restrict__ Pv, double2* __restrict__ devp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    for (int c = 0; c < 3; ++c) {
        const double2 w = cmul(al[i], Pv[3 * i + c]);
        const double aw = sqrt(w.x * w.x + w.y * w.y);
        const double s = P.power == 2 ? 2.0 : (aw > 0 ? P.power * pow_int(aw, P.power - 2) : 0.0);
        devp[3 * i + c] = cscale(cmul(cconj(w), al[i]), s / eint);
    }
}

// grad_i = explicit_i - Re sum_c lambda_ic dal_i P_ic
__global__ void k_grad_assemble(ObjParams P, int integrated, double eint, const double* __restrict__ rho,
                                const double2* __restrict__ al, const double2* __restrict__ Pv,
                                const double2* __restrict__ lam, double* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double2 dal = dal_drho(P, rho[i]);
    double g = 0.0;
    for (int c = 0; c < 3; ++c) {
        const double2 p = Pv[3 * i + c];
        const double2 dp = cmul(dal, p);
        g -= cmul(lam[3 * i + c], dp).x;
        if (integrated) {
            const double2 w = cmul(al[i], p);
            const double aw = sqrt(w.x * w.x + w.y * w.y);
            const double s = P.power == 2 ? 2.0 : (aw > 0 ? P.power * pow_int(aw, P.power - 2) : 0.0);
            g += s * cmul(cconj(w), dp).x / eint;
        }
    }
    grad[i] = g;
}

__device__ inline void pair_tensor(const ObjParams& P, const double* si, const double* ci, double rx, double ry,
                                   double rz, Sym6& t) {
    if (P.method == 0)
        fcd_tensor(P.K, P.d, si, ci, P.nsici, P.delta, rx, ry, rz, t);
    else
        ldr_tensor(P.K, rx, ry, rz, t);
}

// one block per target: E_t = E_inc(r_t) - sum_j A(r_t - r_j) P_j
__global__ void __launch_bounds__(256) k_target_field(ObjParams P, const double* __restrict__ tgt,
                                                      const double2* __restrict__ Pv, const double* si,
                                                      const double* ci, double2* __restrict__ Et) {
    const int t = blockIdx.x;
    const double x = tgt[3 * t], y = tgt[3 * t + 1], z = tgt[3 * t + 2];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int j = threadIdx.x; j < P.n; j += 256) {
        const int gx = j % P.nx, gy = (j / P.nx) % P.ny, gz = j / (P.nx * P.ny);
        Sym6 A;
        pair_tensor(P, si, ci, x - P.d * gx, y - P.d * gy, z - P.d * gz, A);
        const double2 p0 = Pv[3 * j], p1 = Pv[3 * j + 1], p2 = Pv[3 * j + 2];
        const double2 a00 = make_double2(A.re[0], A.im[0]), a01 = make_double2(A.re[1], A.im[1]);
        const double2 a02 = make_double2(A.re[2], A.im[2]), a11 = make_double2(A.re[3], A.im[3]);
        const double2 a12 = make_double2(A.re[4], A.im[4]), a22 = make_double2(A.re[5], A.im[5]);
        const double2 s0 = cadd(cadd(cmul(a00, p0), cmul(a01, p1)), cmul(a02, p2));
        const double2 s1 = cadd(cadd(cmul(a01, p0), cmul(a11, p1)), cmul(a12, p2));
        const double2 s2 = cadd(cadd(cmul(a02, p0), cmul(a12, p1)), cmul(a22, p2));
        acc[0] += s0.x; acc[1] += s0.y; acc[2] += s1.x; acc[3] += s1.y; acc[4] += s2.x; acc[5] += s2.y;
    }
    block_sum_d<6>(acc);
    if (threadIdx.x == 0) {
        double s, c;
        sincos(P.K * (P.nk[0] * x + P.nk[1] * y + P.nk[2] * z), &s, &c);
        for (int k = 0; k < 3; ++k)
            Et[3 * t + k] = make_double2(P.e0 * P.ne0[k] * c - acc[2 * k], P.e0 * P.ne0[k] * s - acc[2 * k + 1]);
    }
}

// devp_j = -(2/T) sum_t A(r_t - r_j) conj(E_t)
__global__ void k_devp_targets(ObjParams P, int T, const double* __restrict__ tgt, const double2* __restrict__ Et,
                               const double* si, const double* ci, double2* __restrict__ devp) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.n) return;
    const int gx = j % P.nx, gy = (j / P.nx) % P.ny, gz = j / (P.nx * P.ny);
    double2 acc[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    for (int t = 0; t < T; ++t) {
        Sym6 A;
        pair_tensor(P, si, ci, tgt[3 * t] - P.d * gx, tgt[3 * t + 1] - P.d * gy, tgt[3 * t + 2] - P.d * gz, A);
        const double2 e0 = cconj(Et[3 * t]), e1 = cconj(Et[3 * t + 1]), e2 = cconj(Et[3 * t + 2]);
        const double2 a00 = make_double2(A.re[0], A.im[0]), a01 = make_double2(A.re[1], A.im[1]);
        const double2 a02 = make_double2(A.re[2], A.im[2]), a11 = make_double2(A.re[3], A.im[3]);
        const double2 a12 = make_double2(A.re[4], A.im[4]), a22 = make_double2(A.re[5], A.im[5]);
        acc[0] = cadd(acc[0], cadd(cadd(cmul(a00, e0), cmul(a01, e1)), cmul(a02, e2)));
        acc[1] = cadd(acc[1], cadd(cadd(cmul(a01, e0), cmul(a11, e1)), cmul(a12, e2)));
        acc[2] = cadd(acc[2], cadd(cadd(cmul(a02, e0), cmul(a12, e1)), cmul(a22, e2)));
    }
    for (int k = 0; k < 3; ++k) devp[3 * j + k] = cscale(acc[k], -2.0 / T);
}

ObjParams params_of(const Problem& p, int power) {
    ObjParams P{};
    P.nx = p.desc.nx;
    P.ny = p.desc.ny;
    P.nz = p.desc.nz;
    P.n = p.ndip;
    P.d = p.desc.d;
    P.K = p.K;
    for (int k = 0; k < 3; ++k) { P.nk[k] = p.desc.nk[k]; P.ne0[k] = p.desc.ne0[k]; }
    P.e0 = p.desc.e0;
    P.m0 = p.m0;
    P.m1 = p.m1;
    for (int k = 0; k < 3; ++k) P.S += (p.desc.ne0[k] * p.desc.nk[k]) * (p.desc.ne0[k] * p.desc.nk[k]);
    P.power = power;
    P.method = p.desc.method;
    P.nsici = p.tables.n;
    P.delta = p.tables.delta;
    return P;
}

}  // namespace

static double integrated_sum(const Problem& p, int power, hipStream_t st) {
    const int nb = (p.ndip + 255) / 256;
    DevBuf<double> part(nb);
    hipLaunchKernelGGL(k_obj_integrated, dim3(nb), dim3(256), 0, st, params_of(p, power), p.sys->al64.get(),
                       p.sys->P.get(), part.get());
    TDDA_HIP(hipGetLastError());
    std::vector<double> h(nb);
    part.download(h.data(), nb, st);
    TDDA_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (double v : h) s += v;
    return s;
}

static void target_field(const Problem& p, const DevBuf<double>& tg, int T, DevBuf<double2>& Et, hipStream_t st) {
    Et.alloc(3 * (size_t)T);
    hipLaunchKernelGGL(k_target_field, dim3(T), dim3(256), 0, st, params_of(p, 2), tg.get(), p.sys->P.get(),
                       p.tables.si, p.tables.ci, Et.get());
    TDDA_HIP(hipGetLastError());
}

double objective_value(Problem& p, const ObjectiveSpec& o, hipStream_t st, std::vector<double2>* fields) {
    if (o.kind == OBJ_INTEGRATED) {
        TDDA_REQUIRE(o.power >= 1, ERR_INVALID, "objective power must be >= 1");
        return std::log(integrated_sum(p, o.power, st));
    }
    TDDA_REQUIRE(o.kind == OBJ_TARGETS, ERR_INVALID, "unknown objective kind");
    const int T = (int)(o.targets.size() / 3);
    TDDA_REQUIRE(T >= 1 && o.targets.size() == 3 * (size_t)T, ERR_INVALID, "targets must be 3*T coordinates");
    DevBuf<double> tg(3 * (size_t)T);
    tg.upload(o.targets.data(), tg.size(), st);
    DevBuf<double2> Et;
    target_field(p, tg, T, Et, st);
    std::vector<double2> h(3 * (size_t)T);
    Et.download(h.data(), h.size(), st);
    TDDA_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (const double2& e : h) s += e.x * e.x + e.y * e.y;
    if (fields) *fields = h;
    return s / T;
}

SolveResult objective_gradient(Problem& p, const ObjectiveSpec& o, int maxit, double tol, double* grad, bool on_device,
                               hipStream_t st, double* value) {
    System& s = *p.sys;
    const size_t n3 = 3ull * p.ndip;
    // keep the forward solution and right-hand side
    DevBuf<double2> Pf(n3), rhs0(n3);
    TDDA_HIP(hipMemcpyAsync(Pf.get(), s.P.get(), n3 * sizeof(double2), hipMemcpyDeviceToDevice, st));
    TDDA_HIP(hipMemcpyAsync(rhs0.get(), s.rhs.get(), n3 * sizeof(double2), hipMemcpyDeviceToDevice, st));
    const int nb = (p.ndip + 255) / 256;
    double eint = 0.0, val = 0.0;
    if (o.kind == OBJ_INTEGRATED) {
        eint = integrated_sum(p, o.power, st);
        val = std::log(eint);
        hipLaunchKernelGGL(k_devp_integrated, dim3(nb), dim3(256), 0, st, params_of(p, o.power), eint,
                           s.al64.get(), s.P.get(), s.rhs.get());
    } else {
        const int T = (int)(o.targets.size() / 3);
        TDDA_REQUIRE(T >= 1 && o.targets.size() == 3 * (size_t)T, ERR_INVALID, "targets must be 3*T coordinates");
        DevBuf<double> tg(3 * (size_t)T);
        tg.upload(o.targets.data(), tg.size(), st);
        DevBuf<double2> Et;
        target_field(p, tg, T, Et, st);
        std::vector<double2> h(3 * (size_t)T);
        Et.download(h.data(), h.size(), st);
        TDDA_HIP(hipStreamSynchronize(st));
        for (const double2& e : h) val += e.x * e.x + e.y * e.y;
        val /= T;
        hipLaunchKernelGGL(k_devp_targets, dim3(nb), dim3(256), 0, st, params_of(p, 2), T, tg.get(), Et.get(),
                           p.tables.si, p.tables.ci, s.rhs.get());
        TDDA_HIP(hipStreamSynchronize(st));
    }
    TDDA_HIP(hipGetLastError());
    s.P.zero(st);
    SolveOptions so;
    so.refined = s.refined;
    so.maxit = maxit;
    so.tol = tol;
    so.err_tol = s.refined ? tol : 0.0;
    std::vector<System*> v{&s};
    SolveResult r = solve_batch(v, so, st)[0];
    DevBuf<double> g;
    double* gd = grad;
    if (!on_device) {
        g.alloc(p.ndip);
        gd = g.get();
    }
    hipLaunchKernelGGL(k_grad_assemble, dim3(nb), dim3(256), 0, st, params_of(p, o.power),
                       o.kind == OBJ_INTEGRATED ? 1 : 0, eint, p.rho.get(), s.al64.get(), Pf.get(), s.P.get(), gd);
    TDDA_HIP(hipGetLastError());
    if (!on_device) g.download(grad, p.ndip, st);
    // restore the forward state
    TDDA_HIP(hipMemcpyAsync(s.P.get(), Pf.get(), n3 * sizeof(double2), hipMemcpyDeviceToDevice, st));
    TDDA_HIP(hipMemcpyAsync(s.rhs.get(), rhs0.get(), n3 * sizeof(double2), hipMemcpyDeviceToDevice, st));
    TDDA_HIP(hipStreamSynchronize(st));
    if (value) *value = val;
    return r;
}

}  // namespace tdda

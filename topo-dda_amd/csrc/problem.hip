// Device setup of a full-box problem: density -> eps -> 1/alpha, incident field.
#include "problem.h"

#include <cmath>
#include <cstring>

namespace tdda {

namespace {

struct SetupParams {
    int nx, ny, nz, n;
    double d, K;
    double nk[3], ne0[3], e0;
    double2 m0, m1;
    double S;        // sum_j (e_j k_j)^2
    double kd;       // K d
    double en[3];    // e0 n_e0
};

__device__ inline double2 cdiv_d(double2 a, double2 b) { return ddiv(a, b); }

// DDAModel.cpp:319-326 with two materials: eps = m0 + x (m1 - m0) for x < 1, m1 at x >= 1
__device__ inline double2 eps_of(const SetupParams& P, double x) {
    x = fmin(fmax(x, 0.0), 1.0);
    if (x >= 1.0) return P.m1;
    return make_double2(P.m0.x + x * (P.m1.x - P.m0.x), P.m0.y + x * (P.m1.y - P.m0.y));
}

// 1/alpha, Tools.cpp:353-366
__device__ inline double2 inv_alpha(const SetupParams& P, double2 eps) {
    const double b1 = -1.891531, b2 = 0.1648469, b3 = -1.7700004;
    const double d3 = P.d * P.d * P.d;
    const double kd = P.K * P.d;
    const double c = 3.0 * d3 / (4.0 * M_PI);
    const double2 a1 = cscale(cdiv_d(make_double2(eps.x - 1.0, eps.y), make_double2(eps.x + 2.0, eps.y)), c);
    const double2 bb = make_double2(b1 + eps.x * (b2 + b3 * P.S), eps.y * (b2 + b3 * P.S));
    const double2 inner = make_double2(bb.x * kd * kd, bb.y * kd * kd - (2.0 / 3.0) * kd * kd * kd);
    const double2 den = cadd(make_double2(1.0, 0.0), cmul(cscale(a1, 1.0 / d3), inner));
    const double2 alpha = cdiv_d(a1, den);
    return cdiv_d(make_double2(1.0, 0.0), alpha);
}

// per-dipole set-up of dipole i (density value r)
__device__ inline void setup_dipole(const SetupParams& P, int i, double r, double2* __restrict__ al64,
                                    float2* __restrict__ al32, double2* __restrict__ ial64,
                                    float2* __restrict__ ial32, double2* __restrict__ rhs,
                                    float2* __restrict__ efac) {
    const double2 eps = eps_of(P, r);
    const double2 al = inv_alpha(P, eps);
    al64[i] = al;
    al32[i] = make_float2((float)al.x, (float)al.y);
    const double2 ia = ddiv(make_double2(1.0, 0.0), al);   // alpha (Jacobi preconditioning)
    ial64[i] = ia;
    ial32[i] = make_float2((float)ia.x, (float)ia.y);
    if (efac != nullptr) {   // E = f P, f = al 3/(eps + 2) (DDAModel.cpp:1185-1192)
        const double2 f = ddiv(al, make_double2((eps.x + 2.0) / 3.0, eps.y / 3.0));
        efac[i] = make_float2((float)f.x, (float)f.y);
    }
    const int gx = i % P.nx, gy = (i / P.nx) % P.ny, gz = i / (P.nx * P.ny);
    const double2 ph = incident_phase(P.kd, P.nk, gx, gy, gz);
    for (int k = 0; k < 3; ++k) rhs[3 * i + k] = incident_wave(ph, P.en[k]);
}

__global__ void k_problem_setup(SetupParams P, const double* __restrict__ rho, double2* __restrict__ al64,
                                float2* __restrict__ al32, double2* __restrict__ ial64, float2* __restrict__ ial32,
                                double2* __restrict__ rhs, float2* __restrict__ efac) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    setup_dipole(P, i, rho[i], al64, al32, ial64, ial32, rhs, efac);
}

// one problem of a round set-up; its blocks are [block0, block0 + ceil(n / 256))
struct RoundJob {
    SetupParams P;
    long rho_off;   // byte offset of its densities in the round buffer
    int block0;
    double* rho;
    double2 *al64, *ial64, *rhs, *sol;
    float2 *al32, *ial32, *efac;
};

constexpr int ROUND_THREADS = 256;

__global__ void __launch_bounds__(ROUND_THREADS)
k_problem_setup_round(const RoundJob* __restrict__ jobs, int njobs, const char* __restrict__ base) {
    const int b = blockIdx.x;
    int lo = 0, hi = njobs - 1;   // last job with block0 <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (jobs[mid].block0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const RoundJob& J = jobs[lo];
    const int i = (b - J.block0) * ROUND_THREADS + threadIdx.x;
    if (i >= J.P.n) return;
    const double r = reinterpret_cast<const double*>(base + J.rho_off)[i];
    J.rho[i] = r;
    setup_dipole(J.P, i, r, J.al64, J.al32, J.ial64, J.ial32, J.rhs, J.efac);
    for (int k = 0; k < 3; ++k) J.sol[3 * i + k] = make_double2(0.0, 0.0);
}

// ial64 non-null: Pv holds Y = al P (a preconditioned solve), P = Y alpha
__global__ void k_problem_field(SetupParams P, const double* __restrict__ rho, const double2* __restrict__ al64,
                                const double2* __restrict__ ial64, const double2* __restrict__ Pv,
                                double2* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double2 eps = eps_of(P, rho[i]);
    const double2 lf = make_double2((eps.x + 2.0) / 3.0, eps.y / 3.0);
    const double2 f = ddiv(al64[i], lf);
    for (int k = 0; k < 3; ++k) {
        const double2 p = ial64 != nullptr ? dcmul(Pv[3 * i + k], ial64[i]) : Pv[3 * i + k];
        out[3 * i + k] = cmul(f, p);
    }
}

SetupParams params_of(const Problem& p) {
    SetupParams P{};
    P.nx = p.desc.nx;
    P.ny = p.desc.ny;
    P.nz = p.desc.nz;
    P.n = p.ndip;
    P.d = p.desc.d;
    P.K = p.K;
    for (int k = 0; k < 3; ++k) {
        P.nk[k] = p.desc.nk[k];
        P.ne0[k] = p.desc.ne0[k];
    }
    P.e0 = p.desc.e0;
    P.m0 = p.m0;
    P.m1 = p.m1;
    P.S = 0.0;
    for (int k = 0; k < 3; ++k) P.S += (p.desc.ne0[k] * p.desc.nk[k]) * (p.desc.ne0[k] * p.desc.nk[k]);
    P.kd = p.K * p.desc.d;
    for (int k = 0; k < 3; ++k) P.en[k] = p.desc.e0 * p.desc.ne0[k];
    return P;
}

}  // namespace

void problem_materials(const ProblemDesc& d, double2* m0, double2* m1) {
    double2 m[2];
    for (int k = 0; k < 2; ++k) {
        std::complex<double> v(d.mat[2 * k], d.mat[2 * k + 1]);
        v /= d.nback * d.nback;
        if (std::abs(v) < 1.01) v += 0.01;
        m[k] = make_double2(v.real(), v.imag());
    }
    *m0 = m[0];
    *m1 = m[1];
}

OperatorKey problem_operator_key(const ProblemDesc& d, const SiciDevice& tables) {
    const double K = 2.0 * M_PI / (d.lam / d.nback);
    return OperatorKey{d.nx, d.ny, d.nz, d.d, K, d.method, d.maxm, d.maxn, d.lm * d.d, d.ln * d.d,
                       d.method == 0 ? tables.id : 0};
}

Problem::Problem(const ProblemDesc& d, const SiciDevice& tables, bool refined, hipStream_t st,
                 std::shared_ptr<Operator> op_, std::unique_ptr<System> reuse, GreenBatch* batch)
    : desc(d), op(std::move(op_)), tables(tables) {
    TDDA_REQUIRE(d.nx >= 1 && d.ny >= 1 && d.nz >= 1, ERR_INVALID, "empty box");
    TDDA_REQUIRE(d.method == 0 || d.method == 1, ERR_METHOD, "method must be FCD (0) or LDR (1)");
    TDDA_REQUIRE(d.lam > 0 && d.nback > 0 && d.d > 0, ERR_INVALID, "lambda, nback and d must be positive");
    K = 2.0 * M_PI / (d.lam / d.nback);
    problem_materials(d, &m0, &m1);
    ndip = d.nx * d.ny * d.nz;
    if (!op) op = OperatorCache::instance().get(problem_operator_key(d, tables), tables, st, &op_hit, batch);
    else op_hit = true;
    if (reuse && reuse->job.lx == op->lx && reuse->job.ly == op->ly && reuse->job.lz == op->lz &&
        reuse->ndip == ndip && reuse->refined == refined && reuse->job.box_map == nullptr &&
        reuse->job.nx == op->nx && reuse->job.ny == op->ny && reuse->job.nz == op->nz) {
        sys = std::move(reuse);
        sys->op = op;
        sys->job.green32 = op->green32.get();
        sys->job.green64 = op->green64.get();
    } else {
        sys.reset(new System(op, ndip, nullptr, refined, st));
    }
    rho.alloc(ndip);
}

void Problem::set_density(const double* r, bool on_device, hipStream_t st) {
    if (on_device)
        TDDA_HIP(hipMemcpyAsync(rho.get(), r, ndip * sizeof(double), hipMemcpyDeviceToDevice, st));
    else
        rho.upload(r, ndip, st);
    SetupParams P = params_of(*this);
    hipLaunchKernelGGL(k_problem_setup, dim3((ndip + 255) / 256), dim3(256), 0, st, P, rho.get(), sys->al64.get(),
                       sys->al32.get(), sys->ial64.get(), sys->ial32.get(), sys->rhs.get(), sys->efac.get());
    TDDA_HIP(hipGetLastError());
    sys->job.efac = sys->efac.get();
}

void RoundSetup::run(const std::vector<Problem*>& probs, const std::vector<const double*>& rho, hipStream_t st) {
    TDDA_REQUIRE(probs.size() == rho.size(), ERR_INVALID, "one density per problem");
    if (probs.empty()) return;
    const int nj = (int)probs.size();
    auto al256 = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t bytes = al256(nj * sizeof(RoundJob));
    for (const Problem* p : probs) bytes += al256(p->ndip * sizeof(double));
    char* h = stage_.reserve(bytes);
    if (dev_.size() < bytes) dev_.alloc(std::max(bytes, 2 * dev_.size()));
    RoundJob* tab = reinterpret_cast<RoundJob*>(h);
    size_t off = al256(nj * sizeof(RoundJob));
    int blocks = 0;
    for (int j = 0; j < nj; ++j) {
        Problem& p = *probs[j];
        std::memcpy(h + off, rho[j], p.ndip * sizeof(double));
        RoundJob& J = tab[j];
        J.P = params_of(p);
        J.rho_off = (long)off;
        J.block0 = blocks;
        J.rho = p.rho.get();
        J.al64 = p.sys->al64.get();
        J.al32 = p.sys->al32.get();
        J.ial64 = p.sys->ial64.get();
        J.ial32 = p.sys->ial32.get();
        J.rhs = p.sys->rhs.get();
        J.efac = p.sys->efac.get();
        J.sol = p.sys->P.get();
        p.sys->job.efac = p.sys->efac.get();
        blocks += (p.ndip + ROUND_THREADS - 1) / ROUND_THREADS;
        off += al256(p.ndip * sizeof(double));
    }
    TDDA_HIP(hipMemcpyAsync(dev_.get(), h, bytes, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_problem_setup_round, dim3(blocks), dim3(ROUND_THREADS), 0, st,
                       reinterpret_cast<const RoundJob*>(dev_.get()), nj, dev_.get());
    TDDA_HIP(hipGetLastError());
}

void Problem::use_incident_r0() {
    TDDA_REQUIRE(sys->job.box_map == nullptr, ERR_INVALID, "the incident shadow residual needs the full box");
    const SetupParams P = params_of(*this);
    GridJob& J = sys->job;
    J.r0_inc = 1;
    J.inc_kd = P.kd;
    for (int k = 0; k < 3; ++k) {
        J.inc_nk[k] = P.nk[k];
        J.inc_en[k] = P.en[k];
    }
}

void Problem::field(double2* out, hipStream_t st, bool from_y) const {
    SetupParams P = params_of(*this);
    hipLaunchKernelGGL(k_problem_field, dim3((ndip + 255) / 256), dim3(256), 0, st, P, rho.get(), sys->al64.get(),
                       from_y ? sys->ial64.get() : nullptr, sys->P.get(), out);
    TDDA_HIP(hipGetLastError());
}

size_t Problem::estimate_bytes(const ProblemDesc& d, bool refined, bool need_operator) {
    const int lx = choose_fft_length(d.nx), ly = choose_fft_length(d.ny), lz = choose_fft_length(d.nz);
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    const size_t kb = refined ? sizeof(float2) : sizeof(double2);
    size_t b = (size_t)3 * d.nz * ly * lx * sizeof(double2);   // work grid
    b += 6 * 3 * n * kb + 3 * 3 * n * sizeof(double2) + n * (2 * sizeof(double2) + 3 * sizeof(float2) + sizeof(double));
    b += ((size_t)d.ny * d.nz + 3 * n / 1024 + 8) * 6 * sizeof(double);
    if (need_operator) {
        const size_t h = (size_t)(lx / 2 + 1) * (ly / 2 + 1) * (lz / 2 + 1);
        b += 6 * h * (sizeof(double2) + sizeof(float2));
        b += 6 * n * sizeof(double2) * 3;   // setup temporaries (upper bound)
    }
    return b;
}

}  // namespace tdda

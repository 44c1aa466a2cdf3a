// DDAModel semantics over the gfx950 runtime.
#include "model.h"

#include <cmath>
#include <cstdio>
#include <map>
#include <tuple>

namespace tdda {

double c_round(double x) { return std::round(x); }   // half away from zero, as C's round()

// Tools.cpp:353-366: alpha = a1 / (1 + (a1/d^3)((b1 + eps b2 + eps b3 S)(Kd)^2 - (2/3) i (Kd)^3)),
// a1 = (3 d^3 / 4 pi)(eps - 1)/(eps + 2).  Returns 1/alpha.
cd ldr_inverse_alpha(double K, double d, cd eps, const double* ne0, const double* nk) {
    const double b1 = -1.891531, b2 = 0.1648469, b3 = -1.7700004;
    const cd a1 = (3.0 * std::pow(d, 3) / (4.0 * M_PI)) * (eps - 1.0) / (eps + 2.0);
    cd rad = cd(0.0, (2.0 / 3.0) * std::pow(K * d, 3));
    const double S = std::pow(ne0[0] * nk[0], 2) + std::pow(ne0[1] * nk[1], 2) + std::pow(ne0[2] * nk[2], 2);
    const cd den = 1.0 + (a1 / std::pow(d, 3)) * ((b1 + eps * b2 + eps * b3 * S) * std::pow(K * d, 2) - rad);
    const cd alpha = a1 / den;
    return 1.0 / alpha;
}

// 1/alpha = 1/a1 + K^2 (b1 + eps (b2 + b3 S))/d - (2/3) i K^3, hence
// d(1/alpha)/d eps = -(4 pi / d^3) / (eps - 1)^2 + K^2 (b2 + b3 S) / d.
cd ldr_inverse_alpha_deps(double K, double d, cd eps, const double* ne0, const double* nk) {
    const double b2 = 0.1648469, b3 = -1.7700004;
    const double S = std::pow(ne0[0] * nk[0], 2) + std::pow(ne0[1] * nk[1], 2) + std::pow(ne0[2] * nk[2], 2);
    const double c = 3.0 * d * d * d / (4.0 * M_PI);
    return -3.0 / (c * (eps - 1.0) * (eps - 1.0)) + K * K * (b2 + b3 * S) / d;
}

// ObjDDAModel.cpp:135 parses `xMin <= g <= xMax` as `(xMin <= g) <= xMax`.
bool objective_region(const int* g, const std::vector<double>& op) {
    const int lo[3] = {(int)c_round(op[1]), (int)c_round(op[3]), (int)c_round(op[5])};
    const int hi[3] = {(int)c_round(op[2]), (int)c_round(op[4]), (int)c_round(op[6])};
    for (int a = 0; a < 3; ++a)
        if (!((int)(lo[a] <= g[a]) <= hi[a])) return false;
    return true;
}

// ---------------------------------------------------------------------------
// construction, DDAModel.cpp:79-333
// ---------------------------------------------------------------------------
Model::Model(const ModelConfig& cfg) : cfg_(cfg) {
    N_ = cfg.n;
    int dividesym;
    if (cfg.symmetry == "None")
        dividesym = 1;
    else if (cfg.symmetry == "4fold")
        dividesym = 4;
    else
        throw Error(ERR_INVALID, "SpacePara: symmetry '" + cfg.symmetry + "' is neither None nor 4fold");
    TDDA_REQUIRE(N_ >= 1 && cfg.nx >= 1 && cfg.ny >= 1 && cfg.nz >= 1, ERR_INVALID, "empty geometry");
    TDDA_REQUIRE((long)cfg.geometry.size() == 3L * N_, ERR_INVALID, "geometry must hold 3*num_total ints");
    TDDA_REQUIRE(cfg.symmetry != "4fold" || cfg.symaxis.size() >= 2, ERR_INVALID, "4fold needs two axes");
    TDDA_REQUIRE(cfg.objpara.size() >= 7, ERR_INVALID, "objPara needs at least 7 entries");
    const int gsize = (int)cfg.geometry.size();
    nfpara_ = gsize / 3 / cfg.nz / dividesym;

    // parameter mapping (DDAModel.cpp:37-64, 179-226)
    std::map<std::pair<int, int>, int> fcur;
    int pos = 0;
    gpara_.assign(N_, 0);
    for (int i = 0; i < N_; ++i) {
        const std::pair<int, int> xy(cfg.geometry[3 * i], cfg.geometry[3 * i + 1]);
        if (!fcur.count(xy)) {
            if (dividesym == 1) {
                fcur[xy] = pos++;
            } else {
                const int x = xy.first, y = xy.second;
                const int mx = (int)c_round(2 * cfg.symaxis[0] - x);
                const int my = (int)c_round(2 * cfg.symaxis[1] - y);
                const std::pair<int, int> s1(mx, y), s2(mx, my), s3(x, my);
                if (fcur.count(s1)) fcur[xy] = fcur[s1];
                else if (fcur.count(s2)) fcur[xy] = fcur[s2];
                else if (fcur.count(s3)) fcur[xy] = fcur[s3];
                else fcur[xy] = pos++;
            }
        }
        gpara_[i] = fcur[xy];
    }
    TDDA_REQUIRE(pos <= nfpara_, ERR_INVALID, "geometry has more free parameters than NFpara");
    p2g_.assign(nfpara_, {});
    for (int i = 0; i < N_; ++i) p2g_[gpara_[i]].push_back(i);
    for (int i = 0; i < nfpara_; ++i)
        TDDA_REQUIRE(!p2g_[i].empty(), ERR_INVALID, "free parameter without dipoles");
    TDDA_REQUIRE(cfg.input_diel.size() == cfg.geometry.size(), ERR_INVALID,
                 "ERROR: SpacePara::SpacePara: (*InputGeo).size() != (*Inputdiel).size()");
    std::map<std::tuple<int, int, int>, double> inmap;
    for (int i = 0; i < gsize / 3; ++i)
        inmap.insert({std::make_tuple(cfg.geometry[3 * i], cfg.geometry[3 * i + 1], cfg.geometry[3 * i + 2]),
                      cfg.input_diel[3 * i]});
    params_.assign(nfpara_, 0.0);
    for (int i = 0; i < nfpara_; ++i) {
        const int p = p2g_[i][0];
        params_[i] = inmap[std::make_tuple(cfg.geometry[3 * p], cfg.geometry[3 * p + 1], cfg.geometry[3 * p + 2])];
    }
    diel_old_.assign(3 * (size_t)N_, 0.0);
    for (int i = 0; i < N_; ++i)
        for (int c = 0; c < 3; ++c) diel_old_[3 * i + c] = params_[gpara_[i]];

    // interaction method (AProductCore.cpp:322-333 throws invalid_argument)
    int method;
    if (cfg.method == "FCD")
        method = 0;
    else if (cfg.method == "LDR")
        method = 1;
    else
        throw Error(ERR_METHOD, "\"" + cfg.method + "\" is not a valid polarizabiliy calculation method.");

    // wavenumber and material (AProductCore.cpp:25-33)
    K_ = 2.0 * M_PI / (cfg.lam / cfg.nback);
    material_.clear();
    for (cd m : cfg.material) {
        m = m / (cfg.nback * cfg.nback);
        if (std::sqrt(std::norm(m)) < 1.01) m = m + 0.01;
        material_.push_back(m);
    }
    TDDA_REQUIRE(!material_.empty(), ERR_INVALID, "material list is empty");

    // dipole -> box cell map
    box_map_.assign((size_t)cfg.nx * cfg.ny * cfg.nz, -1);
    identity_map_ = (long)N_ == (long)cfg.nx * cfg.ny * cfg.nz;
    for (int i = 0; i < N_; ++i) {
        const int x = cfg.geometry[3 * i], y = cfg.geometry[3 * i + 1], z = cfg.geometry[3 * i + 2];
        TDDA_REQUIRE(x >= 0 && x < cfg.nx && y >= 0 && y < cfg.ny && z >= 0 && z < cfg.nz, ERR_INVALID,
                     "dipole position outside the num_x*num_y*num_z box");
        const size_t cell = ((size_t)z * cfg.ny + y) * cfg.nx + x;
        TDDA_REQUIRE(box_map_[cell] < 0, ERR_INVALID, "two dipoles share one grid position");
        box_map_[cell] = i;
        if ((long)cell != i) identity_map_ = false;
    }

    // operator (cached on the device by its key)
    SiciDevice tables = sici_tables(method == 0 ? cfg.si : nullptr, method == 0 ? cfg.ci : nullptr,
                                    method == 0 ? cfg.nsici : 0, cfg.delta);
    OperatorKey key{cfg.nx, cfg.ny, cfg.nz, cfg.d, K_, method, cfg.maxm, cfg.maxn,
                    cfg.lm * cfg.d, cfg.ln * cfg.d, method == 0 ? tables.id : 0};
    op_ = OperatorCache::instance().get(key, tables, default_stream(), &op_hit_);
    op_setup_ms_ = op_->setup_ms();

    // objective (ObjDDAModel.cpp:103-128), incident field (DDAModel.cpp:312-316)
    pownum_ = (int)c_round(cfg.objpara[0]);
    region_.assign(N_, 0);
    for (int i = 0; i < N_; ++i) region_[i] = objective_region(&cfg.geometry[3 * i], cfg.objpara) ? 1 : 0;
    einc_.assign(3 * (size_t)N_, cd(0, 0));
    for (int i = 0; i < N_; ++i) {
        const double ph = K_ * cfg.d *
                          (cfg.nk[0] * cfg.geometry[3 * i] + cfg.nk[1] * cfg.geometry[3 * i + 1] +
                           cfg.nk[2] * cfg.geometry[3 * i + 2]);
        const cd e(std::cos(ph), std::sin(ph));
        for (int c = 0; c < 3; ++c) einc_[3 * i + c] = cfg.e0 * cfg.ne0[c] * e;
    }
    P_.assign(3 * (size_t)N_, cd(0, 0));
    eresult_.assign(3 * (size_t)N_, cd(0, 0));
    upload_obj_maps();
    update_alpha();
    if (cfg.verbose) {
        std::printf("DDAModel: %d dipoles, %d free parameters, FFT grid %dx%dx%d, K=%.9g, operator %s (%.1f ms)\n",
                    N_, nfpara_, op_->lx, op_->ly, op_->lz, K_, op_hit_ ? "cached" : "built", op_setup_ms_);
        std::fflush(stdout);
    }
}

// DDAModel.cpp:319-326
cd Model::eps_of(double x) const {
    const int lf = (int)std::floor(x);
    const int ln = lf >= 1 ? lf : lf + 1;
    TDDA_REQUIRE(lf >= 0 && ln < (int)material_.size(), ERR_INVALID,
                 "parameter value outside the material interpolation range");
    return material_[lf] + (x - (double)lf) * (material_[ln] - material_[lf]);
}

// The objective's static maps on the device: region mask, dipole ->
// parameter and parameter -> dipoles (in p2g_ order), and zeroed buffers for
// the solution and the partials.
void Model::upload_obj_maps() {
    hipStream_t st = default_stream();
    const size_t n3 = 3 * (size_t)N_;
    std::vector<unsigned char> reg(region_.begin(), region_.end());
    std::vector<int> off(nfpara_ + 1, 0), dips;
    dips.reserve(N_);
    for (int i = 0; i < nfpara_; ++i) {
        dips.insert(dips.end(), p2g_[i].begin(), p2g_[i].end());
        off[i + 1] = (int)dips.size();
    }
    d_region_.alloc(N_);
    d_gpara_.alloc(N_);
    d_off_.alloc(nfpara_ + 1);
    d_dips_.alloc(dips.size());
    d_region_.upload(reg.data(), reg.size(), st);
    d_gpara_.upload(gpara_.data(), N_, st);
    d_off_.upload(off.data(), off.size(), st);
    d_dips_.upload(dips.data(), dips.size(), st);
    d_dal_.alloc(N_);
    d_P_.alloc(n3);
    d_P_.zero(st);
    d_w_.alloc(n3);
    d_part_.alloc((N_ + 255) / 256);
    d_E_.alloc(1);
    d_devx_.alloc(nfpara_);
    d_grad_.alloc(nfpara_);
    d_step_.alloc(nfpara_);
    d_devp_.alloc(n3);
    d_adevxp_.alloc(n3);
    d_alt_.alloc(nfpara_);
    TDDA_HIP(hipStreamSynchronize(st));   // the host vectors leave scope
    obj_maps_.region = d_region_.get();
    obj_maps_.gpara = d_gpara_.get();
    obj_maps_.off = d_off_.get();
    obj_maps_.dips = d_dips_.get();
}

// DDAModel.cpp:1122-1135
void Model::update_alpha() {
    diel_.assign(N_, cd(0, 0));
    al_.assign(N_, cd(0, 0));
    std::vector<cd> dal(N_);
    for (int i = 0; i < N_; ++i) {
        diel_[i] = eps_of(diel_old_[3 * i]);
        al_[i] = ldr_inverse_alpha(K_, cfg_.d, diel_[i], cfg_.ne0, cfg_.nk);
        // d(1/alpha)/d rho of the exact gradient: eps is linear in rho up to 1
        const double x = diel_old_[3 * i];
        const cd slope = x <= 1.0 && material_.size() >= 2 ? material_[1] - material_[0] : cd(0, 0);
        dal[i] = ldr_inverse_alpha_deps(K_, cfg_.d, diel_[i], cfg_.ne0, cfg_.nk) * slope;
    }
    d_dal_.upload(reinterpret_cast<const double2*>(dal.data()), N_, default_stream());
    TDDA_HIP(hipStreamSynchronize(default_stream()));
    if (sys64_) sys64_->set_alpha(reinterpret_cast<const double2*>(al_.data()), default_stream());
    if (sys32_) {
        sys32_->set_alpha(reinterpret_cast<const double2*>(al_.data()), default_stream());
        upload_field_factor(*sys32_);
    }
}

// E = f P with f = al 3/(eps + 2) (DDAModel.cpp:1185-1192): the weight of the
// refined solver's field-error estimate
void Model::upload_field_factor(System& s) {
    std::vector<cd> f(N_);
    for (int i = 0; i < N_; ++i) f[i] = al_[i] / ((2.0 + diel_[i]) / 3.0);
    s.set_field_factor(reinterpret_cast<const double2*>(f.data()), default_stream());
}

// DDAModel.cpp:747-772
void Model::set_parameters(const double* v, long n) {
    TDDA_REQUIRE(n == nfpara_, ERR_INVALID,
                 "ERROR: In CoreStructure::UpdateStr(VectorXd step), step.size!=FreePara.size");
    std::vector<double> old = params_;
    for (int i = 0; i < nfpara_; ++i) params_[i] = v[i];
    for (int i = 0; i < N_; ++i)
        for (int c = 0; c < 3; ++c) diel_old_[3 * i + c] = params_[gpara_[i]];
    try {
        update_alpha();
    } catch (...) {
        params_ = old;
        for (int i = 0; i < N_; ++i)
            for (int c = 0; c < 3; ++c) diel_old_[3 * i + c] = params_[gpara_[i]];
        update_alpha();
        throw;
    }
}

void Model::set_precision(int p) {
    TDDA_REQUIRE(p == PREC_FP64 || p == PREC_FP32_REFINED, ERR_INVALID, "unknown precision mode");
    precision_ = p;
}

System& Model::system() {
    std::unique_ptr<System>& s = precision_ == PREC_FP64 ? sys64_ : sys32_;
    if (!s) {
        s.reset(new System(op_, N_, identity_map_ ? nullptr : box_map_.data(), precision_ != PREC_FP64,
                           default_stream()));
        s->set_alpha(reinterpret_cast<const double2*>(al_.data()), default_stream());
        if (s->refined) upload_field_factor(*s);
    }
    return *s;
}

SolveResult Model::run_solve(const std::vector<cd>* rhs, const double2* rhs_dev, const cd* P0, int maxit, double tol,
                             std::vector<cd>& P) {
    System& s = system();
    hipStream_t st = default_stream();
    if (rhs) {
        s.set_rhs(reinterpret_cast<const double2*>(rhs->data()), st);
        rhs_on_device_[precision_] = rhs == &einc_;
    } else {
        TDDA_HIP(hipMemcpyAsync(s.rhs.get(), rhs_dev, 3 * (size_t)N_ * sizeof(double2), hipMemcpyDeviceToDevice, st));
        rhs_on_device_[precision_] = false;
    }
    bool warm = false;
    if (P0) {
        for (long i = 0; i < 3L * N_; ++i)
            if (P0[i] != cd(0, 0)) { warm = true; break; }
    }
    if (warm)
        s.set_solution(reinterpret_cast<const double2*>(P0), st);
    else
        s.P.zero(st);
    SolveOptions o;
    o.refined = precision_ == PREC_FP32_REFINED;
    o.maxit = maxit;
    o.tol = tol;
    o.err_tol = o.refined ? tol : 0.0;   // refined solves also bound the estimated field error by tol
    o.warm = warm;
    std::vector<System*> v{&s};
    SolveResult r = solve_batch(v, o, st)[0];
    TDDA_HIP(hipMemcpyAsync(d_P_.get(), s.P.get(), 3 * (size_t)N_ * sizeof(double2), hipMemcpyDeviceToDevice, st));
    P.assign(3 * (size_t)N_, cd(0, 0));
    s.get_solution(reinterpret_cast<double2*>(P.data()), st);
    TDDA_HIP(hipStreamSynchronize(st));
    if (r.status == JOB_MAXIT) {
        std::printf("                ERROR:does not converge in %d iterations\n", maxit);
        std::fflush(stdout);
    }
    return r;
}

// DDAModel.cpp:874-885 + 1153-1192
void Model::solve_electric_field(const cd* P0, long n, int maxit, double tol) {
    TDDA_REQUIRE(n == 3L * N_, ERR_INVALID, "originPolarization must hold 3*num_total values");
    last_ = run_solve(&einc_, nullptr, P0, maxit, tol, P_);
    for (int i = 0; i < N_; ++i) {
        const cd lf = (2.0 + diel_[i]) / 3.0;
        for (int c = 0; c < 3; ++c) eresult_[3 * i + c] = al_[i] * P_[3 * i + c] / lf;
    }
}

// w = |al P|^p and E = sum of the region's w for the current solution, on the
// device (ObjDDAModel.cpp:130-212)
void Model::objective_terms(hipStream_t st) {
    launch_obj_terms(N_, pownum_, system().al64.get(), d_P_.get(), d_region_.get(), d_w_.get(), d_part_.get(),
                     d_E_.get(), st);
}

// ObjDDAModel.cpp:205-212 (GetVal = Reset + N SingleResponse + log)
double Model::calculate_objective() {
    hipStream_t st = default_stream();
    objective_terms(st);
    double E = 0.0;
    d_E_.download(&E, 1, st);
    TDDA_HIP(hipStreamSynchronize(st));
    return std::log(E);
}

// DDAModel.cpp:398-578: the explicit partials and the assembly run on the
// device (model_obj.hip); the host forms only the perturbed 1/alpha of each
// parameter for the reference's one-sided differences (the material table
// interpolation of update_alpha).
std::vector<double> Model::calculate_gradients(double eps, double origin, int maxit, double tol) {
    hipStream_t st = default_stream();
    objective_terms(st);
    ObjElementParams A{N_, nfpara_, pownum_, grad_mode_ == GRAD_EXACT ? 1 : 0, eps, origin};
    if (!A.exact) {
        // one-sided difference with the reference's sign rule (DDAModel.cpp:482-491)
        std::vector<cd> alt(nfpara_);
        std::vector<double> step(nfpara_);
        for (int i = 0; i < nfpara_; ++i) {
            const double x0 = params_[i];
            step[i] = (x0 >= eps ? -1 : 1) * eps;
            alt[i] = ldr_inverse_alpha(K_, cfg_.d, eps_of(x0 + step[i]), cfg_.ne0, cfg_.nk);
        }
        d_alt_.upload(reinterpret_cast<const double2*>(alt.data()), nfpara_, st);
        d_step_.upload(step.data(), nfpara_, st);
        TDDA_HIP(hipStreamSynchronize(st));   // the host vectors leave scope
    }
    launch_obj_partials(A, obj_maps_, system().al64.get(), d_P_.get(), d_w_.get(), d_E_.get(), d_dal_.get(),
                        d_alt_.get(), d_step_.get(), d_devp_.get(), d_adevxp_.get(), d_devx_.get(), st);
    // adjoint solve (A is complex symmetric): lambda = A^-1 devp, from zero;
    // the reference leaves lambda in P (DDAModel.cpp:433)
    last_adj_ = run_solve(nullptr, d_devp_.get(), nullptr, maxit, tol, P_);
    launch_obj_assemble(nfpara_, obj_maps_, d_devx_.get(), d_P_.get(), d_adevxp_.get(), d_grad_.get(), st);
    std::vector<double> grad(nfpara_, 0.0);
    d_grad_.download(grad.data(), nfpara_, st);
    TDDA_HIP(hipStreamSynchronize(st));
    return grad;
}

SolveResult Model::solve_resident(int maxit, double tol) {
    System& s = system();
    hipStream_t st = default_stream();
    if (!rhs_on_device_[precision_]) {
        s.set_rhs(reinterpret_cast<const double2*>(einc_.data()), st);
        rhs_on_device_[precision_] = true;
    }
    s.P.zero(st);
    SolveOptions o;
    o.refined = precision_ == PREC_FP32_REFINED;
    o.maxit = maxit;
    o.tol = tol;
    o.err_tol = o.refined ? tol : 0.0;
    std::vector<System*> v{&s};
    return solve_batch(v, o, st)[0];
}

void Model::time_passes(int reps, double* ms5) {
    tdda::time_passes(system(), precision_ == PREC_FP32_REFINED, reps, default_stream(), ms5);
}

std::vector<cd> Model::apply(const std::vector<cd>& x) {
    TDDA_REQUIRE(x.size() == 3 * (size_t)N_, ERR_INVALID, "vector length must be 3*num_total");
    System& s = system();
    hipStream_t st = default_stream();
    DevBuf<double2> dx(x.size()), dy(x.size());
    dx.upload(reinterpret_cast<const double2*>(x.data()), x.size(), st);
    apply_operator64(s, dx.get(), dy.get(), st);
    std::vector<cd> y(x.size());
    dy.download(reinterpret_cast<double2*>(y.data()), y.size(), st);
    TDDA_HIP(hipStreamSynchronize(st));
    return y;
}

}  // namespace tdda

// extern "C" entry points (include/topodda.h) over the Model runtime.
#include <algorithm>
#include <array>
#include <cstring>
#include <string>

#include "../../include/topodda.h"
#include "model.h"
#include "noalloc.h"
#include "schedule.h"
#include "objective.h"
#include "serve.h"

namespace tdda {
static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* get_last_error() { return g_last_error.c_str(); }
}  // namespace tdda

using tdda::cd;
using tdda::Model;

struct tdda_model {
    Model* m;
};

template <typename F> static int guarded(F&& f) {
    try {
        tdda::set_last_error("");
        f();
        return TDDA_OK;
    } catch (const tdda::Error& e) {
        tdda::set_last_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        tdda::set_last_error(e.what());
        return TDDA_ERR_INVALID;
    } catch (...) {
        tdda::set_last_error("Caught an unknown exception!");
        return TDDA_ERR_INVALID;
    }
}

#define CHECK_MODEL(m) TDDA_REQUIRE((m) != nullptr && (m)->m != nullptr, tdda::ERR_STATE, "null model handle")

extern "C" {

const char* tdda_last_error(void) { return tdda::get_last_error(); }

long tdda_device_bytes(void) { return (long)tdda::device_bytes_in_use(); }

int tdda_set_graph_capture(int on) {
    return guarded([&] { tdda::set_graph_capture(on != 0); });
}

int tdda_set_fused_yz(int mode) {
    return guarded([&] { tdda::set_fused_yz(mode); });
}

int tdda_set_plane_passes(int on) {
    return guarded([&] { tdda::set_plane_passes(on != 0); });
}

static void copy_text(const std::string& t, char* buf, long cap, long* len) {
    if (len) *len = (long)t.size();
    if (buf && cap > 0) {
        const size_t n = std::min((size_t)(cap - 1), t.size());
        std::memcpy(buf, t.data(), n);
        buf[n] = 0;
    }
}

int tdda_schedule_text(char* buf, long cap, long* len) {
    return guarded([&] { copy_text(tdda::Schedule::instance().text(), buf, cap, len); });
}

int tdda_schedule_parse(const char* text) {
    return guarded([&] {
        TDDA_REQUIRE(text != nullptr, tdda::ERR_INVALID, "null schedule text");
        tdda::Schedule::instance().parse(text);
    });
}

int tdda_schedule_load(const char* path) {
    return guarded([&] {
        TDDA_REQUIRE(path != nullptr, tdda::ERR_INVALID, "null path");
        tdda::Schedule::instance().load(path);
    });
}

int tdda_schedule_save(const char* path) {
    return guarded([&] {
        TDDA_REQUIRE(path != nullptr, tdda::ERR_INVALID, "null path");
        tdda::Schedule::instance().save(path);
    });
}

int tdda_schedule_clear(void) {
    return guarded([&] { tdda::Schedule::instance().clear(); });
}

int tdda_tune(const int* boxes, int n, int count, int reps, char* report, long cap, long* len) {
    return guarded([&] {
        TDDA_REQUIRE(boxes != nullptr && n >= 1, tdda::ERR_INVALID, "no boxes to tune for");
        std::vector<std::array<int, 3>> b(n);
        for (int i = 0; i < n; ++i) b[i] = {boxes[3 * i], boxes[3 * i + 1], boxes[3 * i + 2]};
        const tdda::TuneReport r = tdda::tune(b, count, reps, tdda::default_stream());
        copy_text(tdda::tune_report_text(r), report, cap, len);
    });
}

int tdda_trim_caches(void) {
    return guarded([&] {
        tdda::OperatorCache::instance().trim(~size_t(0));
        TDDA_HIP(hipDeviceSynchronize());
        tdda::device_release_cached();
    });
}

int tdda_model_create(const tdda_model_args* a, tdda_model** out) {
    return guarded([&] {
        TDDA_REQUIRE(a != nullptr && out != nullptr, tdda::ERR_INVALID, "null argument");
        *out = nullptr;
        tdda::ModelConfig c;
        c.symmetry = a->symmetryType ? a->symmetryType : "";
        c.symaxis.assign(a->symmetryAxis, a->symmetryAxis + (a->symmetryAxis ? a->nSymmetryAxis : 0));
        c.objective = a->objectiveName ? a->objectiveName : "";
        c.objpara.assign(a->objPara, a->objPara + (a->objPara ? a->nObjPara : 0));
        TDDA_REQUIRE(a->geometry != nullptr && a->inputDielectrics != nullptr, tdda::ERR_INVALID,
                     "geometry and inputDielectrics are required");
        c.geometry.assign(a->geometry, a->geometry + a->nGeometry);
        c.input_diel.assign(a->inputDielectrics, a->inputDielectrics + a->nInputDielectrics);
        c.nx = a->num_x;
        c.ny = a->num_y;
        c.nz = a->num_z;
        c.n = a->num_total;
        TDDA_REQUIRE(a->n_K && a->n_E0, tdda::ERR_INVALID, "n_K and n_E0 are required");
        for (int i = 0; i < 3; ++i) { c.nk[i] = a->n_K[i]; c.ne0[i] = a->n_E0[i]; }
        c.e0 = a->E0;
        c.lam = a->lam;
        for (long i = 0; i < a->nMaterial; ++i) c.material.emplace_back(a->material[2 * i], a->material[2 * i + 1]);
        c.nback = a->nback;
        c.maxm = a->MAXm;
        c.maxn = a->MAXn;
        c.lm = a->Lm;
        c.ln = a->Ln;
        c.method = a->matrixAMethodName ? a->matrixAMethodName : "";
        c.d = a->d;
        c.si = a->sineIntegralValues;
        c.ci = a->cosineIntegralValues;
        TDDA_REQUIRE(a->nSine == a->nCosine, tdda::ERR_INVALID, "Si and Ci tables differ in length");
        c.nsici = a->nSine;
        c.delta = a->integralDelta;
        c.verbose = a->verbose != 0;
        auto* h = new tdda_model{nullptr};
        try {
            h->m = new Model(c);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int tdda_model_destroy(tdda_model* m) {
    return guarded([&] {
        if (!m) return;
        delete m->m;
        delete m;
    });
}

int tdda_model_solve_electric_field(tdda_model* m, const double* P0, long n, int maxit, double tol) {
    return guarded([&] {
        CHECK_MODEL(m);
        m->m->solve_electric_field(reinterpret_cast<const cd*>(P0), n, maxit, tol);
    });
}

int tdda_model_calculate_objective(tdda_model* m, double* value) {
    return guarded([&] {
        CHECK_MODEL(m);
        *value = m->m->calculate_objective();
    });
}

int tdda_model_calculate_gradients(tdda_model* m, double eps, double origin, int maxit, double tol, double* grad,
                                   long n) {
    return guarded([&] {
        CHECK_MODEL(m);
        TDDA_REQUIRE(n == m->m->nfpara(), tdda::ERR_INVALID, "gradient buffer must hold NFpara values");
        std::vector<double> g = m->m->calculate_gradients(eps, origin, maxit, tol);
        std::memcpy(grad, g.data(), g.size() * sizeof(double));
    });
}

int tdda_model_set_parameters(tdda_model* m, const double* values, long n) {
    return guarded([&] {
        CHECK_MODEL(m);
        m->m->set_parameters(values, n);
    });
}

int tdda_model_get_parameters(tdda_model* m, double* out, long n) {
    return guarded([&] {
        CHECK_MODEL(m);
        const auto& p = m->m->parameters();
        TDDA_REQUIRE(n == (long)p.size(), tdda::ERR_INVALID, "buffer must hold NFpara values");
        std::memcpy(out, p.data(), p.size() * sizeof(double));
    });
}

int tdda_model_get_dielectrics(tdda_model* m, double* out, long n) {
    return guarded([&] {
        CHECK_MODEL(m);
        const auto& p = m->m->dielectrics();
        TDDA_REQUIRE(n == (long)p.size(), tdda::ERR_INVALID, "buffer must hold 3*num_total values");
        std::memcpy(out, p.data(), p.size() * sizeof(double));
    });
}

int tdda_model_get_electric_field(tdda_model* m, double* out, long n) {
    return guarded([&] {
        CHECK_MODEL(m);
        const auto& e = m->m->electric_field();
        TDDA_REQUIRE(n == (long)e.size(), tdda::ERR_INVALID, "buffer must hold 3*num_total complex");
        std::memcpy(out, e.data(), e.size() * sizeof(cd));
    });
}

int tdda_model_sizes(tdda_model* m, long* nfpara, long* ndip) {
    return guarded([&] {
        CHECK_MODEL(m);
        if (nfpara) *nfpara = m->m->nfpara();
        if (ndip) *ndip = m->m->ndip();
    });
}

int tdda_model_set_precision(tdda_model* m, int p) {
    return guarded([&] {
        CHECK_MODEL(m);
        m->m->set_precision(p);
    });
}

int tdda_model_set_gradient_mode(tdda_model* m, int mode) {
    return guarded([&] {
        CHECK_MODEL(m);
        TDDA_REQUIRE(mode == TDDA_GRAD_REFERENCE || mode == TDDA_GRAD_EXACT, tdda::ERR_INVALID,
                     "unknown gradient mode");
        m->m->set_gradient_mode(mode);
    });
}

int tdda_model_solve_info(tdda_model* m, int which, tdda_solve_info* out) {
    return guarded([&] {
        CHECK_MODEL(m);
        const tdda::SolveResult& r = which == 0 ? m->m->last_solve() : m->m->last_adjoint();
        out->status = r.status;
        out->iterations = r.iterations;
        out->error = r.err;
        out->reliable_updates = r.n_ru;
        out->ms = r.ms;
        out->err_est = r.err_est;
    });
}

int tdda_model_get_polarization(tdda_model* m, double* out, long n) {
    return guarded([&] {
        CHECK_MODEL(m);
        const auto& p = m->m->polarization();
        TDDA_REQUIRE(n == (long)p.size(), tdda::ERR_INVALID, "buffer must hold 3*num_total complex");
        std::memcpy(out, p.data(), p.size() * sizeof(cd));
    });
}

int tdda_model_get_inverse_alpha(tdda_model* m, double* out, long n) {
    return guarded([&] {
        CHECK_MODEL(m);
        const auto& p = m->m->inverse_alpha();
        TDDA_REQUIRE(n == (long)p.size(), tdda::ERR_INVALID, "buffer must hold num_total complex");
        std::memcpy(out, p.data(), p.size() * sizeof(cd));
    });
}

int tdda_model_apply(tdda_model* m, const double* x, double* y, long n) {
    return guarded([&] {
        CHECK_MODEL(m);
        std::vector<cd> xv(reinterpret_cast<const cd*>(x), reinterpret_cast<const cd*>(x) + n);
        std::vector<cd> yv = m->m->apply(xv);
        std::memcpy(y, yv.data(), yv.size() * sizeof(cd));
    });
}

int tdda_model_solve_resident(tdda_model* m, int maxit, double tol, tdda_solve_info* out) {
    return guarded([&] {
        CHECK_MODEL(m);
        tdda::SolveResult r = m->m->solve_resident(maxit, tol);
        if (out) {
            out->status = r.status;
            out->iterations = r.iterations;
            out->error = r.err;
            out->reliable_updates = r.n_ru;
            out->ms = r.ms;
            out->err_est = r.err_est;
        }
    });
}

int tdda_model_time_passes(tdda_model* m, int reps, double* ms5) {
    return guarded([&] {
        CHECK_MODEL(m);
        m->m->time_passes(reps, ms5);
    });
}

int tdda_model_grid(tdda_model* m, int* lxyz, double* setup_ms) {
    return guarded([&] {
        CHECK_MODEL(m);
        if (lxyz) {
            lxyz[0] = m->m->op().lx;
            lxyz[1] = m->m->op().ly;
            lxyz[2] = m->m->op().lz;
        }
        if (setup_ms) *setup_ms = m->m->operator_setup_ms();
    });
}

// ---------------------------------------------------------------------------
// problems and server
// ---------------------------------------------------------------------------
struct tdda_problem {
    tdda::Problem* p;
};
struct tdda_server {
    tdda::Server* s;
};

static tdda::ProblemDesc to_desc(const tdda_problem_desc* d) {
    TDDA_REQUIRE(d != nullptr, tdda::ERR_INVALID, "null problem description");
    tdda::ProblemDesc o{};
    o.nx = d->nx; o.ny = d->ny; o.nz = d->nz;
    o.d = d->d; o.lam = d->lam; o.nback = d->nback;
    o.method = d->method; o.maxm = d->maxm; o.maxn = d->maxn; o.lm = d->lm; o.ln = d->ln;
    for (int k = 0; k < 3; ++k) { o.nk[k] = d->n_k[k]; o.ne0[k] = d->n_e0[k]; }
    o.e0 = d->e0;
    for (int k = 0; k < 4; ++k) o.mat[k] = d->material[k];
    return o;
}

int tdda_problem_create(const tdda_problem_desc* d, int precision, const double* si, const double* ci, long nsici,
                        double delta, tdda_problem** out) {
    return guarded([&] {
        TDDA_REQUIRE(out != nullptr, tdda::ERR_INVALID, "null output");
        TDDA_REQUIRE(precision == TDDA_PREC_FP64 || precision == TDDA_PREC_FP32_REFINED, tdda::ERR_INVALID,
                     "unknown precision");
        tdda::ProblemDesc pd = to_desc(d);
        tdda::SiciDevice t = tdda::sici_tables(pd.method == 0 ? si : nullptr, pd.method == 0 ? ci : nullptr,
                                               pd.method == 0 ? nsici : 0, delta);
        auto* h = new tdda_problem{nullptr};
        try {
            h->p = new tdda::Problem(pd, t, precision == TDDA_PREC_FP32_REFINED, tdda::default_stream());
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int tdda_problems_create(const tdda_problem_desc* d, int n, int precision, const double* si, const double* ci,
                         long nsici, double delta, tdda_problem** out) {
    return guarded([&] {
        TDDA_REQUIRE(out != nullptr && (d != nullptr || n == 0) && n >= 0, tdda::ERR_INVALID, "null argument");
        TDDA_REQUIRE(precision == TDDA_PREC_FP64 || precision == TDDA_PREC_FP32_REFINED, tdda::ERR_INVALID,
                     "unknown precision");
        std::vector<std::unique_ptr<tdda::Problem>> made;
        {
            tdda::GreenBatch gb;   // cancelled by its destructor if a creation fails
            for (int i = 0; i < n; ++i) {
                tdda::ProblemDesc pd = to_desc(&d[i]);
                tdda::SiciDevice t = tdda::sici_tables(pd.method == 0 ? si : nullptr, pd.method == 0 ? ci : nullptr,
                                                       pd.method == 0 ? nsici : 0, delta);
                made.emplace_back(new tdda::Problem(pd, t, precision == TDDA_PREC_FP32_REFINED,
                                                    tdda::default_stream(), nullptr, nullptr, &gb));
            }
            gb.flush(tdda::default_stream());
            // the batch's job tables are read by its launches
            TDDA_HIP(hipStreamSynchronize(tdda::default_stream()));
        }
        for (int i = 0; i < n; ++i) out[i] = new tdda_problem{made[i].release()};
    });
}

int tdda_problem_destroy(tdda_problem* p) {
    return guarded([&] {
        if (!p) return;
        TDDA_HIP(hipDeviceSynchronize());
        delete p->p;
        delete p;
    });
}

int tdda_problem_set_density(tdda_problem* p, const double* rho, int on_device) {
    return guarded([&] {
        TDDA_REQUIRE(p && p->p && rho, tdda::ERR_INVALID, "null argument");
        p->p->set_density(rho, on_device != 0, tdda::default_stream());
    });
}

int tdda_problems_solve(tdda_problem** ps, int n, int maxit, double tol, tdda_solve_info* infos) {
    return guarded([&] {
        TDDA_REQUIRE(ps != nullptr && n >= 1, tdda::ERR_INVALID, "no problems");
        std::vector<tdda::System*> sys;
        bool refined = ps[0]->p->sys->refined;
        for (int i = 0; i < n; ++i) {
            TDDA_REQUIRE(ps[i] && ps[i]->p, tdda::ERR_INVALID, "null problem");
            TDDA_REQUIRE(ps[i]->p->sys->refined == refined, tdda::ERR_INVALID, "mixed precisions in one batch");
            sys.push_back(ps[i]->p->sys.get());
            ps[i]->p->sys->P.zero(tdda::default_stream());
        }
        tdda::SolveOptions o;
        o.refined = refined;
        o.maxit = maxit;
        o.tol = tol;
        o.err_tol = refined ? tol : 0.0;
        std::vector<tdda::SolveResult> r = tdda::solve_batch(sys, o, tdda::default_stream());
        if (infos)
            for (int i = 0; i < n; ++i) {
                infos[i].status = r[i].status;
                infos[i].iterations = r[i].iterations;
                infos[i].error = r[i].err;
                infos[i].reliable_updates = r[i].n_ru;
                infos[i].ms = r[i].ms;
                infos[i].err_est = r[i].err_est;
            }
    });
}

int tdda_problem_set_preconditioner(tdda_problem* p, int precond) {
    return guarded([&] {
        TDDA_REQUIRE(p && p->p, tdda::ERR_INVALID, "null problem");
        TDDA_REQUIRE(precond == TDDA_PRECOND_NONE || precond == TDDA_PRECOND_JACOBI, tdda::ERR_INVALID,
                     "unknown preconditioner");
        p->p->sys->set_jacobi(precond == TDDA_PRECOND_JACOBI);
    });
}

int tdda_problem_grid(tdda_problem* p, int* lxyz, double* setup_ms, int* cached) {
    return guarded([&] {
        TDDA_REQUIRE(p && p->p, tdda::ERR_INVALID, "null problem");
        const tdda::Operator& op = *p->p->op;
        if (lxyz) {
            lxyz[0] = op.lx;
            lxyz[1] = op.ly;
            lxyz[2] = op.lz;
        }
        if (setup_ms) *setup_ms = op.setup_ms();
        if (cached) *cached = p->p->op_hit ? 1 : 0;
    });
}

int tdda_problem_field(tdda_problem* p, double* E, int on_device) {
    return guarded([&] {
        TDDA_REQUIRE(p && p->p && E, tdda::ERR_INVALID, "null argument");
        hipStream_t st = tdda::default_stream();
        const size_t n3 = 3ull * p->p->ndip;
        if (on_device) {
            p->p->field(reinterpret_cast<double2*>(E), st);
        } else {
            tdda::DevBuf<double2> b(n3);
            p->p->field(b.get(), st);
            b.download(reinterpret_cast<double2*>(E), n3, st);
        }
        TDDA_HIP(hipStreamSynchronize(st));
    });
}

int tdda_problem_polarization(tdda_problem* p, double* P, int on_device) {
    return guarded([&] {
        TDDA_REQUIRE(p && p->p && P, tdda::ERR_INVALID, "null argument");
        const size_t n3 = 3ull * p->p->ndip;
        TDDA_HIP(hipMemcpy(P, p->p->sys->P.get(), n3 * sizeof(double2),
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost));
    });
}

static tdda::ObjectiveSpec obj_spec(int kind, int power, const double* targets, int nt) {
    tdda::ObjectiveSpec o;
    o.kind = kind;
    o.power = power;
    if (kind == TDDA_OBJ_TARGETS) {
        TDDA_REQUIRE(targets != nullptr && nt >= 1, tdda::ERR_INVALID, "targets required");
        o.targets.assign(targets, targets + 3 * (size_t)nt);
    }
    return o;
}

int tdda_problem_objective(tdda_problem* p, int kind, int power, const double* targets, int nt, double* value,
                           double* fields) {
    return guarded([&] {
        TDDA_REQUIRE(p && p->p && value, tdda::ERR_INVALID, "null argument");
        std::vector<double2> f;
        *value = tdda::objective_value(*p->p, obj_spec(kind, power, targets, nt), tdda::default_stream(),
                                       fields ? &f : nullptr);
        if (fields && !f.empty()) std::memcpy(fields, f.data(), f.size() * sizeof(double2));
    });
}

int tdda_problem_gradient(tdda_problem* p, int kind, int power, const double* targets, int nt, int maxit,
                          double tol, double* grad, int on_device, double* value, tdda_solve_info* adj) {
    return guarded([&] {
        TDDA_REQUIRE(p && p->p && grad, tdda::ERR_INVALID, "null argument");
        tdda::SolveResult r = tdda::objective_gradient(*p->p, obj_spec(kind, power, targets, nt), maxit, tol, grad,
                                                       on_device != 0, tdda::default_stream(), value);
        if (adj) {
            adj->status = r.status;
            adj->iterations = r.iterations;
            adj->error = r.err;
            adj->reliable_updates = r.n_ru;
            adj->ms = r.ms;
            adj->err_est = r.err_est;
        }
    });
}

int tdda_problems_apply(tdda_problem** ps, int n, int c64, const double* x, double* y) {
    return guarded([&] {
        TDDA_REQUIRE(ps != nullptr && n >= 1 && x != nullptr && y != nullptr, tdda::ERR_INVALID, "null argument");
        std::vector<tdda::System*> sys;
        for (int i = 0; i < n; ++i) {
            TDDA_REQUIRE(ps[i] && ps[i]->p, tdda::ERR_INVALID, "null problem");
            sys.push_back(ps[i]->p->sys.get());
        }
        tdda::apply_batch(sys, c64 != 0, reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y),
                          tdda::default_stream());
    });
}

int tdda_problem_apply_noalloc(tdda_problem* p, const double* x, double* y) {
    return guarded([&] {
        TDDA_REQUIRE(p && p->p, tdda::ERR_INVALID, "null problem");
        tdda::apply_noalloc(*p->p->sys, reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y),
                            tdda::default_stream());
    });
}

int tdda_problems_time_passes(tdda_problem** ps, int n, int c64, int reps, double* ms5) {
    return guarded([&] {
        TDDA_REQUIRE(ps != nullptr && n >= 1 && ms5 != nullptr, tdda::ERR_INVALID, "no problems");
        std::vector<tdda::System*> sys;
        for (int i = 0; i < n; ++i) {
            TDDA_REQUIRE(ps[i] && ps[i]->p, tdda::ERR_INVALID, "null problem");
            sys.push_back(ps[i]->p->sys.get());
        }
        tdda::time_passes_batch(sys, c64 != 0, reps, tdda::default_stream(), ms5);
    });
}

int tdda_problems_time_iteration(tdda_problem** ps, int n, int reps, double* ms10) {
    return guarded([&] {
        TDDA_REQUIRE(ps != nullptr && n >= 1 && ms10 != nullptr, tdda::ERR_INVALID, "no problems");
        std::vector<tdda::System*> sys;
        for (int i = 0; i < n; ++i) {
            TDDA_REQUIRE(ps[i] && ps[i]->p, tdda::ERR_INVALID, "null problem");
            sys.push_back(ps[i]->p->sys.get());
        }
        tdda::time_iteration_batch(sys, reps, tdda::default_stream(), ms10);
    });
}

int tdda_server_create(const tdda_server_opts* o, const double* si, const double* ci, long nsici, double delta,
                       tdda_server** out) {
    return guarded([&] {
        TDDA_REQUIRE(o && out, tdda::ERR_INVALID, "null argument");
        tdda::ServerOptions so;
        so.refined = o->precision == TDDA_PREC_FP32_REFINED;
        so.maxit = o->maxit;
        so.tol = o->tol;
        so.ru_delta = o->ru_delta > 0 ? o->ru_delta : 0.01;
        so.err_tol = o->err_tol > 0 ? o->err_tol : (o->err_tol == 0 ? o->tol : 0.0);
        so.hbm_budget = (size_t)(o->hbm_budget_gb * (double)(1ull << 30));
        so.max_slots = o->max_slots;
        so.chunk = o->chunk;
        so.keep_fields = o->keep_fields != 0;
        TDDA_REQUIRE(o->precond == TDDA_PRECOND_NONE || o->precond == TDDA_PRECOND_JACOBI, tdda::ERR_INVALID,
                     "unknown preconditioner");
        so.jacobi = o->precond == TDDA_PRECOND_JACOBI;
        tdda::SiciDevice t = tdda::sici_tables(si, ci, si ? nsici : 0, delta);
        *out = new tdda_server{new tdda::Server(so, t)};
    });
}

int tdda_server_destroy(tdda_server* s) {
    return guarded([&] {
        if (!s) return;
        delete s->s;
        delete s;
    });
}

int tdda_server_submit(tdda_server* s, long id, const tdda_problem_desc* d, const double* rho, long n) {
    return guarded([&] {
        TDDA_REQUIRE(s && s->s && rho, tdda::ERR_INVALID, "null argument");
        tdda::ServeRequest r;
        r.id = id;
        r.desc = to_desc(d);
        TDDA_REQUIRE(r.desc.method == 0 || r.desc.method == 1, tdda::ERR_METHOD, "method must be FCD or LDR");
        r.rho.assign(rho, rho + n);
        s->s->submit(std::move(r));
    });
}

int tdda_server_poll(tdda_server* s, tdda_result* out, int max, int* n_out) {
    return guarded([&] {
        TDDA_REQUIRE(s && s->s && out && n_out, tdda::ERR_INVALID, "null argument");
        const std::string err = s->s->error();
        TDDA_REQUIRE(err.empty(), tdda::ERR_STATE, "server loop failed: " + err);
        std::vector<tdda::ServeResult> v;
        const int k = s->s->poll(v, max);
        for (int i = 0; i < k; ++i) {
            out[i].id = v[i].id;
            out[i].status = v[i].status;
            out[i].iterations = v[i].iterations;
            out[i].reliable_updates = v[i].n_ru;
            out[i].error = v[i].err;
            out[i].t_submit = v[i].t_submit;
            out[i].t_admit = v[i].t_admit;
            out[i].t_done = v[i].t_done;
            out[i].operator_cached = v[i].op_hit;
            out[i].err_est = v[i].err_est;
        }
        *n_out = k;
    });
}

int tdda_server_take_field(tdda_server* s, long id, double* E, long n) {
    return guarded([&] {
        TDDA_REQUIRE(s && s->s && E, tdda::ERR_INVALID, "null argument");
        TDDA_REQUIRE(n >= 0, tdda::ERR_INVALID, "negative field length");
        std::vector<double2> v;
        const int rc = s->s->take_field(id, (size_t)n, v);
        TDDA_REQUIRE(rc != 1, tdda::ERR_STATE, "no field kept for this request id");
        TDDA_REQUIRE(rc != 2, tdda::ERR_INVALID, "field buffer must hold 3*nx*ny*nz complex (the field is kept)");
        std::memcpy(E, v.data(), v.size() * sizeof(double2));
    });
}

int tdda_server_idle(tdda_server* s, int* idle) {
    return guarded([&] {
        TDDA_REQUIRE(s && s->s && idle, tdda::ERR_INVALID, "null argument");
        const std::string err = s->s->error();
        TDDA_REQUIRE(err.empty(), tdda::ERR_STATE, "server loop failed: " + err);
        *idle = s->s->idle() ? 1 : 0;
    });
}

int tdda_server_get_stats(tdda_server* s, tdda_server_stats* out) {
    return guarded([&] {
        TDDA_REQUIRE(s && s->s && out, tdda::ERR_INVALID, "null argument");
        tdda::ServerStats st = s->s->stats();
        out->submitted = st.submitted;
        out->completed = st.completed;
        out->chunks = st.chunks;
        out->rebuilds = st.rebuilds;
        out->launches = st.launches;
        out->max_active = st.max_active;
        out->peak_bytes = (long)st.peak_bytes;
        out->host_admit_ms = st.host_ms[0];
        out->host_wait_ms = st.host_ms[1];
        out->host_retire_ms = st.host_ms[2];
        out->host_build_ms = st.host_ms[3];
        out->host_enqueue_ms = st.enqueue_ms;
        out->graph_chunks = st.graph_chunks;
    });
}

}  // extern "C"

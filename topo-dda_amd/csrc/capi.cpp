// extern "C" entry points (include/topodda.h) over the Model runtime.
#include <cstring>
#include <string>

#include "../../include/topodda.h"
#include "model.h"

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

}  // extern "C"

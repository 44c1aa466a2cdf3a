// _dda_model: the reference's pybind11 module (python_bindings/PythonBindings.cpp:16-70)
// rebuilt over the C ABI of include/topodda.h.  Same class name, constructor
// keywords and order, and method names; solves release the GIL.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/topodda.h"

namespace py = pybind11;
using cd = std::complex<double>;

namespace {

// Map C ABI failures to the exception types the reference module raises:
// `throw 1` -> RuntimeError, std::invalid_argument -> ValueError.
void check(int rc) {
    if (rc == TDDA_OK) return;
    std::string msg = tdda_last_error();
    if (rc == TDDA_ERR_METHOD) throw py::value_error(msg);
    if (rc == TDDA_ERR_MEMORY) throw std::runtime_error("device memory: " + msg);
    throw std::runtime_error(msg.empty() ? std::string("Caught an unknown exception!") : msg);
}

template <typename T> std::vector<T> as_vec(py::handle h) {
    auto a = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(h);
    if (!a) throw std::runtime_error("expected an array-like argument");
    return std::vector<T>(a.data(), a.data() + a.size());
}

class PyDDAModel {
   public:
    PyDDAModel(double betaMin, double betaMax, double ita, const std::string& betaType,
               const std::vector<int>& filterIterations, const std::vector<double>& filterRadii, bool useFilter,
               const std::string& symmetryType, const std::vector<double>& symmetryAxis, bool isPeriodic,
               const std::string& objectiveName, const std::vector<double>& objPara, py::handle geometry,
               py::handle inputDielectrics, int num_x, int num_y, int num_z, int num_total, py::handle n_K,
               double E0, py::handle n_E0, double lam, py::handle material, double nback, int MAXm, int MAXn,
               double Lm, double Ln, const std::string& matrixAMethodName, double d, py::handle sineIntegralValues,
               py::handle cosineIntegralValues, double integralDelta, bool verbose) {
        std::vector<int> geo = as_vec<int>(geometry);
        std::vector<double> diel = as_vec<double>(inputDielectrics);
        std::vector<double> nk = as_vec<double>(n_K), ne0 = as_vec<double>(n_E0);
        if (nk.size() != 3 || ne0.size() != 3) throw std::runtime_error("n_K and n_E0 need 3 components");
        std::vector<cd> mat = as_vec<cd>(material);
        std::vector<double> si = as_vec<double>(sineIntegralValues), ci = as_vec<double>(cosineIntegralValues);
        tdda_model_args a{};
        a.betaMin = betaMin;
        a.betaMax = betaMax;
        a.ita = ita;
        a.betaType = betaType.c_str();
        a.filterIterations = filterIterations.data();
        a.nFilterIterations = (long)filterIterations.size();
        a.filterRadii = filterRadii.data();
        a.nFilterRadii = (long)filterRadii.size();
        a.useFilter = useFilter;
        a.symmetryType = symmetryType.c_str();
        a.symmetryAxis = symmetryAxis.data();
        a.nSymmetryAxis = (long)symmetryAxis.size();
        a.isPeriodic = isPeriodic;
        a.objectiveName = objectiveName.c_str();
        a.objPara = objPara.data();
        a.nObjPara = (long)objPara.size();
        a.geometry = geo.data();
        a.nGeometry = (long)geo.size();
        a.inputDielectrics = diel.data();
        a.nInputDielectrics = (long)diel.size();
        a.num_x = num_x;
        a.num_y = num_y;
        a.num_z = num_z;
        a.num_total = num_total;
        a.n_K = nk.data();
        a.E0 = E0;
        a.n_E0 = ne0.data();
        a.lam = lam;
        a.material = reinterpret_cast<const double*>(mat.data());
        a.nMaterial = (long)mat.size();
        a.nback = nback;
        a.MAXm = MAXm;
        a.MAXn = MAXn;
        a.Lm = Lm;
        a.Ln = Ln;
        a.matrixAMethodName = matrixAMethodName.c_str();
        a.d = d;
        a.sineIntegralValues = si.data();
        a.nSine = (long)si.size();
        a.cosineIntegralValues = ci.data();
        a.nCosine = (long)ci.size();
        a.integralDelta = integralDelta;
        a.verbose = verbose;
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = tdda_model_create(&a, &m_);
        }
        check(rc);
        check(tdda_model_sizes(m_, &nfpara_, &ndip_));
    }
    ~PyDDAModel() { tdda_model_destroy(m_); }

    double calculateObjective() {
        double v = 0.0;
        check(tdda_model_calculate_objective(m_, &v));
        return v;
    }

    py::array_t<double> calculateGradients(double eps, double origin, int maxit, double tol) {
        py::array_t<double> g(nfpara_);
        double* out = g.mutable_data();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = tdda_model_calculate_gradients(m_, eps, origin, maxit, tol, out, nfpara_);
        }
        check(rc);
        return g;
    }

    void solveElectricField(py::handle P0, int maxit, double tol) {
        std::vector<cd> p = as_vec<cd>(P0);
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = tdda_model_solve_electric_field(m_, reinterpret_cast<const double*>(p.data()), (long)p.size(),
                                                 maxit, tol);
        }
        check(rc);
    }

    void setParameters(py::handle values) {
        std::vector<double> v = as_vec<double>(values);
        check(tdda_model_set_parameters(m_, v.data(), (long)v.size()));
    }

    py::array_t<double> getParameters() {
        py::array_t<double> a(nfpara_);
        check(tdda_model_get_parameters(m_, a.mutable_data(), nfpara_));
        return a;
    }

    py::array_t<double> getDielectrics() {
        py::array_t<double> a(3 * ndip_);
        check(tdda_model_get_dielectrics(m_, a.mutable_data(), 3 * ndip_));
        return a;
    }

    py::array_t<cd> getElectricField() {
        py::array_t<cd> a(3 * ndip_);
        check(tdda_model_get_electric_field(m_, reinterpret_cast<double*>(a.mutable_data()), 3 * ndip_));
        return a;
    }

    // ---- additions (no reference counterpart) ----
    void setSolverPrecision(const std::string& p) {
        int v;
        if (p == "fp64") v = TDDA_PREC_FP64;
        else if (p == "fp32_refined") v = TDDA_PREC_FP32_REFINED;
        else throw py::value_error("precision must be 'fp64' or 'fp32_refined'");
        check(tdda_model_set_precision(m_, v));
    }
    void setGradientMode(const std::string& mode) {
        int v;
        if (mode == "reference") v = TDDA_GRAD_REFERENCE;
        else if (mode == "exact") v = TDDA_GRAD_EXACT;
        else throw py::value_error("gradient mode must be 'reference' or 'exact'");
        check(tdda_model_set_gradient_mode(m_, v));
    }
    py::dict solveInfo(int which) {
        tdda_solve_info s{};
        check(tdda_model_solve_info(m_, which, &s));
        static const char* names[] = {"active", "converged", "breakdown", "maxit", "idle"};
        py::dict d;
        d["status"] = names[(s.status >= 0 && s.status <= 4) ? s.status : 4];
        d["iterations"] = s.iterations;
        d["error"] = s.error;
        d["reliable_updates"] = s.reliable_updates;
        d["ms"] = s.ms;
        d["err_est"] = s.err_est;
        return d;
    }
    py::array_t<cd> getPolarization() {
        py::array_t<cd> a(3 * ndip_);
        check(tdda_model_get_polarization(m_, reinterpret_cast<double*>(a.mutable_data()), 3 * ndip_));
        return a;
    }
    py::array_t<cd> getInverseAlpha() {
        py::array_t<cd> a(ndip_);
        check(tdda_model_get_inverse_alpha(m_, reinterpret_cast<double*>(a.mutable_data()), ndip_));
        return a;
    }
    py::array_t<cd> applyOperator(py::handle x) {
        std::vector<cd> xv = as_vec<cd>(x);
        py::array_t<cd> y((py::ssize_t)xv.size());
        int rc;
        double* yp = reinterpret_cast<double*>(y.mutable_data());
        {
            py::gil_scoped_release nogil;
            rc = tdda_model_apply(m_, reinterpret_cast<const double*>(xv.data()), yp, (long)xv.size());
        }
        check(rc);
        return y;
    }
    py::dict solveResident(int maxit, double tol) {
        tdda_solve_info s{};
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = tdda_model_solve_resident(m_, maxit, tol, &s);
        }
        check(rc);
        static const char* names[] = {"active", "converged", "breakdown", "maxit", "idle"};
        py::dict d;
        d["status"] = names[(s.status >= 0 && s.status <= 4) ? s.status : 4];
        d["iterations"] = s.iterations;
        d["error"] = s.error;
        d["reliable_updates"] = s.reliable_updates;
        d["ms"] = s.ms;
        d["err_est"] = s.err_est;
        return d;
    }
    std::vector<double> timePasses(int reps) {
        std::vector<double> ms(5, 0.0);
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = tdda_model_time_passes(m_, reps, ms.data());
        }
        check(rc);
        return ms;
    }
    py::tuple fftGrid() {
        int l[3] = {0, 0, 0};
        double ms = 0.0;
        check(tdda_model_grid(m_, l, &ms));
        return py::make_tuple(py::make_tuple(l[0], l[1], l[2]), ms);
    }

   private:
    tdda_model* m_ = nullptr;
    long nfpara_ = 0, ndip_ = 0;
};

}  // namespace

PYBIND11_MODULE(_dda_model, module) {
    py::class_<PyDDAModel, std::unique_ptr<PyDDAModel>>(module, "DDAModel")
        .def(py::init<double, double, double, const std::string&, const std::vector<int>&, const std::vector<double>&,
                      bool, const std::string&, const std::vector<double>&, bool, const std::string&,
                      const std::vector<double>&, py::handle, py::handle, int, int, int, int, py::handle, double,
                      py::handle, double, py::handle, double, int, int, double, double, const std::string&, double,
                      py::handle, py::handle, double, bool>(),
             py::arg("betaMin"), py::arg("betaMax"), py::arg("ita"), py::arg("betaType"),
             py::arg("filterIterations"), py::arg("filterRadii"), py::arg("useFilter"), py::arg("symmetryType"),
             py::arg("symmetryAxis"), py::arg("isPeriodic"), py::arg("objectiveName"), py::arg("objPara"),
             py::arg("geometry"), py::arg("inputDielectrics"), py::arg("num_x"), py::arg("num_y"),
             py::arg("num_z"), py::arg("num_total"), py::arg("n_K"), py::arg("E0"), py::arg("n_E0"),
             py::arg("lam"), py::arg("material"), py::arg("nback"), py::arg("MAXm"), py::arg("MAXn"),
             py::arg("Lm"), py::arg("Ln"), py::arg("matrixAMethodName"), py::arg("d"),
             py::arg("sineIntegralValues"), py::arg("cosineIntegralValues"), py::arg("integralDelta"),
             py::arg("verbose"))
        .def("calculateObjective", &PyDDAModel::calculateObjective)
        .def("calculateGradients", &PyDDAModel::calculateGradients)
        .def("solveElectricField", &PyDDAModel::solveElectricField)
        .def("setParameters", &PyDDAModel::setParameters)
        .def("getParameters", &PyDDAModel::getParameters)
        .def("getDielectrics", &PyDDAModel::getDielectrics)
        .def("getElectricField", &PyDDAModel::getElectricField)
        .def("setSolverPrecision", &PyDDAModel::setSolverPrecision)
        .def("setGradientMode", &PyDDAModel::setGradientMode)
        .def("solveInfo", &PyDDAModel::solveInfo, py::arg("which") = 0)
        .def("getPolarization", &PyDDAModel::getPolarization)
        .def("getInverseAlpha", &PyDDAModel::getInverseAlpha)
        .def("applyOperator", &PyDDAModel::applyOperator)
        .def("fftGrid", &PyDDAModel::fftGrid)
        .def("solveResident", &PyDDAModel::solveResident)
        .def("timePasses", &PyDDAModel::timePasses, py::arg("reps") = 20);
    module.def("device_bytes", []() { return tdda_device_bytes(); });
}

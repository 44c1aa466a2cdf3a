// DDAModel: the reference's model/solver/objective/gradient semantics
// (src/DDAModel.cpp, src/ObjDDAModel.cpp, src/Tools.cpp) over the gfx950 runtime.
#pragma once

#include <complex>
#include <string>
#include <vector>

#include "model_obj.h"
#include "runtime.h"

namespace tdda {

using cd = std::complex<double>;

struct ModelConfig {
    std::string symmetry;          // "None" | "4fold"
    std::vector<double> symaxis;
    std::string objective;         // name is accepted, IntegratedE is always built (DDAModel.cpp:23-35)
    std::vector<double> objpara;
    std::vector<int> geometry;     // 3N
    std::vector<double> input_diel;   // 3N
    int nx, ny, nz, n;
    double nk[3], e0, ne0[3];
    double lam;
    std::vector<cd> material;
    double nback;
    int maxm, maxn;
    double lm, ln;                 // in cells (multiplied by d for the images)
    std::string method;            // "FCD" | "LDR"
    double d;
    const double* si = nullptr;
    const double* ci = nullptr;
    long nsici = 0;
    double delta = 0.1;
    bool verbose = false;
};

enum Precision : int { PREC_FP64 = 0, PREC_FP32_REFINED = 1 };
enum GradMode : int { GRAD_REFERENCE = 0, GRAD_EXACT = 1 };

// 1/alpha for the LDR polarizability, Tools.cpp:353-366
cd ldr_inverse_alpha(double K, double d, cd eps, const double* ne0, const double* nk);
// d(1/alpha)/d eps in closed form
cd ldr_inverse_alpha_deps(double K, double d, cd eps, const double* ne0, const double* nk);
// ((xMin <= g) <= xMax) && ... (ObjDDAModel.cpp:135)
bool objective_region(const int* g, const std::vector<double>& objpara);
double c_round(double x);

class Model {
   public:
    explicit Model(const ModelConfig& cfg);

    // bound API (python_bindings/PythonBindings.cpp:61-69)
    void solve_electric_field(const cd* P0, long n, int maxit, double tol);
    double calculate_objective();
    std::vector<double> calculate_gradients(double eps, double origin, int maxit, double tol);
    void set_parameters(const double* v, long n);
    const std::vector<double>& parameters() const { return params_; }
    const std::vector<double>& dielectrics() const { return diel_old_; }
    const std::vector<cd>& electric_field() const { return eresult_; }

    // additions
    void set_precision(int p);
    void set_gradient_mode(int m) { grad_mode_ = m; }
    const SolveResult& last_solve() const { return last_; }
    const SolveResult& last_adjoint() const { return last_adj_; }
    int nfpara() const { return nfpara_; }
    int ndip() const { return N_; }
    const std::vector<cd>& polarization() const { return P_; }
    const std::vector<cd>& inverse_alpha() const { return al_; }
    double wavenumber() const { return K_; }
    double operator_setup_ms() const { return op_setup_ms_; }
    const Operator& op() const { return *op_; }
    bool operator_cache_hit() const { return op_hit_; }
    // Re-solve the forward system from P = 0 with everything already on the
    // device and nothing copied back (timing); returns the solver record.
    SolveResult solve_resident(int maxit, double tol);
    // average ms of the five matvec passes for the current precision
    void time_passes(int reps, double* ms5);
    // y = (diag(al) + A) x in FP64 on the device (tests)
    std::vector<cd> apply(const std::vector<cd>& x);

   private:
    void update_alpha();
    void upload_field_factor(System& s);
    cd eps_of(double x) const;
    // solve from the host right-hand side rhs, or from rhs_dev on the device
    // when rhs is null; P receives the solution, and so does d_P_
    SolveResult run_solve(const std::vector<cd>* rhs, const double2* rhs_dev, const cd* P0, int maxit, double tol,
                          std::vector<cd>& P);
    // w, E of the current solution on the device (objective_terms) and the
    // model's static maps there (upload_obj_maps)
    void objective_terms(hipStream_t st);
    void upload_obj_maps();
    System& system();

    ModelConfig cfg_;
    int N_, nfpara_;
    std::vector<int> gpara_;                 // dipole -> parameter
    std::vector<std::vector<int>> p2g_;      // parameter -> dipoles
    std::vector<double> params_;
    std::vector<double> diel_old_;           // 3N
    std::vector<cd> material_;               // nudged, divided by nback^2
    std::vector<cd> diel_;                   // per dipole
    std::vector<cd> al_;                     // per dipole
    std::vector<cd> einc_;                   // 3N
    std::vector<cd> P_;                      // 3N
    std::vector<cd> eresult_;                // 3N
    std::vector<char> region_;               // per dipole
    int pownum_;
    double K_;
    std::vector<int> box_map_;
    bool identity_map_;
    std::shared_ptr<Operator> op_;
    std::unique_ptr<System> sys64_, sys32_;
    int precision_ = PREC_FP64;
    int grad_mode_ = GRAD_REFERENCE;
    SolveResult last_, last_adj_;
    double op_setup_ms_ = 0.0;
    bool op_hit_ = false;
    bool rhs_on_device_[2] = {false, false};
    // device side of the objective and the gradient partials
    ObjDeviceMaps obj_maps_;
    DevBuf<unsigned char> d_region_;
    DevBuf<int> d_gpara_, d_off_, d_dips_;
    DevBuf<double2> d_dal_;       // exact mode: d(1/alpha)/d rho per dipole
    DevBuf<double2> d_P_;         // the last solution (forward or adjoint), 3N
    DevBuf<double> d_w_, d_part_, d_E_, d_devx_, d_grad_, d_step_;
    DevBuf<double2> d_devp_, d_adevxp_, d_alt_;
};

}  // namespace tdda

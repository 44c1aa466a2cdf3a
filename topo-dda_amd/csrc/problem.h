// A device-resident DDA problem on a full Nx*Ny*Nz box (x-fastest dipoles)
// with one density value per voxel: the unit of work of the batched solver,
// the request server and the autograd path.
#pragma once

#include <complex>
#include <memory>
#include <vector>

#include "runtime.h"

namespace tdda {

struct ProblemDesc {
    int nx, ny, nz;
    double d;           // dipole spacing, nm
    double lam;         // wavelength, nm
    double nback;       // background refractive index
    int method;         // 0 FCD, 1 LDR
    int maxm, maxn;     // image lattice range
    double lm, ln;      // image periods in cells
    double nk[3];       // propagation direction
    double ne0[3];      // polarisation
    double e0;          // amplitude
    double mat[4];      // two permittivities (re, im): background, foreground
};

// Permittivities after AProductCore.cpp:25-33 (divided by nback^2, +0.01 when |m| < 1.01).
void problem_materials(const ProblemDesc& d, double2* m0, double2* m1);
OperatorKey problem_operator_key(const ProblemDesc& d, const SiciDevice& tables);

class Problem {
   public:
    // op may be null: it is then taken from the operator cache (built if
    // absent, through `batch` when one is given: the problem is usable once
    // the batch has been flushed).
    Problem(const ProblemDesc& desc, const SiciDevice& tables, bool refined, hipStream_t st,
            std::shared_ptr<Operator> op = nullptr, std::unique_ptr<System> reuse = nullptr,
            GreenBatch* batch = nullptr);
    // rho: N densities in [0, 1] (clipped), host or device memory; computes
    // 1/alpha (DDAModel.cpp:317-328, Tools.cpp:353-366) and the incident field
    // (DDAModel.cpp:312-316) on the device.
    void set_density(const double* rho, bool on_device, hipStream_t st);
    // Later solves from P = 0 with the incident field as right-hand side (as
    // the server runs them) take that field as BiCGSTAB's shadow residual r0
    // (the initial residual, DDAModel.cpp:928) and evaluate it where it is
    // used instead of loading it.  Not for solves with another right-hand
    // side or a starting guess.
    void use_incident_r0();
    // E = al P 3/(eps + 2) (DDAModel.cpp:1185-1192) into out (device, 3N
    // complex); from_y: the solution vector holds Y = al P (the end of a
    // preconditioned solve whose scaling back was skipped)
    void field(double2* out, hipStream_t st, bool from_y = false) const;
    static size_t estimate_bytes(const ProblemDesc& d, bool refined, bool need_operator);

    ProblemDesc desc;
    double K;
    double2 m0, m1;
    int ndip;
    std::shared_ptr<Operator> op;
    std::unique_ptr<System> sys;
    DevBuf<double> rho;
    SiciDevice tables;   // FCD tables (device), for field evaluations off the grid
    bool op_hit = false;
};

// set_density for a round of problems with one upload and one launch: the
// densities (host) and a table of the problems' set-up parameters go through
// a pinned stage into a device buffer, and one kernel sets every problem's
// density, 1/alpha, incident field and efac and clears its solution vector.
// The caller must not run a round on the same RoundSetup before the previous
// round's kernel has completed (the stage and the device buffer are reused).
class RoundSetup {
   public:
    void run(const std::vector<Problem*>& probs, const std::vector<const double*>& rho, hipStream_t st);

   private:
    PinnedStage stage_;
    DevBuf<char> dev_;
};

}  // namespace tdda

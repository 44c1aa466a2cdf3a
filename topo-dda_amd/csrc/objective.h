// Objectives of a solved problem and their adjoint gradients (device).
#pragma once

#include <vector>

#include "problem.h"

namespace tdda {

enum ObjectiveKind : int {
    OBJ_INTEGRATED = 0,   // log sum |al P|^p over all dipoles (ObjDDAModel.cpp:130-212)
    OBJ_TARGETS = 1,      // mean |E|^2 over external target points
};

struct ObjectiveSpec {
    int kind = OBJ_INTEGRATED;
    int power = 2;
    std::vector<double> targets;   // 3T coordinates in nm (OBJ_TARGETS)
};

// value of the objective for the current solution P (fields: E at the targets)
double objective_value(Problem& p, const ObjectiveSpec& o, hipStream_t st, std::vector<double2>* fields = nullptr);
// d objective / d rho (N values) by one adjoint solve; grad is a device or host
// pointer.  The forward P and right-hand side are restored afterwards.
SolveResult objective_gradient(Problem& p, const ObjectiveSpec& o, int maxit, double tol, double* grad,
                               bool on_device, hipStream_t st, double* value);

}  // namespace tdda

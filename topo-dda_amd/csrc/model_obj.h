// Device objective and explicit gradient partials of the compat model
// (model_obj.hip).
#pragma once

#include "common.h"

namespace tdda {

struct ObjElementParams {
    int n;         // dipoles
    int nfpara;    // free parameters
    int p;         // power of |al P|
    int exact;     // 1: closed-form partials, 0: the reference's one-sided differences
    double eps;    // difference step (reference mode)
    double origin; // objective offset of the reference's differences
};

// Static maps of a model on the device: objective-region mask per dipole,
// dipole -> parameter, and parameter -> dipoles (CSR, off[nfpara + 1]).
struct ObjDeviceMaps {
    const unsigned char* region = nullptr;
    const int* gpara = nullptr;
    const int* off = nullptr;
    const int* dips = nullptr;
};

// w_k = |al_i P_k|^p and E = sum of the region's w (part: scratch of
// ceil(n/256) doubles; E: one double)
void launch_obj_terms(int n, int p, const double2* al, const double2* P, const unsigned char* region, double* w,
                      double* part, double* E, hipStream_t st);
// devp, adevxp (3N) and devx (nfpara) of the mode; dal (exact mode, per
// dipole), alt and step (reference mode, per parameter)
void launch_obj_partials(const ObjElementParams& A, const ObjDeviceMaps& m, const double2* al, const double2* P,
                         const double* w, const double* E, const double2* dal, const double2* alt,
                         const double* step, double2* devp, double2* adevxp, double* devx, hipStream_t st);
// grad_j = devx_j - Re sum over the parameter's dipoles of lambda_k adevxp_k
void launch_obj_assemble(int nfpara, const ObjDeviceMaps& m, const double* devx, const double2* lam,
                         const double2* adevxp, double* grad, hipStream_t st);

}  // namespace tdda

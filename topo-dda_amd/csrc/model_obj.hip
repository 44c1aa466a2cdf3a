// IntegratedE objective and the explicit partials of the compat model's
// gradient on the device (ObjDDAModel.cpp:103-222, DDAModel.cpp:398-578).
//
// w_k = |al_i P_k|^p for the 3N elements; E = sum of w over the dipoles of
// the objective region; f = log E.  The gradient's explicit parts:
//   exact mode      devp_k  = p |w|^(p-2) conj(al P) al / E (region), 0 outside
//                   devx_j  = sum over the parameter's region dipoles of
//                             p |w|^(p-2) Re(conj(al P) P dal) / E
//                   adevxp_k = dal_i P_k
//   reference mode  one-sided differences of log E (DDAModel.cpp:471-578):
//                   devx_j  = (log E - origin + log1p(D_j / E)) / step_j with
//                             D_j = sum (|alt_j P|^p - w) over the parameter's
//                             region dipoles, alt_j the 1/alpha of the
//                             perturbed parameter
//                   devp_k  = ((b + log1p(dr/E))/eps, -(b + log1p(di/E))/eps),
//                             b = log E - origin, dr/di the changes of w for
//                             P_k + eps and P_k + i eps (region only)
//                   adevxp_k = (alt_j - al_i) / step_j P_k
// and after the adjoint solve lambda = S^-1 devp:
//   grad_j = devx_j - Re sum over the parameter's dipoles of lambda_k adevxp_k.
// Per-parameter sums run over the parameter's dipoles in a fixed order (the
// CSR map), one thread per parameter.
#include "model_obj.h"

namespace tdda {

namespace {

constexpr int OBJ_THREADS = 256;

__device__ inline double mag(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }

__device__ inline double pow_p(double x, int p) {
    double r = 1.0;
    for (int k = 0; k < p; ++k) r *= x;
    return r;
}

// p |w|^(p-2) with |w| = aw (0 for aw = 0 unless p = 2)
__device__ inline double dpow_scale(double aw, int p) {
    if (p == 2) return 2.0;
    return aw > 0.0 ? p * pow_p(aw, p - 2) : 0.0;
}

// w_k and the block sums of the region's w (fixed order within a block)
__global__ void __launch_bounds__(OBJ_THREADS)
k_obj_terms(int n, int p, const double2* __restrict__ al, const double2* __restrict__ P,
            const unsigned char* __restrict__ region, double* __restrict__ w, double* __restrict__ part) {
    __shared__ double red[OBJ_THREADS];
    const int i = blockIdx.x * OBJ_THREADS + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        for (int c = 0; c < 3; ++c) {
            const double v = pow_p(mag(cmul(al[i], P[3 * i + c])), p);
            w[3 * i + c] = v;
            if (region[i]) s += v;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = OBJ_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// E = sum of the block sums in block order (one block)
__global__ void __launch_bounds__(OBJ_THREADS) k_obj_total(int nb, const double* __restrict__ part,
                                                          double* __restrict__ E) {
    __shared__ double red[OBJ_THREADS];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += OBJ_THREADS) s += part[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = OBJ_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *E = red[0];
}

// per element: devp and adevxp
__global__ void k_obj_elements(ObjElementParams A, const double2* __restrict__ al, const double2* __restrict__ P,
                               const unsigned char* __restrict__ region, const double* __restrict__ w,
                               const double* __restrict__ E, const int* __restrict__ gpara,
                               const double2* __restrict__ dal, const double2* __restrict__ alt,
                               const double* __restrict__ step, double2* __restrict__ devp,
                               double2* __restrict__ adevxp) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 3L * A.n) return;
    const int i = (int)(k / 3);
    const double eint = *E;
    const double2 pk = P[k];
    const double2 a = al[i];
    if (A.exact) {
        adevxp[k] = cmul(dal[i], pk);
        double2 d = make_double2(0.0, 0.0);
        if (region[i]) {
            const double2 wk = cmul(a, pk);
            const double aw = mag(wk);
            if (!(aw == 0.0 && A.p < 2)) d = cscale(cmul(cconj(wk), a), dpow_scale(aw, A.p) / eint);
        }
        devp[k] = d;
    } else {
        const int j = gpara[i];
        const double2 dd = csub(alt[j], a);
        const double2 change = make_double2(dd.x / step[j], dd.y / step[j]);
        adevxp[k] = cmul(change, pk);
        const double base = log(eint) - A.origin;
        double dr = 0.0, di = 0.0;
        if (region[i]) {
            dr = pow_p(mag(cmul(a, make_double2(pk.x + A.eps, pk.y))), A.p) - w[k];
            di = pow_p(mag(cmul(a, make_double2(pk.x, pk.y + A.eps))), A.p) - w[k];
        }
        const double ar = (base + log1p(dr / eint)) / A.eps;
        const double ai = (base + log1p(di / eint)) / A.eps;
        devp[k] = make_double2(ar, -ai);
    }
}

// per parameter: the explicit partial devx_j
__global__ void k_obj_params(ObjElementParams A, const double2* __restrict__ al, const double2* __restrict__ P,
                             const unsigned char* __restrict__ region, const double* __restrict__ w,
                             const double* __restrict__ E, const int* __restrict__ off,
                             const int* __restrict__ dips, const double2* __restrict__ dal,
                             const double2* __restrict__ alt, const double* __restrict__ step,
                             double* __restrict__ devx) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.nfpara) return;
    const double eint = *E;
    double acc = 0.0;
    if (A.exact) {
        for (int q = off[j]; q < off[j + 1]; ++q) {
            const int i = dips[q];
            if (!region[i]) continue;
            const double2 a = al[i];
            for (int c = 0; c < 3; ++c) {
                const double2 pk = P[3 * i + c];
                const double2 wk = cmul(a, pk);
                const double aw = mag(wk);
                if (aw == 0.0) continue;
                acc += dpow_scale(aw, A.p) * cmul(cmul(cconj(wk), pk), dal[i]).x / eint;
            }
        }
        devx[j] = acc;
    } else {
        const double2 at = alt[j];
        for (int q = off[j]; q < off[j + 1]; ++q) {
            const int i = dips[q];
            if (!region[i]) continue;
            for (int c = 0; c < 3; ++c) acc += pow_p(mag(cmul(at, P[3 * i + c])), A.p) - w[3 * i + c];
        }
        const double base = log(eint) - A.origin;
        devx[j] = (base + log1p(acc / eint)) / step[j];
    }
}

// per parameter: grad_j = devx_j - Re sum lambda_k adevxp_k
__global__ void k_obj_assemble(int nfpara, const int* __restrict__ off, const int* __restrict__ dips,
                               const double* __restrict__ devx, const double2* __restrict__ lam,
                               const double2* __restrict__ adevxp, double* __restrict__ grad) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nfpara) return;
    double2 m = make_double2(0.0, 0.0);
    for (int q = off[j]; q < off[j + 1]; ++q) {
        const int i = dips[q];
        for (int c = 0; c < 3; ++c) m = cadd(m, cmul(lam[3 * i + c], adevxp[3 * i + c]));
    }
    grad[j] = devx[j] - m.x;
}

}  // namespace

void launch_obj_terms(int n, int p, const double2* al, const double2* P, const unsigned char* region, double* w,
                      double* part, double* E, hipStream_t st) {
    const int nb = (n + OBJ_THREADS - 1) / OBJ_THREADS;
    hipLaunchKernelGGL(k_obj_terms, dim3(nb), dim3(OBJ_THREADS), 0, st, n, p, al, P, region, w, part);
    hipLaunchKernelGGL(k_obj_total, dim3(1), dim3(OBJ_THREADS), 0, st, nb, part, E);
    TDDA_HIP(hipGetLastError());
}

void launch_obj_partials(const ObjElementParams& A, const ObjDeviceMaps& m, const double2* al, const double2* P,
                         const double* w, const double* E, const double2* dal, const double2* alt,
                         const double* step, double2* devp, double2* adevxp, double* devx, hipStream_t st) {
    const long n3 = 3L * A.n;
    hipLaunchKernelGGL(k_obj_elements, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, A, al, P, m.region, w,
                       E, m.gpara, dal, alt, step, devp, adevxp);
    hipLaunchKernelGGL(k_obj_params, dim3((A.nfpara + 255) / 256), dim3(256), 0, st, A, al, P, m.region, w, E, m.off,
                       m.dips, dal, alt, step, devx);
    TDDA_HIP(hipGetLastError());
}

void launch_obj_assemble(int nfpara, const ObjDeviceMaps& m, const double* devx, const double2* lam,
                         const double2* adevxp, double* grad, hipStream_t st) {
    hipLaunchKernelGGL(k_obj_assemble, dim3((nfpara + 255) / 256), dim3(256), 0, st, nfpara, m.off, m.dips, devx, lam,
                       adevxp, grad);
    TDDA_HIP(hipGetLastError());
}

}  // namespace tdda

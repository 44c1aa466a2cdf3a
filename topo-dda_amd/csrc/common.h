// Common device/host definitions for the gfx950 DDA solve path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace tdda {

// ---------------------------------------------------------------------------
// complex helpers on HIP vector types (float2 / double2)
// ---------------------------------------------------------------------------
template <typename R> struct cx;
template <> struct cx<float> { using type = float2; };
template <> struct cx<double> { using type = double2; };
template <typename R> using cx_t = typename cx<R>::type;

template <typename C> __host__ __device__ inline C mk(double re, double im) {
    C c; c.x = re; c.y = im; return c;
}
template <typename C> __host__ __device__ inline C cadd(C a, C b) { C c; c.x = a.x + b.x; c.y = a.y + b.y; return c; }
template <typename C> __host__ __device__ inline C csub(C a, C b) { C c; c.x = a.x - b.x; c.y = a.y - b.y; return c; }
template <typename C> __host__ __device__ inline C cmul(C a, C b) {
    C c; c.x = a.x * b.x - a.y * b.y; c.y = a.x * b.y + a.y * b.x; return c;
}
// a * conj(b)
template <typename C> __host__ __device__ inline C cmulc(C a, C b) {
    C c; c.x = a.x * b.x + a.y * b.y; c.y = a.y * b.x - a.x * b.y; return c;
}
template <typename C, typename S> __host__ __device__ inline C cscale(C a, S s) { C c; c.x = a.x * s; c.y = a.y * s; return c; }
// multiply by -i (forward) or +i (inverse): DIR = -1 -> -i
template <int DIR, typename C> __host__ __device__ inline C mul_i(C a) {
    C c;
    if (DIR < 0) { c.x = a.y; c.y = -a.x; } else { c.x = -a.y; c.y = a.x; }
    return c;
}
template <typename C> __host__ __device__ inline C cconj(C a) { C c; c.x = a.x; c.y = -a.y; return c; }

__host__ __device__ inline double2 to_d(float2 a) { return make_double2(a.x, a.y); }
__host__ __device__ inline double2 to_d(double2 a) { return a; }
template <typename C> __host__ __device__ inline C from_d(double2 a);
template <> __host__ __device__ inline double2 from_d<double2>(double2 a) { return a; }
template <> __host__ __device__ inline float2 from_d<float2>(double2 a) { return make_float2((float)a.x, (float)a.y); }

__host__ __device__ inline double2 dcmul(double2 a, double2 b) { return cmul(a, b); }
__host__ __device__ inline double2 ddiv(double2 a, double2 b) {
    double den = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}

#if defined(__HIPCC__)
// Incident plane wave at cell (gx, gy, gz) (DDAModel.cpp:312-316): the
// phase factor exp(i kd (nk . g)) with kd = K d, and a component
// en exp(i ...) with en = e0 n_e0[c].  Every product and sum is rounded on
// its own (no contraction), so each kernel that evaluates them gets the same
// bits.
__device__ inline double2 incident_phase(double kd, const double (&nk)[3], int gx, int gy, int gz) {
    const double q = __dadd_rn(__dadd_rn(__dmul_rn(nk[0], (double)gx), __dmul_rn(nk[1], (double)gy)),
                               __dmul_rn(nk[2], (double)gz));
    double s, c;
    sincos(__dmul_rn(kd, q), &s, &c);
    return make_double2(c, s);
}
__device__ inline double2 incident_wave(double2 phase, double en) {
    return make_double2(__dmul_rn(en, phase.x), __dmul_rn(en, phase.y));
}
#endif

// ---------------------------------------------------------------------------
// error handling (host)
// ---------------------------------------------------------------------------
void set_last_error(const std::string& msg);
const char* get_last_error();

struct Error : public std::exception {
    int code;
    std::string msg;
    Error(int c, std::string m) : code(c), msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

enum Status : int {
    ST_OK = 0,
    ERR_INVALID = 1,       // bad argument (reference: throw 1 / invalid_argument)
    ERR_METHOD = 2,        // unknown interaction method (reference: std::invalid_argument)
    ERR_HIP = 3,           // HIP runtime failure
    ERR_MEMORY = 4,        // device memory budget exceeded
    ERR_STATE = 5,         // call order violated
};

#define TDDA_HIP(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::tdda::Error(::tdda::ERR_HIP, std::string("HIP error ") +          \
                                hipGetErrorString(e_) + " at " + __FILE__ + ":" +          \
                                std::to_string(__LINE__) + " in " #call);                  \
    } while (0)

#define TDDA_REQUIRE(cond, code, msg)                                                      \
    do {                                                                                   \
        if (!(cond)) throw ::tdda::Error((code), (msg));                                   \
    } while (0)

}  // namespace tdda

// Shared device helpers for the gfx950 kernels: complex arithmetic on double2 / float2,
// launch-error checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace xtr {

template <typename T> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };

template <typename T> using cplx = typename vec2<T>::type;

template <typename T> __host__ __device__ inline cplx<T> mk(T re, T im) { cplx<T> r; r.x = re; r.y = im; return r; }
template <typename T> __host__ __device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __host__ __device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b)
{
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// acc + a*b
template <typename T> __host__ __device__ inline cplx<T> cfma(cplx<T> acc, cplx<T> a, cplx<T> b)
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
    return acc;
}
// acc - a*b
template <typename T> __host__ __device__ inline cplx<T> cfms(cplx<T> acc, cplx<T> a, cplx<T> b)
{
    acc.x = fma(-a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(-a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
    return acc;
}
// acc + conj(a)*b
template <typename T> __host__ __device__ inline cplx<T> cfma_conj(cplx<T> acc, cplx<T> a, cplx<T> b)
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
    return acc;
}
template <typename T> __host__ __device__ inline cplx<T> cscale(cplx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
template <typename T> __host__ __device__ inline cplx<T> cconj(cplx<T> a) { return mk<T>(a.x, -a.y); }

// (double2 <-> float2) conversions used by the mixed-precision paths
__host__ __device__ inline float2 to_f(double2 a) { return make_float2((float)a.x, (float)a.y); }
__host__ __device__ inline double2 to_d(float2 a) { return make_double2((double)a.x, (double)a.y); }
template <typename T> __host__ __device__ inline cplx<T> from_d(double2 a);
template <> __host__ __device__ inline double2 from_d<double>(double2 a) { return a; }
template <> __host__ __device__ inline float2 from_d<float>(double2 a) { return to_f(a); }

struct hip_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw hip_error(std::string(what) + ": " + hipGetErrorString(e));
}

#define XTR_CHECK(call) ::xtr::check((call), #call)
#define XTR_LAUNCH_CHECK(name) ::xtr::check(hipGetLastError(), name)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace xtr

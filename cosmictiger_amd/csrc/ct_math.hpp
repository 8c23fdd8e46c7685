// Scalar math helpers shared by host C++ (double, tree build) and gfx950
// device code (float, kick kernels).  Device versions call the ROCm device
// library; host versions call libm.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CT_HD __host__ __device__
#else
#define CT_HD
#endif

namespace ctfmm {

// Lane select and "any lane" test; for scalar types a plain conditional.
// Short-vector types (the CPU SIMD kick) provide their own overloads.
template<class T>
CT_HD inline T ct_select(bool c, T a, T b) {
	return c ? a : b;
}
CT_HD inline bool ct_any(bool c) {
	return c;
}

#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline float ct_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double ct_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ inline float ct_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double ct_rsqrt(double x) { return 1.0 / sqrt(x); }
__device__ inline float ct_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ inline double ct_sqrt(double x) { return sqrt(x); }
__device__ inline float ct_exp(float x) { return __expf(x); }
__device__ inline double ct_exp(double x) { return exp(x); }
__device__ inline float ct_erfc(float x) { return erfcf(x); }
__device__ inline double ct_erfc(double x) { return erfc(x); }
__device__ inline void ct_sincos(float x, float* s, float* c) { sincosf(x, s, c); }
__device__ inline void ct_sincos(double x, double* s, double* c) { sincos(x, s, c); }
#else
CT_HD inline float ct_fma(float a, float b, float c) { return std::fma(a, b, c); }
CT_HD inline double ct_fma(double a, double b, double c) { return std::fma(a, b, c); }
CT_HD inline float ct_rsqrt(float x) { return 1.0f / std::sqrt(x); }
CT_HD inline double ct_rsqrt(double x) { return 1.0 / std::sqrt(x); }
CT_HD inline float ct_sqrt(float x) { return std::sqrt(x); }
CT_HD inline double ct_sqrt(double x) { return std::sqrt(x); }
CT_HD inline float ct_exp(float x) { return std::exp(x); }
CT_HD inline double ct_exp(double x) { return std::exp(x); }
CT_HD inline float ct_erfc(float x) { return std::erfc(x); }
CT_HD inline double ct_erfc(double x) { return std::erfc(x); }
CT_HD inline void ct_sincos(float x, float* s, float* c) { *s = std::sin(x); *c = std::cos(x); }
CT_HD inline void ct_sincos(double x, double* s, double* c) { *s = std::sin(x); *c = std::cos(x); }
#endif

}  // namespace ctfmm

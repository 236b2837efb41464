// mep_common.hpp -- shared device helpers for the MEPinterp k-space kernels
// (gfx950 / CDNA4, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MEP_WAVE 64

// Complex FP64 value, interleaved (re, im): binary-compatible with Fortran
// complex(c_double_complex) and C double _Complex.
struct zd {
  double re, im;
};

__host__ __device__ inline zd zmk(double r, double i) { return zd{r, i}; }
__host__ __device__ inline zd zadd(zd a, zd b) { return zd{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline zd zsub(zd a, zd b) { return zd{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline zd zmul(zd a, zd b) {
  return zd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// conj(a) * b
__host__ __device__ inline zd zcmul(zd a, zd b) {
  return zd{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
__host__ __device__ inline zd zconj(zd a) { return zd{a.re, -a.im}; }
__host__ __device__ inline zd zscale(zd a, double s) { return zd{a.re * s, a.im * s}; }
__host__ __device__ inline zd zmuli(zd a) { return zd{-a.im, a.re}; }  // i*a
__host__ __device__ inline double zabs2(zd a) { return a.re * a.re + a.im * a.im; }
// One 16-byte access of a complex value at a 16-byte aligned address.  zd
// is only 8-byte aligned, so a plain zd load becomes two 8-byte accesses;
// over LDS at a 16-byte lane stride each of those is a 2-way bank conflict,
// where one 16-byte access per lane is conflict-free.
typedef double mep_d2 __attribute__((ext_vector_type(2)));
__device__ inline zd ldz(const zd* p) {
  const mep_d2 v = *reinterpret_cast<const mep_d2*>(p);
  return zd{v.x, v.y};
}
__device__ inline void stz(zd* p, zd a) {
  mep_d2 v;
  v.x = a.re;
  v.y = a.im;
  *reinterpret_cast<mep_d2*>(p) = v;
}

// a / b  (textbook formula; denominators here are well scaled)
__host__ __device__ inline zd zdiv(zd a, zd b) {
  double d = b.re * b.re + b.im * b.im;
  return zd{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// Complex FP32 value, interleaved (re, im).
struct zf {
  float re, im;
};
__host__ __device__ inline zf zadd(zf a, zf b) { return zf{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline zf zsub(zf a, zf b) { return zf{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline zf zmul(zf a, zf b) {
  return zf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline zf zcmul(zf a, zf b) {
  return zf{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
__host__ __device__ inline zf zconj(zf a) { return zf{a.re, -a.im}; }
__host__ __device__ inline zf zscale(zf a, float s) { return zf{a.re * s, a.im * s}; }
__host__ __device__ inline zf zmuli(zf a) { return zf{-a.im, a.re}; }
__host__ __device__ inline float zabs2(zf a) { return a.re * a.re + a.im * a.im; }

// complex type and machine epsilon of a real type
template <class R>
struct cplx_of;
template <>
struct cplx_of<double> {
  using type = zd;
  static constexpr double eps = 2.220446049250313e-16;
};
template <>
struct cplx_of<float> {
  using type = zf;
  static constexpr float eps = 1.1920929e-07f;
};
// reciprocal of a normal, nonzero x: hardware estimate refined by Newton
// steps (a shorter dependency chain than the IEEE division sequence)
__device__ inline double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}
// 1/sqrt(x) for normal x > 0: the hardware estimate and two Newton steps
__device__ inline double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}
// sqrt(x), x >= 0: x / sqrt(x) from rsq_nr in the normal range (about half
// the instructions of the IEEE sequence), the IEEE sqrt below it and at 0
__device__ inline double sqrt_nr(double x) {
  if (x > 1e-290) return x * rsq_nr(x);
  return sqrt(x);
}
__device__ inline float rcp_nr(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  r = fmaf(r, fmaf(-x, r, 1.0f), r);
  return r;
}
__device__ inline void sin_cos(double a, double* s, double* c) { sincos(a, s, c); }
__device__ inline void sin_cos(float a, float* s, float* c) { sincosf(a, s, c); }

// Wave-wide FP64 sum (all 64 lanes receive the result).
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, MEP_WAVE);
  return v;
}

// Index helpers in the reference's column-major layout (0-based).
__host__ __device__ inline int64_t ix3(int a, int n, int m, int N) {
  return a + 3 * ((int64_t)n + (int64_t)N * m);
}
__host__ __device__ inline int64_t ix33(int a, int b, int n, int m, int N) {
  return a + 3 * (b + 3 * ((int64_t)n + (int64_t)N * m));
}

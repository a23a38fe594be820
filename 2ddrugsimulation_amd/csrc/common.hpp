// common.hpp -- shared device helpers for the cardiac cell kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CEP_HD __host__ __device__ __forceinline__
#define CEP_D __device__ __forceinline__

namespace cep {

// Per-cell arrays are structure-of-arrays in HBM: element k of cell c lives at
// base[k * pitch + c], pitch = cell count rounded up to a multiple of 64, so the 64
// lanes of a wavefront touch one contiguous 512-byte line per element.
template <typename T>
struct Soa {
  T *base;
  int pitch;
  CEP_D T &operator()(int k, int c) const { return base[(size_t)k * pitch + c]; }
};

// Precision traits for the rate/gating arithmetic.  double: IEEE division and the
// ocml exp/log.  float: hardware transcendental (v_exp_f32 / v_log_f32 through the
// fast intrinsics) and reciprocal-multiply division on the hardware reciprocal
// (v_rcp_f32, 1 ulp): two instructions where the IEEE-rounded quotient (what __fdividef and
// __frcp_rn compile to here) is a dependent sequence of about ten; state stays in double.
template <typename R>
struct M;

template <>
struct M<double> {
  static CEP_D double ex(double x) { return exp(x); }
  // e^(x s) for a constant scale s, as exp(x * s)
  static CEP_D double exs(double x, double s) { return exp(x * s); }
  static CEP_D double lg(double x) { return log(x); }
  static CEP_D double dv(double a, double b) { return a / b; }
  static CEP_D double rcp(double a) { return 1.0 / a; }
  static CEP_D double pw(double a, double b) { return pow(a, b); }
  static CEP_D double sq(double a) { return sqrt(a); }
};

template <>
struct M<float> {
  static CEP_D float ex(float x) { return __expf(x); }
  // e^(x s) for a constant scale s: 2^(x (s log2 e)), the two constants folded into one
  // multiply ahead of v_exp_f32 (__expf(x * s) multiplies by s, then by log2 e)
  static CEP_D float exs(float x, float s) { return __builtin_amdgcn_exp2f(x * (s * 1.44269504088896341f)); }
  static CEP_D float lg(float x) { return __logf(x); }
  static CEP_D float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  static CEP_D float rcp(float a) { return __builtin_amdgcn_rcpf(a); }
  static CEP_D float pw(float a, float b) { return __powf(a, b); }
  static CEP_D float sq(float a) { return __fsqrt_rn(a); }
};

// per-cell run status; ST_IDLE: a parked cell (no job in it), which every run skips
enum : int { ST_OK = 0, ST_CONV_FAIL = 1, ST_ERR_FAIL = 2, ST_NONFINITE = 3, ST_MAX_STEPS = 4, ST_IDLE = 5 };

}  // namespace cep

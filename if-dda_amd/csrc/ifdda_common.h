// ifdda_common.h -- shared device helpers and launch descriptors for the
// gfx950 DDA solve library (libifdda_hip.so).  Complex values are float2 /
// double2 (re, im), the same bytes as Fortran COMPLEX / DOUBLE COMPLEX.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ifdda {

template <typename R> struct cx;
template <> struct cx<float> { using T = float2; };
template <> struct cx<double> { using T = double2; };

__device__ __forceinline__ float2 mk(float a, float b) { return make_float2(a, b); }
__device__ __forceinline__ double2 mk(double a, double b) { return make_double2(a, b); }

template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return mk(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return mk(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
template <typename V> __device__ __forceinline__ V cmulc(V a, V b) {
  return mk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
template <typename V, typename S> __device__ __forceinline__ V cscal(V a, S s) { return mk(a.x * s, a.y * s); }
// a * (sgn * i)
template <int SGN, typename V> __device__ __forceinline__ V mul_si(V a) {
  return SGN > 0 ? mk(-a.y, a.x) : mk(a.y, -a.x);
}
template <typename V> __device__ __forceinline__ V czero() { return mk(decltype(V::x)(0), decltype(V::x)(0)); }

// Mixed-radix plan of one FFT axis of length L: stage radices in
// execution order (Stockham, product of radices == L).
constexpr int kMaxStages = 12;
struct AxisPlan {
  int L;
  int nst;
  int rad[kMaxStages];
};

// Polarisability storage modes for the scatter pass.
enum AlphaMode : int { kAlphaNone = 0, kAlphaScalar = 1, kAlphaTensor = 2 };

}  // namespace ifdda

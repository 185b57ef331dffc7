// Dual numbers for the tangent-linear step: value + one directional
// derivative, with branch-frozen semantics -- comparisons, min/max, integer
// conversions and table indices act on the value only, so every threshold
// takes the forward branch and the derivative is that branch's derivative.
//
// Instantiating the step kernels with R = Dual<double> gives the FP64
// tangent-linear model of exactly the arithmetic the FP64 twin performs
// (same operations on the value parts, so identical values).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace tempo {

template <class T>
struct Dual {
  T v, d;
  __host__ __device__ constexpr Dual() : v(0), d(0) {}
  __host__ __device__ constexpr Dual(T x) : v(x), d(0) {}  // constants carry no derivative
  __host__ __device__ constexpr Dual(T x, T dx) : v(x), d(dx) {}
  // narrowing/widening between dual precisions keeps the derivative
  template <class U>
  __host__ __device__ explicit constexpr Dual(const Dual<U>& o) : v(T(o.v)), d(T(o.d)) {}
  // integer conversions take the value (branch-frozen)
  __host__ __device__ explicit operator int() const { return int(v); }
  __host__ __device__ explicit operator unsigned() const { return unsigned(v); }
};

// scalar type of a working precision, its wide (FP64 accumulation) type and
// the type the sedimented state and tendencies are stored in between kernels:
// the working precision for a plain step (the values are working-precision
// numbers when phase A writes them and are rounded to it once, when the
// sedimentation writes its tendencies back), FP64 duals for derivatives
template <class R>
struct Num {
  using S = R;
  using W = double;
  using D = R;
};
template <class T>
struct Num<Dual<T>> {
  using S = T;
  using W = Dual<double>;
  using D = Dual<double>;
};
template <class R>
using S_t = typename Num<R>::S;
template <class R>
using W_t = typename Num<R>::W;
template <class R>
using Sd_t = typename Num<R>::D;
template <class R>
struct IsDual {
  static constexpr bool value = false;
};
template <class T>
struct IsDual<Dual<T>> {
  static constexpr bool value = true;
};

template <class R>
__host__ __device__ __forceinline__ S_t<R> value(const R& x) { return x; }
template <class T>
__host__ __device__ __forceinline__ T value(const Dual<T>& x) { return x.v; }

// conversion to another working/wide type (keeps derivatives between duals)
template <class To, class From>
__host__ __device__ __forceinline__ To conv(const From& x) { return To(x); }
template <class To, class T>
__host__ __device__ __forceinline__ To conv_d(const Dual<T>& x) { return To(x); }

template <class T>
__device__ __forceinline__ Dual<T> operator+(Dual<T> a, Dual<T> b) { return {a.v + b.v, a.d + b.d}; }
template <class T>
__device__ __forceinline__ Dual<T> operator-(Dual<T> a, Dual<T> b) { return {a.v - b.v, a.d - b.d}; }
template <class T>
__device__ __forceinline__ Dual<T> operator-(Dual<T> a) { return {-a.v, -a.d}; }
template <class T>
__device__ __forceinline__ Dual<T> operator*(Dual<T> a, Dual<T> b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
template <class T>
__device__ __forceinline__ Dual<T> operator/(Dual<T> a, Dual<T> b) {
  const T q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
template <class T>
__device__ __forceinline__ Dual<T>& operator+=(Dual<T>& a, Dual<T> b) { return a = a + b; }
template <class T>
__device__ __forceinline__ Dual<T>& operator-=(Dual<T>& a, Dual<T> b) { return a = a - b; }
template <class T>
__device__ __forceinline__ Dual<T>& operator*=(Dual<T>& a, Dual<T> b) { return a = a * b; }
template <class T>
__device__ __forceinline__ Dual<T>& operator/=(Dual<T>& a, Dual<T> b) { return a = a / b; }

// mixed with plain arithmetic scalars (constants)
#define TEMPO_DUAL_MIXED(OP)                                                            \
  template <class T, class U>                                                           \
  __device__ __forceinline__ Dual<T> operator OP(Dual<T> a, U b) { return a OP Dual<T>(T(b)); } \
  template <class T, class U>                                                           \
  __device__ __forceinline__ Dual<T> operator OP(U a, Dual<T> b) { return Dual<T>(T(a)) OP b; }
TEMPO_DUAL_MIXED(+)
TEMPO_DUAL_MIXED(-)
TEMPO_DUAL_MIXED(*)
TEMPO_DUAL_MIXED(/)
#undef TEMPO_DUAL_MIXED

#define TEMPO_DUAL_CMP(OP)                                                                            \
  template <class T>                                                                                  \
  __device__ __forceinline__ bool operator OP(Dual<T> a, Dual<T> b) { return a.v OP b.v; }            \
  template <class T, class U>                                                                         \
  __device__ __forceinline__ bool operator OP(Dual<T> a, U b) { return a.v OP T(b); }                 \
  template <class T, class U>                                                                         \
  __device__ __forceinline__ bool operator OP(U a, Dual<T> b) { return T(a) OP b.v; }
TEMPO_DUAL_CMP(<)
TEMPO_DUAL_CMP(>)
TEMPO_DUAL_CMP(<=)
TEMPO_DUAL_CMP(>=)
TEMPO_DUAL_CMP(==)
TEMPO_DUAL_CMP(!=)
#undef TEMPO_DUAL_CMP

// wavefront permutes of both parts (the plain overloads stay visible)
using ::__shfl;
using ::__shfl_down;
template <class T>
__device__ __forceinline__ Dual<T> __shfl(Dual<T> x, int src, int width = 64) {
  return {__shfl(x.v, src, width), __shfl(x.d, src, width)};
}
template <class T>
__device__ __forceinline__ Dual<T> __shfl_down(Dual<T> x, unsigned delta, int width = 64) {
  return {__shfl_down(x.v, delta, width), __shfl_down(x.d, delta, width)};
}

}  // namespace tempo

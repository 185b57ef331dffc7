// Branch signatures of the level-local physics.
//
// Sig<T> is a pass-through working type: arithmetic and the library
// functions act on the value exactly as on T, while every threshold decision
// the step takes on a value -- each comparison in an `if`, and each
// conversion to an integer (table indices, nint, bin counts) -- is folded
// into a per-lane 32-bit hash.  Instantiating the level-local kernels with
// R = Sig<float> and R = Sig<double> runs the production step and records,
// per point, which branches it took; two runs whose signatures at a point
// agree took the same branches there.  min/max and |x| are continuous in
// their operands, so the selections they make are not recorded.
//
// The hash lives in LDS, one word per thread of the block (at most 256
// threads); a kernel instantiated with Sig opens it with sig_begin() and
// reads it with sig_end().
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace tempo {

__device__ __forceinline__ unsigned& sig_word() {
  __shared__ unsigned acc[256];
  return acc[threadIdx.x];
}
__device__ __forceinline__ void sig_begin() { sig_word() = 2166136261u; }
__device__ __forceinline__ unsigned sig_end() { return sig_word(); }
// FNV-1a step: order-sensitive, so a different sequence of decisions (a
// flipped branch changes which decisions follow) changes the hash too
__device__ __forceinline__ void sig_note(unsigned x) {
  unsigned h = sig_word();
  h = (h ^ (x & 0xffu)) * 16777619u;
  h = (h ^ ((x >> 8) & 0xffu)) * 16777619u;
  h = (h ^ ((x >> 16) & 0xffu)) * 16777619u;
  h = (h ^ (x >> 24)) * 16777619u;
  sig_word() = h;
}
__device__ __forceinline__ bool sig_cmp(bool r) {
  sig_note(r ? 0x5bd1e995u : 0x1b873593u);
  return r;
}

template <class T>
struct Sig {
  T v;
  __host__ __device__ constexpr Sig() : v(0) {}
  __host__ __device__ constexpr Sig(T x) : v(x) {}
  template <class U, class = typename std::enable_if<std::is_arithmetic<U>::value>::type>
  __host__ __device__ constexpr Sig(U x) : v(T(x)) {}
  __device__ explicit operator int() const {
    const int i = int(v);
    sig_note(unsigned(i) ^ 0x2545f491u);
    return i;
  }
  __device__ explicit operator unsigned() const {
    const unsigned i = unsigned(v);
    sig_note(i ^ 0x9e3779b9u);
    return i;
  }
  __host__ __device__ explicit operator T() const { return v; }
};

template <class T>
__device__ __forceinline__ Sig<T> operator+(Sig<T> a, Sig<T> b) { return a.v + b.v; }
template <class T>
__device__ __forceinline__ Sig<T> operator-(Sig<T> a, Sig<T> b) { return a.v - b.v; }
template <class T>
__device__ __forceinline__ Sig<T> operator-(Sig<T> a) { return -a.v; }
template <class T>
__device__ __forceinline__ Sig<T> operator*(Sig<T> a, Sig<T> b) { return a.v * b.v; }
template <class T>
__device__ __forceinline__ Sig<T> operator/(Sig<T> a, Sig<T> b) { return a.v / b.v; }
template <class T>
__device__ __forceinline__ Sig<T>& operator+=(Sig<T>& a, Sig<T> b) { return a = a + b; }
template <class T>
__device__ __forceinline__ Sig<T>& operator-=(Sig<T>& a, Sig<T> b) { return a = a - b; }
template <class T>
__device__ __forceinline__ Sig<T>& operator*=(Sig<T>& a, Sig<T> b) { return a = a * b; }
template <class T>
__device__ __forceinline__ Sig<T>& operator/=(Sig<T>& a, Sig<T> b) { return a = a / b; }

#define TEMPO_SIG_MIXED(OP)                                                                    \
  template <class T, class U, class = typename std::enable_if<std::is_arithmetic<U>::value>::type> \
  __device__ __forceinline__ Sig<T> operator OP(Sig<T> a, U b) { return a OP Sig<T>(T(b)); }  \
  template <class T, class U, class = typename std::enable_if<std::is_arithmetic<U>::value>::type> \
  __device__ __forceinline__ Sig<T> operator OP(U a, Sig<T> b) { return Sig<T>(T(a)) OP b; }
TEMPO_SIG_MIXED(+)
TEMPO_SIG_MIXED(-)
TEMPO_SIG_MIXED(*)
TEMPO_SIG_MIXED(/)
#undef TEMPO_SIG_MIXED

#define TEMPO_SIG_CMP(OP)                                                                             \
  template <class T>                                                                                  \
  __device__ __forceinline__ bool operator OP(Sig<T> a, Sig<T> b) { return sig_cmp(a.v OP b.v); }     \
  template <class T, class U, class = typename std::enable_if<std::is_arithmetic<U>::value>::type>    \
  __device__ __forceinline__ bool operator OP(Sig<T> a, U b) { return sig_cmp(a.v OP T(b)); }         \
  template <class T, class U, class = typename std::enable_if<std::is_arithmetic<U>::value>::type>    \
  __device__ __forceinline__ bool operator OP(U a, Sig<T> b) { return sig_cmp(T(a) OP b.v); }
TEMPO_SIG_CMP(<)
TEMPO_SIG_CMP(>)
TEMPO_SIG_CMP(<=)
TEMPO_SIG_CMP(>=)
TEMPO_SIG_CMP(==)
TEMPO_SIG_CMP(!=)
#undef TEMPO_SIG_CMP

}  // namespace tempo

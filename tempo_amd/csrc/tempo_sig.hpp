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
// Two hashes per point separate the kinds of decision: class 0 the
// comparisons of `if` conditions, class 1 the discrete outcomes -- integer
// conversions (table indices, nint, bin counts) and the checks' presence
// flags, clamps and density index.  They live in LDS, one pair of words per
// thread of the block (at most 256 threads); a kernel instantiated with Sig
// opens them with sig_begin() and reads them with sig_end(cls).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace tempo {

__device__ __forceinline__ unsigned& sig_word(int cls) {
  __shared__ unsigned acc[2][256];
  return acc[cls][threadIdx.x];
}
__device__ __forceinline__ void sig_begin() { sig_word(0) = sig_word(1) = 2166136261u; }
__device__ __forceinline__ unsigned sig_end(int cls) { return sig_word(cls); }
// FNV-1a step: order-sensitive, so a different sequence of decisions (a
// flipped branch changes which decisions follow) changes the hash too
__device__ __forceinline__ void sig_mix(int cls, unsigned x) {
  unsigned h = sig_word(cls);
  h = (h ^ (x & 0xffu)) * 16777619u;
  h = (h ^ ((x >> 8) & 0xffu)) * 16777619u;
  h = (h ^ ((x >> 16) & 0xffu)) * 16777619u;
  h = (h ^ (x >> 24)) * 16777619u;
  sig_word(cls) = h;
}
// a discrete outcome (class 1)
__device__ __forceinline__ void sig_note(unsigned x) { sig_mix(1, x); }
// a comparison (class 0)
__device__ __forceinline__ bool sig_cmp(bool r) {
  sig_mix(0, r ? 0x5bd1e995u : 0x1b873593u);
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

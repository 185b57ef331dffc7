// TEMPO column step on CDNA4: device code templated on the working precision
// R (float = production path, double = FP64 twin).
//
// The reference step (tempo_main, src/module_mp_tempo_main.F90:73-864) is
// split at its column couplings into four launches over one batch:
//
//   k_prepass  one lane per column: the early-exit test (any hydrometeor or
//              ice supersaturation, :416-418), first-step graupel
//              initialisation test (:387-393) and the height prefix of the
//              default aerosol profiles (src/module_mp_tempo_aerosols.F90:24-28)
//   k_phase_a  one lane per (column, level): everything from the initial
//              checks to rain evaporation (:320-588) is level-local; writes
//              the post-check state in place and the sedimentation and
//              final-update operands to scratch
//   k_phase_b  one lane per column: fall speeds (top-down fill, CFL
//              substeps, ktop) and flux-form sedimentation (:590-727), with
//              the mass/number state and tendencies carried in FP64
//   k_phase_c  one lane per (column, level): freeze/melt, final update,
//              final checks and diagnostics (:729-863)
//
// Lanes of a wavefront take consecutive columns at one level, so every
// field access is a coalesced 64-lane row of the (level, column) layout.
#pragma once

#include <hip/hip_runtime.h>

#include "tempo_const.hpp"
#include "tempo_crmath.hpp"
#include "tempo_dual.hpp"
#include "tempo_sig.hpp"

namespace tempo {

// ---------------------------------------------------------------- math ----
// FP64: double-double evaluations rounded once (tempo_crmath.hpp), so the
// twin rounds as the reference's host libm does; FP32: the hardware log2 and
// exp2 behind an exact range reduction for x**y, the device libm otherwise.

// x**y in FP32 (x >= 0, as every power of the step): x = m 2^e with m in
// [sqrt(1/2), sqrt(2)), so log2 x = e + log2 m where e is exact and the
// hardware log2 sees |log2 m| < 1/2 (absolute error near 2^-24); y*e is
// carried as a rounded product plus its exact fma residual, split into an
// integer n and a fraction |f| <= 1/2 + small, and x**y = 2^f 2^n with the
// hardware exp2 on f.  Error: a few units in the last place, against the
// 1-ulp device powf at several times the instruction count.
__device__ __forceinline__ float pow_f32(float x, float y) {
  if (y == 0.0f) return 1.0f;
  int e = __builtin_amdgcn_frexp_expf(x);
  float m = __builtin_amdgcn_frexp_mantf(x);  // x = m 2^e, m in [1/2, 1) (0 for x = 0)
  if (m < 0.70710678f) {
    m = m + m;
    e = e - 1;
  }
  const float lm = __builtin_amdgcn_logf(m);  // log2 m (-inf for x = 0)
  const float fe = float(e);
  const float p = y * fe;
  const float perr = __builtin_fmaf(y, fe, -p);  // y*e = p + perr exactly
  const float n = __builtin_rintf(p);
  const float f = (p - n) + __builtin_fmaf(y, lm, perr);  // p - n is exact
  return __builtin_ldexpf(__builtin_amdgcn_exp2f(f), int(n));
}

__device__ __forceinline__ double mpow(double x, double y) { return crm::pow_cr(x, y); }
__device__ __forceinline__ float mpow(float x, float y) { return pow_f32(x, y); }
__device__ __forceinline__ double mexp(double x) { return crm::exp_cr(x); }
__device__ __forceinline__ float mexp(float x) { return expf(x); }
__device__ __forceinline__ double mlog(double x) { return crm::log_cr(x); }
__device__ __forceinline__ float mlog(float x) { return logf(x); }
__device__ __forceinline__ double mlog10(double x) { return crm::log10_cr(x); }
__device__ __forceinline__ float mlog10(float x) { return log10f(x); }
__device__ __forceinline__ double msqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float msqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double mcos(double x) { return cos(x); }
__device__ __forceinline__ float mcos(float x) { return cosf(x); }
template <class T>
__device__ __forceinline__ T mabs(T x) { return x < T(0) ? -x : x; }
template <class T>
__device__ __forceinline__ T mx(T a, T b) { return a > b ? a : b; }
template <class T>
__device__ __forceinline__ T mn(T a, T b) { return a < b ? a : b; }
// Fortran NINT: round half away from zero
template <class T>
__device__ __forceinline__ int nint_(T x) { return int(x < T(0) ? x - T(0.5) : x + T(0.5)); }
template <class T>
__device__ __forceinline__ int nint_exact(T x) { return int(round(double(x))); }
template <class T>
__device__ __forceinline__ int nint_exact(Dual<T> x) { return nint_exact(x.v); }
using ::copysign;
template <class T>
__device__ __forceinline__ Dual<T> copysign(Dual<T> a, Dual<T> b) {
  const T v = copysign(a.v, b.v);
  return {v, (v == a.v) ? a.d : -a.d};
}

// derivatives of the library functions (tangent-linear step)
template <class T>
__device__ __forceinline__ Dual<T> mpow(Dual<T> x, Dual<T> y) {
  const T p = mpow(x.v, y.v);
  T d = T(0);
  if (x.v > T(0)) {
    d = p * (y.v * x.d / x.v);
    if (y.d != T(0)) d = d + p * mlog(x.v) * y.d;
  }
  return {p, d};
}
template <class T, class U>
__device__ __forceinline__ Dual<T> mpow(Dual<T> x, U y) { return mpow(x, Dual<T>(T(y))); }
template <class T, class U>
__device__ __forceinline__ Dual<T> mpow(U x, Dual<T> y) { return mpow(Dual<T>(T(x)), y); }
template <class T>
__device__ __forceinline__ Dual<T> mexp(Dual<T> x) {
  const T e = mexp(x.v);
  return {e, e * x.d};
}
template <class T>
__device__ __forceinline__ Dual<T> mlog(Dual<T> x) { return {mlog(x.v), x.d / x.v}; }
template <class T>
__device__ __forceinline__ Dual<T> mlog10(Dual<T> x) {
  return {mlog10(x.v), x.d / (x.v * T(2.302585092994045684))};
}
template <class T>
__device__ __forceinline__ Dual<T> msqrt(Dual<T> x) {
  const T r = msqrt(x.v);
  return {r, r > T(0) ? x.d / (T(2) * r) : T(0)};
}
template <class T>
__device__ __forceinline__ Dual<T> mcos(Dual<T> x) { return {mcos(x.v), -sin(x.v) * x.d}; }

// branch-signature runs (tempo_sig.hpp): the library functions on the value,
// exactly as for the plain type; min, max and |x| record nothing
template <class T>
struct Num<Sig<T>> {
  using S = T;
  using W = double;
  using D = T;
};
template <class T>
__host__ __device__ __forceinline__ T value(const Sig<T>& x) { return x.v; }
template <class T>
__device__ __forceinline__ Sig<T> mpow(Sig<T> x, Sig<T> y) { return mpow(x.v, y.v); }
template <class T, class U, class = typename std::enable_if<std::is_arithmetic<U>::value>::type>
__device__ __forceinline__ Sig<T> mpow(Sig<T> x, U y) { return mpow(x.v, T(y)); }
template <class T, class U, class = typename std::enable_if<std::is_arithmetic<U>::value>::type>
__device__ __forceinline__ Sig<T> mpow(U x, Sig<T> y) { return mpow(T(x), y.v); }
template <class T>
__device__ __forceinline__ Sig<T> mexp(Sig<T> x) { return mexp(x.v); }
template <class T>
__device__ __forceinline__ Sig<T> mlog(Sig<T> x) { return mlog(x.v); }
template <class T>
__device__ __forceinline__ Sig<T> mlog10(Sig<T> x) { return mlog10(x.v); }
template <class T>
__device__ __forceinline__ Sig<T> msqrt(Sig<T> x) { return msqrt(x.v); }
template <class T>
__device__ __forceinline__ Sig<T> mcos(Sig<T> x) { return mcos(x.v); }
template <class T>
__device__ __forceinline__ Sig<T> mabs(Sig<T> x) { return mabs(x.v); }
template <class T>
__device__ __forceinline__ Sig<T> mx(Sig<T> a, Sig<T> b) { return mx(a.v, b.v); }
template <class T>
__device__ __forceinline__ Sig<T> mn(Sig<T> a, Sig<T> b) { return mn(a.v, b.v); }
template <class T>
__device__ __forceinline__ Sig<T> copysign(Sig<T> a, Sig<T> b) { return copysign(a.v, b.v); }
template <class T>
__device__ __forceinline__ int nint_exact(Sig<T> x) {
  const int i = nint_exact(x.v);
  sig_note(unsigned(i) ^ 0x2545f491u);
  return i;
}

// real**integer as flang lowers it (repeated squaring, reciprocal for n<0)
template <class R>
__device__ __forceinline__ R powi(R a, int b) {
  bool recip = b < 0;
  R r = R(1);
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? R(1) / r : r;
}

// ------------------------------------------------------------ operands ----
// state fields (pointer slots)
enum Field {
  F_T, F_P, F_W, F_DZ, F_QV, F_QC, F_QR, F_QI, F_QS, F_QG, F_NI, F_NR, F_NC, F_NWFA, F_NIFA, F_NG, F_QB,
  F_TH, F_PII, F_NFIELDS
};
enum Diag2 { D2_RAIN, D2_ICE, D2_SNOW, D2_GRAUPEL, D2_FROZEN, D2_FRZRAIN, D2_HAIL_SFC, D2_HAIL_COL, D2_N };
enum Diag3 { D3_REFL, D3_RE_C, D3_RE_I, D3_RE_S, D3_MVD_R, D3_MVD_G, D3_NC_MIX, D3_N };
// working-precision scratch (per point)
enum ScrR {
  SR_TEMP, SR_QV, SR_RHO, SR_TTEN, SR_QVTEN, SR_NWFATEN, SR_NIFATEN, SR_VTRR, SR_VTNR, SR_VTRG, SR_VTNG, SR_VTS,
  SR_VTRI, SR_VTNI, SR_VTRC, SR_VTNC, SR_HGT,
  SR_RHOF, SR_VISCO,  // density factor and viscosity (written only for semi-Lagrangian sedimentation)
  SR_NCSAVE,          // one-moment cloud number of the network (written only with ml_for_nc and no nc)
  SR_N
};
// scratch of the sedimented state and tendencies (per point, Sd_t<R>)
enum ScrD {
  SD_RR, SD_NR, SD_RG, SD_NG, SD_RB, SD_RS, SD_RI, SD_NI, SD_RC, SD_NC,
  SD_QRTEN, SD_NRTEN, SD_QGTEN, SD_NGTEN, SD_QBTEN, SD_QSTEN, SD_QITEN, SD_NITEN, SD_QCTEN, SD_NCTEN, SD_N
};
// per point flag bits
enum : unsigned { PF_QR = 1u, PF_QG = 2u, PF_QS = 4u, PF_QI = 8u, PF_QC = 16u, PF_SNOWMIX = 32u };
// per column flag bits
enum : unsigned { CF_EXIT = 1u, CF_GINIT = 2u };

constexpr int kCcnSlice = 7 * 9 * 7;  // floats of the CCN activation table the step reads

template <class R>
struct Tables {
  const R* qg;       // [(i,j,n3,k,m) Fortran index][5]
  const R* qs;       // [(i,j,k,m)][12]
  const R* fzr;      // [(i,j,k,m)][4]
  const R* fzc;      // [(i,j,k,m)][2]
  const float* ccn;  // tnccn_act(:,:,:,3,2): [k][j][i] 7x9x7 (kCcnSlice)
  const R* efrw;     // (100,100)
  const R* efsw;     // (100,100)
  const R* tnc_wev;  // (100,37,100)
  const R* tps_iaus; // (64,55)
  const R* tni_iaus;
  const R* tpi_ide;
};

// rows of the per-column sedimentation record (tempo_hip_substeps): substep
// count and top of rain, graupel, snow, ice and cloud, then hashes of the
// semi-Lagrangian rain and graupel sedimentation's decisions (0 in flux form)
constexpr int kSubRows = 12;
// one more decision folded into a 32-bit hash (FNV-1a step)
__host__ __device__ constexpr unsigned decision(unsigned h, unsigned v) { return (h ^ v) * 16777619u; }

template <class R>
struct StepArgs {
  R* f[F_NFIELDS];   // nullptr = absent (optional fields); the step's inputs
  R* fa[F_NFIELDS];  // state after the initial checks: written by phase A, read by the final update
  R* fo[F_NFIELDS];  // new state written by the final update (fa = fo = f for a step in place)
  R* d2[D2_N];       // per column diagnostics (nullptr = not requested)
  R* d3[D3_N];       // per level diagnostics
  const int* land;   // per column land flag or nullptr
  const int* nlev;   // per column active level count or nullptr
  const R* qc_bl;      // boundary-layer cloud water and its cloud fraction (state layout), or nullptr
  const R* qcfrac_bl;
  const Radar* radar;  // melting-layer reflectivity constants (device copy)
  int* kmelt;          // per column melting level (1-based), written when the melting reflectivity is on
  int ni, nj, nz, ncol;
  int nzs;           // level stride of a scratch column group (scratch_levels(nz))
  long long ldk, ldj, dldk, dldj, d2ldj;
  R* sr;             // SR_N planes
  Sd_t<R>* sd;       // SD_N planes (working precision, or FP64 duals)
  size_t plane;      // elements per scratch plane: scratch_levels(nz) x ncol rounded up to kScrW columns
  unsigned* pf;      // nz x ncol point flags
  unsigned* cf;      // ncol column flags
  W_t<R>* precip;    // 5 x ncol: rain, graupel, snow, ice, cloud
  int* ksub;         // kSubRows x ncol: substeps and ktop per species, semi-Lagrangian decision hashes
  W_t<R>* semiw;     // semi-Lagrangian sedimentation work: kSemiArrays x (nz + 2) x ncol, or nullptr
  const Const<S_t<R>>* cst;  // derived constants of this context (wave-uniform reads)
  // level-local kernels: bit 0 (phase A) / bit 1 (final update) set = each
  // block hands its level's points with microphysics to do to its first
  // lanes, clear = lane i takes column i
  int compact;
  int ccn_lds;  // phase A reads the CCN table's slice from an LDS copy (1) or from global memory (0)
  Tables<S_t<R>> tab;
  RunFlags fl;
  double dt;
  // 1/dt in the working precision, divided on the host: the FP32 unit's
  // reciprocal arithmetic would leave a one-ulp error that keeps the
  // x + (-x/dt)*dt of a fully removed mass from cancelling
  double odt;
  int first_step;
  // work queue of the fused step (a counter of column groups handed out)
  int* queue;
  // adjoint passes of the final update: levels whose output cotangents are
  // all zero (lvl_on[k] == 0) are skipped, their contributions being zero;
  // nullptr = every level
  const int* lvl_on;
  // an adjoint's local passes of phase A: points (c, k) with
  // pt_on[c + k * ncol] == 0 (no sedimentation cotangent, final update
  // skipped) are skipped; nullptr = every point
  const int* pt_on;
  // an adjoint's local passes of the final update: the post-sedimentation
  // tendency values, 10 planes (SD_QRTEN order) at post[m * plane + s],
  // replacing the scratch's values; nullptr = the scratch's values
  const double* post;
  // branch signatures (signature runs only): 6 x nz x ncol words, phase p
  // (0 the level-local physics, 1 the final update, 2 the sedimentation),
  // class q (0 comparisons, 1 discrete outcomes), point (c, k) at
  // (2p + q)*nz*ncol + k*ncol + c
  unsigned* sig;
};

// operands of the public cloud/ice/snow_check_and_update over n points
// (src/module_mp_tempo_main.F90:18-20): q the mass mixing ratio, nn its
// number (cloud, ice; nullptr = absent for cloud), r and num the per-volume
// mass and number, l the presence flag (0/1), ilam the inverse slope, mvd
// the cloud's median volume diameter
template <class R>
struct CheckArgs {
  int n;
  R dt, odt;
  const R* rho;
  int* l;
  R *q, *nn;
  const R* ncsave;  // cloud: one-moment numbers when nn is absent (nullptr = the land value)
  R *r, *num, *qten, *nten, *ilam, *mvd;
  const Const<S_t<R>>* cst;
};

// geometry of a caller array: element (i, k, j) at i + k*ldk + j*ldj,
// column c = i + j*ni (2-D arrays: nk = 1, ldk = 0)
struct DualGeom {
  int ni, nj, nk;
  long long ldk, ldj;
};

// launch shape of one step (host-side choice, see tempo_hip_tune)
struct LaunchCfg {
  int fused = 0;      // 1: the whole step in one persistent launch where it applies (k_step_fused)
  int sed_wave = 1;   // sedimentation: 1 = one wavefront per column (nz <= 128), 0 = one lane per column,
                      // 2 = the wavefront kernel issuing every staging pass (check form)
  int order = 0;      // point grouping within blocks: bit 0 phase A, bit 1 final update
  int block_a = 256;  // threads per block of the level-local kernels (64, 128 or 256)
  int block_c = 256;
  int waves_a = 0;       // phase A compiled for the default waves per SIMD (0) or one fewer (1)
  int ccn_lds = 0;       // phase A stages the CCN table's slice in LDS (1)
  int host_slabs = 0;    // host-array steps: slabs of columns (0: about 128 MB of uploads each)
  int host_threads = 8;  // ... and host threads copying between the caller's arrays and pinned memory
};

#define TEMPO_IDX_STATE(a, c, k) (size_t((c) % (a).ni) + size_t(k) * (a).ldk + size_t((c) / (a).ni) * (a).ldj)
// scratch planes are column-tiled: groups of kScrW consecutive columns hold
// their levels contiguously, (c / kScrW, k, c % kScrW), so that a column
// kernel reads one contiguous block per group.  The level stride of a group
// is nz rounded up to kScrLevels, so that groups start on 128-byte lines and
// the 128-byte line holding levels 4m..4m+3 (FP32) or 2m, 2m+1 (FP64) of a
// group is used whole by one block of a level-local kernel (64 columns x 2-4
// levels).
constexpr int kScrW = 8;
constexpr int kScrLevels = 16;
__host__ __device__ constexpr int scratch_levels(int nz) { return (nz + kScrLevels - 1) / kScrLevels * kScrLevels; }
// elements of one plane of a fused-step block's scratch slot: one column tile
__host__ __device__ constexpr size_t fused_slot_plane(int nz) { return size_t(kScrW) * size_t(scratch_levels(nz)); }
#define TEMPO_IDX_SCR(a, c, k) \
  ((size_t((c) / kScrW) * size_t((a).nzs) + size_t(k)) * kScrW + size_t((c) % kScrW))

// Where the per-point scratch of a kernel lives: the batch's column-tiled
// planes (c0 = 0; element of (c, k) at TEMPO_IDX_SCR), or the tile of one
// column group starting at column c0 (the fused step's block-private slot),
// indexed the same way relative to c0.
template <class R>
struct Scr {
  R* sr;
  Sd_t<R>* sd;
  unsigned* pf;
  size_t plane;  // elements per plane
  int c0, nzs;
  __device__ __forceinline__ size_t at(int c, int k) const {
    const int cc = c - c0;
    return (size_t(cc / kScrW) * size_t(nzs) + size_t(k)) * kScrW + size_t(cc % kScrW);
  }
};
template <class R>
__device__ __forceinline__ Scr<R> batch_scr(const StepArgs<R>& a) {
  return Scr<R>{a.sr, a.sd, a.pf, a.plane, 0, a.nzs};
}

// ------------------------------------------------ level-local physics ----
// Per-level working state: the scalars of one (column, level) point.  The
// process functions that take it (and the tendencies, Tend) are forced
// inline: a call that stays out of line needs the structure in memory, and
// in FP64, where the double-double arithmetic makes the functions large
// enough for the inliner to keep some out of line, that put about 1 KB per
// lane in scratch (phase A 12.2 against 8.3 ms at 100,000 x 90).
template <class R>
struct Lvl {
  // thermodynamics
  R temp, pres, qv, rho, rhof, rhof2, qvs, qvsi, delqvs, ssatw, ssati, diffu, visco, vsc2, ocp, lvap, tcond, lvt2;
  // hydrometeor contents
  R rc, ri, rr, rs, rg, rb, ni, nr, nc, ng, nwfa, nifa, ncsave;
  R ilamc, ilami, ilamr, ilamg, mvd_r, mvd_c, mvd_g;
  R smob, smoc, smo0, smo1, smo2, smoe, smof, smoz;
  int idx_bg;
  bool l_qc, l_qi, l_qr, l_qs, l_qg, supersat;
  // tendencies of the 1-d state
  R tten, qvten, qcten, qiten, qrten, qsten, qgten, qbten, niten, nrten, ncten, ngten, nwfaten, nifaten;
};

// process rates (ty_tend, src/module_mp_tempo_main.F90:47-64)
template <class R>
struct Tend {
  R prr_wau, pnr_wau, pnc_wau, prr_rcw, pnc_rcw, pnr_rcr;
  R prs_scw, pnc_scw, png_scw, pbg_scw, prg_gcw, pnc_gcw, pbg_gcw, pri_ihm, pni_ihm, prs_ihm, prg_ihm, prg_scw;
  R prr_rcs, pnr_rcs, prg_rcs, png_rcs, prs_rcs, pbg_rcs;
  R prr_rcg, pnr_rcg, prg_rcg, png_rcg, pbg_rcg;
  R pri_inu, pni_inu, pri_iha, pni_iha;
  R pri_wfz, pni_wfz;
  R prg_rfz, png_rfz, pnr_rfz, pri_rfz, pni_rfz, pbg_rfz;
  R prs_sde, pri_ide, pni_ide, prs_ide, prg_gde, png_gde;
  R pni_iau, prs_iau;
  R prr_sml, prr_gml, pbg_sml, pbg_gml, pnr_sml, pnr_gml;
  R prr_rci, pnr_rci, pri_rci, pni_rci, prg_rci, png_rci, pbg_rci;
  R pni_sci, prs_sci;
  R prw_vcd, pnc_wcd, prv_rev, pnr_rev;
  R pna_rca, pna_sca, pna_gca, pnd_rcd, pnd_scd, pnd_gcd;
};

template <class R>
__device__ __forceinline__ R rho_of(const Const<S_t<R>>& C, R pres, R temp, R qv) {
  return C.roverrv * pres / (R(kRdry) * temp * (qv + C.roverrv));
}

// calc_rslf / calc_rsif (src/module_mp_tempo_utils.F90:285-331)
template <class R>
__device__ __forceinline__ R rslf(R p, R t) {
  R x = mx(R(-80.0), t - R(273.16));
  R esl = R(.611583699E03) +
          x * (R(.444606896E02) +
               x * (R(.143177157e01) +
                    x * (R(.264224321e-1) +
                         x * (R(.299291081e-3) +
                              x * (R(.203154182e-5) +
                                   x * (R(.702620698e-8) + x * (R(.379534310e-11) + x * R(-.321582393e-13))))))));
  esl = mn(esl, p * R(0.15));
  return R(.622) * esl / (p - esl);
}
template <class R>
__device__ __forceinline__ R rsif(R p, R t) {
  R x = mx(R(-80.0), t - R(273.16));
  R esi = R(.609868993e03) +
          x * (R(.499320233e02) +
               x * (R(.184672631e01) +
                    x * (R(.402737184e-1) +
                         x * (R(.565392987e-3) +
                              x * (R(.521693933e-5) +
                                   x * (R(.307839583e-7) + x * (R(.105785160e-9) + x * R(.161444444e-12))))))));
  esi = mn(esi, p * R(0.15));
  return R(.622) * esi / mx(R(1.e-4), (p - esi));
}

template <class R>
__device__ __forceinline__ int get_nuc(R nc) {  // src/module_mp_tempo_utils.F90:334-346
  if ((R(kNuCScale) / nc) >= R(12.5)) return 15;
  return min(15, nint_exact(R(kNuCScale) / nc) + 2);
}

// snow_moments polynomial (src/module_mp_tempo_utils.F90:148-282)
template <class R, class S>
__device__ __forceinline__ R snow_poly(const S* s, R tc, R cc) {
  return s[0] + s[1] * tc + s[2] * cc + s[3] * tc * cc + s[4] * tc * tc + s[5] * cc * cc + s[6] * tc * tc * cc +
         s[7] * tc * cc * cc + s[8] * tc * tc * tc + s[9] * cc * cc * cc;
}
// want: 1 smo0, 2 smo1, 4 smo2, 8 smoe, 16 smof, 32 smoz ; smob, smoc always
template <class R>
__device__ __forceinline__ void snow_moments(const Const<S_t<R>>& C, R rs, R tc, int want, Lvl<R>& L) {
  const S_t<R>* sa = C.sa;
  const S_t<R>* sb = C.sb;
  L.smob = rs * C.oams;
  const R bms = R(kBmS);
  R loga = sa[0] + sa[1] * tc + sa[2] * bms + sa[3] * tc * bms + sa[4] * tc * tc + sa[5] * bms * bms +
           sa[6] * tc * tc * bms + sa[7] * tc * bms * bms + sa[8] * tc * tc * tc + sa[9] * bms * bms * bms;
  R a = mpow(R(10.0), loga);
  R b = sb[0] + sb[1] * tc + sb[2] * bms + sb[3] * tc * bms + sb[4] * tc * tc + sb[5] * bms * bms +
        sb[6] * tc * tc * bms + sb[7] * tc * bms * bms + sb[8] * tc * tc * tc + sb[9] * bms * bms * bms;
  const R smo2 = mpow(L.smob / a, R(1.0) / b);
  if (want & 4) L.smo2 = smo2;
  a = mpow(R(10.0), snow_poly(sa, tc, R(C.cse[0])));
  L.smoc = a * mpow(smo2, snow_poly(sb, tc, R(C.cse[0])));
  if (want & 1) {
    R la = sa[0] + sa[1] * tc + sa[4] * tc * tc + sa[8] * tc * tc * tc;
    R lb = sb[0] + sb[1] * tc + sb[4] * tc * tc + sb[8] * tc * tc * tc;
    L.smo0 = mpow(R(10.0), la) * mpow(smo2, lb);
  }
  if (want & 2) {
    R la = sa[0] + sa[1] * tc + sa[2] + sa[3] * tc + sa[4] * tc * tc + sa[5] + sa[6] * tc * tc + sa[7] * tc +
           sa[8] * tc * tc * tc + sa[9];
    R lb = sb[0] + sb[1] * tc + sb[2] + sb[3] * tc + sb[4] * tc * tc + sb[5] + sb[6] * tc * tc + sb[7] * tc +
           sb[8] * tc * tc * tc + sb[9];
    L.smo1 = mpow(R(10.0), la) * mpow(smo2, lb);
  }
  if (want & 8) L.smoe = mpow(R(10.0), snow_poly(sa, tc, R(C.cse[12]))) * mpow(smo2, snow_poly(sb, tc, R(C.cse[12])));
  if (want & 16) L.smof = mpow(R(10.0), snow_poly(sa, tc, R(C.cse[15]))) * mpow(smo2, snow_poly(sb, tc, R(C.cse[15])));
  if (want & 32) L.smoz = mpow(R(10.0), snow_poly(sa, tc, R(C.cse[2]))) * mpow(smo2, snow_poly(sb, tc, R(C.cse[2])));
}

// thermo_vars, one level (src/module_mp_tempo_main.F90:1263-1310)
template <class R>
__device__ __forceinline__ void thermo(const Const<S_t<R>>& C, Lvl<R>& L) {
  const R otemp = R(1.0) / L.temp, tempc = L.temp - R(kT0);
  L.rhof = msqrt(C.rho_not / L.rho);
  L.rhof2 = msqrt(L.rhof);
  L.qvs = rslf(L.pres, L.temp);
  L.delqvs = mx(R(0.0), rslf(L.pres, R(kT0)) - L.qv);
  L.qvsi = (tempc <= R(0.0)) ? rsif(L.pres, L.temp) : L.qvs;
  const R satw = L.qv / L.qvs, sati = L.qv / L.qvsi;
  L.ssatw = satw - R(1.0);
  L.ssati = sati - R(1.0);
  if (mabs(L.ssatw) < R(kEps)) L.ssatw = R(0.0);
  if (mabs(L.ssati) < R(kEps)) L.ssati = R(0.0);
  L.supersat = L.ssati > R(0.0);
  L.diffu = R(2.11e-5) * mpow(L.temp / R(kT0), R(1.94)) * (R(101325.0) / L.pres);
  if (tempc >= R(0.0))
    L.visco = (R(1.718) + R(0.0049) * tempc) * R(1.0e-5);
  else
    L.visco = (R(1.718) + R(0.0049) * tempc - R(1.2e-5) * tempc * tempc) * R(1.0e-5);
  L.ocp = R(1.0) / (R(kCp) * (R(1.0) + R(0.887) * L.qv));
  L.vsc2 = msqrt(L.rho / L.visco);
  L.lvap = R(kLvap0) + R(2106.0 - 4218.0) * tempc;
  L.tcond = (R(5.69) + R(0.0168) * tempc) * R(1.0e-5) * R(418.936);
  L.lvt2 = L.lvap * L.lvap * L.ocp * C.orv * otemp * otemp;
}

// ---- cloud number from the trained network (src/module_mp_tempo_ml.F90) ----
#include "tempo_ml_weights.inc"

template <class S>
__device__ __forceinline__ S ml_par(const double* d, const float* f, int i) {
  if constexpr (sizeof(S) == 4)
    return f[i];
  else
    return d[i];
}

// tempo_ml_predict_cloud_number (src/module_mp_tempo_ml.F90:104-174) at one
// level, per kg: standard-scaled inputs (x - mean) / sqrt(var), 24 ReLU
// nodes, one ReLU output clamped to [-6, 9.3010299957] (log10 of the number),
// the quintic bias correction on [0, 3), and 10**output.  The products of
// each layer are summed in input order onto the first, then the bias added,
// as the reference's matmul and bias arrays do.
template <class R>
__device__ R ml_cloud_number(R qc, R qr, R qi, R qs, R pres, R temp, R w) {
  using S = S_t<R>;
  const R x[7] = {qc, qr, qi, qs, pres, temp, w};
  R xt[7];
#pragma unroll
  for (int j = 0; j < 7; ++j)
    xt[j] = (x[j] - R(ml_par<S>(kMlMeanD, kMlMeanF, j))) / msqrt(R(ml_par<S>(kMlVarD, kMlVarF, j)));
  R out = R(0.0);
#pragma unroll
  for (int n = 0; n < 24; ++n) {
    R h = R(ml_par<S>(kMlW0D, kMlW0F, n)) * xt[0];
#pragma unroll
    for (int j = 1; j < 7; ++j) h = h + R(ml_par<S>(kMlW0D, kMlW0F, n + 24 * j)) * xt[j];
    h = mx(h + R(ml_par<S>(kMlB0D, kMlB0F, n)), R(0.0));
    out = (n == 0) ? R(ml_par<S>(kMlW1D, kMlW1F, 0)) * h : out + R(ml_par<S>(kMlW1D, kMlW1F, n)) * h;
  }
  out = mx(out + R(ml_par<S>(kMlB1D, kMlB1F, 0)), R(0.0));
  const R p = mn(R(S(9.3010299957)), mx(R(-6.0), out));
  R corr = R(1.0);
  if (p >= R(0.0) && p < R(3.0))
    corr = -R(S(0.2704)) * powi(p, 5) + R(S(1.838)) * powi(p, 4) - R(S(5.127)) * powi(p, 3) +
           R(S(8.547)) * powi(p, 2) - R(S(8.439)) * p + R(S(4.297));
  return corr * mpow(R(10.0), p);
}

// ---- *_check_and_update for one level (src/module_mp_tempo_main.F90:891-1231) ----
// ncv: the column's one-moment cloud number (used when nc is absent, has_nc false)
template <class R>
__device__ __forceinline__ void cloud_check(const Const<S_t<R>>& C, Lvl<R>& L, R& qc1d, R& nc1d, bool has_nc, R ncv, R dt, R odt) {
  // x + ten*dt rounded as two operations, as the reference forms it: where the
  // tendency removes the whole mass the product rounds to -x and nothing is
  // left, whereas a fused multiply-add would leave the rounding error of
  // the tendency, above the r1 presence threshold for typical masses
#pragma clang fp contract(off)
  bool hit = false;
  if (qc1d + L.qcten * dt > R(kR1)) {
    L.l_qc = true;
    L.rc = (qc1d + L.qcten * dt) * L.rho;
    qc1d = qc1d + L.qcten * dt;
    int nu;
    R lamc;
    if (has_nc) {
      L.nc = mx(R(kNtCMin), (nc1d + L.ncten * dt) * L.rho);
      if (L.nc <= R(kNtCMin)) { hit = true; L.nc = R(kNtCMin); }
      if (L.nc > R(kNtCMax)) { hit = true; L.nc = R(kNtCMax); }
      nu = get_nuc(L.nc);
      lamc = mpow(L.nc * C.am_r * C.ccg[1][nu] * C.ocg1[nu] / L.rc, C.obmr);
      const R xdc = (R(kBmR) + R(nu) + R(1.0)) / lamc;
      if (xdc < R(kD0C)) {
        lamc = C.cce[1][nu] / R(kD0C);
        hit = true;
      } else if (xdc > R(kD0R) * R(2.0)) {
        lamc = C.cce[1][nu] / (R(kD0R) * R(2.0));
        hit = true;
      }
      L.nc = C.ccg[0][nu] * C.ocg2[nu] * L.rc / C.am_r * mpow(lamc, R(kBmR));
      if (hit) L.ncten = (L.nc / L.rho - nc1d) * odt;
      nc1d = mx(R(kNtCMin) / L.rho,
                 mn(C.ccg[0][nu] * C.ocg2[nu] * qc1d / C.am_r * mpow(lamc, R(kBmR)), R(kNtCMax) / L.rho));
    } else {
      L.nc = ncv;
    }
    nu = get_nuc(L.nc);
    lamc = mpow(L.nc * C.am_r * C.ccg[1][nu] * C.ocg1[nu] / L.rc, C.obmr);
    L.ilamc = R(1.0) / lamc;
    L.mvd_c = mx(mn((R(3.0) + R(nu) + R(0.672)) * L.ilamc, R(kD0R)), R(kD0C));
  } else {
    L.l_qc = false;
    L.rc = R(kR1);
    L.nc = R(kNtCMin);
    L.mvd_c = R(kD0C);
    L.ilamc = R(0.0);
    L.qcten = -qc1d * odt;
    qc1d = R(0.0);
    if (has_nc) {
      L.ncten = -nc1d * odt;
      nc1d = R(0.0);
    }
  }
}

template <class R>
__device__ __forceinline__ void rain_check(const Const<S_t<R>>& C, Lvl<R>& L, R& qr1d, R& nr1d, R dt, R odt) {
#pragma clang fp contract(off)
  bool hit = false;
  if (qr1d + L.qrten * dt > R(kR1)) {
    L.l_qr = true;
    L.rr = (qr1d + L.qrten * dt) * L.rho;
    qr1d = qr1d + L.qrten * dt;
    L.nr = mx(R(kR2), (nr1d + L.nrten * dt) * L.rho);
    const R c3 = R(3.0) + R(kMuR) + R(0.672);
    R lamr;
    if (L.nr <= R(kR2)) {
      hit = true;
      L.mvd_r = R(1.0e-3);
      lamr = c3 / L.mvd_r;
      L.nr = C.crg[1] * C.org3 * L.rr * mpow(lamr, R(kBmR)) / C.am_r;
    }
    lamr = mpow(C.am_r * C.crg[2] * C.org2 * L.nr / L.rr, C.obmr);
    L.mvd_r = c3 / lamr;
    if (L.mvd_r > R(kD0RMax)) {
      hit = true;
      L.mvd_r = R(kD0RMax);
      lamr = c3 / L.mvd_r;
      L.nr = C.crg[1] * C.org3 * L.rr * mpow(lamr, R(kBmR)) / C.am_r;
    } else if (L.mvd_r < R(kD0R) * R(0.75)) {
      hit = true;
      L.mvd_r = R(kD0R) * R(0.75);
      lamr = c3 / L.mvd_r;
      L.nr = C.crg[1] * C.org3 * L.rr * mpow(lamr, R(kBmR)) / C.am_r;
    }
    if (hit) L.nrten = (L.nr / L.rho - nr1d) * odt;
    nr1d = C.crg[1] * C.org3 * qr1d * mpow(lamr, R(kBmR)) / C.am_r;
    L.ilamr = R(1.0) / lamr;
  } else {
    L.l_qr = false;
    L.rr = R(kR1);
    L.nr = R(kR2);
    L.mvd_r = R(kD0R);
    L.ilamr = R(0.0);
    L.qrten = -qr1d * odt;
    L.nrten = -nr1d * odt;
    qr1d = R(0.0);
    nr1d = R(0.0);
  }
}

template <class R>
__device__ __forceinline__ void ice_check(const Const<S_t<R>>& C, Lvl<R>& L, R& qi1d, R& ni1d, R dt, R odt) {
#pragma clang fp contract(off)
  bool hit = false;
  if (qi1d + L.qiten * dt > R(kR1)) {
    L.l_qi = true;
    L.ri = (qi1d + L.qiten * dt) * L.rho;
    qi1d = qi1d + L.qiten * dt;
    L.ni = mx(R(kR2), (ni1d + L.niten * dt) * L.rho);
    R lami;
    if (L.ni <= R(kR2)) {
      hit = true;
      lami = C.cie[1] / R(5.e-6);
      L.ni = mn(R(kMaxNi), C.cig[0] * C.oig2 * L.ri / C.am_i * mpow(lami, R(kBmI)));
    }
    lami = mpow(C.am_i * C.cig[1] * C.oig1 * L.ni / L.ri, C.obmi);
    const R xdi = (R(kBmI) + R(kMuI) + R(1.0)) / lami;
    if (xdi < R(5.e-6)) {
      hit = true;
      lami = C.cie[1] / R(5.e-6);
      L.ni = mn(R(kMaxNi), C.cig[0] * C.oig2 * L.ri / C.am_i * mpow(lami, R(kBmI)));
    } else if (xdi > R(kD0S)) {
      hit = true;
      lami = C.cie[1] / R(kD0S);
      L.ni = C.cig[0] * C.oig2 * L.ri / C.am_i * mpow(lami, R(kBmI));
    }
    if (hit) L.niten = (L.ni / L.rho - ni1d) * odt;
    ni1d = mx(R(kR2) / L.rho, mn(C.cig[0] * C.oig2 * qi1d / C.am_i * mpow(lami, R(kBmI)), R(kMaxNi) / L.rho));
    L.ilami = R(1.0) / lami;
  } else {
    L.l_qi = false;
    L.ri = R(kR1);
    L.ni = R(kR2);
    L.ilami = R(0.0);
    L.qiten = -qi1d * odt;
    L.niten = -ni1d * odt;
    qi1d = R(0.0);
    ni1d = R(0.0);
  }
}

template <class R>
__device__ __forceinline__ void snow_check(Lvl<R>& L, R& qs1d, R dt, R odt) {
#pragma clang fp contract(off)
  if (qs1d + L.qsten * dt > R(kR1)) {
    L.l_qs = true;
    L.rs = (qs1d + L.qsten * dt) * L.rho;
    qs1d = qs1d + L.qsten * dt;
  } else {
    L.l_qs = false;
    L.rs = R(kR1);
    L.qsten = -qs1d * odt;
    qs1d = R(0.0);
  }
}

template <class R>
__device__ R rho_g(int idx) {  // densities of graupel (src/module_mp_tempo_params.F90:143-144)
  return R(idx <= 3 ? (idx == 1 ? 50.0 : (idx == 2 ? 100.0 : 200.0)) : 100.0 * (idx - 1));
}

template <class R>
__device__ __forceinline__ void graupel_check(const Const<S_t<R>>& C, Lvl<R>& L, R& qg1d, R& ng1d, R& qb1d, bool has_ngqb, R dt, R odt) {
#pragma clang fp contract(off)
  bool hit = false;
  if (qg1d + L.qgten * dt > R(kR1)) {
    L.l_qg = true;
    L.rg = (qg1d + L.qgten * dt) * L.rho;
    qg1d = qg1d + L.qgten * dt;
    R lamg;
    const R c3 = R(3.0) + R(kMuG) + R(0.672);
    if (has_ngqb) {
      L.ng = mx(R(kR2), (ng1d + L.ngten * dt) * L.rho);
      L.rb = mn(mx(L.rg * R(kM3ToL) / rho_g<R>(kNrhg), (qb1d + L.qbten * dt) * L.rho),
                L.rg * R(kM3ToL) / rho_g<R>(1));
      L.idx_bg = max(1, min(nint_exact(R(10.0) * L.rg / L.rb) + 1, kNrhg));
      if (L.ng <= R(kR2)) {
        hit = true;
        L.mvd_g = R(1.5e-3);
        lamg = c3 / L.mvd_g;
        L.ng = C.cgg[1][0] * C.ogg3 * L.rg * mpow(lamg, R(kBmG)) / C.am_g[L.idx_bg - 1];
      }
      lamg = mpow(C.am_g[L.idx_bg - 1] * C.cgg[2][0] * C.ogg2 * L.ng / L.rg, C.obmg);
      L.mvd_g = c3 / lamg;
      if (L.mvd_g > R(25.4e-3)) {
        hit = true;
        L.mvd_g = R(25.4e-3);
        lamg = c3 / L.mvd_g;
        L.ng = C.cgg[1][0] * C.ogg3 * L.rg * mpow(lamg, R(kBmG)) / C.am_g[L.idx_bg - 1];
      } else if (L.mvd_g < R(kD0R)) {
        hit = true;
        L.mvd_g = R(kD0R);
        lamg = c3 / L.mvd_g;
        L.ng = C.cgg[1][0] * C.ogg3 * L.rg * mpow(lamg, R(kBmG)) / C.am_g[L.idx_bg - 1];
      }
      if (hit) L.ngten = (L.ng / L.rho - ng1d) * odt;
      ng1d = C.cgg[1][0] * C.ogg3 * qg1d * mpow(lamg, R(kBmG)) / C.am_g[L.idx_bg - 1];
      qb1d = mn(mx(qg1d * R(kM3ToL) / rho_g<R>(kNrhg), qb1d + L.qbten * dt), R(kM3ToL) * qg1d / rho_g<R>(1));
      L.idx_bg = max(1, min(nint_exact(R(10.0) * qg1d / qb1d) + 1, kNrhg));
    } else {
      L.idx_bg = kIdxBg1;
      const R ygra1 = mlog10(mx(R(1.e-9), L.rg));
      const R zans1 = R(3.4) + R(2.0 / 7.0) * (ygra1 + R(8.0));
      const R n0_exp = mx(R(kGonvMin), mn(mpow(R(10.0), zans1), R(kGonvMax)));
      const R lam_exp = mpow(n0_exp * C.am_g[L.idx_bg - 1] * C.cgg[0][0] / L.rg, C.oge1);
      lamg = lam_exp * mpow(C.cgg[2][0] * C.ogg2 * C.ogg1, C.obmg);
      L.ng = C.cgg[1][0] * C.ogg3 * L.rg * mpow(lamg, R(kBmG)) / C.am_g[L.idx_bg - 1];
      L.rb = R(kM3ToL) * L.rg / rho_g<R>(L.idx_bg);
    }
    L.ilamg = R(1.0) / lamg;
    L.mvd_g = c3 * L.ilamg;
  } else {
    L.l_qg = false;
    L.rg = R(kR1);
    L.ng = R(kR2);
    L.mvd_g = R(kD0G);
    L.ilamg = R(0.0);
    L.idx_bg = kIdxBg1;
    L.rb = R(kM3ToL) * R(kR1) / rho_g<R>(L.idx_bg);
    L.qgten = -qg1d * odt;
    qg1d = R(0.0);
    if (has_ngqb) {
      L.ngten = -ng1d * odt;
      L.qbten = -qb1d * odt;
      ng1d = R(0.0);
      qb1d = R(0.0);
    }
  }
}

// The checks above round x + ten*dt as two operations (their contraction
// pragma).  In a signature run they execute on the plain type, compiled
// exactly as in the production step, and their decisions are recorded from
// their outcome: the presence flag, the cloud's shape index, the graupel
// density index and the number clamps (a changed number marks a hit limit).
template <class T>
__device__ __forceinline__ Lvl<T> sig_unwrap(const Lvl<Sig<T>>& L) {
  static_assert(sizeof(Lvl<Sig<T>>) == sizeof(Lvl<T>) && sizeof(Sig<T>) == sizeof(T), "Sig layout");
  Lvl<T> M;
  __builtin_memcpy(&M, &L, sizeof(M));
  return M;
}
template <class T>
__device__ __forceinline__ void sig_wrap(Lvl<Sig<T>>& L, const Lvl<T>& M) {
  __builtin_memcpy(&L, &M, sizeof(M));
}
template <class T>
__device__ void cloud_check(const Const<T>& C, Lvl<Sig<T>>& L, Sig<T>& qc1d, Sig<T>& nc1d, bool has_nc, Sig<T> ncv,
                            Sig<T> dt, Sig<T> odt) {
  Lvl<T> M = sig_unwrap(L);
  cloud_check<T>(C, M, qc1d.v, nc1d.v, has_nc, ncv.v, dt.v, odt.v);
  sig_wrap(L, M);
  sig_note(M.l_qc ? 0x11u + 256u * unsigned(get_nuc(M.nc)) : 0x12u);
}
template <class T>
__device__ void rain_check(const Const<T>& C, Lvl<Sig<T>>& L, Sig<T>& qr1d, Sig<T>& nr1d, Sig<T> dt, Sig<T> odt) {
  Lvl<T> M = sig_unwrap(L);
  const T nten = M.nrten;
  rain_check<T>(C, M, qr1d.v, nr1d.v, dt.v, odt.v);
  sig_wrap(L, M);
  sig_note(M.l_qr ? (M.nrten != nten ? 0x21u : 0x22u) : 0x23u);
}
template <class T>
__device__ void ice_check(const Const<T>& C, Lvl<Sig<T>>& L, Sig<T>& qi1d, Sig<T>& ni1d, Sig<T> dt, Sig<T> odt) {
  Lvl<T> M = sig_unwrap(L);
  const T nten = M.niten;
  ice_check<T>(C, M, qi1d.v, ni1d.v, dt.v, odt.v);
  sig_wrap(L, M);
  sig_note(M.l_qi ? (M.niten != nten ? 0x31u : 0x32u) : 0x33u);
}
template <class T>
__device__ void snow_check(Lvl<Sig<T>>& L, Sig<T>& qs1d, Sig<T> dt, Sig<T> odt) {
  Lvl<T> M = sig_unwrap(L);
  snow_check<T>(M, qs1d.v, dt.v, odt.v);
  sig_wrap(L, M);
  sig_note(M.l_qs ? 0x41u : 0x42u);
}
template <class T>
__device__ void graupel_check(const Const<T>& C, Lvl<Sig<T>>& L, Sig<T>& qg1d, Sig<T>& ng1d, Sig<T>& qb1d,
                              bool has_ngqb, Sig<T> dt, Sig<T> odt) {
  Lvl<T> M = sig_unwrap(L);
  const T nten = M.ngten;
  graupel_check<T>(C, M, qg1d.v, ng1d.v, qb1d.v, has_ngqb, dt.v, odt.v);
  sig_wrap(L, M);
  sig_note(M.l_qg ? 0x51u + 256u * unsigned(M.idx_bg) + (M.ngten != nten ? 0x10000u : 0u) : 0x52u);
}

// ---- table indices (src/module_mp_tempo_main.F90:2493-2635, 2839-2858) ----
template <class R>
__device__ __forceinline__ int decade_index(R x, int base, int nmax) {
  const int nx = nint_exact(mlog10(x));
  int n = nx + 1;
  for (int nn = nx - 1; nn <= nx + 1; ++nn) {
    n = nn;
    const R q = x / powi(R(10.0), nn);
    if (q >= R(1.0) && q < R(10.0)) break;
  }
  const int idx = int(x / powi(R(10.0), n)) + 10 * (n - base) - (n - base);
  return max(1, min(idx, nmax));
}

template <class R>
__device__ __forceinline__ int temperature_index(R tempk) {
  const R tempc = tempk - R(kT0);
  int idx = int((tempc - R(2.5)) / R(5.0)) - 1;
  idx = max(1, -idx);
  return min(idx, kNtbT);
}

template <class R>
__device__ void rain_index(const Const<S_t<R>>& C, R rr, R ilamr, int& ir, int& ir1) {
  ir = decade_index(rr, C.nir2, kNtbR);
  const R lamr = R(1.) / ilamr;
  const R lam_exp = lamr * mpow(C.crg[2] * C.org2 * C.org1, R(kBmR));
  const R n0_exp = C.org1 * rr / C.am_r * mpow(lam_exp, C.cre[0]);
  ir1 = decade_index(n0_exp, C.nir3, kNtbR1);
}

template <class R>
__device__ void graupel_index(const Const<S_t<R>>& C, R rg, R ilamg, int idx, int& ig, int& ig1) {
  ig = decade_index(rg, C.nig2, kNtbG);
  const R lamg = R(1.) / ilamg;
  const R lam_exp = lamg * mpow(C.cgg[2][0] * C.ogg2 * C.ogg1, R(kBmG));
  const R n0_exp = C.ogg1 * rg / C.am_g[idx - 1] * mpow(lam_exp, C.cge[0][0]);
  ig1 = decade_index(n0_exp, C.nig3, kNtbG1);
}

template <class R>
__device__ void cloud_index(const Const<S_t<R>>& C, R rc, R nc, int& ic, int& in) {
  ic = decade_index(rc, C.nic2, kNtbC);
  const int n = nint_exact(R(1.0) + R(kNbins) * mlog(nc / C.t_nc1) / C.nic1);
  in = max(1, min(n, kNbins));
}

// ------------------------------------------------------ process rates ----
// warm_rain (src/module_mp_tempo_main.F90:2279-2355)
template <class R>
__device__ __forceinline__ void warm_rain(const Const<S_t<R>>& C, const Tables<S_t<R>>& T, const Lvl<R>& L, Tend<R>& t, R odt) {
  if (L.l_qr && L.mvd_r > R(kD0R)) {
    const R ef_rr = mx(R(-0.1), R(1.0) - mexp(R(2300.0) * (L.mvd_r - R(1950.0e-6))));
    t.pnr_rcr = ef_rr * R(2.0) * L.nr * L.rr;
  }
  if (L.l_qc && L.rc > R(0.01e-3)) {
    const int nu = get_nuc(L.nc);
    const R lamc = R(1.0) / L.ilamc;
    const R xdc = mx(R(kD0C) * R(1.e6), (mpow(L.rc / (C.am_r * L.nc), C.obmr)) * R(1.e6));
    const R dc_g = (mpow(C.ccg[2][nu] * C.ocg2[nu], C.obmr) / lamc) * R(1.e6);
    const R dc_b = mpow(xdc * xdc * xdc * dc_g * dc_g * dc_g - xdc * xdc * xdc * xdc * xdc * xdc, R(1.0 / 6.0));
    const R z0 = R(6.25e-6) * xdc * dc_b * dc_b * dc_b - R(0.4);
    const R zeta1 = R(0.5) * (z0 + mabs(z0));
    const R zeta = R(0.027) * L.rc * zeta1;
    const R t0 = R(0.5) * dc_b - R(7.5);
    const R taud = R(0.5) * (t0 + mabs(t0)) + R(kR1);
    const R tau = R(3.72) / (L.rc * taud);
    t.prr_wau = zeta / tau;
    t.prr_wau = mn(L.rc * odt, t.prr_wau);
    t.pnr_wau = t.prr_wau / (C.am_r * R(nu) * R(kAutoconNrFactor) * R(kD0R) * R(kD0R) * R(kD0R));
    t.pnc_wau = mn(L.nc * odt, t.prr_wau / (C.am_r * L.mvd_c * L.mvd_c * L.mvd_c));
  }
  if (L.l_qr && L.l_qc && L.mvd_r > R(kD0R) && L.mvd_c > R(kD0C)) {
    const R lamr = (R(3.0) + R(kMuR) + R(0.672)) / L.mvd_r;
    int idx = 1 + int(R(kNbins) * mlog(L.mvd_r / C.dr1) / C.log_dr_ratio);
    idx = min(idx, kNbins);
    const int jc = int(L.mvd_c * R(1.e6));
    const R ef_rw = T.efrw[(idx - 1) + kNbins * (jc - 1)];
    const R n0_r = L.nr * C.org2 * mpow(lamr, C.cre[1]);
    const R f = mpow(lamr + R(kFvR), -C.cre[8]);
    t.prr_rcw = L.rhof * C.t1_qr_qc * ef_rw * L.rc * n0_r * f;
    t.prr_rcw = mn(L.rc * odt, t.prr_rcw);
    t.pnc_rcw = L.rhof * C.t1_qr_qc * ef_rw * L.nc * n0_r * f;
    t.pnc_rcw = mn(L.nc * odt, t.pnc_rcw);
  }
}

// rain_snow_rain_graupel (:2638-2727)
template <class R>
__device__ __forceinline__ void rain_snow_graupel(const Const<S_t<R>>& C, const Tables<S_t<R>>& T, const Lvl<R>& L, Tend<R>& t, R odt) {
  if (L.l_qr && L.l_qs && L.rr >= R(1.e-6) && L.rs >= R(1.e-6)) {
    const int it = temperature_index(L.temp);
    int ir, ir1;
    rain_index(C, L.rr, L.ilamr, ir, ir1);
    const int is = decade_index(L.rs, C.nis2, kNtbS);
    const size_t e = size_t(is - 1) + kNtbS * (size_t(it - 1) + kNtbT * (size_t(ir1 - 1) + kNtbR1 * size_t(ir - 1)));
    const S_t<R>* q = T.qs + e * 12;
    // file order: 0 tcs_racs1 1 tmr_racs1 2 tcs_racs2 3 tmr_racs2 4 tcr_sacr1 5 tms_sacr1
    //             6 tcr_sacr2 7 tms_sacr2 8 tnr_racs1 9 tnr_racs2 10 tnr_sacr1 11 tnr_sacr2
    if (L.temp < R(kT0)) {
      t.prr_rcs = -(q[3] + q[6] + q[1] + q[4]);
      t.prs_rcs = q[3] + q[6] - q[0] - q[5];
      t.prg_rcs = q[1] + q[4] + q[0] + q[5];
      t.prr_rcs = mx(-L.rr * odt, t.prr_rcs);
      t.prs_rcs = mx(-L.rs * odt, t.prs_rcs);
      t.prg_rcs = mn((L.rr + L.rs) * odt, t.prg_rcs);
      t.pnr_rcs = q[8] + q[9] + q[10] + q[11];
      t.pnr_rcs = mn(L.nr * odt, t.pnr_rcs);
      t.png_rcs = t.pnr_rcs;
      t.pbg_rcs = R(kM3ToL) * t.prg_rcs / R(kRhoI);
    } else {
      t.prs_rcs = -q[0] - q[5] + q[3] + q[6];
      t.prs_rcs = mx(-L.rs * odt, t.prs_rcs);
      t.prr_rcs = -t.prs_rcs;
    }
  }
  if (L.l_qr && L.l_qg && L.rr >= R(1.e-6) && L.rg >= R(1.e-6)) {
    int ir, ir1, ig, ig1;
    const int idx = L.idx_bg;
    rain_index(C, L.rr, L.ilamr, ir, ir1);
    graupel_index(C, L.rg, L.ilamg, idx, ig, ig1);
    const size_t e = size_t(ig1 - 1) +
                     kNtbG1 * (size_t(ig - 1) + kNtbG * (size_t(idx - 1) + kNrhg * (size_t(ir1 - 1) + kNtbR1 * size_t(ir - 1))));
    const S_t<R>* q = T.qg + e * 5;  // tcg_racg tmr_racg tcr_gacr tnr_racg tnr_gacr
    if (L.temp < R(kT0)) {
      t.prg_rcg = q[1] + q[2];
      t.prg_rcg = mn(L.rr * odt, t.prg_rcg);
      t.prr_rcg = -t.prg_rcg;
      t.pnr_rcg = q[3] + q[4];
      t.pnr_rcg = mn(L.nr * odt, t.pnr_rcg);
      t.pbg_rcg = R(kM3ToL) * t.prg_rcg / R(kRhoI);
    } else {
      t.prr_rcg = q[0];
      t.prr_rcg = mn(L.rg * odt, t.prr_rcg);
      t.prg_rcg = -t.prr_rcg;
      t.png_rcg = q[3];
      t.png_rcg = mn(L.ng * odt, t.png_rcg);
      t.pbg_rcg = R(kM3ToL) * t.prg_rcg / rho_g<R>(idx);
      t.pnr_rcg = R(-1.5) * q[4];
    }
  }
}

// demott_nucleation (:2823-2836) and koop_nucleation (:2168-2195)
template <class R>
__device__ R demott(const Const<S_t<R>>& C, R tempc, R rho, R nifa) {
  const R nifa_cc = mx(R(0.5), nifa * C.rho_not0 * R(1.e-6) / rho);
  R xni = (R(5.94e-5) * mpow(-tempc, R(3.33))) * (mpow(nifa_cc, (R(-0.0264) * (tempc)) + R(0.0033)));
  xni = xni * rho / C.rho_not0 * R(1000.0);
  return mx(R(0.0), xni);
}
template <class R>
__device__ R koop(const Const<S_t<R>>& C, R temp, R satw, R naero, R dt) {
  R xni = R(0.0);
  const R mu_diff = R(210368.0) + (R(131.438) * temp) - (R(3.32373e6) / temp) - (R(41729.1) * mlog(temp));
  const R a_w_i = mexp(mu_diff / (R(kRUni) * temp));
  const R d = satw - a_w_i;
  R lj = R(-906.7) + (R(8502.0) * d) - (R(26924.0) * d * d) + (R(29180.0) * d * d * d);
  lj = mn(R(20.0), lj);
  const R j_rate = mpow(R(10.0), lj);
  const R prob_h = mn(R(1.0) - mexp(-j_rate * C.ar_volume * dt), R(1.0));
  if (prob_h > R(0.0)) xni = mn(prob_h * naero, R(1000.e3));
  return mx(R(0.0), xni);
}

// ice_nucleation (:2730-2820)
template <class R>
__device__ __forceinline__ void ice_nucleation(const Const<S_t<R>>& C, const Tables<S_t<R>>& T, const Lvl<R>& L, Tend<R>& t, bool has_nwfa,
                               bool has_nifa, R w, R dt, R odt) {
  if (!(L.temp < R(kT0))) return;
  const R tempc = L.temp - R(kT0);
  const int idx_tc = max(1, min(nint_exact(-tempc), kNtbT1));
  const R rate_max = (L.qv - L.qvsi) * L.rho * odt * R(0.999);
  R xni = has_nifa ? demott(C, tempc, L.rho, L.nifa) : R(1.0) * R(1000.0);
  const int idx_in = (xni > R(1.0)) ? decade_index(xni, C.niin2, kNtbIn) : 1;
  if (L.rr > R(1.e-6)) {
    int ir, ir1;
    rain_index(C, L.rr, L.ilamr, ir, ir1);
    const size_t e = size_t(ir - 1) + kNtbR * (size_t(ir1 - 1) + kNtbR1 * (size_t(idx_tc - 1) + kNtbT1 * size_t(idx_in - 1)));
    const S_t<R>* q = T.fzr + e * 4;  // tpi_qrfz tni_qrfz tpg_qrfz tnr_qrfz
    t.prg_rfz = q[2] * odt;
    t.pri_rfz = q[0] * odt;
    t.pni_rfz = q[1] * odt;
    t.pnr_rfz = q[3] * odt;
    t.prg_rfz = mn(L.rr * odt, t.prg_rfz);
    t.pnr_rfz = mn(L.nr * odt, t.pnr_rfz);
    t.png_rfz = t.pnr_rfz * mx(mn((mpow(R(10.0), R(-0.1) * w) + R(0.1)), R(1.0)), R(0.1));
  } else if (L.rr > R(kR1) && L.temp < R(kHgFrz)) {
    t.pri_rfz = L.rr * odt;
    t.pni_rfz = L.nr * odt;
  }
  t.pbg_rfz = R(kM3ToL) * t.prg_rfz / R(kRhoI);
  if (L.rc > R(1.e-6)) {
    int ic, in;
    cloud_index(C, L.rc, L.nc, ic, in);
    const size_t e = size_t(ic - 1) + kNtbC * (size_t(in - 1) + kNbins * (size_t(idx_tc - 1) + kNtbT1 * size_t(idx_in - 1)));
    const S_t<R>* q = T.fzc + e * 2;  // tpi_qcfz tni_qcfz
    t.pri_wfz = q[0] * odt;
    t.pri_wfz = mn(L.rc * odt, t.pri_wfz);
    t.pni_wfz = q[1] * odt;
    t.pni_wfz = mn(mn(L.nc * odt, t.pri_wfz / (R(2.0) * R(kXm0I))), t.pni_wfz);
  } else if (L.rc > R(kR1) && L.temp < R(kHgFrz)) {
    t.pri_wfz = L.rc * odt;
    t.pni_wfz = L.nc * odt;
  }
  if ((L.ssati >= R(kDemottSsati)) || (L.ssatw > R(kEps) && tempc < R(kDemottTempc))) {
    const R xnc = has_nifa ? demott(C, tempc, L.rho, L.nifa)
                           : mn(R(kIcenucMax), R(kTno) * mexp(R(kAto) * (R(kT0) - L.temp)));
    xni = L.ni + (t.pni_rfz + t.pni_wfz) * dt;
    t.pni_inu = R(0.5) * (xnc - xni + mabs(xnc - xni)) * odt;
    t.pri_inu = mn(rate_max, R(kXm0I) * t.pni_inu);
    t.pni_inu = t.pri_inu / R(kXm0I);
  }
  xni = L.smo0 + L.ni + (t.pni_rfz + t.pni_wfz + t.pni_inu) * dt;
  if (has_nwfa && (xni <= R(kMaxNi)) && (L.temp < R(238.0)) && (L.ssati >= R(0.4))) {
    const R xnc = koop(C, L.temp, L.ssatw, L.nwfa, dt);
    t.pni_iha = xnc * odt;
    t.pri_iha = mn(rate_max, R(kXm0I) * R(0.1) * t.pni_iha);
    t.pni_iha = t.pri_iha / (R(kXm0I) * R(0.1));
  }
}

// get_t1_subl (:2861-2888)
template <class R>
__device__ R t1_subl(const Const<S_t<R>>& C, const Lvl<R>& L) {
  const R otemp = R(1.0) / L.temp, rvs = L.rho * L.qvsi;
  const R a = R(kLsub) * otemp * C.orv - R(1.0);
  const R rvs_p = rvs * otemp * a;
  const R rvs_pp = rvs * (otemp * a * otemp * a + (R(-2.) * R(kLsub) * otemp * otemp * otemp * C.orv) + otemp * otemp);
  const R gamsc = R(kLsub) * L.diffu / L.tcond * rvs_p;
  R alphsc = R(0.5) * (gamsc / (R(1.0) + gamsc)) * (gamsc / (R(1.0) + gamsc)) * rvs_pp / rvs_p * rvs / rvs_p;
  alphsc = mx(R(1.e-9), alphsc);
  R xs = L.ssati;
  if (mabs(xs) < R(1.e-9)) xs = R(0.0);
  return R(4.0) * R(kPi) *
         (R(1.0) - alphsc * xs + R(2.0) * alphsc * alphsc * xs * xs - R(5.0) * alphsc * alphsc * alphsc * xs * xs * xs) /
         (R(1.0) + gamsc);
}

// ice_processes (:2891-3018)
template <class R>
__device__ __forceinline__ void ice_processes(const Const<S_t<R>>& C, const Tables<S_t<R>>& T, const Lvl<R>& L, Tend<R>& t, R t1s, R w,
                              R odt) {
  if (!(L.temp < R(kT0))) return;
  const R rvs = L.rho * L.qvsi;
  const R rate_max = (L.qv - L.qvsi) * L.rho * odt * R(0.999);
  if (L.l_qi) {
    const int ii = decade_index(L.ri, C.nii2, kNtbI);
    const int ii1 = decade_index(L.ni, C.nii3, kNtbI1);
    const R xdi = mx(R(C.d0i), (R(kBmI) + R(kMuI) + R(1.)) * L.ilami);
    const R xmi = C.am_i * mpow(xdi, R(kBmI));
    const R oxmi = R(1.0) / xmi;
    const size_t it = size_t(ii - 1) + size_t(kNtbI) * (ii1 - 1);
    t.pri_ide = R(kCCube) * t1s * L.diffu * L.ssati * rvs * C.oig1 * C.cig[4] * L.ni * L.ilami;
    if (t.pri_ide < R(0.0)) {
      t.pri_ide = mx(mx(-L.ri * odt, t.pri_ide), rate_max);
      t.pni_ide = t.pri_ide * oxmi;
      t.pni_ide = mx(-L.ni * odt, t.pni_ide);
    } else {
      t.pri_ide = mn(t.pri_ide, rate_max);
      t.prs_ide = (R(1.0) - T.tpi_ide[it]) * t.pri_ide;
      t.pri_ide = T.tpi_ide[it] * t.pri_ide;
    }
    if ((ii == kNtbI) || (xdi > R(5.0) * R(kD0S))) {
      t.prs_iau = L.ri * R(.99) * odt;
      t.pni_iau = L.ni * R(.95) * odt;
    } else if (xdi < R(0.1) * R(kD0S)) {
      t.prs_iau = R(0.0);
      t.pni_iau = R(0.0);
    } else {
      t.prs_iau = T.tps_iaus[it] * odt;
      t.prs_iau = mn(L.ri * R(.99) * odt, t.prs_iau);
      t.pni_iau = T.tni_iaus[it] * odt;
      t.pni_iau = mn(L.ni * R(.95) * odt, t.pni_iau);
    }
    if (L.rs >= R(1.e-6)) {
      t.prs_sci = C.t1_qs_qi * L.rhof * R(kEfSI) * L.ri * L.smoe;
      t.pni_sci = t.prs_sci * oxmi;
    }
    if (L.rr >= R(1.e-6) && L.mvd_r > R(4.0) * xdi) {
      const R lamr = R(1.0) / L.ilamr;
      const R n0_r = L.nr * C.org2 * mpow(lamr, C.cre[1]);
      const R f9 = mpow(lamr + R(kFvR), -C.cre[8]);
      t.pri_rci = L.rhof * C.t1_qr_qi * R(kEfRI) * L.ri * n0_r * f9;
      t.pnr_rci = L.rhof * C.t1_qr_qi * R(kEfRI) * L.ni * n0_r * f9;
      t.pnr_rci = mn(L.nr * odt, t.pnr_rci);
      t.png_rci = t.pnr_rci * mx(mn((mpow(R(10.0), R(-0.1) * w) + R(0.1)), R(1.0)), R(0.1));
      t.pni_rci = t.pri_rci * oxmi;
      t.prr_rci = L.rhof * C.t2_qr_qi * R(kEfRI) * L.ni * n0_r * mpow(lamr + R(kFvR), -C.cre[7]);
      t.prr_rci = mn(L.rr * odt, t.prr_rci);
      t.prg_rci = t.pri_rci + t.prr_rci;
      t.pbg_rci = t.prg_rci / R(kRhoI);
    }
  }
  if (L.l_qs) {
    R c_snow = R(kCSqrd) + (L.temp - R(kT0) + R(1.5)) * (R(kCCube) - R(kCSqrd)) / (R(-30.0) + R(1.5));
    c_snow = mx(R(kCSqrd), mn(c_snow, R(kCCube)));
    t.prs_sde = c_snow * t1s * L.diffu * L.ssati * rvs * (C.t1_qs_sd * L.smo1 + C.t2_qs_sd * L.rhof2 * L.vsc2 * L.smof);
    if (t.prs_sde < R(0.0))
      t.prs_sde = mx(mx(-L.rs * odt, t.prs_sde), rate_max);
    else
      t.prs_sde = mn(t.prs_sde, rate_max);
  }
  if (L.l_qg && L.ssati < R(-kEps)) {
    const int idx = L.idx_bg;
    const R n0_g = L.ng * C.ogg2 * mpow(R(1.0) / L.ilamg, C.cge[1][0]);
    const R t2_qg_sd = R(0.28) * C.sc3 * msqrt(C.av_g[idx - 1]) * C.cgg[10][idx - 1];
    t.prg_gde = R(kCCube) * t1s * L.diffu * L.ssati * rvs * n0_g *
                (C.t1_qg_sd * mpow(L.ilamg, C.cge[9][0]) + t2_qg_sd * L.vsc2 * L.rhof2 * mpow(L.ilamg, C.cge[10][idx - 1]));
    if (t.prg_gde < R(0.0)) {
      t.prg_gde = mx(mx(-L.rg * odt, t.prg_gde), rate_max);
      t.png_gde = t.prg_gde * L.ng / L.rg;
    } else {
      t.prg_gde = mn(t.prg_gde, rate_max);
    }
  }
}

// riming (:2358-2490); returns vtboost
template <class R>
__device__ __forceinline__ R riming(const Const<S_t<R>>& C, const Tables<S_t<R>>& T, const Lvl<R>& L, Tend<R>& t, R odt) {
  const R tempc = L.temp - R(kT0);
  R vtboost = R(1.0);
  if (L.l_qc && L.l_qs) {
    const R xds = L.smoc / L.smob;
    if ((L.mvd_c > R(kD0C)) && (xds > R(kD0S))) {
      int idxs = 1 + int(R(kNbins) * mlog(xds / C.ds1) / C.log_ds_ratio);
      idxs = min(idxs, kNbins);
      const int jc = int(L.mvd_c * R(1.e6));
      const R ef_sw = T.efsw[(idxs - 1) + kNbins * (jc - 1)];
      t.prs_scw = L.rhof * C.t1_qs_qc * ef_sw * L.rc * L.smoe;
      t.prs_scw = mn(L.rc * odt, t.prs_scw);
      t.pnc_scw = L.rhof * C.t1_qs_qc * ef_sw * L.nc * L.smoe;
      t.pnc_scw = mn(L.nc * odt, t.pnc_scw);
      if (L.temp < R(kT0) && t.prs_scw > R(kRimeThreshold) * t.prs_sde && t.prs_sde > R(kEps)) {
        const R r_frac = mn(R(30.0), t.prs_scw / t.prs_sde);
        R g_frac = mn(R(kRimeConversion), R(0.15) + (r_frac - R(2.0)) * R(.028));
        vtboost = mn(R(1.5), R(1.1) + (r_frac - R(2.)) * R(.014));
        t.prg_scw = g_frac * t.prs_scw;
        t.png_scw = t.prg_scw * L.smo0 / L.rs;
        const R vts = R(kAvS) * mpow(xds, R(kBvS)) * mexp(-R(kFvS) * xds);
        R const_ri = R(-1.0) * (L.mvd_c * R(0.5e6)) * vts / mn(R(-0.1), tempc);
        const_ri = mx(R(0.1), mn(const_ri, R(10.0)));
        const R rime_dens = (R(0.051) + R(0.114) * const_ri - R(0.0055) * const_ri * const_ri) * R(1000.0);
        if (rime_dens < R(150.0)) {
          g_frac = R(0.0);
          t.prg_scw = R(0.0);
          t.png_scw = R(0.0);
        }
        const R sdf = mn(R(1.0), mx(R(0.0), L.rs * odt / (L.rs * odt + t.prg_scw)));
        t.pbg_scw = R(kM3ToL) * t.prg_scw / (R(kRhoS) * sdf + rime_dens * (R(1.0) - sdf));
        t.prs_scw = (R(1.0) - g_frac) * t.prs_scw;
      }
    }
  }
  if (L.l_qc && L.l_qg && L.rg >= R(1.e-6) && L.mvd_c > R(kD0C)) {
    const int idx = L.idx_bg;
    const R xdg = (R(kBmG) + R(kMuG) + R(1.0)) * L.ilamg;
    const R vtg = L.rhof * C.av_g[idx - 1] * C.cgg[5][idx - 1] * C.ogg3 * mpow(L.ilamg, C.bv_g[idx - 1]);
    const R stoke_g = L.mvd_c * L.mvd_c * vtg * R(kRhoW) / (R(9.0) * L.visco * xdg);
    R const_ri = R(-1.0) * (L.mvd_c * R(0.5e6)) * vtg / mn(R(-0.1), tempc);
    const_ri = mx(R(0.1), mn(const_ri, R(10.0)));
    const R rime_dens = (R(0.051) + R(0.114) * const_ri - R(0.0055) * const_ri * const_ri) * R(1000.0);
    if (xdg > R(kD0G)) {
      R ef_gw = R(0.0);
      if (stoke_g >= R(0.4) && stoke_g <= R(10.0))
        ef_gw = R(0.55) * mlog10(R(2.51) * stoke_g);
      else if (stoke_g > R(10.0))
        ef_gw = R(0.77);
      if (L.temp > R(kT0)) ef_gw = ef_gw * R(0.1);
      const R t1_qg_qc = R(kPi) * R(.25) * C.av_g[idx - 1] * C.cgg[8][idx - 1];
      const R n0_g = L.ng * C.ogg2 * mpow(R(1.0) / L.ilamg, C.cge[1][0]);
      const R f9 = mpow(L.ilamg, C.cge[8][idx - 1]);
      t.prg_gcw = L.rhof * t1_qg_qc * ef_gw * L.rc * n0_g * f9;
      t.pnc_gcw = L.rhof * t1_qg_qc * ef_gw * L.nc * n0_g * f9;
      t.pnc_gcw = mn(L.nc * odt, t.pnc_gcw);
      if (L.temp < R(kT0)) {
        t.pbg_gcw = R(kM3ToL) * t.prg_gcw / rime_dens;
        if (t.prg_gcw > R(kEps) && tempc > R(-8.0)) {
          R tf = R(0.0);
          if (tempc >= R(-5.0) && tempc < R(-3.0))
            tf = R(0.5) * (R(-3.0) - tempc);
          else if (tempc > R(-8.0) && tempc < R(-5.0))
            tf = R(0.33333333) * (R(8.0) + tempc);
          t.pni_ihm = R(3.5e8) * tf * t.prg_gcw;
          t.pri_ihm = R(kXm0I) * t.pni_ihm;
          t.prs_ihm = t.prs_scw / (t.prs_scw + t.prg_gcw) * t.pri_ihm;
          t.prg_ihm = t.prg_gcw / (t.prs_scw + t.prg_gcw) * t.pri_ihm;
        }
      }
    }
  }
  return vtboost;
}

// melting (:3021-3119)
template <class R>
__device__ __forceinline__ void melting(const Const<S_t<R>>& C, const Lvl<R>& L, Tend<R>& t, R t1s, R dt, R odt) {
  if (!(L.temp > R(kT0))) return;
  const R tempc = L.temp - R(kT0), rvs = L.rho * L.qvsi;
  const R drive = tempc * L.tcond - R(kLvap0) * L.diffu * L.delqvs;
  if (L.l_qs) {
    t.prr_sml = drive * (C.t1_qs_me * L.smo1 + C.t2_qs_me * L.rhof2 * L.vsc2 * L.smof);
    if (t.prr_sml > R(0.0)) {
      t.prr_sml = t.prr_sml + R(4218.0) * C.olfus * tempc * (t.prr_rcs + t.prs_scw);
      t.prr_sml = mn(L.rs * odt, mx(R(0.0), t.prr_sml));
      t.pnr_sml = L.smo0 / L.rs * t.prr_sml * mpow(R(10.0), R(-0.25) * tempc);
      t.pnr_sml = mn(L.smo0 * odt, t.pnr_sml);
    } else {
      t.prr_sml = R(0.0);
      t.pnr_sml = R(0.0);
      if (L.ssati < R(0.0)) {
        t.prs_sde = R(kCCube) * t1s * L.diffu * L.ssati * rvs * (C.t1_qs_sd * L.smo1 + C.t2_qs_sd * L.rhof2 * L.vsc2 * L.smof);
        t.prs_sde = mx(-L.rs * odt, t.prs_sde);
      }
    }
  }
  if (L.l_qg) {
    const int idx = L.idx_bg;
    const R n0_g = L.ng * C.ogg2 * mpow(R(1.0) / L.ilamg, C.cge[1][0]);
    R n0_melt = n0_g;
    if ((L.rg * L.ng) < R(1.e-4)) {
      const R lamg = R(1.) / L.ilamg;
      n0_melt = (R(1.e-4) / L.rg) * C.ogg2 * mpow(lamg, C.cge[1][0]);
    }
    const R t2_qg_me = R(kPi) * R(4.0) * R(kCCube) * C.olfus * R(0.28) * C.sc3 * msqrt(C.av_g[idx - 1]) * C.cgg[10][idx - 1];
    const R f10 = mpow(L.ilamg, C.cge[9][0]), f11 = mpow(L.ilamg, C.cge[10][idx - 1]);
    t.prr_gml = drive * n0_melt * (C.t1_qg_me * f10 + t2_qg_me * L.rhof2 * L.vsc2 * f11);
    t.prr_gml = mn(L.rg * odt, mx(R(0.0), t.prr_gml));
    if (t.prr_gml > R(0.0)) {
      const R melt_f = mx(R(0.05), mn(t.prr_gml * dt / L.rg, R(1.0)));
      t.pbg_gml = R(kM3ToL) * t.prr_gml / mx(mn(melt_f * rho_g<R>(idx), R(kRhoW)), R(50.0));
      t.pnr_gml = t.prr_gml * L.ng / L.rg * mpow(R(10.0), R(-0.33) * (L.temp - R(kT0)));
    } else {
      t.prr_gml = R(0.0);
      t.pnr_gml = R(0.0);
      t.pbg_gml = R(0.0);
      if (L.ssati < R(0.0)) {
        const R t2_qg_sd = R(0.28) * C.sc3 * msqrt(C.av_g[idx - 1]) * C.cgg[10][idx - 1];
        t.prg_gde = R(kCCube) * t1s * L.diffu * L.ssati * rvs * n0_g * (C.t1_qg_sd * f10 + t2_qg_sd * L.vsc2 * L.rhof2 * f11);
        t.prg_gde = mx(-L.rg * odt, t.prg_gde);
        t.png_gde = t.prg_gde * L.ng / L.rg;
      }
    }
  }
  if (dt > R(kDtRimeToRain)) {
    t.prr_rcw = t.prr_rcw + t.prs_scw + t.prg_gcw;
    t.prs_scw = R(0.0);
    t.prg_gcw = R(0.0);
  }
}

// aerosol_collection_efficiency (src/module_mp_tempo_aerosols.F90:76-118)
template <class R>
__device__ R aero_eff(const Const<S_t<R>>& C, R d, R da, R visc, R rhoa, R temp, char species) {
  const R boltzman = R(1.3806503e-23), meanpath = R(0.0256e-6);
  R vt = R(1.0), rho_p = R(kRhoW);
  if (species == 'r') {
    vt = R(-0.1021) + R(4.932e3) * d - R(0.9551e6) * d * d + R(0.07934e9) * d * d * d - R(0.002362e12) * d * d * d * d;
  } else if (species == 's') {
    vt = R(kAvS) * mpow(d, R(kBvS));
    rho_p = R(kRhoS);
  } else {
    vt = C.av_g[kIdxBg1 - 1] * mpow(d, C.bv_g[kIdxBg1 - 1]);
    rho_p = rho_g<R>(kIdxBg1);
  }
  const R cc = R(1.0) + R(2.0) * meanpath / da * (R(1.257) + R(0.4) * mexp(R(-0.55) * da / meanpath));
  const R diff = boltzman * temp * cc / (R(3.0) * R(kPi) * visc * da);
  const R re = R(0.5) * rhoa * d * vt / visc;
  const R sc = visc / (rhoa * diff);
  const R st = (rho_p - rhoa) * da * da * vt * cc / (R(9.0) * visc * d);
  const R aval = mlog(R(1.0) + re);
  const R st2 = (R(1.2) + R(1.0 / 12.0) * aval) / (R(1.0) + aval);
  const R sre = msqrt(re);
  R eff = R(4.0) / (re * sc) * (R(1.0) + R(0.4) * sre * mpow(sc, R(0.3333)) + R(0.16) * sre * msqrt(sc)) +
          R(4.0) * da / d * (R(0.02) + da / d * (R(1.0) + R(2.0) * sre));
  if (st > st2) eff = eff + mpow((st - st2) / (st - st2 + R(0.666667)), R(1.5));
  return mx(R(1.e-5), mn(eff, R(1.0)));
}

// aerosol_scavenging (src/module_mp_tempo_main.F90:3122-3193)
template <class R>
__device__ __forceinline__ void aerosol_scavenging(const Const<S_t<R>>& C, const Lvl<R>& L, Tend<R>& t, R odt) {
  const R wf = R(0.04e-6), iff = R(0.8e-6);
  if (L.l_qr && L.mvd_r > R(kD0R)) {
    R ef = aero_eff(C, L.mvd_r, wf, L.visco, L.rho, L.temp, 'r');
    const R lamr = R(1.0) / L.ilamr;
    const R n0_r = L.nr * C.org2 * mpow(lamr, C.cre[1]);
    const R f9 = mpow(lamr + R(kFvR), -C.cre[8]);
    t.pna_rca = L.rhof * C.t1_qr_qc * ef * L.nwfa * n0_r * f9;
    t.pna_rca = mn(L.nwfa * odt, t.pna_rca);
    ef = aero_eff(C, L.mvd_r, iff, L.visco, L.rho, L.temp, 'r');
    t.pnd_rcd = L.rhof * C.t1_qr_qc * ef * L.nifa * n0_r * f9;
    t.pnd_rcd = mn(L.nifa * odt, t.pnd_rcd);
  }
  if (L.l_qs && L.rs > R(1.e-6)) {
    const R xds = L.smoc / L.smob;
    R ef = aero_eff(C, xds, wf, L.visco, L.rho, L.temp, 's');
    t.pna_sca = L.rhof * C.t1_qs_qc * ef * L.nwfa * L.smoe;
    t.pna_sca = mn(L.nwfa * odt, t.pna_sca);
    ef = aero_eff(C, xds, iff, L.visco, L.rho, L.temp, 's');
    t.pnd_scd = L.rhof * C.t1_qs_qc * ef * L.nifa * L.smoe;
    t.pnd_scd = mn(L.nifa * odt, t.pnd_scd);
  }
  if (L.l_qg && L.rg > R(1.e-6)) {
    const int idx = L.idx_bg;
    const R xdg = (R(kBmG) + R(kMuG) + R(1.0)) * L.ilamg;
    R ef = aero_eff(C, xdg, wf, L.visco, L.rho, L.temp, 'g');
    const R t1_qg_qc = R(kPi) * R(.25) * C.av_g[idx - 1] * C.cgg[8][idx - 1];
    const R n0_g = L.ng * C.ogg2 * mpow(R(1.0) / L.ilamg, C.cge[1][0]);
    const R f9 = mpow(L.ilamg, C.cge[8][idx - 1]);
    t.pna_gca = L.rhof * t1_qg_qc * ef * L.nwfa * n0_g * f9;
    t.pna_gca = mn(L.nwfa * odt, t.pna_gca);
    ef = aero_eff(C, xdg, iff, L.visco, L.rho, L.temp, 'g');
    t.pnd_gcd = L.rhof * t1_qg_qc * ef * L.nifa * n0_g * f9;
    t.pnd_gcd = mn(L.nifa * odt, t.pnd_gcd);
  }
}

// check_over_depletion (:1313-1438)
template <class R>
__device__ __forceinline__ void check_over_depletion(const Lvl<R>& L, Tend<R>& t, R odt) {
  R sump, rate_max, ratio;
  sump = t.pri_inu + t.pri_ide + t.prs_ide + t.prs_sde + t.prg_gde + t.pri_iha;
  rate_max = (L.qv - L.qvsi) * L.rho * odt * R(0.999);
  if ((sump > R(kEps) && sump > rate_max) || (sump < R(-kEps) && sump < rate_max)) {
    ratio = rate_max / sump;
    t.pri_inu *= ratio; t.pri_ide *= ratio; t.pni_ide *= ratio; t.prs_ide *= ratio;
    t.prs_sde *= ratio; t.prg_gde *= ratio; t.pri_iha *= ratio;
  }
  sump = -t.prr_wau - t.pri_wfz - t.prr_rcw - t.prs_scw - t.prg_scw - t.prg_gcw;
  rate_max = -L.rc * odt;
  if (L.l_qc && sump < rate_max) {
    ratio = rate_max / sump;
    t.prr_wau *= ratio; t.pri_wfz *= ratio; t.prr_rcw *= ratio; t.prs_scw *= ratio;
    t.prg_scw *= ratio; t.prg_gcw *= ratio;
  }
  sump = t.pri_ide - t.prs_iau - t.prs_sci - t.pri_rci;
  rate_max = -L.ri * odt;
  if (L.l_qi && sump < rate_max) {
    ratio = rate_max / sump;
    t.pri_ide *= ratio; t.prs_iau *= ratio; t.prs_sci *= ratio; t.pri_rci *= ratio;
  }
  sump = -t.prg_rfz - t.pri_rfz - t.prr_rci + t.prr_rcs + t.prr_rcg;
  rate_max = -L.rr * odt;
  if (L.l_qr && sump < rate_max) {
    ratio = rate_max / sump;
    t.prg_rfz *= ratio; t.pbg_rfz *= ratio; t.pri_rfz *= ratio; t.prr_rci *= ratio;
    t.prr_rcs *= ratio; t.prr_rcg *= ratio;
  }
  sump = t.prs_sde - t.prs_ihm - t.prr_sml + t.prs_rcs;
  rate_max = -L.rs * odt;
  if (L.l_qs && sump < rate_max) {
    ratio = rate_max / sump;
    t.prs_sde *= ratio; t.prs_ihm *= ratio; t.prr_sml *= ratio; t.prs_rcs *= ratio;
  }
  sump = t.prg_gde - t.prg_ihm - t.prr_gml + t.prg_rcg;
  rate_max = -L.rg * odt;
  if (L.l_qg && sump < rate_max) {
    ratio = rate_max / sump;
    t.prg_gde *= ratio; t.prg_ihm *= ratio; t.prr_gml *= ratio; t.prg_rcg *= ratio; t.pbg_rcg *= ratio;
  }
  t.pri_ihm = t.prs_ihm + t.prg_ihm;
  ratio = mn(mabs(t.prr_rcg), mabs(t.prg_rcg));
  t.prr_rcg = ratio * copysign(R(1.0), t.prr_rcg);
  t.prg_rcg = -t.prr_rcg;
  t.pbg_rcg = R(kM3ToL) * t.prg_rcg / R(kRhoI);
  if (L.temp > R(kT0)) {
    ratio = mn(mabs(t.prr_rcs), mabs(t.prs_rcs));
    t.prr_rcs = ratio * copysign(R(1.0), t.prr_rcs);
    t.prs_rcs = -t.prr_rcs;
  }
}

// sum_tendencies (:1441-1513)
template <class R>
__device__ __forceinline__ void sum_tendencies(const Const<S_t<R>>& C, Lvl<R>& L, const Tend<R>& t) {
  const R orho = R(1.) / L.rho, lfus2 = R(kLsub) - L.lvap;
  L.qvten = L.qvten + (-t.pri_inu - t.pri_iha - t.pri_ide - t.prs_ide - t.prs_sde - t.prg_gde) * orho;
  L.qcten = L.qcten + (-t.prr_wau - t.pri_wfz - t.prr_rcw - t.prs_scw - t.prg_scw - t.prg_gcw) * orho;
  L.ncten = L.ncten + (-t.pnc_wau - t.pnc_rcw - t.pni_wfz - t.pnc_scw - t.pnc_gcw) * orho;
  L.qiten = L.qiten + (t.pri_inu + t.pri_iha + t.pri_ihm + t.pri_wfz + t.pri_rfz + t.pri_ide - t.prs_iau -
                       t.prs_sci - t.pri_rci) * orho;
  L.niten = L.niten + (t.pni_inu + t.pni_iha + t.pni_ihm + t.pni_wfz + t.pni_rfz + t.pni_ide - t.pni_iau -
                       t.pni_sci - t.pni_rci) * orho;
  L.qrten = L.qrten + (t.prr_wau + t.prr_rcw + t.prr_sml + t.prr_gml + t.prr_rcs + t.prr_rcg - t.prg_rfz -
                       t.pri_rfz - t.prr_rci) * orho;
  L.nrten = L.nrten + (t.pnr_wau + t.pnr_sml + t.pnr_gml -
                       (t.pnr_rfz + t.pnr_rcr + t.pnr_rcg + t.pnr_rcs + t.pnr_rci + t.pni_rfz)) * orho;
  L.qsten = L.qsten + (t.prs_iau + t.prs_sde + t.prs_sci + t.prs_scw + t.prs_rcs + t.prs_ide - t.prs_ihm -
                       t.prr_sml) * orho;
  L.qgten = L.qgten + (t.prg_scw + t.prg_rfz + t.prg_gde + t.prg_rcg + t.prg_gcw + t.prg_rci + t.prg_rcs -
                       t.prg_ihm - t.prr_gml) * orho;
  L.ngten = L.ngten + (t.png_scw + t.png_rfz - t.png_rcg + t.png_rci + t.png_rcs + t.png_gde - t.pnr_gml) * orho;
  L.qbten = L.qbten + (t.pbg_scw + t.pbg_rfz + t.pbg_gcw + t.pbg_rci + t.pbg_rcs + t.pbg_rcg + t.pbg_sml -
                       t.pbg_gml + R(kM3ToL) * (t.prg_gde - t.prg_ihm) / rho_g<R>(L.idx_bg)) * orho;
  if (L.temp < R(kT0)) {
    L.tten = L.tten + (R(kLsub) * L.ocp * (t.pri_inu + t.pri_ide + t.prs_ide + t.prs_sde + t.prg_gde + t.pri_iha) +
                       lfus2 * L.ocp * (t.pri_wfz + t.pri_rfz + t.prg_rfz + t.prs_scw + t.prg_scw + t.prg_gcw +
                                        t.prg_rcs + t.prs_rcs + t.prr_rci + t.prg_rcg)) * orho;
  } else {
    L.tten = L.tten + (C.lfus * L.ocp * (-t.prr_sml - t.prr_gml - t.prr_rcg - t.prr_rcs) +
                       R(kLsub) * L.ocp * (t.prs_sde + t.prg_gde)) * orho;
  }
}

// the activation table's aerosol-number and vertical-velocity breakpoints
// (ta_na, ta_ww, src/module_mp_tempo_params.F90:266-270) by index, as
// selects rather than a private array
template <class R>
__device__ __forceinline__ R ccn_na(int i) {
  return i == 0 ? R(10.) : i == 1 ? R(31.6) : i == 2 ? R(100.) : i == 3 ? R(316.) : i == 4 ? R(1000.)
       : i == 5 ? R(3160.) : R(10000.);
}
template <class R>
__device__ __forceinline__ R ccn_ww(int i) {
  return i == 0 ? R(0.01) : i == 1 ? R(0.0316) : i == 2 ? R(0.1) : i == 3 ? R(0.316) : i == 4 ? R(1.)
       : i == 5 ? R(3.16) : i == 6 ? R(10.) : i == 7 ? R(31.6) : R(100.);
}

// activate_cloud_number (:2198-2276), land never passed by the caller
template <class R>
__device__ R activate_cloud_number(const Tables<S_t<R>>& T, R temp, R w1d, R nwfa) {
  auto na = [](int i) { return ccn_na<R>(i); };
  auto ww = [](int i) { return ccn_ww<R>(i); };
  R n_local = nwfa * R(1.e-6);
  if (n_local >= na(6)) n_local = na(6) - R(1.0);
  else if (n_local <= na(0)) n_local = na(0) + R(1.0);
  int i = 7;
  for (int n = 2; n <= 7; ++n)
    if (n_local >= na(n - 2) && n_local < na(n - 1)) { i = n; break; }
  R w_local = w1d;
  if (w_local >= ww(8)) w_local = ww(8) - R(1.0);
  else if (w_local <= ww(0)) w_local = ww(0) + R(0.001);
  int j = 9;
  for (int n = 2; n <= 9; ++n)
    if (w_local >= ww(n - 2) && w_local < ww(n - 1)) { j = n; break; }
  const R x1 = mlog(na(i - 2)), x2 = mlog(na(i - 1)), y1 = mlog(ww(j - 2)), y2 = mlog(ww(j - 1));
  const int k = max(1, min(nint_exact((temp - R(243.15)) * R(0.1)) + 1, kNtbArt));
  auto at = [&](int ii, int jj) { return R(T.ccn[(ii - 1) + 7 * ((jj - 1) + 9 * (k - 1))]); };
  const R a = at(i - 1, j - 1), b = at(i, j - 1), c = at(i, j), d = at(i - 1, j);
  const R t = (mlog(n_local) - x1) / (x2 - x1), u = (mlog(w_local) - y1) / (y2 - y1);
  R fraction = (R(1.0) - t) * (R(1.0) - u) * a + t * (R(1.0) - u) * b + t * u * c + (R(1.0) - t) * u * d;
  fraction = mx(fraction, R(0.0));
  return nwfa * fraction;
}

// cloud_condensation (:1980-2056)
template <class R>
__device__ __forceinline__ void cloud_condensation(const Const<S_t<R>>& C, const Tables<S_t<R>>& T, const Lvl<R>& L, Tend<R>& t, bool has_nwfa,
                                   bool has_nc, R w, R dt, R odt) {
  if (mabs(L.ssatw) < R(kEps)) return;
  const R orho = R(1.0) / L.rho;
  R clap = (L.qv - L.qvs) / (R(1.0) + L.lvt2 * L.qvs);
  for (int n = 0; n < 3; ++n) {
    const R e = mexp(L.lvt2 * clap);
    const R fcd = L.qvs * e - L.qv + clap;
    const R dfcd = L.qvs * L.lvt2 * e + R(1.0);
    clap = clap - fcd / dfcd;
  }
  const R xrc = L.rc + clap * L.rho;
  if (xrc > R(kR1)) {
    t.prw_vcd = clap * odt;
    if (clap > R(kEps)) {
      R xnc;
      if (has_nwfa) xnc = mx(R(kNtCMin), activate_cloud_number(T, L.temp, w, L.nwfa));
      else if (!has_nc) xnc = L.ncsave;
      else xnc = R(kNtCL);
      t.pnc_wcd = R(0.5) * (xnc - L.nc + mabs(xnc - L.nc)) * odt * orho;
    } else if (L.l_qc && L.ssatw < R(-1.e-6) && clap < R(-kEps)) {
      const R otemp = R(1.0) / L.temp, rvs = L.rho * L.qvs;
      const R a = L.lvap * otemp * C.orv - R(1.0);
      const R rvs_p = rvs * otemp * a;
      const R rvs_pp = rvs * (otemp * a * otemp * a + (R(-2.0) * L.lvap * otemp * otemp * otemp * C.orv) + otemp * otemp);
      const R gamsc = L.lvap * L.diffu / L.tcond * rvs_p;
      R alphsc = R(0.5) * (gamsc / (R(1.0) + gamsc)) * (gamsc / (R(1.0) + gamsc)) * rvs_pp / rvs_p * rvs / rvs_p;
      alphsc = mx(R(1.e-9), alphsc);
      R xs = L.ssatw;
      if (mabs(xs) < R(1.e-9)) xs = R(0.0);
      const R t1_evap = R(2.0) * R(kPi) *
                        (R(1.0) - alphsc * xs + R(2.0) * alphsc * alphsc * xs * xs -
                         R(5.0) * alphsc * alphsc * alphsc * xs * xs * xs) /
                        (R(1.0) + gamsc);
      const R dc_star = msqrt(R(-2.0) * dt * t1_evap / (R(2.0) * R(kPi)) * R(4.0) * L.diffu * L.ssatw * rvs / R(kRhoW));
      const int idx_d = max(1, min(int(R(1.e6) * dc_star), kNbins));
      int ic, in;
      cloud_index(C, L.rc, L.nc, ic, in);
      t.prw_vcd = mx(-L.rc * R(0.99) * orho * odt, t.prw_vcd);
      const R wev = T.tnc_wev[(idx_d - 1) + kNbins * (size_t(ic - 1) + kNtbC * size_t(in - 1))];
      t.pnc_wcd = mx(-L.nc * R(0.99) * orho * odt, -wev * orho * odt);
    }
  } else {
    t.prw_vcd = -L.rc * orho * odt;
    t.pnc_wcd = -L.nc * orho * odt;
  }
}

// rain_evaporation (:2059-2121)
template <class R>
__device__ __forceinline__ void rain_evaporation(const Const<S_t<R>>& C, const Lvl<R>& L, Tend<R>& t, R odt) {
  if (!L.l_qr || !((L.ssatw < R(-kEps)) && t.prw_vcd <= R(0.0))) return;
  const R orho = R(1.0) / L.rho, tempc = L.temp - R(kT0), otemp = R(1.0) / L.temp;
  const R rvs = L.rho * L.qvs;
  const R a = L.lvap * otemp * C.orv - R(1.0);
  const R rvs_p = rvs * otemp * a;
  const R rvs_pp = rvs * (otemp * a * otemp * a + (R(-2.0) * L.lvap * otemp * otemp * otemp * C.orv) + otemp * otemp);
  const R gamsc = L.lvap * L.diffu / L.tcond * rvs_p;
  R alphsc = R(0.5) * (gamsc / (R(1.0) + gamsc)) * (gamsc / (R(1.0) + gamsc)) * rvs_pp / rvs_p * rvs / rvs_p;
  alphsc = mx(R(1.e-9), alphsc);
  const R xs = mn(R(-1.e-9), L.ssatw);
  const R t1_evap = R(2.0) * R(kPi) *
                    (R(1.0) - alphsc * xs + R(2.0) * alphsc * alphsc * xs * xs - R(5.0) * alphsc * alphsc * alphsc * xs * xs * xs) /
                    (R(1.0) + gamsc);
  if (L.qv / L.qvs < R(0.95) && L.rr * orho <= R(1.e-8)) {
    t.prv_rev = L.rr * orho * odt;
  } else {
    const R lamr = R(1.0) / L.ilamr;
    const R n0_r = L.nr * C.org2 * mpow(lamr, C.cre[1]);
    t.prv_rev = t1_evap * L.diffu * (-L.ssatw) * n0_r * rvs *
                (C.t1_qr_ev * mpow(L.ilamr, C.cre[9]) +
                 C.t2_qr_ev * L.vsc2 * L.rhof2 * (mpow(lamr + R(0.5) * R(kFvR), -C.cre[10])));
    const R rate_max = mn((L.rr * orho * odt), (L.qvs - L.qv) * odt);
    t.prv_rev = mn(rate_max, t.prv_rev * orho);
    if (t.prr_gml > R(0.0)) {
      const R eva_factor = mn(R(1.0), R(0.01) + R(0.99 - 0.01) * (tempc / R(20.0)));
      t.prv_rev = t.prv_rev * eva_factor;
    }
  }
  t.pnr_rev = mn(L.nr * R(0.99) * orho * odt, t.prv_rev * L.nr / L.rr);
}

}  // namespace tempo

// Launchable kernels of the column step (see tempo_step.hpp for the
// phase split).  Included by tempo_step_f32.hip and tempo_step_f64.hip.
#pragma once

#include "tempo_step.hpp"
#include "tempo_radar.hpp"

namespace tempo {

// Waves per SIMD the level-local kernels are compiled for, per working type.
// The FP32 level physics needs 130 VGPRs unconstrained (3 waves per SIMD);
// held to 128 it runs 4 waves with a few bytes more of private memory and
// is measured faster.  Tighter budgets spill heavily and are slower.  The
// FP64 level physics is held to three waves (168 VGPRs, a few hundred
// bytes of spills per lane), measured faster than two waves without
// spills.  The dual instantiations keep the compiler's choice (the FP32
// dual phase A takes 244 VGPRs, two waves; held to three waves it spilled
// and made the adjoint 7-13 % slower; its final update at four waves
// instead of three gained nothing).  One wave fewer (a_alt: 3 in FP32,
// 2 in FP64) is a tuning option of phase A (TEMPO_HIP_TUNE_WAVES_A).
template <class R>
struct WavesPerEu {
  static constexpr int a = 1, a_alt = 1, c = 1;
};
template <>
struct WavesPerEu<float> {
  static constexpr int a = 4, a_alt = 3, c = 1;
};
template <>
struct WavesPerEu<double> {
  static constexpr int a = 3, a_alt = 2, c = 1;
};
}  // namespace tempo

// register-budget attributes of the final update (occupancy against
// spills); a build may override it
#ifndef TEMPO_ATTR_C
#define TEMPO_ATTR_C __attribute__((amdgpu_waves_per_eu(WavesPerEu<R>::c)))
#endif

namespace tempo {

template <class R>
__device__ __forceinline__ R load_t(const StepArgs<R>& a, size_t o) {
  return a.f[F_T] ? a.f[F_T][o] : a.f[F_TH][o] * a.f[F_PII][o];
}

template <class R>
__device__ __forceinline__ R column_nc(const StepArgs<R>& a, int c) {
  // one-moment cloud number when nc is absent (src/module_mp_tempo_main.F90:363-380,
  // src/module_mp_tempo_utils.F90:12-24)
  if (a.land) return a.land[c] != 1 ? R(kNtCO) : R(kNtCL);
  return R(a.fl.nc_const);
}

// ------------------------------------------------------------ pre-pass ----
// Blocks of 64 columns x 4 waves: wave w tests levels w, w + 4, ... of its
// lane's column, so that every load is one coalesced row of 64 columns and
// the chip holds four times as many waves as columns/64; the waves' answers
// are combined through LDS.  The tests are ORs over levels, so their order
// does not matter; the first-step sums and the height prefix keep the
// reference's sequential order in wave 0.
constexpr int kPrepassCols = 64;
template <class R>
__global__ void __launch_bounds__(256) k_prepass(StepArgs<R> a) {
  __shared__ unsigned char flag[4][kPrepassCols];
  const int lane = int(threadIdx.x & 63u), w = int(threadIdx.x >> 6);
  const int c = int(blockIdx.x) * kPrepassCols + lane;
  const bool in = c < a.ncol;
  const int nzc = !in ? 0 : a.nlev ? min(max(a.nlev[c], 1), a.nz) : a.nz;
  // early exit (:416-418): any hydrometeor above r1 after the initial checks
  // (tendencies are still zero, so l_q == q1d > r1), or ice supersaturation
  bool q = false;
  for (int k = w; k < nzc && !q; k += 4) {
    const size_t o = TEMPO_IDX_STATE(a, c, k);
    q = a.f[F_QC][o] > R(kR1) || a.f[F_QR][o] > R(kR1) || a.f[F_QI][o] > R(kR1) || a.f[F_QS][o] > R(kR1) ||
        a.f[F_QG][o] > R(kR1);
  }
  flag[w][lane] = q;
  __syncthreads();
  const bool any_q = flag[0][lane] | flag[1][lane] | flag[2][lane] | flag[3][lane];
  bool sup = false;
  for (int k = w; !any_q && k < nzc && !sup; k += 4) {
    const size_t o = TEMPO_IDX_STATE(a, c, k);
    const R temp = load_t(a, o), pres = a.f[F_P][o], qv = mx(R(kMinQv), a.f[F_QV][o]);
    const R qvs = rslf(pres, temp);
    const R qvsi = (temp - R(kT0) <= R(0.0)) ? rsif(pres, temp) : qvs;
    R ssati = qv / qvsi - R(1.0);
    if (mabs(ssati) < R(kEps)) ssati = R(0.0);
    sup = ssati > R(0.0);
  }
  __syncthreads();  // every wave has read the first answers
  flag[w][lane] = sup;
  __syncthreads();
  if (w != 0 || !in) return;
  unsigned cf = 0;
  if (!any_q && !(flag[0][lane] | flag[1][lane] | flag[2][lane] | flag[3][lane])) cf |= CF_EXIT;
  // first-step graupel initialisation test (:387-393)
  if (a.first_step && a.f[F_NG] && a.f[F_QB]) {
    R sqg = R(0), sng = R(0), sqb = R(0);
    for (int k = 0; k < nzc; ++k) {
      const size_t o = TEMPO_IDX_STATE(a, c, k);
      sqg += a.f[F_QG][o];
      sng += a.f[F_NG][o];
      sqb += a.f[F_QB][o];
    }
    if (sqg > R(kR1) && sng < R(kEps) && sqb < R(kEps)) cf |= CF_GINIT;
  }
  // height prefix of the default aerosol profile (src/module_mp_tempo_aerosols.F90:24-28)
  if (!a.f[F_NWFA] || !a.f[F_NIFA]) {
    R h = R(0);
    for (int k = 0; k < nzc; ++k) {
      if (k > 0) h = h + a.f[F_DZ][TEMPO_IDX_STATE(a, c, k)];
      a.sr[size_t(SR_HGT) * a.plane + TEMPO_IDX_SCR(a, c, k)] = h;
    }
  }
  a.cf[c] = cf;
}
template <class R>
dim3 prepass_grid(const StepArgs<R>& a) {
  return dim3(unsigned((a.ncol + kPrepassCols - 1) / kPrepassCols));
}

// ------------------------------------------------ point grouping ----
// The level-local kernels give one lane per point, and at one level a wave
// of 64 consecutive columns usually mixes points that run the microphysics
// (hydrometeors or supersaturated vapour) with points that only pass the
// checks: the wave then pays the full path at a fraction of its lanes.  A
// block (up to 256 consecutive columns of one level) therefore partitions
// its points, active ones first, each group in column order, and lane i
// takes the i-th point of that order.  The block's loads and stores stay in
// its own window of columns (shared by its waves through the CU's caches);
// every point is still computed on its own, so results do not depend on the
// grouping.

// phase A: hydrometeors, or vapour above liquid or (below 0 C) ice saturation
template <class R>
__device__ bool point_active_a(const StepArgs<R>& a, int c, int k) {
  if ((a.nlev && k >= a.nlev[c]) || (a.cf[c] & CF_EXIT)) return false;
  const size_t o = TEMPO_IDX_STATE(a, c, k);
  if (a.f[F_QC][o] > R(kR1) || a.f[F_QR][o] > R(kR1) || a.f[F_QI][o] > R(kR1) || a.f[F_QS][o] > R(kR1) ||
      a.f[F_QG][o] > R(kR1))
    return true;
  const R temp = load_t(a, o), pres = a.f[F_P][o], qv = a.f[F_QV][o];
  if (qv > rslf(pres, temp)) return true;
  return temp < R(kT0) && qv > rsif(pres, temp);
}

// final update: a hydrometeor present after the tendencies of phases A and B
template <class R>
__device__ bool point_active_c(const StepArgs<R>& a, int c, int k) {
  if ((a.nlev && k >= a.nlev[c]) || (a.cf[c] & CF_EXIT)) return false;
  const size_t o = TEMPO_IDX_STATE(a, c, k), s = TEMPO_IDX_SCR(a, c, k);
  const R dt = R(a.dt);
  auto present = [&](int f, int ten) { return a.fa[f][o] + R(a.sd[size_t(ten) * a.plane + s]) * dt > R(kR1); };
  return present(F_QC, SD_QCTEN) || present(F_QR, SD_QRTEN) || present(F_QI, SD_QITEN) ||
         present(F_QS, SD_QSTEN) || present(F_QG, SD_QGTEN);
}

// The level-local kernels run blocks of 64 columns x blockDim.x/64 levels,
// one level per wavefront (blockDim.x = 64, 128 or 256), on a grid of
// ceil(ncol/64) x ceil(nz/(blockDim.x/64)) blocks.  Point (column, level) of
// this thread; every thread of the block calls it, and the point may lie
// outside the batch (c >= ncol or k >= nz) in the last blocks.
template <class R, bool kFinal>
__device__ __forceinline__ int2 block_point(const StepArgs<R>& a) {
  const int lane = int(threadIdx.x & 63u), w = int(threadIdx.x >> 6);
  const int c0 = int(blockIdx.x) * 64, k0 = int(blockIdx.y) * int(blockDim.x >> 6);
  if (!(a.compact & (kFinal ? 2 : 1))) return make_int2(c0 + lane, k0 + w);
  __shared__ unsigned short slot_pt[256];
  __shared__ unsigned wave_act[4];
  const int c = c0 + lane, k = k0 + w;
  const bool act = c < a.ncol && k < a.nz && (kFinal ? point_active_c(a, c, k) : point_active_a(a, c, k));
  const unsigned long long b = __ballot(act);
  const unsigned below = __builtin_amdgcn_mbcnt_hi(unsigned(b >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(b), 0u));
  const unsigned nw = blockDim.x >> 6;
  if (lane == 0) wave_act[w] = unsigned(__popcll(b));
  __syncthreads();
  unsigned before = below, nact = 0;  // active points before this thread, and in the block
  for (unsigned i = 0; i < nw; ++i) {
    if (int(i) < w) before += wave_act[i];
    nact += wave_act[i];
  }
  const unsigned pos = act ? before : nact + (threadIdx.x - before);
  slot_pt[pos] = static_cast<unsigned short>(threadIdx.x);
  __syncthreads();
  const unsigned t = slot_pt[threadIdx.x];
  return make_int2(c0 + int(t & 63u), k0 + int(t >> 6));
}

// grid of a level-local kernel for blocks of `threads` threads
template <class R>
dim3 level_grid(const StepArgs<R>& a, int threads) {
  const int lv = threads / 64;
  return dim3(unsigned((a.ncol + 63) / 64), unsigned((a.nz + lv - 1) / lv));
}

// ----------------------------------------------------------- phase A ----
template <class R>
__device__ __forceinline__ R default_aerosol(const Scr<R>& scr, int c, int k, R n0, R n1) {
  // init_water/ice_friendly_aerosols (src/module_mp_tempo_aerosols.F90:14-73):
  // hgt(1) = 0, so h_01 = 0.8; level 1 uses hgt(2)
  const R* hgt = scr.sr + size_t(SR_HGT) * scr.plane;
  const R h1 = hgt[scr.at(c, 0)];
  const R hk = hgt[scr.at(c, (k == 0 ? 1 : k))];
  const R ni3 = R(-1.0) * mlog(n1 / n0) / R(0.8);
  return n1 + n0 * mexp(-((hk - h1) / R(1000.0)) * ni3);
}

__device__ __forceinline__ bool same_bits(float x, float y) { return __float_as_uint(x) == __float_as_uint(y); }
__device__ __forceinline__ bool same_bits(double x, double y) {
  return __double_as_longlong(x) == __double_as_longlong(y);
}
template <class T>
__device__ __forceinline__ bool same_bits(const Dual<T>& x, const Dual<T>& y) {
  return same_bits(x.v, y.v) && same_bits(x.d, y.d);
}
template <class T>
__device__ __forceinline__ bool same_bits(const Sig<T>& x, const Sig<T>& y) {
  return same_bits(x.v, y.v);
}
// post-check value of field f at o; where the step runs in place the input
// is still in memory (the read is a cache hit) and an unchanged value is not
// stored again
template <class R>
__device__ __forceinline__ void put_post_check(const StepArgs<R>& a, int f, size_t o, const R& v) {
  if (a.fa[f] != a.f[f] || !same_bits(a.f[f][o], v)) a.fa[f][o] = v;
}

// branch signature of a level-local kernel's point: opened on entry, stored
// at every exit (the plain types record nothing)
template <class R>
struct SigScope {
  __device__ SigScope(const StepArgs<R>&, int, int, int) {}
};
template <class T>
struct SigScope<Sig<T>> {
  unsigned* out = nullptr;
  size_t n3 = 0;
  __device__ SigScope(const StepArgs<Sig<T>>& a, int c, int k, int phase) {
    if (!a.sig) return;
    n3 = size_t(a.ncol) * size_t(a.nz);
    out = a.sig + size_t(2 * phase) * n3 + size_t(c) + size_t(k) * size_t(a.ncol);
    sig_begin();
  }
  __device__ ~SigScope() {
    if (!out) return;
    out[0] = sig_end(0);
    out[n3] = sig_end(1);
  }
};

// the level-local physics of point (c, k), from the initial checks to the
// fall speeds (src/module_mp_tempo_main.F90:320-588); scratch in scr
template <class R>
__device__ __forceinline__ void phase_a_point(const StepArgs<R>& a, const Scr<R>& scr, int c, int k) {
  const Const<S_t<R>>& C = *a.cst;
  const size_t o = TEMPO_IDX_STATE(a, c, k);
  const size_t s = scr.at(c, k);
  const size_t plane = scr.plane;
  const R dt = R(a.dt), odt = R(a.odt);
  const unsigned cf = a.cf[c];
  const bool has_nc = a.f[F_NC], has_nwfa = a.f[F_NWFA], has_nifa = a.f[F_NIFA];
  const bool has_ngqb = a.f[F_NG] && a.f[F_QB];

  Lvl<R> L;
  const R t1d = load_t(a, o), qv1d = a.f[F_QV][o], w = a.f[F_W][o];
  R qc1d = a.f[F_QC][o], qr1d = a.f[F_QR][o], qi1d = a.f[F_QI][o], qs1d = a.f[F_QS][o], qg1d = a.f[F_QG][o];
  R ni1d = a.f[F_NI][o], nr1d = a.f[F_NR][o];
  R nc1d = has_nc ? a.f[F_NC][o] : R(0), ng1d = has_ngqb ? a.f[F_NG][o] : R(0), qb1d = has_ngqb ? a.f[F_QB][o] : R(0);
  const R nwfa1d = has_nwfa ? a.f[F_NWFA][o] : R(0), nifa1d = has_nifa ? a.f[F_NIFA][o] : R(0);

  L.tten = L.qvten = L.qcten = L.qiten = L.qrten = L.qsten = L.qgten = L.qbten = R(0);
  L.niten = L.nrten = L.ncten = L.ngten = L.nwfaten = L.nifaten = R(0);
  L.smob = L.smoc = L.smo0 = L.smo1 = L.smo2 = L.smoe = L.smof = L.smoz = R(0);
  L.temp = t1d;
  L.qv = mx(R(kMinQv), qv1d);
  L.pres = a.f[F_P][o];
  L.rho = rho_of(C, L.pres, L.temp, L.qv);
  // aerosols: a present field overrides the default profile (:328-345)
  L.nwfa = has_nwfa ? nwfa1d * L.rho : default_aerosol(scr, c, k, R(kNaccn0), R(kNaccn1));
  L.nifa = has_nifa ? nifa1d * L.rho : default_aerosol(scr, c, k, R(kNain0), R(kNain1));
  L.nwfa = mx(R(kNwfaDefault) * L.rho, mn(R(kAeroMax) * L.rho, L.nwfa));
  L.nifa = mx(R(kNifaDefault) * L.rho, mn(R(kAeroMax) * L.rho, L.nifa));

  // initial checks (:347-396)
  rain_check(C, L, qr1d, nr1d, dt, odt);
  ice_check(C, L, qi1d, ni1d, dt, odt);
  snow_check(L, qs1d, dt, odt);
  if (L.l_qs) snow_moments(C, L.rs, mn(R(-0.1), L.temp - R(kT0)), 1 | 2 | 4 | 8 | 16, L);
  // one-moment cloud number (:363-380): the network's prediction from this
  // point of the step (the final update reads it back), else the column's
  // constant
  R ncv;
  if (a.fl.ml_for_nc && !has_nc) {
    const R xrx = qc1d <= R(S_t<R>(1.e-12)) ? R(0.0) : qc1d;
    ncv = ml_cloud_number(xrx, qr1d, qi1d, qs1d, L.pres, L.temp, w) * L.rho;
    scr.sr[SR_NCSAVE * plane + s] = ncv;
  } else {
    ncv = column_nc(a, c);
  }
  L.ncsave = ncv;
  cloud_check(C, L, qc1d, nc1d, has_nc, ncv, dt, odt);
  if ((cf & CF_GINIT) && qg1d > R(kR1)) {  // graupel_init (:1234-1260)
    const R rg = qg1d * L.rho;
    const R rb = rg * R(kM3ToL) / rho_g<R>(kIdxBg1);
    const int idx = max(1, min(nint_exact(R(10.0) * rg / rb) + 1, kNrhg));
    const R lamg = (R(3.0) + R(kMuG) + R(0.672)) / R(5.e-3);
    const R ng = C.cgg[1][0] * C.ogg3 * rg * mpow(lamg, R(kBmG)) / C.am_g[idx - 1];
    ng1d = ng / L.rho;
    qb1d = rb / L.rho;
  }
  graupel_check(C, L, qg1d, ng1d, qb1d, has_ngqb, dt, odt);
  // the post-check state is the 1-d state for the rest of the step; a step
  // in place writes back only the values the checks changed
  put_post_check(a, F_QC, o, qc1d);
  put_post_check(a, F_QR, o, qr1d);
  put_post_check(a, F_NR, o, nr1d);
  put_post_check(a, F_QI, o, qi1d);
  put_post_check(a, F_NI, o, ni1d);
  put_post_check(a, F_QS, o, qs1d);
  put_post_check(a, F_QG, o, qg1d);
  if (has_nc) put_post_check(a, F_NC, o, nc1d);
  if (has_ngqb) {
    put_post_check(a, F_NG, o, ng1d);
    put_post_check(a, F_QB, o, qb1d);
  }
  if (cf & CF_EXIT) return;
  L.qcten = L.qiten = L.qrten = L.qsten = L.qgten = L.ngten = L.qbten = L.niten = L.nrten = L.ncten = R(0);
  thermo(C, L);

  const bool micro = !a.fl.turn_off_micro;
  Tend<R> t{};  // every process rate starts at zero
  R vtboost = R(1.0);
  if (micro) {
    warm_rain(C, a.tab, L, t, odt);
    rain_snow_graupel(C, a.tab, L, t, odt);
    ice_nucleation(C, a.tab, L, t, has_nwfa, has_nifa, w, dt, odt);
    const R t1s = t1_subl(C, L);
    ice_processes(C, a.tab, L, t, t1s, w, odt);
    vtboost = riming(C, a.tab, L, t, odt);
    melting(C, L, t, t1s, dt, odt);
    if (has_nwfa || has_nifa) aerosol_scavenging(C, L, t, odt);
  }
  check_over_depletion(L, t, odt);
  sum_tendencies(C, L, t);

  // provisional update and re-checks on copies (:459-526)
  L.temp = t1d + L.tten * dt;
  L.qv = mx(R(kMinQv), qv1d + L.qvten * dt);
  L.rho = rho_of(C, L.pres, L.temp, L.qv);
  L.nwfaten = L.nwfaten - (t.pna_rca + t.pna_sca + t.pna_gca + t.pni_iha) / L.rho;
  L.nifaten = L.nifaten - (t.pnd_rcd + t.pnd_scd + t.pnd_gcd) / L.rho;
  if (has_nwfa) L.nwfa = (nwfa1d + L.nwfaten * dt) * L.rho;
  if (has_nifa) L.nifa = (nifa1d + L.nifaten * dt) * L.rho;
  L.nwfa = mx(R(kNwfaDefault) * L.rho, mn(R(kAeroMax) * L.rho, L.nwfa));
  L.nifa = mx(R(kNifaDefault) * L.rho, mn(R(kAeroMax) * L.rho, L.nifa));
  {
    R x = qc1d, xn = nc1d;
    cloud_check(C, L, x, xn, has_nc, ncv, dt, odt);
    x = qr1d;
    xn = nr1d;
    rain_check(C, L, x, xn, dt, odt);
    x = qi1d;
    xn = ni1d;
    ice_check(C, L, x, xn, dt, odt);
    x = qs1d;
    snow_check(L, x, dt, odt);
    L.smob = L.smoc = L.smo2 = R(0);
    if (L.l_qs) snow_moments(C, L.rs, mn(R(-0.1), L.temp - R(kT0)), 4, L);
    x = qg1d;
    R xg = ng1d, xb = qb1d;
    graupel_check(C, L, x, xg, xb, has_ngqb, dt, odt);
  }
  thermo(C, L);
  if (micro && a.fl.cloud_condensation) {  // (:528-560)
    cloud_condensation(C, a.tab, L, t, has_nwfa, has_nc, w, dt, odt);
    L.qvten = L.qvten - t.prw_vcd;
    L.qcten = L.qcten + t.prw_vcd;
    L.ncten = L.ncten + t.pnc_wcd;
    L.nwfaten = L.nwfaten - t.pnc_wcd;
    L.tten = L.tten + L.lvap * L.ocp * t.prw_vcd;
    R x = qc1d, xn = nc1d;
    cloud_check(C, L, x, xn, has_nc, ncv, dt, odt);
    L.qv = mx(R(kMinQv), qv1d + L.qvten * dt);
    L.temp = t1d + L.tten * dt;
    L.rho = rho_of(C, L.pres, L.temp, L.qv);
    thermo(C, L);
  }
  if (micro) {  // (:562-588)
    rain_evaporation(C, L, t, odt);
    L.qrten = L.qrten - t.prv_rev;
    L.qvten = L.qvten + t.prv_rev;
    L.nrten = L.nrten - t.pnr_rev;
    L.nwfaten = L.nwfaten + t.pnr_rev;
    L.tten = L.tten - L.lvap * L.ocp * t.prv_rev;
    R x = qr1d, xn = nr1d;
    rain_check(C, L, x, xn, dt, odt);
    L.qv = mx(R(kMinQv), qv1d + L.qvten * dt);
    L.temp = t1d + L.tten * dt;
    L.rho = rho_of(C, L.pres, L.temp, L.qv);
    thermo(C, L);
  }

  // level-local fall speeds (the top-down fill is done in phase B)
  R vtrr = R(0), vtnr = R(0), vtrg = R(0), vtng = R(0), vts = R(0), vtri = R(0), vtni = R(0), vtrc = R(0), vtnc = R(0);
  unsigned pf = (L.l_qr ? PF_QR : 0u) | (L.l_qg ? PF_QG : 0u) | (L.l_qs ? PF_QS : 0u) | (L.l_qi ? PF_QI : 0u) |
                (L.l_qc ? PF_QC : 0u);
  if (L.rr > R(kR1)) {  // rain_fallspeed (:1771-1774)
    const R lamr = R(1.0) / L.ilamr;
    vtrr = L.rhof * R(kAvR) * C.crg[5] * C.org3 * mpow(lamr, C.cre[2]) * mpow(lamr + R(kFvR), -C.cre[5]);
    vtnr = L.rhof * R(kAvR) * C.crg[6] / C.crg[11] * mpow(lamr, C.cre[11]) * mpow(lamr + R(kFvR), -C.cre[6]);
  }
  if (L.rg > R(kR1)) {  // graupel_fallspeed (:1815-1832)
    R afall, bfall;
    if (a.f[F_QB]) {
      const R dens_g = mx(rho_g<R>(1), mn(R(kM3ToL) * L.rg / L.rb, rho_g<R>(kNrhg)));
      afall = R(kACoeff) * mpow((R(4.0) * dens_g * R(kGravity)) / (R(3.0) * L.rho), R(kBCoeff));
      afall = afall * mpow(L.visco, R(1.0) - R(2.0) * R(kBCoeff));
      bfall = R(3.0) * R(kBCoeff) - R(1.0);
    } else {
      afall = R(kAvGOld);
      bfall = R(kBvGOld);
    }
    const int idx = L.idx_bg;
    vtrg = L.rhof * afall * C.cgg[5][idx - 1] * C.ogg3 * mpow(L.ilamg, bfall);
    vtng = L.rhof * afall * C.cgg[6][idx - 1] / C.cgg[11][idx - 1] * mpow(L.ilamg, bfall);
  }
  if (L.rs > R(kR1)) {  // snow_fallspeed (:1873-1891)
    const R xds = L.smoc / L.smob;
    const R mrat = R(1.0) / xds;
    R ils1 = R(1.0) / (mrat * R(kLam0) + R(kFvS)), ils2 = R(1.0) / (mrat * R(kLam1) + R(kFvS));
    const R t1v = R(kKap0) * C.csg[3] * mpow(ils1, C.cse[3]);
    const R mm = mpow(mrat, R(kMuS));
    const R t2v = R(kKap1) * mm * C.csg[9] * mpow(ils2, C.cse[9]);
    ils1 = R(1.0) / (mrat * R(kLam0));
    ils2 = R(1.0) / (mrat * R(kLam1));
    const R t3v = R(kKap0) * C.csg[0] * mpow(ils1, C.cse[0]);
    const R t4v = R(kKap1) * mm * C.csg[6] * mpow(ils2, C.cse[6]);
    vts = L.rhof * R(kAvS) * (t1v + t2v) / (t3v + t4v);
    if (t.prr_sml > R(0.0))
      pf |= PF_SNOWMIX;
    else
      vts = vts * vtboost;
  }
  if (L.ri > R(kR1)) {  // ice_fallspeed (:1928-1930)
    vtri = L.rhof * R(kAvI) * C.cig[2] * C.oig2 * mpow(L.ilami, R(kBvI));
    vtni = L.rhof * R(kAvI) * C.cig[5] / C.cig[6] * mpow(L.ilami, R(kBvI));
  }
  if (L.rc > R(kR1) && w < R(0.1)) {  // cloud_fallspeed (:1971-1974)
    const int nu = get_nuc(L.nc);
    vtrc = L.rhof * R(kAvC) * C.ccg[4][nu] * C.ocg2[nu] * mpow(L.ilamc, R(kBvC));
    vtnc = L.rhof * R(kAvC) * C.ccg[3][nu] * C.ocg1[nu] * mpow(L.ilamc, R(kBvC));
  }

  R* sr = scr.sr;
  sr[SR_TEMP * plane + s] = L.temp;
  sr[SR_QV * plane + s] = L.qv;
  sr[SR_RHO * plane + s] = L.rho;
  sr[SR_TTEN * plane + s] = L.tten;
  sr[SR_QVTEN * plane + s] = L.qvten;
  sr[SR_NWFATEN * plane + s] = L.nwfaten;
  sr[SR_NIFATEN * plane + s] = L.nifaten;
  sr[SR_VTRR * plane + s] = vtrr;
  sr[SR_VTNR * plane + s] = vtnr;
  sr[SR_VTRG * plane + s] = vtrg;
  sr[SR_VTNG * plane + s] = vtng;
  sr[SR_VTS * plane + s] = vts;
  sr[SR_VTRI * plane + s] = vtri;
  sr[SR_VTNI * plane + s] = vtni;
  sr[SR_VTRC * plane + s] = vtrc;
  sr[SR_VTNC * plane + s] = vtnc;
  if (a.fl.semi_sedi) {  // the fall-speed refresh between semi-Lagrangian substeps reads these
    sr[SR_RHOF * plane + s] = L.rhof;
    sr[SR_VISCO * plane + s] = L.visco;
  }
  using D = Sd_t<R>;
  D* sd = scr.sd;
  sd[SD_RR * plane + s] = D(L.rr);
  sd[SD_NR * plane + s] = D(L.nr);
  sd[SD_RG * plane + s] = D(L.rg);
  sd[SD_NG * plane + s] = D(L.ng);
  sd[SD_RB * plane + s] = D(L.rb);
  sd[SD_RS * plane + s] = D(L.rs);
  sd[SD_RI * plane + s] = D(L.ri);
  sd[SD_NI * plane + s] = D(L.ni);
  sd[SD_RC * plane + s] = D(L.rc);
  sd[SD_NC * plane + s] = D(L.nc);
  sd[SD_QRTEN * plane + s] = D(L.qrten);
  sd[SD_NRTEN * plane + s] = D(L.nrten);
  sd[SD_QGTEN * plane + s] = D(L.qgten);
  sd[SD_NGTEN * plane + s] = D(L.ngten);
  sd[SD_QBTEN * plane + s] = D(L.qbten);
  sd[SD_QSTEN * plane + s] = D(L.qsten);
  sd[SD_QITEN * plane + s] = D(L.qiten);
  sd[SD_NITEN * plane + s] = D(L.niten);
  sd[SD_QCTEN * plane + s] = D(L.qcten);
  sd[SD_NCTEN * plane + s] = D(L.ncten);
  scr.pf[s] = pf;
}

// kWaves: waves per SIMD the kernel is compiled for.  With a.ccn_lds the
// block first copies the CCN activation table's slice (kCcnSlice floats)
// into LDS and the activation reads it there.
template <class R, int kWaves = WavesPerEu<R>::a>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWaves))) k_phase_a(StepArgs<R> a) {
  __shared__ float ccn_s[kCcnSlice];
  if (a.ccn_lds) {
    for (int i = int(threadIdx.x); i < kCcnSlice; i += int(blockDim.x)) ccn_s[i] = a.tab.ccn[i];
    __syncthreads();
    a.tab.ccn = ccn_s;
  }
  const int2 pt = block_point<R, false>(a);
  const int c = pt.x, k = pt.y;
  if (c >= a.ncol || k >= a.nz) return;
  if (a.nlev && k >= a.nlev[c]) return;
  if (IsDual<R>::value && a.pt_on && !a.pt_on[size_t(c) + size_t(k) * a.ncol]) return;  // an adjoint's local pass
  SigScope<R> scope(a, c, k, 0);
  phase_a_point(a, batch_scr(a), c, k);
}

// ----------------------------------------------------------- phase B ----
// flux-form sedimentation of one field over one substep (:1516-1559), FP64.
// xr, vt and xten are distinct planes (and distinct from the density and the
// layer depths), so the loads of several levels may be issued ahead of the
// stores of the level being updated.
template <class R>
__device__ void sediment(const StepArgs<R>& a, int c, int nzc, W_t<R>* __restrict__ xr,
                         const W_t<R>* __restrict__ vt, W_t<R>* __restrict__ xten, double limit, int steps, int ktop,
                         W_t<R>* precip) {
  // every operation rounded on its own, in every sedimentation kernel, so
  // that the kernels agree bit for bit also where the working precision's
  // build lets the compiler fuse multiply-adds
#pragma clang fp contract(off)
  using W = W_t<R>;
  const size_t plane = a.plane;
  const R* __restrict__ rho = a.sr + SR_RHO * plane;
  const R* __restrict__ dzf = a.f[F_DZ];
  const double ost = 1.0 / double(steps), dt = a.dt;
  W above = W(0.0), s0 = W(0.0);
#pragma unroll 4
  for (int k = nzc - 1; k >= 0; --k) {
    const size_t s = TEMPO_IDX_SCR(a, c, k);
    const W sed = vt[s] * xr[s];
    if (k == nzc - 1 || k <= ktop - 1) {
      const W odz = 1.0 / W(dzf[TEMPO_IDX_STATE(a, c, k)]);
      const W orho = 1.0 / W(rho[s]);
      const W flux = (k == nzc - 1) ? -sed : (above - sed);
      xten[s] = xten[s] + flux * odz * ost * orho;
      const W v = xr[s] + flux * odz * dt * ost;
      xr[s] = v > limit ? v : W(limit);
    }
    above = sed;
    if (k == 0) s0 = sed;
  }
  if (precip && xr[TEMPO_IDX_SCR(a, c, 0)] > kLowLimitPrecip) *precip = *precip + s0 * dt * ost;
}

// top-down fill of a fall-speed profile with CFL substeps and ktop
// (:1768-1789 and the species variants); vt raw values are valid where the
// mass exceeds r1.  Writes the filled profile to out (FP64).
template <class R>
__device__ void fill_fallspeed(const StepArgs<R>& a, int c, int nzc, const W_t<R>* __restrict__ mass,
                               const R* __restrict__ raw1, const R* __restrict__ raw2, W_t<R>* __restrict__ out1,
                               W_t<R>* __restrict__ out2, bool use_max, int& substeps, int& ktop) {
#pragma clang fp contract(off)
  using W = W_t<R>;
  const double dt = a.dt;
  W v1 = W(0.0), v2 = W(0.0);
  substeps = 1;
  ktop = 1;
#pragma unroll 4
  for (int k = nzc - 1; k >= 0; --k) {
    const size_t s = TEMPO_IDX_SCR(a, c, k);
    if (mass[s] > kR1) {
      v1 = W(raw1[s]);
      if (raw2) v2 = W(raw2[s]);
    }
    out1[s] = v1;
    if (out2) out2[s] = v2;
    // the CFL test runs in the working precision, as the reference's vt is wp
    const R vr1 = R(v1), vr2 = R(v2);
    const R vmax = use_max ? mx(vr1, vr2) : vr1;
    if (vmax > R(1.e-3)) {
      ktop = max(ktop, k + 1);
      const R dz_by_vt = a.f[F_DZ][TEMPO_IDX_STATE(a, c, k)] / vmax;
      substeps = max(substeps, int(R(dt) / dz_by_vt + R(1.0)));
    }
  }
  if (ktop == nzc) ktop = nzc - 1;
}

// Decisions of the semi-Lagrangian sedimentation of one column: a 32-bit
// hash per column of its structural decisions (h), and in a signature run
// per level, in the sedimentation's planes of the signature array (w0:
// comparisons, w1: discrete outcomes, point (c, k) at k*ncol + c).
struct SemiRec {
  unsigned h = 0;
  unsigned *w0 = nullptr, *w1 = nullptr;
  size_t ncol = 0;
  int c = 0;
  __device__ void cmp(int k, unsigned v) {
    if (w0) w0[size_t(k) * ncol + size_t(c)] = decision(w0[size_t(k) * ncol + size_t(c)], v);
  }
  __device__ void outcome(int k, unsigned v, bool column = true) {
    if (column) h = decision(h, v);
    if (w1) w1[size_t(k) * ncol + size_t(c)] = decision(w1[size_t(k) * ncol + size_t(c)], v);
  }
};

// semi-Lagrangian sedimentation of one field over one substep
// (semilagrangian_sedimentation, :1562-1749, after Juang and Hong 2010) for
// the column c of nzc levels: interface speeds by third-order interpolation
// of vt (vt above the top is 0), limited so that the departure cells do not
// shrink by more than con1 per substep; departure points za and their
// spacing; monotone piecewise-parabolic interface values; the remap of the
// departure cells onto the model levels; the surface flux from the cells
// that leave the column.  Work arrays in semiw ([array][level][column]).
constexpr int kSemiArrays = 8;
template <class R>
__device__ void semi_sediment(const StepArgs<R>& a, int c, int nzc, W_t<R>* xr, W_t<R>* xten, const W_t<R>* vt,
                              int steps, double limit, W_t<R>* precip, SemiRec& rec) {
  // rec: the decisions, per column (structural: each level's source cells
  // kb and kt, where the remap ends early, which cells feed the surface
  // flux) and, in a signature run, per level (also the interface-speed
  // replacements and limiters, the monotone limiter and the clamps)
  // The remap and the surface flux branch on differences of nearly equal
  // positions (departure point against interface, za against 0), where the
  // reference's algorithm is discontinuous: its rounding is kept exactly,
  // without fused multiply-adds, in either precision of the step.
#pragma clang fp contract(off)
  using W = W_t<R>;
  const size_t plane = a.plane, nk = size_t(a.nz) + 2, nc = size_t(a.ncol);
  auto arr = [&](int q, int k) -> W& { return a.semiw[(size_t(q) * nk + size_t(k)) * nc + size_t(c)]; };
  enum { ZI, WI, ZA, DZA, QA, QMI, QPI, SAVE };
  auto X = [&](int k) -> W& { return xr[TEMPO_IDX_SCR(a, c, k)]; };
  auto V = [&](int k) -> W { return k < nzc ? vt[TEMPO_IDX_SCR(a, c, k)] : W(0.0); };
  auto DZ = [&](int k) -> W { return W(a.f[F_DZ][TEMPO_IDX_STATE(a, c, k)]); };
  const int nz = nzc;
  const W dt = W(a.dt), odt = W(1.0) / dt, ost = W(1.0) / W(double(steps));
  for (int k = 0; k < nz; ++k) arr(SAVE, k) = X(k);
  arr(ZI, 0) = W(0.0);
  for (int k = 0; k < nz; ++k) arr(ZI, k + 1) = arr(ZI, k) + DZ(k);
  // interface speeds (:1599-1612)
  const W fa1 = W(9.0 / 16.0), fa2 = W(1.0 / 16.0);
  arr(WI, 0) = V(0);
  arr(WI, 1) = W(0.5) * (V(1) + V(0));
  for (int k = 2; k < nz - 1; ++k) arr(WI, k) = fa1 * (V(k) + V(k - 1)) - fa2 * (V(k + 1) + V(k - 2));
  arr(WI, nz - 1) = W(0.5) * (V(nz - 1) + V(nz - 2));
  arr(WI, nz) = V(nz);
  for (int k = 1; k < nz; ++k)
    if (V(k) == W(0.0)) {
      arr(WI, k) = V(k - 1);
      rec.cmp(k, 1u);
    }
  // diffusivity limit, top-down (:1614-1621)
  const W con1 = W(0.05);
  for (int k = nz - 1; k >= 0; --k) {
    const W decfl = (arr(WI, k + 1) - arr(WI, k)) * dt * ost / DZ(k);
    if (decfl > con1) {
      arr(WI, k) = arr(WI, k + 1) - con1 * DZ(k) * odt * W(double(steps));
      rec.cmp(k, 2u);
    }
  }
  // departure points and cell sizes (:1623-1635)
  for (int k = 0; k < nz + 1; ++k) arr(ZA, k) = arr(ZI, k) - arr(WI, k) * dt * ost;
  arr(ZA, nz + 1) = arr(ZI, nz);
  for (int k = 0; k < nz + 1; ++k) arr(DZA, k) = arr(ZA, k + 1) - arr(ZA, k);
  for (int k = 0; k < nz; ++k) arr(QA, k) = X(k) * DZ(k) / arr(DZA, k);
  arr(QA, nz) = W(0.0);
  // monotone interface values (:1638-1658)
  for (int k = 1; k < nz; ++k) {
    const W dip = (arr(QA, k + 1) - arr(QA, k)) / (arr(DZA, k + 1) + arr(DZA, k));
    const W dim = (arr(QA, k) - arr(QA, k - 1)) / (arr(DZA, k - 1) + arr(DZA, k));
    if (dip * dim <= W(0.0)) {
      arr(QMI, k) = arr(QA, k);
      arr(QPI, k) = arr(QA, k);
      rec.cmp(k, 3u);
    } else {
      arr(QPI, k) = arr(QA, k) + W(0.5) * (dip + dim) * arr(DZA, k);
      arr(QMI, k) = W(2.0) * arr(QA, k) - arr(QPI, k);
      if (arr(QPI, k) > W(0.0) || arr(QMI, k) < W(0.0)) {
        arr(QPI, k) = arr(QA, k);
        arr(QMI, k) = arr(QA, k);
        rec.cmp(k, 4u);
      }
    }
  }
  arr(QPI, 0) = arr(QA, 0);
  arr(QMI, 0) = arr(QA, 0);
  arr(QMI, nz) = arr(QA, nz);
  arr(QPI, nz) = arr(QA, nz);
  // remap onto the model levels (:1660-1727); kb, kt 1-based as in the reference
  int kb = 1, kt = 1;
  const R* rho = a.sr + SR_RHO * plane;
  for (int k = 1; k <= nz; ++k) {
    kb = max(kb - 1, 1);
    kt = max(kt - 1, 1);
    if (arr(ZI, k - 1) >= arr(ZA, nz)) {
      rec.outcome(k - 1, 0x100u + unsigned(k));
      break;
    }
    for (int kk = kb; kk <= nz; ++kk)
      if (arr(ZI, k - 1) <= arr(ZA, kk)) {
        kb = kk;
        break;
      }
    for (int kk = kt; kk <= nz + 2; ++kk)
      if (arr(ZI, k) <= arr(ZA, kk - 1)) {
        kt = kk;
        break;
      }
    kt = kt - 1;
    rec.outcome(k - 1, (unsigned(kb) << 16) ^ unsigned(kt));
    W& x = X(k - 1);
    if (kt == kb) {
      const W tl = (arr(ZI, k - 1) - arr(ZA, kb - 1)) / arr(DZA, kb - 1);
      const W th = (arr(ZI, k) - arr(ZA, kb - 1)) / arr(DZA, kb - 1);
      const W tl2 = tl * tl, th2 = th * th;
      const W qqd = W(0.5) * (arr(QPI, kb - 1) - arr(QMI, kb - 1));
      const W qqh = qqd * th2 + arr(QMI, kb - 1) * th;
      const W qql = qqd * tl2 + arr(QMI, kb - 1) * tl;
      x = (qqh - qql) / (th - tl);
    } else if (kt > kb) {
      const W tl = (arr(ZI, k - 1) - arr(ZA, kb - 1)) / arr(DZA, kb - 1);
      const W tl2 = tl * tl;
      W qqd = W(0.5) * (arr(QPI, kb - 1) - arr(QMI, kb - 1));
      const W qql = qqd * tl2 + arr(QMI, kb - 1) * tl;
      const W dql = arr(QA, kb - 1) - qql;
      W zsum = (W(1.0) - tl) * arr(DZA, kb - 1);
      W qsum = dql * arr(DZA, kb - 1);
      if (kt - kb > 1)
        for (int m = kb + 1; m <= kt - 1; ++m) {
          zsum = zsum + arr(DZA, m - 1);
          qsum = qsum + arr(QA, m - 1) * arr(DZA, m - 1);
        }
      const W th = (arr(ZI, k) - arr(ZA, kt - 1)) / arr(DZA, kt - 1);
      const W th2 = th * th;
      qqd = W(0.5) * (arr(QPI, kt - 1) - arr(QMI, kt - 1));
      const W dqh = qqd * th2 + arr(QMI, kt - 1) * th;
      zsum = zsum + th * arr(DZA, kt - 1);
      qsum = qsum + dqh * arr(DZA, kt - 1);
      x = qsum / zsum;
    }
    const W orho = W(1.0) / W(rho[TEMPO_IDX_SCR(a, c, k - 1)]);
    if (!(x > W(limit))) rec.cmp(k - 1, 5u);
    x = x > W(limit) ? x : W(limit);
    W& ten = xten[TEMPO_IDX_SCR(a, c, k - 1)];
    ten = ten + (x - arr(SAVE, k - 1)) * orho * odt;
  }
  // surface precipitation (:1729-1744)
  if (!precip) return;
  for (int k = 0; k < nz; ++k) {
    if (arr(ZA, k) < W(0.0) && arr(ZA, k + 1) <= W(0.0)) {
      *precip = *precip + arr(QA, k) * arr(DZA, k);
      rec.outcome(k, 0x200u + unsigned(k));
    } else if (arr(ZA, k) < W(0.0) && arr(ZA, k + 1) > W(0.0)) {
      rec.outcome(k, 0x300u + unsigned(k));
      const W th = (W(0.0) - arr(ZA, k)) / arr(DZA, k);
      const W th2 = th * th;
      const W qqd = W(0.5) * (arr(QPI, k) - arr(QMI, k));
      const W qqh = qqd * th2 + arr(QMI, k) * th;
      *precip = *precip + qqh * arr(DZA, k);
      break;
    }
  }
}

// between semi-Lagrangian substeps (:606-614, 651-666): fall speeds reset,
// rain_check_and_update / graupel_check_and_update on copies of the 1-d
// state with the tendencies so far, then rain_fallspeed / graupel_fallspeed
// (top-down fill from 0 above the column).  A tendency keeps its FP64
// accumulator unless the check rewrites it.
template <class R>
__device__ void semi_rain_refresh(const StepArgs<R>& a, int c, int nzc, W_t<R>* sd, W_t<R>* vt1, W_t<R>* vt2,
                                  SemiRec& rec) {
  using W = W_t<R>;
  const Const<S_t<R>>& C = *a.cst;
  const size_t plane = a.plane;
  const R* sr = a.sr;
  const R dt = R(a.dt), odt = R(a.odt);
  R v1 = R(0.0), v2 = R(0.0);
  for (int k = nzc - 1; k >= 0; --k) {
    const size_t s = TEMPO_IDX_SCR(a, c, k), o = TEMPO_IDX_STATE(a, c, k);
    Lvl<R> L;
    L.rho = sr[SR_RHO * plane + s];
    W& qrten = sd[SD_QRTEN * plane + s];
    W& nrten = sd[SD_NRTEN * plane + s];
    L.qrten = R(qrten);
    L.nrten = R(nrten);
    const R q0 = L.qrten, n0 = L.nrten;
    R x = a.fa[F_QR][o], xn = a.fa[F_NR][o];
    rain_check(C, L, x, xn, dt, odt);
    rec.outcome(k, L.l_qr ? 1u : 0u, false);
    if (L.qrten != q0) qrten = W(L.qrten);
    if (L.nrten != n0) nrten = W(L.nrten);
    sd[SD_RR * plane + s] = W(L.rr);
    sd[SD_NR * plane + s] = W(L.nr);
    if (L.rr > R(kR1)) {  // rain_fallspeed (:1771-1774)
      const R rhof = sr[SR_RHOF * plane + s], lamr = R(1.0) / L.ilamr;
      v1 = rhof * R(kAvR) * C.crg[5] * C.org3 * mpow(lamr, C.cre[2]) * mpow(lamr + R(kFvR), -C.cre[5]);
      v2 = rhof * R(kAvR) * C.crg[6] / C.crg[11] * mpow(lamr, C.cre[11]) * mpow(lamr + R(kFvR), -C.cre[6]);
    }
    vt1[s] = W(v1);
    vt2[s] = W(v2);
  }
}

template <class R>
__device__ void semi_graupel_refresh(const StepArgs<R>& a, int c, int nzc, W_t<R>* sd, W_t<R>* vt1, W_t<R>* vt2,
                                     SemiRec& rec) {
  using W = W_t<R>;
  const Const<S_t<R>>& C = *a.cst;
  const size_t plane = a.plane;
  const R* sr = a.sr;
  const R dt = R(a.dt), odt = R(a.odt);
  const bool has_ngqb = a.f[F_NG] && a.f[F_QB];
  R v1 = R(0.0), v2 = R(0.0);
  for (int k = nzc - 1; k >= 0; --k) {
    const size_t s = TEMPO_IDX_SCR(a, c, k), o = TEMPO_IDX_STATE(a, c, k);
    Lvl<R> L;
    L.rho = sr[SR_RHO * plane + s];
    W& qgten = sd[SD_QGTEN * plane + s];
    W& ngten = sd[SD_NGTEN * plane + s];
    W& qbten = sd[SD_QBTEN * plane + s];
    L.qgten = R(qgten);
    L.ngten = R(ngten);
    L.qbten = R(qbten);
    const R g0 = L.qgten, n0 = L.ngten, b0 = L.qbten;
    R x = a.fa[F_QG][o];
    R xg = has_ngqb ? a.fa[F_NG][o] : R(0.0), xb = has_ngqb ? a.fa[F_QB][o] : R(0.0);
    graupel_check(C, L, x, xg, xb, has_ngqb, dt, odt);
    rec.outcome(k, (L.l_qg ? 1u : 0u) | (unsigned(L.idx_bg) << 1), false);
    if (L.qgten != g0) qgten = W(L.qgten);
    if (L.ngten != n0) ngten = W(L.ngten);
    if (L.qbten != b0) qbten = W(L.qbten);
    sd[SD_RG * plane + s] = W(L.rg);
    sd[SD_NG * plane + s] = W(L.ng);
    sd[SD_RB * plane + s] = W(L.rb);
    if (L.rg > R(kR1)) {  // graupel_fallspeed (:1815-1832)
      const R rhof = sr[SR_RHOF * plane + s];
      R afall, bfall;
      if (a.f[F_QB]) {
        const R dens_g = mx(rho_g<R>(1), mn(R(kM3ToL) * L.rg / L.rb, rho_g<R>(kNrhg)));
        afall = R(kACoeff) * mpow((R(4.0) * dens_g * R(kGravity)) / (R(3.0) * L.rho), R(kBCoeff));
        afall = afall * mpow(sr[SR_VISCO * plane + s], R(1.0) - R(2.0) * R(kBCoeff));
        bfall = R(3.0) * R(kBCoeff) - R(1.0);
      } else {
        afall = R(kAvGOld);
        bfall = R(kBvGOld);
      }
      const int idx = L.idx_bg;
      v1 = rhof * afall * C.cgg[5][idx - 1] * C.ogg3 * mpow(L.ilamg, bfall);
      v2 = rhof * afall * C.cgg[6][idx - 1] / C.cgg[11][idx - 1] * mpow(L.ilamg, bfall);
    }
    vt1[s] = W(v1);
    vt2[s] = W(v2);
  }
}

template <class R>
__global__ void __launch_bounds__(128) k_phase_b(StepArgs<R> a, W_t<R>* vtbuf) {
  using W = W_t<R>;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncol) return;
  const int nzc = a.nlev ? min(max(a.nlev[c], 1), a.nz) : a.nz;
  const size_t plane = a.plane;
  W pr[5] = {W(0.0), W(0.0), W(0.0), W(0.0), W(0.0)};  // rain graupel snow ice cloud
  const unsigned cf = a.cf[c];
  int kst[kSubRows] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0};
  // decisions of the semi-Lagrangian rain and graupel sedimentation (per
  // level in the signature array's sedimentation planes, cleared here, in a
  // signature run)
  SemiRec hr, hg;
  if (a.sig) {
    const size_t n3 = size_t(a.ncol) * size_t(a.nz);
    hr.w0 = hg.w0 = a.sig + 4 * n3;
    hr.w1 = hg.w1 = a.sig + 5 * n3;
    hr.ncol = hg.ncol = size_t(a.ncol);
    hr.c = hg.c = c;
    for (int k = 0; k < nzc; ++k) hr.w0[size_t(k) * a.ncol + c] = hr.w1[size_t(k) * a.ncol + c] = 0u;
  }
  if (!(cf & CF_EXIT)) {
    unsigned any = 0;
    for (int k = 0; k < nzc; ++k) any |= a.pf[TEMPO_IDX_SCR(a, c, k)];
    // the column's masses and tendencies in FP64 while they sediment: the
    // scratch planes themselves when they are wide, else a wide copy
    constexpr bool narrow = !std::is_same<Sd_t<R>, W>::value;
    W* sd = narrow ? vtbuf + 3 * plane : reinterpret_cast<W*>(a.sd);
    if constexpr (narrow)
      for (int q = 0; q < SD_N; ++q)
#pragma unroll 8
        for (int k = 0; k < nzc; ++k) {
          const size_t s = size_t(q) * plane + TEMPO_IDX_SCR(a, c, k);
          sd[s] = W(a.sd[s]);
        }
    const R* sr = a.sr;
    W* v1 = vtbuf;           // filled fall speeds, FP64
    W* v2 = vtbuf + plane;
    W* vr = vtbuf + 2 * plane;  // filled rain mass fall speed kept for snow
    int steps, ktop;
    if (any & PF_QR) {
      fill_fallspeed(a, c, nzc, sd + SD_RR * plane, sr + SR_VTRR * plane, sr + SR_VTNR * plane, vr, v2, true, steps, ktop);
      if (a.fl.semi_sedi) {  // (:599-615): far fewer substeps, fall speeds refreshed after each
        steps = max(int(R(steps) / R(10.0)) + 1, 1);
        for (int n = 0; n < steps; ++n) {
          semi_sediment(a, c, nzc, sd + SD_RR * plane, sd + SD_QRTEN * plane, vr, steps, kR1, &pr[0], hr);
          semi_sediment(a, c, nzc, sd + SD_NR * plane, sd + SD_NRTEN * plane, v2, steps, kR2, (W*)nullptr, hr);
          semi_rain_refresh(a, c, nzc, sd, vr, v2, hr);
        }
      } else {
        for (int n = 0; n < steps; ++n) {
          sediment(a, c, nzc, sd + SD_RR * plane, vr, sd + SD_QRTEN * plane, kR1, steps, ktop, &pr[0]);
          sediment(a, c, nzc, sd + SD_NR * plane, v2, sd + SD_NRTEN * plane, kR2, steps, ktop, (W*)nullptr);
        }
      }
      kst[0] = steps;
      kst[1] = ktop;
    } else {
      for (int k = 0; k < nzc; ++k) vr[TEMPO_IDX_SCR(a, c, k)] = W(0.0);
    }
    if (any & PF_QG) {
      fill_fallspeed(a, c, nzc, sd + SD_RG * plane, sr + SR_VTRG * plane, sr + SR_VTNG * plane, v1, v2, false, steps, ktop);
      const double rb_lim = kM3ToL * kR1 / 800.0;
      if (a.fl.semi_sedi) {  // (:641-667)
        steps = max(int(R(steps) / R(10.0)) + 1, 1);
        for (int n = 0; n < steps; ++n) {
          semi_sediment(a, c, nzc, sd + SD_RG * plane, sd + SD_QGTEN * plane, v1, steps, kR1, &pr[1], hg);
          semi_sediment(a, c, nzc, sd + SD_NG * plane, sd + SD_NGTEN * plane, v2, steps, kR2, (W*)nullptr, hg);
          semi_sediment(a, c, nzc, sd + SD_RB * plane, sd + SD_QBTEN * plane, v1, steps, rb_lim, (W*)nullptr, hg);
          semi_graupel_refresh(a, c, nzc, sd, v1, v2, hg);
        }
      } else {
        for (int n = 0; n < steps; ++n) {
          sediment(a, c, nzc, sd + SD_RG * plane, v1, sd + SD_QGTEN * plane, kR1, steps, ktop, &pr[1]);
          sediment(a, c, nzc, sd + SD_NG * plane, v2, sd + SD_NGTEN * plane, kR2, steps, ktop, (W*)nullptr);
          sediment(a, c, nzc, sd + SD_RB * plane, v1, sd + SD_QBTEN * plane, rb_lim, steps, ktop, (W*)nullptr);
        }
      }
      kst[2] = steps;
      kst[3] = ktop;
    }
    if (any & PF_QS) {
      // snow_fallspeed mixes in rain where snow melts (:1886-1891), using the
      // rain mass after rain sedimentation and the filled rain fall speed
#pragma clang fp contract(off)
      const double dt = a.dt;
      W vprev = W(0.0);
      steps = 1;
      ktop = 1;
      for (int k = nzc - 1; k >= 0; --k) {
        const size_t s = TEMPO_IDX_SCR(a, c, k);
        const W rs = sd[SD_RS * plane + s];
        if (rs > kR1) {
          R v = sr[SR_VTS * plane + s];
          if (a.pf[s] & PF_SNOWMIX) {
            const R sr_ = R(rs / (rs + sd[SD_RR * plane + s]));
            v = v * sr_ + (R(1.0) - sr_) * R(vr[s]);
          }
          vprev = W(v);
        }
        v1[s] = vprev;
        const R vv = R(vprev);
        if (vv > R(1.e-3)) {
          ktop = max(ktop, k + 1);
          const R dz_by_vt = a.f[F_DZ][TEMPO_IDX_STATE(a, c, k)] / vv;
          steps = max(steps, int(R(dt) / dz_by_vt + R(1.0)));
        }
      }
      if (ktop == nzc) ktop = nzc - 1;
      for (int n = 0; n < steps; ++n)
        sediment(a, c, nzc, sd + SD_RS * plane, v1, sd + SD_QSTEN * plane, kR1, steps, ktop, &pr[2]);
      kst[4] = steps;
      kst[5] = ktop;
    }
    if (any & PF_QI) {
      // one call with steps = substeps (the reference applies 1/steps of the flux)
      fill_fallspeed(a, c, nzc, sd + SD_RI * plane, sr + SR_VTRI * plane, sr + SR_VTNI * plane, v1, v2, false, steps, ktop);
      sediment(a, c, nzc, sd + SD_RI * plane, v1, sd + SD_QITEN * plane, kR1, steps, ktop, &pr[3]);
      sediment(a, c, nzc, sd + SD_NI * plane, v2, sd + SD_NITEN * plane, kR2, steps, ktop, (W*)nullptr);
      kst[6] = steps;
      kst[7] = ktop;
    }
    if (any & PF_QC) {
      // cloud_fallspeed (:1945-1977): settling only below 500 m
      int kt = 1;
      R hgt = R(0.0);
      for (int k = 0; k < nzc - 1; ++k) {
        if (sd[SD_RC * plane + TEMPO_IDX_SCR(a, c, k)] > kR2) kt = k + 1;
        hgt = hgt + a.f[F_DZ][TEMPO_IDX_STATE(a, c, k)];
        if (hgt > R(500.0)) break;
      }
      for (int k = 0; k < nzc; ++k) {
        const size_t s = TEMPO_IDX_SCR(a, c, k);
        v1[s] = (k < kt) ? W(sr[SR_VTRC * plane + s]) : W(0.0);
        v2[s] = (k < kt) ? W(sr[SR_VTNC * plane + s]) : W(0.0);
      }
      sediment(a, c, nzc, sd + SD_RC * plane, v1, sd + SD_QCTEN * plane, kR1, 1, kt, &pr[4]);
      sediment(a, c, nzc, sd + SD_NC * plane, v2, sd + SD_NCTEN * plane, kR2, 1, kt, (W*)nullptr);
      kst[8] = 1;
      kst[9] = kt;
    }
    if constexpr (narrow)  // the tendencies, rounded once to the working precision
      for (int q = SD_QRTEN; q < SD_N; ++q)
#pragma unroll 8
        for (int k = 0; k < nzc; ++k) {
          const size_t s = size_t(q) * plane + TEMPO_IDX_SCR(a, c, k);
          a.sd[s] = Sd_t<R>(sd[s]);
        }
  }
  for (int q = 0; q < 5; ++q) a.precip[size_t(q) * a.ncol + c] = pr[q];
  kst[10] = int(hr.h);
  kst[11] = int(hg.h);
  if (a.ksub)
    for (int q = 0; q < kSubRows; ++q) a.ksub[size_t(q) * a.ncol + c] = kst[q];
  // column diagnostics (src/module_mp_tempo_main.F90:784-794, driver :452-457)
  const size_t d2 = size_t(c % a.ni) + size_t(c / a.ni) * a.d2ldj;
  const bool live = !(cf & CF_EXIT);
  const R rain = R(pr[0]), graupel = R(pr[1]), snow = R(pr[2]), ice = R(pr[3]), cloud = R(pr[4]);
  if (a.d2[D2_RAIN]) a.d2[D2_RAIN][d2] = rain;
  if (a.d2[D2_ICE]) a.d2[D2_ICE][d2] = ice;
  if (a.d2[D2_SNOW]) a.d2[D2_SNOW][d2] = snow;
  if (a.d2[D2_GRAUPEL]) a.d2[D2_GRAUPEL][d2] = graupel;
  if (a.d2[D2_FROZEN])
    a.d2[D2_FROZEN][d2] = live ? (ice + snow + graupel) / (ice + snow + graupel + rain + R(kR1)) : R(0);
  if (a.d2[D2_FRZRAIN]) {
    R frz = R(0);
    if (live && (a.sr[SR_TEMP * plane + TEMPO_IDX_SCR(a, c, 0)] - R(kT0)) < R(-0.5)) frz = rain + cloud;
    a.d2[D2_FRZRAIN][d2] = frz;
  }
  if (a.d2[D2_HAIL_SFC]) a.d2[D2_HAIL_SFC][d2] = R(0);
  if (a.d2[D2_HAIL_COL]) a.d2[D2_HAIL_COL][d2] = R(0);
}

// ------------------------------------------------------ melting level ----
// find_melting_level (src/module_mp_tempo_diags.F90:124-141) on the state
// the final update will check: rain present at a level above freezing with
// snow or graupel present at the level above; presence as the final
// checks decide it (mass after the step's tendencies above r1).  One lane
// per column, before the final update.
template <class R>
__global__ void __launch_bounds__(256) k_melt_level(StepArgs<R> a) {
  const int c = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= a.ncol || (a.cf[c] & CF_EXIT)) return;
  const int nzc = a.nlev ? min(max(a.nlev[c], 1), a.nz) : a.nz;
  const R dt = R(a.dt);
  const size_t plane = a.plane;
  auto present = [&](int f, int ten, int k) {
    const size_t o = TEMPO_IDX_STATE(a, c, k), s = TEMPO_IDX_SCR(a, c, k);
    return a.fa[f][o] + R(a.sd[size_t(ten) * plane + s]) * dt > R(kR1);
  };
  int k_melt = 1;
  for (int k = nzc - 2; k >= 0; --k) {
    const R temp = a.sr[SR_TEMP * plane + TEMPO_IDX_SCR(a, c, k)];
    if (temp > R(kT0) && present(F_QR, SD_QRTEN, k) &&
        (present(F_QS, SD_QSTEN, k + 1) || present(F_QG, SD_QGTEN, k + 1))) {
      k_melt = max(k + 2, k_melt);
      break;
    }
  }
  a.kmelt[c] = k_melt;
}

// a level below the column's melting level, where reflectivity_10cm takes
// snow and graupel as melting (:92-112; k_melt > 2 and k < k_melt - 1, 1-based)
template <class R>
__device__ __forceinline__ bool melt_level(const StepArgs<R>& a, int c, int k) {
  return a.fl.refl_from_melting && a.fl.refl10cm && a.kmelt[c] > 2 && k + 1 < a.kmelt[c] - 1;
}

// reflectivity_10cm at the levels with melting snow or graupel: the final
// rain, snow and graupel checks of the final update (on copies, from the same
// operands), dry rain and the binned wet-particle reflectivity of snow and
// graupel.  Its own launch, before the final update (which leaves these
// levels' reflectivity to it), so that the final update keeps its register
// budget.
template <class R>
__global__ void __launch_bounds__(256) k_melt_refl(StepArgs<R> a) {
  const int c = int(blockIdx.x) * 64 + int(threadIdx.x & 63u);
  const int k = int(blockIdx.y) * int(blockDim.x >> 6) + int(threadIdx.x >> 6);
  if (c >= a.ncol || k >= a.nz) return;
  if ((a.nlev && k >= a.nlev[c]) || (a.cf[c] & CF_EXIT) || !melt_level(a, c, k)) return;
  const Const<S_t<R>>& C = *a.cst;
  const size_t o = TEMPO_IDX_STATE(a, c, k), s = TEMPO_IDX_SCR(a, c, k), plane = a.plane;
  const R dt = R(a.dt), odt = R(a.odt);
  const bool has_ngqb = a.f[F_NG] && a.f[F_QB];
  Lvl<R> L;
  L.temp = a.sr[SR_TEMP * plane + s];
  L.qv = a.sr[SR_QV * plane + s];
  L.pres = a.f[F_P][o];
  L.rho = rho_of(C, L.pres, L.temp, L.qv);
  L.qrten = R(a.sd[SD_QRTEN * plane + s]);
  L.nrten = R(a.sd[SD_NRTEN * plane + s]);
  L.qsten = R(a.sd[SD_QSTEN * plane + s]);
  L.qgten = R(a.sd[SD_QGTEN * plane + s]);
  L.ngten = R(a.sd[SD_NGTEN * plane + s]);
  L.qbten = R(a.sd[SD_QBTEN * plane + s]);
  R qr1d = a.fa[F_QR][o], nr1d = a.fa[F_NR][o], qs1d = a.fa[F_QS][o], qg1d = a.fa[F_QG][o];
  R ng1d = has_ngqb ? a.fa[F_NG][o] : R(0), qb1d = has_ngqb ? a.fa[F_QB][o] : R(0);
  rain_check(C, L, qr1d, nr1d, dt, odt);
  snow_check(L, qs1d, dt, odt);
  L.smob = L.smoc = R(0);
  if (L.l_qs) snow_moments(C, L.rs, mn(R(-0.1), L.temp - R(kT0)), 0, L);
  graupel_check(C, L, qg1d, ng1d, qb1d, has_ngqb, dt, odt);
  if (!(L.l_qs || L.l_qg)) return;
  constexpr bool kCR = sizeof(S_t<R>) == 8;
  R ze_r = R(1.e-22), ze_s = R(1.e-22), ze_g = R(1.e-22);
  if (L.l_qr) {
    const R lamr = R(1.0) / L.ilamr;
    const R n0_r = L.nr * C.org2 * mpow(lamr, C.cre[1]);
    ze_r = n0_r * C.crg[3] * mpow(L.ilamr, C.cre[3]);
  }
  if (L.l_qs)
    ze_s = R(S_t<R>(radar::melting_snow<kCR>(*a.radar, double(value(L.rs)), double(value(L.smob)),
                                             double(value(L.smoc)), double(value(L.rr)))));
  if (L.l_qg) {
    const int idx = L.idx_bg;
    ze_g = R(S_t<R>(radar::melting_graupel<kCR>(*a.radar, double(C.am_g[idx - 1]), double(C.ogg2),
                                                double(C.cge[1][0]), double(value(L.rg)), double(value(L.ng)),
                                                double(value(L.ilamg)), idx, double(value(L.rr)))));
  }
  const size_t d3 = size_t(c % a.ni) + size_t(k) * a.dldk + size_t(c / a.ni) * a.dldj;
  a.d3[D3_REFL][d3] = mx(R(-35.0), R(10.0) * mlog10((ze_r + ze_s + ze_g) * R(1.e18)));
}

// ----------------------------------------------------------- phase C ----
__device__ __forceinline__ void atomic_max_pos(float* p, float v) {
  atomicMax(reinterpret_cast<unsigned int*>(p), __float_as_uint(v));
}
__device__ __forceinline__ void atomic_max_pos(double* p, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}
// tangent-linear runs: the column maximum of the value only (no derivative)
template <class T>
__device__ __forceinline__ void atomic_max_pos(Dual<T>* p, Dual<T> v) {
  atomic_max_pos(&p->v, v.v);
}

// freeze_cloud_melt_ice (src/module_mp_tempo_main.F90:2124-2165), on the
// tendencies of the final update
template <class R>
__device__ __forceinline__ void freeze_cloud_melt_ice(const Const<S_t<R>>& C, Lvl<R>& L, R qc1d, R qi1d, R ni1d,
                                                      R nc1d, bool has_nc, R ncv, R lvap, R ocp, R dt, R odt) {
  // uncontracted, as the final checks: a fully frozen or melted mass then
  // cancels to zero as in the reference
#pragma clang fp contract(off)
  const R xri = mx(R(0.0), qi1d + L.qiten * dt);
  if ((L.temp > R(kT0)) && (xri > R(0.0))) {
    L.qcten = L.qcten + xri * odt;
    L.ncten = L.ncten + ni1d * odt;
    L.qiten = L.qiten - xri * odt;
    L.niten = -ni1d * odt;
    L.tten = L.tten - C.lfus * ocp * xri * odt;
  }
  const R xrc = mx(R(0.0), qc1d + L.qcten * dt);
  if ((L.temp < R(kHgFrz)) && (xrc > R(0.0))) {
    const R lfus2 = R(kLsub) - lvap;
    const R xnc = has_nc ? nc1d + L.ncten * dt : ncv / L.rho + L.ncten * dt;
    L.qiten = L.qiten + xrc * odt;
    L.niten = L.niten + xnc * odt;
    L.qcten = L.qcten - xrc * odt;
    L.ncten = L.ncten - xnc * odt;
    L.tten = L.tten + lfus2 * ocp * xrc * odt;
  }
}
// a signature run executes it on the plain type (its contraction pragma
// applies to the plain type only) and records its two decisions
template <class T>
__device__ __forceinline__ void freeze_cloud_melt_ice(const Const<T>& C, Lvl<Sig<T>>& L, Sig<T> qc1d, Sig<T> qi1d,
                                                      Sig<T> ni1d, Sig<T> nc1d, bool has_nc, Sig<T> ncv, Sig<T> lvap,
                                                      Sig<T> ocp, Sig<T> dt, Sig<T> odt) {
  Lvl<T> M = sig_unwrap(L);
  const T qiten = M.qiten, qcten = M.qcten;
  freeze_cloud_melt_ice<T>(C, M, qc1d.v, qi1d.v, ni1d.v, nc1d.v, has_nc, ncv.v, lvap.v, ocp.v, dt.v, odt.v);
  sig_wrap(L, M);
  sig_note(0x61u + (M.qiten != qiten ? 2u : 0u) + (M.qcten != qcten ? 4u : 0u));
}

template <class T>
__device__ __forceinline__ void atomic_max_pos(Sig<T>* p, Sig<T> v) {
  atomic_max_pos(&p->v, v.v);
}

// freeze/melt, final update, final checks and diagnostics of point (c, k)
// (src/module_mp_tempo_main.F90:729-863); scratch in scr
template <class R>
__device__ __forceinline__ void phase_c_point(const StepArgs<R>& a, const Scr<R>& scr, int c, int k) {
  const Const<S_t<R>>& C = *a.cst;
  const size_t o = TEMPO_IDX_STATE(a, c, k);
  const size_t s = scr.at(c, k);
  const size_t plane = scr.plane;
  const size_t d3 = size_t(c % a.ni) + size_t(k) * a.dldk + size_t(c / a.ni) * a.dldj;
  const R dt = R(a.dt), odt = R(a.odt);
  const bool live = !(a.cf[c] & CF_EXIT);
  const bool has_nc = a.f[F_NC], has_nwfa = a.f[F_NWFA], has_nifa = a.f[F_NIFA];
  const bool has_ngqb = a.f[F_NG] && a.f[F_QB];
  if (!live) {  // early exit: the driver's reset values (src/module_mp_tempo_driver.F90:304-386)
    if (a.d3[D3_REFL]) a.d3[D3_REFL][d3] = R(-35.0);
    for (int q = D3_RE_C; q < D3_N; ++q)
      if (a.d3[q]) a.d3[q][d3] = R(0);
    return;
  }
  Lvl<R> L;
  const R* sr = scr.sr;
  const Sd_t<R>* sd = scr.sd;
  L.temp = sr[SR_TEMP * plane + s];
  L.qv = sr[SR_QV * plane + s];
  L.rho = sr[SR_RHO * plane + s];
  L.tten = sr[SR_TTEN * plane + s];
  L.qvten = sr[SR_QVTEN * plane + s];
  L.nwfaten = sr[SR_NWFATEN * plane + s];
  L.nifaten = sr[SR_NIFATEN * plane + s];
  // the post-sedimentation tendencies; an adjoint's local pass takes their
  // values from a.post and their (local) tangents from the scratch
  auto ten = [&](int m) {
    const Sd_t<R>& x = sd[size_t(SD_QRTEN + m) * plane + s];
    if constexpr (IsDual<R>::value) {
      if (a.post) return R(Sd_t<R>(a.post[size_t(m) * plane + s], x.d));
    }
    return R(x);
  };
  L.qrten = ten(0);
  L.nrten = ten(1);
  L.qgten = ten(2);
  L.ngten = ten(3);
  L.qbten = ten(4);
  L.qsten = ten(5);
  L.qiten = ten(6);
  L.niten = ten(7);
  L.qcten = ten(8);
  L.ncten = ten(9);
  L.pres = a.f[F_P][o];
  const R lvap = R(kLvap0) + R(2106.0 - 4218.0) * (L.temp - R(kT0));
  const R ocp = R(1.0) / (R(kCp) * (R(1.0) + R(0.887) * L.qv));
  R qc1d = a.fa[F_QC][o], qr1d = a.fa[F_QR][o], qi1d = a.fa[F_QI][o], qs1d = a.fa[F_QS][o], qg1d = a.fa[F_QG][o];
  R ni1d = a.fa[F_NI][o], nr1d = a.fa[F_NR][o];
  R nc1d = has_nc ? a.fa[F_NC][o] : R(0), ng1d = has_ngqb ? a.fa[F_NG][o] : R(0);
  R qb1d = has_ngqb ? a.fa[F_QB][o] : R(0);
  const R ncv = (a.fl.ml_for_nc && !has_nc) ? sr[SR_NCSAVE * plane + s] : column_nc(a, c);

  if (!a.fl.turn_off_micro) freeze_cloud_melt_ice(C, L, qc1d, qi1d, ni1d, nc1d, has_nc, ncv, lvap, ocp, dt, odt);
  // final update (:738-748)
  const R t_new = load_t(a, o) + L.tten * dt;
  const R qv_new = mx(R(kMinQv), (a.f[F_QV][o] + L.qvten * dt));
  L.rho = rho_of(C, L.pres, L.temp, L.qv);
  if (has_nwfa) a.fo[F_NWFA][o] = mx(R(kNwfaDefault), mn(R(kAeroMax), (a.f[F_NWFA][o] + L.nwfaten * dt)));
  if (has_nifa) a.fo[F_NIFA][o] = mx(R(kNifaDefault), mn(R(kAeroMax), (a.f[F_NIFA][o] + L.nifaten * dt)));
  // final checks (:750-781)
  cloud_check(C, L, qc1d, nc1d, has_nc, ncv, dt, odt);
  rain_check(C, L, qr1d, nr1d, dt, odt);
  ice_check(C, L, qi1d, ni1d, dt, odt);
  snow_check(L, qs1d, dt, odt);
  L.smob = L.smoc = L.smo2 = L.smoz = R(0);
  if (L.l_qs) snow_moments(C, L.rs, mn(R(-0.1), L.temp - R(kT0)), 4 | 32, L);
  graupel_check(C, L, qg1d, ng1d, qb1d, has_ngqb, dt, odt);

  if (a.f[F_T])
    a.fo[F_T][o] = t_new;
  else
    a.fo[F_TH][o] = t_new / a.f[F_PII][o];
  a.fo[F_QV][o] = qv_new;
  a.fo[F_QC][o] = qc1d;
  a.fo[F_QR][o] = qr1d;
  a.fo[F_NR][o] = nr1d;
  a.fo[F_QI][o] = qi1d;
  a.fo[F_NI][o] = ni1d;
  a.fo[F_QS][o] = qs1d;
  a.fo[F_QG][o] = qg1d;
  if (has_nc) a.fo[F_NC][o] = nc1d;
  if (has_ngqb) {
    a.fo[F_NG][o] = ng1d;
    a.fo[F_QB][o] = qb1d;
  }

  // diagnostics (:796-863, src/module_mp_tempo_diags.F90)
  if (a.d3[D3_NC_MIX]) a.d3[D3_NC_MIX][d3] = a.fl.nc_mixing_ratio ? L.nc * L.rho : R(0);
  if (a.d3[D3_MVD_R]) a.d3[D3_MVD_R][d3] = a.fl.mvd_r ? L.mvd_r : R(0);
  if (a.d3[D3_MVD_G]) a.d3[D3_MVD_G][d3] = a.fl.mvd_g ? L.mvd_g : R(0);
  if (a.fl.max_hail) {  // max_hail_diam (:398-438)
    R hail = R(0.0);
    if (L.rg / L.rho >= R(1.e-6) && rho_g<R>(L.idx_bg) >= R(350.0)) {
      const R lamg = R(1.0) / L.ilamg;
      const R n0_g = L.ng * C.ogg2 * mpow(lamg, C.cge[1][0]);
      R sum_nh = R(0), sum_t = R(0);
      const R thr = R(0.0005);
      int n;
      for (n = kNhbins; n >= 1; --n) {
        const R f_d = n0_g * mpow(C.hbins[n - 1], R(kMuG)) * mexp(-lamg * C.hbins[n - 1]) * C.dhbins[n - 1];
        sum_nh = sum_nh + f_d;
        if (sum_nh > thr) break;
        sum_t = sum_nh;
      }
      R hmax;
      if (n >= kNhbins)
        hmax = C.hbins[kNhbins - 1];
      else if (C.hbins[n] > R(1.e-3))
        hmax = C.hbins[n - 1] - (sum_nh - thr) / (sum_nh - sum_t) * (C.hbins[n - 1] - C.hbins[n]);
      else
        hmax = R(1.e-4);
      hail = R(1000.0) * hmax;
    }
    const size_t d2 = size_t(c % a.ni) + size_t(c / a.ni) * a.d2ldj;
    if (k == 0 && a.d2[D2_HAIL_SFC]) a.d2[D2_HAIL_SFC][d2] = hail;
    if (a.d2[D2_HAIL_COL] && hail > R(0)) atomic_max_pos(&a.d2[D2_HAIL_COL][d2], hail);
  }
  if (a.d3[D3_REFL]) {  // reflectivity_10cm (:59-121)
    R dbz = R(-35.0);
    if (a.fl.refl10cm) {
      R ze_r = R(1.e-22), ze_s = R(1.e-22), ze_g = R(1.e-22);
      if (L.l_qr) {
        const R lamr = R(1.0) / L.ilamr;
        const R n0_r = L.nr * C.org2 * mpow(lamr, C.cre[1]);
        ze_r = n0_r * C.crg[3] * mpow(L.ilamr, C.cre[3]);
      }
      const R k6 = R(0.176 / 0.93) * (R(6.0) / R(kPi)) * (R(6.0) / R(kPi));
      if (L.l_qs) ze_s = k6 * (R(kAmS) / R(900.0)) * (R(kAmS) / R(900.0)) * L.smoz;
      if (L.l_qg) {
        const int idx = L.idx_bg;
        const R n0_g = L.ng * C.ogg2 * mpow(R(1.0) / L.ilamg, C.cge[1][0]);
        ze_g = k6 * (C.am_g[idx - 1] / R(900.0)) * (C.am_g[idx - 1] / R(900.0)) * n0_g * C.cgg[3][0] *
               mpow(L.ilamg, C.cge[3][0]);
      }
      dbz = mx(R(-35.0), R(10.0) * mlog10((ze_r + ze_s + ze_g) * R(1.e18)));
    }
    // levels with melting snow or graupel are k_melt_refl's
    if (!(melt_level(a, c, k) && (L.l_qs || L.l_qg))) a.d3[D3_REFL][d3] = dbz;
  }
  const bool re = a.fl.re_cloud && a.fl.re_ice && a.fl.re_snow;
  R re_c = R(0), re_i = R(0), re_s = R(0);
  if (re && a.qc_bl && a.qcfrac_bl) {
    // boundary-layer cloud in the cloud radius (:839-857): in-cloud PBL water
    // where the resolved cloud is absent, the network's number for the sum,
    // and the cloud check redone on them (state and earlier diagnostics keep
    // the resolved cloud)
    R xrx = qc1d;
    const R qbl = a.qc_bl[o], fbl = a.qcfrac_bl[o];
    if (xrx <= R(kR1) && qbl > R(S_t<R>(1.e-9)) && fbl > R(0.0)) xrx = xrx + qbl / fbl;
    if (xrx <= R(S_t<R>(1.e-12))) xrx = R(0.0);
    R xnc = ml_cloud_number(xrx, qr1d, qi1d, qs1d, L.pres, L.temp, a.f[F_W][o]);
    L.qcten = L.ncten = R(0.0);
    cloud_check(C, L, xrx, xnc, true, ncv, dt, odt);
  }
  if (re) {  // effective_radius (:13-56)
    if (L.l_qc) re_c = mx(R(2.51e-6), mn(R(0.5) * (R(3.0) + R(get_nuc(L.nc))) * L.ilamc, R(50.e-6)));
    if (L.l_qi) re_i = mx(R(2.51e-6), mn(R(0.5) * (R(3.0) + R(kMuI)) * L.ilami, R(125.e-6)));
    if (L.l_qs) {
      Lvl<R> M;
      snow_moments(C, L.rs, mn(R(-0.1), L.temp - R(kT0)), 0, M);
      re_s = mx(R(5.01e-6), mn(R(0.5) * (M.smoc / M.smob), R(999.e-6)));
    }
  }
  if (a.d3[D3_RE_C]) a.d3[D3_RE_C][d3] = re_c;
  if (a.d3[D3_RE_I]) a.d3[D3_RE_I][d3] = re_i;
  if (a.d3[D3_RE_S]) a.d3[D3_RE_S][d3] = re_s;
}

template <class R>
__global__ void __launch_bounds__(256) TEMPO_ATTR_C k_phase_c(StepArgs<R> a) {
  const int2 pt = block_point<R, true>(a);
  const int c = pt.x, k = pt.y;
  if (c >= a.ncol || k >= a.nz) return;
  if (a.nlev && k >= a.nlev[c]) return;
  if (a.lvl_on && !a.lvl_on[k]) return;
  SigScope<R> scope(a, c, k, 1);
  phase_c_point(a, batch_scr(a), c, k);
}

// ------------------------------------ public check-and-update routines ----
// cloud_check_and_update (:891-973), ice_check_and_update (:1045-1107) and
// snow_check_and_update (:1110-1134) on n independent points, the device
// functions the step uses; which = 0 cloud, 1 ice, 2 snow.
template <class R>
__global__ void __launch_bounds__(256) k_check_update(CheckArgs<R> g, int which) {
  const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= g.n) return;
  const Const<S_t<R>>& C = *g.cst;
  Lvl<R> L;
  L.rho = g.rho[i];
  R q = g.q[i];
  if (which == 0) {
    R nn = g.nn ? g.nn[i] : R(0);
    L.qcten = g.qten[i];
    L.ncten = g.nten[i];
    L.rc = g.r[i];
    L.nc = g.num[i];
    const R ncv = g.ncsave ? g.ncsave[i] : R(kNtCL);
    cloud_check(C, L, q, nn, g.nn != nullptr, ncv, g.dt, g.odt);
    if (g.nn) g.nn[i] = nn;
    g.l[i] = L.l_qc ? 1 : 0;
    g.r[i] = L.rc;
    g.num[i] = L.nc;
    g.qten[i] = L.qcten;
    g.nten[i] = L.ncten;
    g.ilam[i] = L.ilamc;
    g.mvd[i] = L.mvd_c;
  } else if (which == 1) {
    R nn = g.nn[i];
    L.qiten = g.qten[i];
    L.niten = g.nten[i];
    L.ri = g.r[i];
    L.ni = g.num[i];
    ice_check(C, L, q, nn, g.dt, g.odt);
    g.nn[i] = nn;
    g.l[i] = L.l_qi ? 1 : 0;
    g.r[i] = L.ri;
    g.num[i] = L.ni;
    g.qten[i] = L.qiten;
    g.nten[i] = L.niten;
    g.ilam[i] = L.ilami;
  } else {
    L.qsten = g.qten[i];
    L.rs = g.r[i];
    snow_check(L, q, g.dt, g.odt);
    g.l[i] = L.l_qs ? 1 : 0;
    g.r[i] = L.rs;
    g.qten[i] = L.qsten;
  }
  g.q[i] = q;
}

template <class R>
hipError_t launch_check_update(const CheckArgs<R>& g, int which, hipStream_t s) {
  if (g.n > 0)
    hipLaunchKernelGGL(k_check_update<R>, dim3(unsigned((g.n + 255) / 256)), dim3(256), 0, s, g, which);
  return hipGetLastError();
}

// -------------------------------------------- surface emissions ----
// tempo_aerosol_surface_emissions (src/module_mp_tempo_driver.F90:516-529):
// nwfa(i, k, j) += nwfa2d(i, j) * dt at one level, one lane per column.
template <class R>
__global__ void __launch_bounds__(256) k_surface_emissions(R* __restrict__ nwfa, const R* __restrict__ nwfa2d,
                                                           int ni, int ncol, long long ldj, long long d2ldj,
                                                           size_t koff, R dt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  const int i = c % ni, j = c / ni;
  R& x = nwfa[size_t(i) + koff + size_t(j) * size_t(ldj)];
  x = x + nwfa2d[size_t(i) + size_t(j) * size_t(d2ldj)] * dt;
}

template <class R>
hipError_t launch_surface_emissions(R* nwfa, const R* nwfa2d, int ni, int nj, long long ldk, long long ldj,
                                    long long d2ldj, int k, double dt, hipStream_t s) {
  const int ncol = ni * nj;
  hipLaunchKernelGGL(k_surface_emissions<R>, dim3(unsigned((ncol + 255) / 256)), dim3(256), 0, s, nwfa, nwfa2d, ni,
                     ncol, ldj, d2ldj, size_t(k) * size_t(ldk), R(dt));
  return hipGetLastError();
}

// ------------------------------------------------ tangent-linear I/O ----
// caller arrays (value, tangent) <-> one compact dual array, element (c, k)
// at c + k*ncol
template <class T>
__global__ void __launch_bounds__(256) k_dual_pack(const T* __restrict__ v, const T* __restrict__ t, DualGeom g,
                                                   Dual<T>* __restrict__ out) {
  const size_t ncol = size_t(g.ni) * g.nj;
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ncol * size_t(g.nk)) return;
  const size_t c = e % ncol, k = e / ncol;
  const size_t o = (c % g.ni) + k * g.ldk + (c / g.ni) * g.ldj;
  out[e] = Dual<T>(v[o], t ? t[o] : T(0));
}

template <class T>
__global__ void __launch_bounds__(256) k_dual_unpack(const Dual<T>* __restrict__ in, DualGeom g, T* __restrict__ v,
                                                     T* __restrict__ t) {
  const size_t ncol = size_t(g.ni) * g.nj;
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ncol * size_t(g.nk)) return;
  const size_t c = e % ncol, k = e / ncol;
  const size_t o = (c % g.ni) + k * g.ldk + (c / g.ni) * g.ldj;
  if (v) v[o] = in[e].v;
  if (t) t[o] = in[e].d;
}

template <class T>
hipError_t launch_dual_pack(const T* v, const T* t, const DualGeom& g, Dual<T>* out, hipStream_t s) {
  const size_t n = size_t(g.ni) * g.nj * g.nk;
  hipLaunchKernelGGL(k_dual_pack<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, v, t, g, out);
  return hipGetLastError();
}

template <class T>
hipError_t launch_dual_unpack(const Dual<T>* in, const DualGeom& g, T* v, T* t, hipStream_t s) {
  const size_t n = size_t(g.ni) * g.nj * g.nk;
  hipLaunchKernelGGL(k_dual_unpack<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, in, g, v, t);
  return hipGetLastError();
}

// the four launches of one tangent-linear step (sedimentation one lane per column)
template <class D>
hipError_t launch_step_tlm_impl(const StepArgs<D>& a, Dual<double>* vtbuf, hipStream_t s) {
  hipLaunchKernelGGL(k_prepass<D>, prepass_grid(a), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_phase_a<D>, level_grid(a, 256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_phase_b<D>, dim3(unsigned((a.ncol + 127) / 128)), dim3(128), 0, s, a, vtbuf);
  hipLaunchKernelGGL(k_phase_c<D>, level_grid(a, 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace tempo

// Phase B on CDNA4: one wavefront per column, levels across lanes.
//
// Level k of a column lives in lane k % 64 of register chunk k / 64
// (kChunks chunks, so nz <= 64 * kChunks).  The column couplings of the
// reference's sedimentation (src/module_mp_tempo_main.F90:590-727) become
// wavefront primitives:
//   top-down fill of a fall-speed profile (:1768-1789 and species variants)
//       -> ballot of the levels holding mass, lowest set bit at or above
//          each lane, value fetched by a lane permute;
//   CFL substep count and ktop -> wavefront max reductions;
//   one flux-form substep (:1516-1559) -> every level at once, the flux
//          from the level above by a one-lane shift (all fluxes of a
//          substep use the fields at its start, so the levels are
//          independent within it).
// Every per-level operand is staged through LDS: a block takes one column
// group of the column-tiled scratch (kScrW columns whose levels are
// contiguous), reads it as one contiguous run and transposes it to
// (column, level) rows that one wavefront reads with unit stride; all planes
// of a species are requested together.
//
// The arithmetic per level is the serial kernel's, operation for operation,
// so both give identical results.
#pragma once

#include <cstring>
#include <type_traits>

#include "tempo_step.hpp"

namespace tempo {

constexpr int kSedCols = kScrW;           // columns (wavefronts) per block: one scratch column group
constexpr int kSedSlots = 8;              // LDS planes staged at once
constexpr int kSedChunks = 2;             // register chunks of 64 levels
constexpr int kSedLevels = 64 * kSedChunks;
// LDS row of one column: nz + 1 doubles (odd, so that the 8 columns of a
// level fall in different banks); the block's LDS is sized at launch
__host__ __device__ constexpr int sed_row(int nz) { return nz + 1; }
__host__ __device__ constexpr size_t sed_lds_bytes(int nz) { return size_t(kSedSlots) * kSedCols * sed_row(nz) * sizeof(double); }

namespace wv {
__device__ __forceinline__ int lane() { return int(threadIdx.x & 63u); }
__device__ __forceinline__ int imax(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ unsigned uor(unsigned v) {
  for (int o = 32; o > 0; o >>= 1) v |= unsigned(__shfl_xor(int(v), o, 64));
  return v;
}
// value of x at level k (k uniform or per lane), chunks held in x[]
template <class T>
__device__ __forceinline__ T at_level(const T (&x)[kSedChunks], int k) {
  T v = T(0);
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) {
    const T t = __shfl(x[j], k & 63, 64);
    if ((k >> 6) == j) v = t;
  }
  return v;
}
// value of x at level k+1 for every lane's level k in chunk j (0 above the top chunk)
template <class T>
__device__ __forceinline__ T above(const T (&x)[kSedChunks], int j) {
  const T up = __shfl_down(x[j], 1, 64);
  T nxt = T(0);
  if (j + 1 < kSedChunks) nxt = __shfl(x[j + 1 < kSedChunks ? j + 1 : j], 0, 64);
  return lane() == 63 ? nxt : up;
}
}  // namespace wv

// the value of a stored number as a double (the staging carries values
// only: a dual plane's tangent is not staged)
__device__ __forceinline__ double sval(double x) { return x; }
__device__ __forceinline__ double sval(float x) { return double(x); }
template <class T>
__device__ __forceinline__ double sval(const Dual<T>& x) {
  return double(x.v);
}

// Block-wide staging of (level, column) planes through LDS slots.  Loads
// and stores issue without barriers so that all planes of a stage are in
// flight together; the caller separates the phases with __syncthreads().
// kEveryPass: issue every load pass, also those past the group's end (all
// lanes then load the group's last element and write nothing) -- a check
// form of the staging, selected with sed_wave = 2; results are identical.
template <bool kEveryPass = false>
struct SedTile {
  double* lds;  // kSedSlots x kSedCols x row
  int c0, ncol, nz;
  int nzs;      // level stride of a scratch column group
  int row;      // LDS row of one column (sed_row(nz))
  size_t sbase; // scratch offset of the group's first element

  __device__ double* slot(int s) const { return lds + s * (kSedCols * row); }
  // Several scratch planes of one stage (the block's column group of each is
  // one contiguous run).  Every global load of the stage is issued into
  // registers before the first LDS write, so the stage waits for one memory
  // latency instead of one per plane and pass (a load-convert-write loop per
  // plane waits for each load before its write).  The block has
  // 64 * kSedCols threads and nz <= kSedLevels, so kSedChunks passes cover
  // the column group; a pass runs only when the group reaches it (a
  // block-uniform test), and its elements past the group's end load the last
  // one (the index is clamped, not branched on per lane, so that the loads
  // issue back to back) and are not written.
  template <class TA, int NA, class TB, int NB>
  __device__ void load_many(const int (&sa)[NA], const TA* const (&pa)[NA], const int (&sb)[NB],
                            const TB* const (&pb)[NB]) const {
    TA va[kSedChunks][NA];
    TB vb[kSedChunks][NB];
    const int n = kSedCols * nz;
    const size_t base = sbase;
#pragma unroll
    for (int it = 0; it < kSedChunks; ++it) {
      if (!kEveryPass && it * 64 * kSedCols >= n) break;
      const int e = min(int(threadIdx.x) + it * 64 * kSedCols, n - 1);
#pragma unroll
      for (int p = 0; p < NA; ++p) va[it][p] = pa[p][base + e];
#pragma unroll
      for (int p = 0; p < NB; ++p) vb[it][p] = pb[p][base + e];
    }
#pragma unroll
    for (int it = 0; it < kSedChunks; ++it) {
      const int e = int(threadIdx.x) + it * 64 * kSedCols;
      if (e < n) {
        const int cc = e % kSedCols, k = e / kSedCols;
#pragma unroll
        for (int p = 0; p < NA; ++p) slot(sa[p])[cc * row + k] = sval(va[it][p]);
#pragma unroll
        for (int p = 0; p < NB; ++p) slot(sb[p])[cc * row + k] = sval(vb[it][p]);
      }
    }
  }
  // the first stage in the same way: per-point flags (slot 0), layer depths
  // from the state (slot 1), density (slot 2) and rain mass (slot 3)
  template <class R, class H, class M>
  __device__ void load_head(const StepArgs<R>& a, const unsigned* pf, const H* rho, const M* rr) const {
    unsigned vf[kSedChunks];
    R vz[kSedChunks];
    H vh[kSedChunks];
    M vr[kSedChunks];
    const int n = kSedCols * nz;
    const size_t base = sbase;
#pragma unroll
    for (int it = 0; it < kSedChunks; ++it) {
      if (!kEveryPass && it * 64 * kSedCols >= n) break;
      const int e = min(int(threadIdx.x) + it * 64 * kSedCols, n - 1);
      const int cc = e % kSedCols, k = e / kSedCols, c = min(c0 + cc, ncol - 1);
      vf[it] = pf[base + e];
      vh[it] = rho[base + e];
      vr[it] = rr[base + e];
      vz[it] = a.f[F_DZ][TEMPO_IDX_STATE(a, c, k)];
    }
#pragma unroll
    for (int it = 0; it < kSedChunks; ++it) {
      const int e = int(threadIdx.x) + it * 64 * kSedCols;
      if (e < n) {
        const int cc = e % kSedCols, k = e / kSedCols, x = cc * row + k;
        slot(0)[x] = double(vf[it]);
        slot(1)[x] = c0 + cc < ncol ? sval(vz[it]) : 0.0;
        slot(2)[x] = sval(vh[it]);
        slot(3)[x] = sval(vr[it]);
      }
    }
  }
  template <class T>
  __device__ void store(int s, T* plane) const {
    const double* d = slot(s);
    T* dst = plane + sbase;
    for (int e = threadIdx.x; e < kSedCols * nz; e += blockDim.x) {
      const int cc = e % kSedCols, k = e / kSedCols;
      if (c0 + cc < ncol) dst[e] = T(d[cc * row + k]);
    }
  }
  // this wavefront's column <-> registers
  template <class T>
  __device__ void get(int s, int w, T (&x)[kSedChunks]) const {
    const double* d = slot(s);
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const int k = j * 64 + wv::lane();
      x[j] = (k < nz) ? T(d[w * row + k]) : T(0);
    }
  }
  template <class T>
  __device__ void put(int s, int w, const T (&x)[kSedChunks]) const {
    double* d = slot(s);
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const int k = j * 64 + wv::lane();
      if (k < nz) d[w * row + k] = double(x[j]);
    }
  }
};

// Per-column level geometry held by every lane.
struct SedCol {
  int nzc;                          // active levels
  double odz[kSedChunks], orho[kSedChunks];
};

// top-down fill (:1768-1789): out(k) = raw(k') at the lowest k' >= k whose
// mass exceeds r1 (0 when none); CFL substeps and ktop from the filled
// speeds in the working precision.
template <class R>
__device__ __forceinline__ void wave_fill(const double (&mass)[kSedChunks], const R (&raw1)[kSedChunks],
                                          const R (&raw2)[kSedChunks], bool two, bool use_max,
                                          const R (&dz)[kSedChunks], int nzc, double dt, double (&out1)[kSedChunks],
                                          double (&out2)[kSedChunks], int& substeps, int& ktop) {
#pragma clang fp contract(off)  // as the lane-per-column kernel: every operation rounded on its own
  unsigned long long m[kSedChunks];
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) {
    const int k = j * 64 + wv::lane();
    m[j] = __ballot(k < nzc && mass[j] > kR1);
  }
  int st = 1, kt = 1;
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) {
    const int l = wv::lane(), k = j * 64 + l;
    int src = -1;
    const unsigned long long mm = m[j] & (~0ull << l);
    if (mm) {
      src = j * 64 + __ffsll((long long)mm) - 1;
    } else {
#pragma unroll
      for (int j2 = j + 1; j2 < kSedChunks; ++j2)
        if (src < 0 && m[j2]) src = j2 * 64 + __ffsll((long long)m[j2]) - 1;
    }
    const R r1 = wv::at_level(raw1, src < 0 ? 0 : src);
    const R r2 = two ? wv::at_level(raw2, src < 0 ? 0 : src) : R(0);
    const double v1 = src < 0 ? 0.0 : double(r1);
    const double v2 = src < 0 ? 0.0 : double(r2);
    out1[j] = v1;
    out2[j] = v2;
    if (k < nzc) {
      const R vr1 = R(v1), vr2 = R(v2);
      const R vmax = use_max ? mx(vr1, vr2) : vr1;
      if (vmax > R(1.e-3)) {
        kt = max(kt, k + 1);
        const R dz_by_vt = dz[j] / vmax;
        st = max(st, int(R(dt) / dz_by_vt + R(1.0)));
      }
    }
  }
  substeps = wv::imax(st);
  ktop = wv::imax(kt);
  if (ktop == nzc) ktop = nzc - 1;
}

// one flux-form substep of one field (:1516-1559); returns the surface
// precipitation increment (0 when the caller does not accumulate it)
__device__ __forceinline__ double wave_sediment(const SedCol& g, double (&xr)[kSedChunks],
                                                const double (&vt)[kSedChunks], double (&xten)[kSedChunks],
                                                double limit, int steps, int ktop, double dt, bool precip) {
#pragma clang fp contract(off)  // as the lane-per-column kernel: every operation rounded on its own
  const double ost = 1.0 / double(steps);
  double sed[kSedChunks];
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) sed[j] = vt[j] * xr[j];
  const double s0 = __shfl(sed[0], 0, 64);
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) {
    const double up = wv::above(sed, j);
    const int k = j * 64 + wv::lane();
    if (k < g.nzc && (k == g.nzc - 1 || k <= ktop - 1)) {
      const double flux = (k == g.nzc - 1) ? -sed[j] : (up - sed[j]);
      xten[j] = xten[j] + flux * g.odz[j] * ost * g.orho[j];
      const double v = xr[j] + flux * g.odz[j] * dt * ost;
      xr[j] = v > limit ? v : limit;
    }
  }
  if (!precip) return 0.0;
  const double x0 = __shfl(xr[0], 0, 64);
  return x0 > kLowLimitPrecip ? s0 * dt * ost : 0.0;
}

// Sedimentation of the column group c0 .. c0 + kSedCols - 1, one wavefront
// per column (64 * kSedCols threads); scr holds the group's scratch,
// tile_mem sed_lds_bytes(nz) of LDS.  For a dual working type it runs on the
// values (V, the scalar type) and writes tendencies with zero tangents: the
// tangent-free forward of an adjoint.
template <class R, bool kEveryPass = false>
__device__ __forceinline__ void phase_b_group(const StepArgs<R>& a, const Scr<R>& scr, int c0, double* tile_mem) {
  __shared__ unsigned blk_any;
  const int w = int(threadIdx.x >> 6);
  const int c = c0 + w;
  const bool col_ok = c < a.ncol;
  const size_t plane = scr.plane;
  const SedTile<kEveryPass> T{tile_mem, c0, a.ncol, a.nz, scr.nzs, sed_row(a.nz), scr.at(c0, 0)};
  const double dt = a.dt;
  Sd_t<R>* sd = scr.sd;
  const R* sr = scr.sr;
  using V = typename Num<R>::S;

  SedCol g;
  g.nzc = col_ok ? (a.nlev ? min(max(a.nlev[c], 1), a.nz) : a.nz) : 0;
  if (a.sig && col_ok) {  // signature run: this kernel's decisions are per column (tempo_hip_substeps)
    const size_t n3 = size_t(a.ncol) * size_t(a.nz);
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const int k = j * 64 + wv::lane();
      if (k < g.nzc) a.sig[4 * n3 + size_t(k) * a.ncol + c] = a.sig[5 * n3 + size_t(k) * a.ncol + c] = 0u;
    }
  }
  const unsigned cf = col_ok ? a.cf[c] : CF_EXIT;
  const bool live = col_ok && !(cf & CF_EXIT);

  // column flags (which species hold mass) and the level geometry
  if (threadIdx.x == 0) blk_any = 0;
  T.load_head(a, scr.pf, sr + SR_RHO * plane, sd + SD_RR * plane);
  __syncthreads();
  unsigned pfl[kSedChunks];
  V dz[kSedChunks], rho[kSedChunks];
  double rr[kSedChunks], vr[kSedChunks];
  T.get(0, w, pfl);
  T.get(1, w, dz);
  T.get(2, w, rho);
  T.get(3, w, rr);
  unsigned any = 0;
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) {
    if (j * 64 + wv::lane() < g.nzc) any |= pfl[j];
    g.odz[j] = 1.0 / double(dz[j]);
    g.orho[j] = 1.0 / double(rho[j]);
    vr[j] = 0.0;
  }
  any = live ? wv::uor(any) : 0u;
  if (wv::lane() == 0 && any) atomicOr(&blk_any, any);
  __syncthreads();
  const unsigned bany = blk_any;

  double pr[5] = {0, 0, 0, 0, 0};
  int kst[10] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1};

  // ---- rain (:603-630) ----
  if (bany & PF_QR) {
    double xn[kSedChunks], tr[kSedChunks], tn[kSedChunks], v2[kSedChunks];
    V raw1[kSedChunks], raw2[kSedChunks];
    T.load_many({0, 1, 2}, {sd + SD_NR * plane, sd + SD_QRTEN * plane, sd + SD_NRTEN * plane}, {4, 5},
                {sr + SR_VTRR * plane, sr + SR_VTNR * plane});
    __syncthreads();
    T.get(0, w, xn);
    T.get(1, w, tr);
    T.get(2, w, tn);
    T.get(4, w, raw1);
    T.get(5, w, raw2);
    if (any & PF_QR) {
      int steps, ktop;
      wave_fill(rr, raw1, raw2, true, true, dz, g.nzc, dt, vr, v2, steps, ktop);
      for (int n = 0; n < steps; ++n) {
        pr[0] += wave_sediment(g, rr, vr, tr, kR1, steps, ktop, dt, true);
        wave_sediment(g, xn, v2, tn, kR2, steps, ktop, dt, false);
      }
      kst[0] = steps;
      kst[1] = ktop;
    }
    T.put(1, w, tr);
    T.put(2, w, tn);
    __syncthreads();
    T.store(1, sd + SD_QRTEN * plane);
    T.store(2, sd + SD_NRTEN * plane);
    __syncthreads();
  }

  // ---- graupel (:632-660) ----
  if (bany & PF_QG) {
    double xg[kSedChunks], xn[kSedChunks], xb[kSedChunks], tg[kSedChunks], tn[kSedChunks], tb[kSedChunks];
    double v1[kSedChunks], v2[kSedChunks];
    V raw1[kSedChunks], raw2[kSedChunks];
    T.load_many({0, 1, 2, 4, 5, 6},
                {sd + SD_RG * plane, sd + SD_NG * plane, sd + SD_RB * plane, sd + SD_QGTEN * plane,
                 sd + SD_NGTEN * plane, sd + SD_QBTEN * plane},
                {7, 3}, {sr + SR_VTRG * plane, sr + SR_VTNG * plane});
    __syncthreads();
    T.get(0, w, xg);
    T.get(1, w, xn);
    T.get(2, w, xb);
    T.get(4, w, tg);
    T.get(5, w, tn);
    T.get(6, w, tb);
    T.get(7, w, raw1);
    T.get(3, w, raw2);
    if (any & PF_QG) {
      int steps, ktop;
      wave_fill(xg, raw1, raw2, true, false, dz, g.nzc, dt, v1, v2, steps, ktop);
      const double rb_lim = kM3ToL * kR1 / 800.0;
      for (int n = 0; n < steps; ++n) {
        pr[1] += wave_sediment(g, xg, v1, tg, kR1, steps, ktop, dt, true);
        wave_sediment(g, xn, v2, tn, kR2, steps, ktop, dt, false);
        wave_sediment(g, xb, v1, tb, rb_lim, steps, ktop, dt, false);
      }
      kst[2] = steps;
      kst[3] = ktop;
    }
    T.put(4, w, tg);
    T.put(5, w, tn);
    T.put(6, w, tb);
    __syncthreads();
    T.store(4, sd + SD_QGTEN * plane);
    T.store(5, sd + SD_NGTEN * plane);
    T.store(6, sd + SD_QBTEN * plane);
    __syncthreads();
  }

  // ---- snow (:662-697): mixes in the filled rain speed where snow melts,
  // with the rain mass after rain sedimentation ----
  if (bany & PF_QS) {
    double xs[kSedChunks], ts[kSedChunks], v1[kSedChunks];
    V raw[kSedChunks];
    T.load_many({0, 1}, {sd + SD_RS * plane, sd + SD_QSTEN * plane}, {2}, {sr + SR_VTS * plane});
    __syncthreads();
    T.get(0, w, xs);
    T.get(1, w, ts);
    T.get(2, w, raw);
    if (any & PF_QS) {
#pragma clang fp contract(off)  // the snow-rain speed mix, as the lane-per-column kernel
      V vraw[kSedChunks];
#pragma unroll
      for (int j = 0; j < kSedChunks; ++j) {
        V v = raw[j];
        if (pfl[j] & PF_SNOWMIX) {
          const V sr_ = V(xs[j] / (xs[j] + rr[j]));
          v = v * sr_ + (V(1.0) - sr_) * V(vr[j]);
        }
        vraw[j] = v;
      }
      double unused[kSedChunks];
      int steps, ktop;
      wave_fill(xs, vraw, vraw, false, false, dz, g.nzc, dt, v1, unused, steps, ktop);
      for (int n = 0; n < steps; ++n) pr[2] += wave_sediment(g, xs, v1, ts, kR1, steps, ktop, dt, true);
      kst[4] = steps;
      kst[5] = ktop;
    }
    T.put(1, w, ts);
    __syncthreads();
    T.store(1, sd + SD_QSTEN * plane);
    __syncthreads();
  }

  // ---- ice (:699-715): one call with steps = substeps ----
  if (bany & PF_QI) {
    double xi[kSedChunks], xn[kSedChunks], ti[kSedChunks], tn[kSedChunks], v1[kSedChunks], v2[kSedChunks];
    V raw1[kSedChunks], raw2[kSedChunks];
    T.load_many({0, 1, 2, 4}, {sd + SD_RI * plane, sd + SD_NI * plane, sd + SD_QITEN * plane, sd + SD_NITEN * plane},
                {5, 6}, {sr + SR_VTRI * plane, sr + SR_VTNI * plane});
    __syncthreads();
    T.get(0, w, xi);
    T.get(1, w, xn);
    T.get(2, w, ti);
    T.get(4, w, tn);
    T.get(5, w, raw1);
    T.get(6, w, raw2);
    if (any & PF_QI) {
      int steps, ktop;
      wave_fill(xi, raw1, raw2, true, false, dz, g.nzc, dt, v1, v2, steps, ktop);
      pr[3] += wave_sediment(g, xi, v1, ti, kR1, steps, ktop, dt, true);
      wave_sediment(g, xn, v2, tn, kR2, steps, ktop, dt, false);
      kst[6] = steps;
      kst[7] = ktop;
    }
    T.put(2, w, ti);
    T.put(4, w, tn);
    __syncthreads();
    T.store(2, sd + SD_QITEN * plane);
    T.store(4, sd + SD_NITEN * plane);
    __syncthreads();
  }

  // ---- cloud (:717-727, cloud_fallspeed :1945-1977): below 500 m only ----
  if (bany & PF_QC) {
    double xc[kSedChunks], xn[kSedChunks], tc[kSedChunks], tn[kSedChunks], v1[kSedChunks], v2[kSedChunks];
    V raw1[kSedChunks], raw2[kSedChunks];
    T.load_many({0, 1, 2, 4}, {sd + SD_RC * plane, sd + SD_NC * plane, sd + SD_QCTEN * plane, sd + SD_NCTEN * plane},
                {5, 6}, {sr + SR_VTRC * plane, sr + SR_VTNC * plane});
    __syncthreads();
    T.get(0, w, xc);
    T.get(1, w, xn);
    T.get(2, w, tc);
    T.get(4, w, tn);
    T.get(5, w, raw1);
    T.get(6, w, raw2);
    if (any & PF_QC) {
      int kt = 1;
      V hgt = V(0.0);
      for (int k = 0; k < g.nzc - 1; ++k) {  // uniform loop: every lane follows level k
        if (wv::at_level(xc, k) > kR2) kt = k + 1;
        hgt = hgt + wv::at_level(dz, k);
        if (hgt > V(500.0)) break;
      }
#pragma unroll
      for (int j = 0; j < kSedChunks; ++j) {
        const int k = j * 64 + wv::lane();
        v1[j] = (k < kt) ? double(raw1[j]) : 0.0;
        v2[j] = (k < kt) ? double(raw2[j]) : 0.0;
      }
      pr[4] += wave_sediment(g, xc, v1, tc, kR1, 1, kt, dt, true);
      wave_sediment(g, xn, v2, tn, kR2, 1, kt, dt, false);
      kst[8] = 1;
      kst[9] = kt;
    }
    T.put(2, w, tc);
    T.put(4, w, tn);
    __syncthreads();
    T.store(2, sd + SD_QCTEN * plane);
    T.store(4, sd + SD_NCTEN * plane);
  }

  if (!col_ok || wv::lane() != 0) return;
  for (int q = 0; q < 5; ++q) a.precip[size_t(q) * a.ncol + c] = pr[q];
  if (a.ksub)
    for (int q = 0; q < kSubRows; ++q) a.ksub[size_t(q) * a.ncol + c] = q < 10 ? kst[q] : 0;
  // column diagnostics (src/module_mp_tempo_main.F90:784-794, driver :452-457)
  const size_t d2 = size_t(c % a.ni) + size_t(c / a.ni) * a.d2ldj;
  const R rain = R(pr[0]), graupel = R(pr[1]), snow = R(pr[2]), ice = R(pr[3]), cloud = R(pr[4]);
  if (a.d2[D2_RAIN]) a.d2[D2_RAIN][d2] = rain;
  if (a.d2[D2_ICE]) a.d2[D2_ICE][d2] = ice;
  if (a.d2[D2_SNOW]) a.d2[D2_SNOW][d2] = snow;
  if (a.d2[D2_GRAUPEL]) a.d2[D2_GRAUPEL][d2] = graupel;
  if (a.d2[D2_FROZEN])
    a.d2[D2_FROZEN][d2] = live ? (ice + snow + graupel) / (ice + snow + graupel + rain + R(kR1)) : R(0);
  if (a.d2[D2_FRZRAIN]) {
    R frz = R(0);
    if (live && (sr[SR_TEMP * plane + scr.at(c, 0)] - R(kT0)) < R(-0.5)) frz = rain + cloud;
    a.d2[D2_FRZRAIN][d2] = frz;
  }
  if (a.d2[D2_HAIL_SFC]) a.d2[D2_HAIL_SFC][d2] = R(0);
  if (a.d2[D2_HAIL_COL]) a.d2[D2_HAIL_COL][d2] = R(0);
}

template <class R, bool kEveryPass = false>
__global__ void __launch_bounds__(64 * kSedCols) __attribute__((amdgpu_waves_per_eu(6))) k_phase_b_wave(StepArgs<R> a) {
  extern __shared__ double tile_mem[];  // sed_lds_bytes(nz)
  phase_b_group<R, kEveryPass>(a, batch_scr(a), int(blockIdx.x) * kSedCols, tile_mem);
}

// ------------------------------------------------------------- launch ----
// The four launches of one step.  ga carries the same arguments as a in the
// level-local kernels' working type: a itself, or its Sig<T> view for a run
// that records branch signatures (StepArgs<T> and StepArgs<Sig<T>> have the
// same layout).
template <class T, class RA>
hipError_t launch_step_impl(const StepArgs<T>& a, const StepArgs<RA>& ga, const LaunchCfg& lc, double* vtbuf,
                            hipStream_t s, hipEvent_t* ev) {
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_prepass<T>, prepass_grid(a), dim3(256), 0, s, a);
  if (ev) (void)hipEventRecord(ev[1], s);
  if (lc.waves_a && std::is_same<RA, T>::value)
    hipLaunchKernelGGL((k_phase_a<RA, WavesPerEu<RA>::a_alt>), level_grid(a, lc.block_a), dim3(unsigned(lc.block_a)),
                       0, s, ga);
  else
    hipLaunchKernelGGL(k_phase_a<RA>, level_grid(a, lc.block_a), dim3(unsigned(lc.block_a)), 0, s, ga);
  if (ev) (void)hipEventRecord(ev[2], s);
  if (lc.sed_wave == 2 && a.nz <= kSedLevels && !a.fl.semi_sedi)
    hipLaunchKernelGGL((k_phase_b_wave<T, true>), dim3(unsigned((a.ncol + kSedCols - 1) / kSedCols)),
                       dim3(64 * kSedCols), sed_lds_bytes(a.nz), s, a);
  else if (lc.sed_wave && a.nz <= kSedLevels && !a.fl.semi_sedi)
    hipLaunchKernelGGL(k_phase_b_wave<T>, dim3(unsigned((a.ncol + kSedCols - 1) / kSedCols)), dim3(64 * kSedCols),
                       sed_lds_bytes(a.nz), s, a);
  else
    hipLaunchKernelGGL(k_phase_b<T>, dim3(unsigned((a.ncol + 127) / 128)), dim3(128), 0, s, a, vtbuf);
  if (a.fl.refl_from_melting && a.fl.refl10cm && a.d3[D3_REFL]) {
    hipLaunchKernelGGL(k_melt_level<T>, dim3(unsigned((a.ncol + 255) / 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_melt_refl<T>, level_grid(a, 256), dim3(256), 0, s, a);
  }
  if (ev) (void)hipEventRecord(ev[3], s);
  hipLaunchKernelGGL(k_phase_c<RA>, level_grid(a, lc.block_c), dim3(unsigned(lc.block_c)), 0, s, ga);
  if (ev) (void)hipEventRecord(ev[4], s);
  return hipGetLastError();
}

template <class T>
StepArgs<Sig<T>> sig_args(const StepArgs<T>& a) {
  static_assert(sizeof(StepArgs<Sig<T>>) == sizeof(StepArgs<T>), "StepArgs layout");
  StepArgs<Sig<T>> g;
  std::memcpy(&g, &a, sizeof(a));
  return g;
}

}  // namespace tempo

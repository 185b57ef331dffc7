// Adjoint of the column step.
//
// The step maps the inputs x of every level through three parts:
//   A (level-local physics)   z_k = A(x_k)
//   B (column sedimentation)  u = T + S(z),  T = the pre-sedimentation
//                             tendencies inside z, S the flux-form update
//                             (src/module_mp_tempo_main.F90:590-727)
//   C (level-local update)    y_k = C(z_k, u_k)
// so the vector-Jacobian product splits into
//   xbar_k = (dy_k/dx_k |local)^T ybar_k           [through A, T and C]
//          + (dz_k/dx_k)^T S^T ubar,   ubar_k = (dy_k/du_k)^T ybar_k
// The local Jacobian columns come from tangent passes of the dual-number
// step (one per input field, S's dependence frozen out), dy/du from tangent
// passes seeded on u, and S^T is written out below: the sedimentation
// substeps run backwards, each recomputed from the column's initial masses
// instead of being stored.  Branches (fills, substep counts, ktop, clamps,
// the surface-precipitation threshold) are frozen at their forward values.
#pragma once

#include "tempo_kernels.hpp"

namespace tempo {

// sedimentation inputs whose derivatives couple levels: masses and numbers,
// raw fall speeds, density
enum AdjZ {
  AZ_RR, AZ_NR, AZ_RG, AZ_NG, AZ_RB, AZ_RS, AZ_RI, AZ_NI, AZ_RC, AZ_NC,
  AZ_VTRR, AZ_VTNR, AZ_VTRG, AZ_VTNG, AZ_VTS, AZ_VTRI, AZ_VTNI, AZ_VTRC, AZ_VTNC, AZ_RHO, AZ_N
};
// work planes of the sedimentation adjoint (FP64)
enum AdjW { AW_XR, AW_XBAR, AW_FBAR, AW_VB1, AW_VB2, AW_VR, AW_RRPOST, AW_VRBAR, AW_RRBAR, AW_RHOBAR, AW_N };
// output fields whose cotangents enter the local passes (order of the
// cotangent array handed to tempo_hip_adjoint)
enum AdjY { AY_T, AY_QV, AY_QC, AY_QR, AY_QI, AY_QS, AY_QG, AY_NI, AY_NR, AY_NC, AY_NWFA, AY_NIFA, AY_NG, AY_QB,
            AY_REFL, AY_N };
// differentiable inputs (every prognostic field, so that steps chain)
enum AdjX { AX_T, AX_QV, AX_QC, AX_QR, AX_QI, AX_QS, AX_QG, AX_NC, AX_NR, AX_NI, AX_NWFA, AX_NIFA, AX_NG, AX_QB, AX_N };

__host__ __device__ constexpr int adj_x_field(int i) {
  return i == AX_T ? F_T : i == AX_QV ? F_QV : i == AX_QC ? F_QC : i == AX_QR ? F_QR : i == AX_QI ? F_QI
       : i == AX_QS ? F_QS : i == AX_QG ? F_QG : i == AX_NC ? F_NC : i == AX_NR ? F_NR : i == AX_NI ? F_NI
       : i == AX_NWFA ? F_NWFA : i == AX_NIFA ? F_NIFA : i == AX_NG ? F_NG : F_QB;
}
__host__ __device__ constexpr int adj_y_field(int i) {
  return i == AY_T ? F_T : i == AY_QV ? F_QV : i == AY_QC ? F_QC : i == AY_QR ? F_QR : i == AY_QI ? F_QI
       : i == AY_QS ? F_QS : i == AY_QG ? F_QG : i == AY_NI ? F_NI : i == AY_NR ? F_NR : i == AY_NC ? F_NC
       : i == AY_NWFA ? F_NWFA : i == AY_NIFA ? F_NIFA : i == AY_NG ? F_NG : F_QB;
}

struct AdjArgs {
  const double* ybar;   // AY_N planes (compact c + k*ncol), absent outputs all zero
  int* lvl_on;          // nz: 1 where some output cotangent of the level is non-zero
  int* pt_loc;          // compact plane: 1 where a local pass's phase A runs (see k_adj_point_flags)
  const double* pbar;   // 4 x ncol: rain, graupel, snow, ice precipitation
  double* xbar;         // AX_N planes (compact)
  double* ubar;         // 10 planes (compact): tendency cotangents, SD_QRTEN order
  double* zbar;         // AZ_N planes (compact)
  double* work;         // AW_N scratch planes (tiled like the step scratch)
};

// ---- tangent seeding and collection (one lane per (column, level)) ------
// seed the tangent of one input field, clearing the previous pass's seed
// (prev < 0: none)
template <class D>
__global__ void __launch_bounds__(256) k_adj_seed_input(StepArgs<D> a, int field, int prev) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (c >= a.ncol) return;
  const size_t o = TEMPO_IDX_STATE(a, c, k);
  if (prev >= 0 && a.f[prev]) a.f[prev][o].d = 0;
  if (a.f[field]) a.f[field][o].d = 1;
}

// the values of the ten post-sedimentation tendency planes after the
// tangent-free forward's sedimentation, plane m at post[m * plane + s]
template <class D>
__global__ void __launch_bounds__(256) k_adj_keep_post(StepArgs<D> a, double* __restrict__ post) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (c >= a.ncol) return;
  const size_t s = TEMPO_IDX_SCR(a, c, k);
#pragma unroll
  for (int m = 0; m < 10; ++m) post[size_t(m) * a.plane + s] = a.sd[size_t(SD_QRTEN + m) * a.plane + s].v;
}

// after phase A of a local pass: the S^T part of xbar_i, (dz/dx_i) . zbar,
// from the tangents of the sedimentation inputs (masses, numbers, raw fall
// speeds, density) and the cotangents zbar that S^T has already produced.
// The final update then reads the forward's post-sedimentation tendency
// values (k_adj_keep_post, a.post) with the local tangents phase A left in
// the scratch (S frozen out); it reads neither the sedimenting masses nor
// the fall speeds, so their tangents need not be cleared.
template <class D>
__global__ void __launch_bounds__(256) k_adj_capture(StepArgs<D> a, AdjArgs g, double* __restrict__ xb) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (c >= a.ncol) return;
  const size_t s = TEMPO_IDX_SCR(a, c, k), e = size_t(c) + size_t(k) * a.ncol, n3 = size_t(a.ncol) * a.nz;
  if (!g.pt_loc[e]) {  // skipped by phase A: zero tangents against zero zbar, no final update
    xb[e] = 0.0;
    return;
  }
  const int sdf[10] = {SD_RR, SD_NR, SD_RG, SD_NG, SD_RB, SD_RS, SD_RI, SD_NI, SD_RC, SD_NC};
  const int srf[9] = {SR_VTRR, SR_VTNR, SR_VTRG, SR_VTNG, SR_VTS, SR_VTRI, SR_VTNI, SR_VTRC, SR_VTNC};
  // a tangent is read only where its cotangent is non-zero (a finite
  // tangent times zero adds nothing)
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 10; ++q) {
    const double z = g.zbar[size_t(AZ_RR + q) * n3 + e];
    if (z != 0.0) acc += double(a.sd[size_t(sdf[q]) * a.plane + s].d) * z;
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const double z = g.zbar[size_t(AZ_VTRR + q) * n3 + e];
    if (z != 0.0) acc += double(a.sr[size_t(srf[q]) * a.plane + s].d) * z;
  }
  const double zr = g.zbar[size_t(AZ_RHO) * n3 + e];
  if (zr != 0.0) acc += double(a.sr[size_t(SR_RHO) * a.plane + s].d) * zr;
  xb[e] = acc;
}

// seed the tangent of one post-sedimentation tendency plane (u pass),
// clearing the previous pass's seed (m_prev < 0: none)
template <class D>
__global__ void __launch_bounds__(256) k_adj_seed_ten(StepArgs<D> a, int m, int m_prev) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (c >= a.ncol) return;
  const size_t s = TEMPO_IDX_SCR(a, c, k);
  if (m_prev >= 0) a.sd[size_t(SD_QRTEN + m_prev) * a.plane + s].d = 0;
  a.sd[size_t(SD_QRTEN + m) * a.plane + s].d = 1;
}

// out[0, n) = 0 (a kernel rather than a memset, so that the sequence may be
// captured in a graph)
template <class T>
__global__ void __launch_bounds__(256) k_adj_zero(T* __restrict__ out, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = T(0);
}

// the levels whose output cotangents are not all zero (g.lvl_on, cleared
// by k_adj_level_zero): the final update's passes contribute nothing at
// the others
template <class D>
__global__ void __launch_bounds__(256) k_adj_level_zero(StepArgs<D> a, AdjArgs g) {
  const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (k < a.nz) g.lvl_on[k] = 0;
}
template <class D>
__global__ void __launch_bounds__(256) k_adj_level_flags(StepArgs<D> a, AdjArgs g) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  bool nz = false;
  if (c < a.ncol) {
    const size_t e = size_t(c) + size_t(k) * a.ncol, n3 = size_t(a.ncol) * a.nz;
    for (int q = 0; q < AY_N; ++q) nz = nz || g.ybar[size_t(q) * n3 + e] != 0.0;
  }
  if (__ballot(nz) != 0ull && (threadIdx.x & 63u) == 0) atomicOr(&g.lvl_on[k], 1);
}
// after S^T: the points where the local passes' phase A runs (g.pt_loc).
// At a point whose zbar are all zero, on a level whose final update is
// skipped, the capture adds tangent * 0 and nothing else reads phase A's
// results, so every local pass skips it; its tangents stay those of the
// tangent-free forward (zero).
template <class D>
__global__ void __launch_bounds__(256) k_adj_point_flags(StepArgs<D> a, AdjArgs g) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (c >= a.ncol) return;
  const size_t e = size_t(c) + size_t(k) * a.ncol, n3 = size_t(a.ncol) * a.nz;
  bool on = g.lvl_on[k] != 0;
  for (int q = 0; q < AZ_N && !on; ++q) on = g.zbar[size_t(q) * n3 + e] != 0.0;
  g.pt_loc[e] = on ? 1 : 0;
}

// the seed of the next pass, moved by the contraction that ends a pass
// (the seeding runs on the same grid): kind 1 an input field's tangent
// (a.f), kind 2 a post-sedimentation tendency plane's (scratch); the
// field or plane `clear` loses its seed, `set` gets one (< 0: none)
struct AdjSeed {
  int kind, set, clear;
};

// out[e] (+)= sum over outputs of ybar * tangent (0 at a level without
// output cotangents, whose final update a.lvl_on skipped); then the next
// pass's seed at this point.  A point the step leaves as it is (an
// early-exit column, a level at or above the column's nlev) has no final
// update writing its new state: its output tangents are its input tangents,
// and its reflectivity is no function of the inputs (a constant in an
// early-exit column, left as it is above nlev).
template <class D>
__global__ void __launch_bounds__(256) k_adj_contract(StepArgs<D> a, AdjArgs g, double* out, int accumulate,
                                                      AdjSeed nx) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (c >= a.ncol) return;
  const size_t o = TEMPO_IDX_STATE(a, c, k), e = size_t(c) + size_t(k) * a.ncol, n3 = size_t(a.ncol) * a.nz;
  const bool kept = (a.cf[c] & CF_EXIT) || (a.nlev && k >= a.nlev[c]);
  if (a.lvl_on && !a.lvl_on[k]) {
    if (!accumulate) out[e] = 0.0;
  } else {
    double acc = 0.0;  // tangents read only where their cotangent is non-zero
    for (int q = 0; q < AY_REFL; ++q) {
      const D* f = kept ? a.f[adj_y_field(q)] : a.fo[adj_y_field(q)];
      if (!f) continue;
      const double y = g.ybar[size_t(q) * n3 + e];
      if (y != 0.0) acc += y * double(f[o].d);
    }
    if (a.d3[D3_REFL] && !kept) {
      const double y = g.ybar[size_t(AY_REFL) * n3 + e];
      if (y != 0.0) acc += y * double(a.d3[D3_REFL][o].d);
    }
    out[e] = accumulate ? out[e] + acc : acc;
  }
  if (nx.kind == 1) {
    if (nx.clear >= 0 && a.f[nx.clear]) a.f[nx.clear][o].d = 0;
    if (nx.set >= 0 && a.f[nx.set]) a.f[nx.set][o].d = 1;
  } else if (nx.kind == 2) {
    const size_t s = TEMPO_IDX_SCR(a, c, k);
    if (nx.clear >= 0) a.sd[size_t(SD_QRTEN + nx.clear) * a.plane + s].d = 0;
    if (nx.set >= 0) a.sd[size_t(SD_QRTEN + nx.set) * a.plane + s].d = 1;
  }
}


// ---- S^T: sedimentation adjoint, one lane per column --------------------
struct SedAdjCol {
  int c, nzc;
  size_t plane;
  double dt;
};

// One field's substeps backwards.  xr0: initial masses (read), vt: filled
// speeds; tenbar: cotangent of this field's tendency (all substeps);
// pbar: surface-precipitation cotangent (0: not accumulated);
// xbar (AW_XBAR plane) holds the cotangent of the final masses on entry and
// of the initial masses on exit; vbar (a work plane) accumulates the
// filled-speed cotangents, rhobar (a work plane) the density cotangents.
template <class D, class FXR, class FVT, class FODZ, class FRHO>
__device__ void sed_adjoint(const StepArgs<D>& a, const SedAdjCol& q, const AdjArgs& g, FXR xr0, FVT vt, FODZ odz,
                            FRHO orho, const double* tenbar, double limit, int nsub, int steps, int ktop,
                            double pbar, double* vbar, double* rhobar) {
  // every operation rounded on its own, as in the forward sedimentation it
  // recomputes (so that the recomputed masses and clamps are the forward's)
#pragma clang fp contract(off)
  double* xr = g.work + AW_XR * q.plane;
  double* xb = g.work + AW_XBAR * q.plane;
  double* fb = g.work + AW_FBAR * q.plane;
  const double ost = 1.0 / double(steps), dt = q.dt;
  const int top = q.nzc - 1;
  auto updated = [&](int k) { return k == top || k <= ktop - 1; };
  const size_t n3 = size_t(a.ncol) * a.nz;
  for (int n = nsub - 1; n >= 0; --n) {
    // masses at the start of substep n, recomputed from the initial ones
    for (int k = 0; k < q.nzc; ++k) xr[TEMPO_IDX_SCR(a, q.c, k)] = xr0(k);
    for (int r = 0; r < n; ++r) {
      double above = 0.0;
      for (int k = top; k >= 0; --k) {
        const size_t s = TEMPO_IDX_SCR(a, q.c, k);
        const double sed = vt(k) * xr[s];
        if (updated(k)) {
          const double flux = (k == top) ? -sed : (above - sed);
          const double v = xr[s] + flux * odz(k) * dt * ost;
          xr[s] = v > limit ? v : limit;
        }
        above = sed;
      }
    }
    // cotangents of the fluxes of substep n
    double above = 0.0, x0post = 0.0;
    for (int k = top; k >= 0; --k) {
      const size_t s = TEMPO_IDX_SCR(a, q.c, k);
      const double sed = vt(k) * xr[s];
      double fbar = 0.0;
      if (updated(k)) {
        const double flux = (k == top) ? -sed : (above - sed);
        const double v = xr[s] + flux * odz(k) * dt * ost;
        const double vb = v > limit ? xb[s] : 0.0;
        const double tb = tenbar ? tenbar[size_t(q.c) + size_t(k) * a.ncol] : 0.0;
        fbar = vb * odz(k) * dt * ost + tb * odz(k) * ost * orho(k);
        if (tb != 0.0) rhobar[s] += -orho(k) * orho(k) * tb * flux * odz(k) * ost;
        xb[s] = vb;  // cotangent of the pre-substep mass through v (identity part)
        if (k == 0) x0post = v > limit ? v : limit;
      } else if (k == 0) {
        x0post = xr[s];
      }
      fb[s] = fbar;
      above = sed;
    }
    // sed_k feeds flux_k (-) and flux_{k-1} (+ when k-1 is an interior update)
    for (int k = top; k >= 0; --k) {
      const size_t s = TEMPO_IDX_SCR(a, q.c, k);
      double sb = updated(k) ? -fb[s] : 0.0;
      if (k >= 1 && updated(k - 1) && k - 1 != top) sb += fb[TEMPO_IDX_SCR(a, q.c, k - 1)];
      if (k == 0 && pbar != 0.0 && x0post > kLowLimitPrecip) sb += pbar * dt * ost;
      xb[s] += sb * vt(k);
      vbar[s] += sb * xr[s];
    }
  }
  (void)n3;
}

template <class D>
__global__ void __launch_bounds__(64) k_adj_sedimentation(StepArgs<D> a, AdjArgs g) {
  // every operation rounded on its own, as in the forward sedimentation it
  // recomputes (so that the recomputed masses and clamps are the forward's)
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncol) return;
  const int nzc = a.nlev ? min(max(a.nlev[c], 1), a.nz) : a.nz;
  const size_t plane = a.plane, n3 = size_t(a.ncol) * a.nz;
  auto zb = [&](int z, int k) -> double& { return g.zbar[size_t(z) * n3 + size_t(c) + size_t(k) * a.ncol]; };
  for (int z = 0; z < AZ_N; ++z)
    for (int k = 0; k < a.nz; ++k) zb(z, k) = 0.0;
  if (a.cf[c] & CF_EXIT) return;
  unsigned any = 0;
  for (int k = 0; k < nzc; ++k) any |= a.pf[TEMPO_IDX_SCR(a, c, k)];
  const SedAdjCol q{c, nzc, plane, a.dt};
  const double dt = a.dt;
  auto S = [&](int k) { return TEMPO_IDX_SCR(a, c, k); };
  auto mass = [&](int f, int k) { return double(a.sd[size_t(f) * plane + S(k)].v); };
  auto raw = [&](int f, int k) { return double(a.sr[size_t(f) * plane + S(k)].v); };
  auto odz = [&](int k) { return 1.0 / double(a.f[F_DZ][TEMPO_IDX_STATE(a, c, k)].v); };
  auto orho = [&](int k) { return 1.0 / double(a.sr[size_t(SR_RHO) * plane + S(k)].v); };
  const double* ub = g.ubar;
  auto ubp = [&](int m) { return ub + size_t(m) * n3; };
  double* w = g.work;
  double* vb1 = w + AW_VB1 * plane;
  double* vb2 = w + AW_VB2 * plane;
  double* xb = w + AW_XBAR * plane;
  double* vr = w + AW_VR * plane;
  double* rrpost = w + AW_RRPOST * plane;
  double* vrbar = w + AW_VRBAR * plane;
  double* rrbar = w + AW_RRBAR * plane;
  double* rhob = w + AW_RHOBAR * plane;
  const double pb_rain = g.pbar[c], pb_graupel = g.pbar[a.ncol + c], pb_snow = g.pbar[2 * size_t(a.ncol) + c],
               pb_ice = g.pbar[3 * size_t(a.ncol) + c];
  using R = D;
  auto zero = [&](double* p) {
    for (int k = 0; k < nzc; ++k) p[S(k)] = 0.0;
  };
  // forward fill (:1768-1789): src(k) = lowest level >= k with mass > r1
  auto fill_src = [&](int fm, int k) {
    for (int kk = k; kk < nzc; ++kk)
      if (mass(fm, kk) > kR1) return kk;
    return -1;
  };
  auto cfl = [&](auto vmaxf, int& steps, int& ktop) {
    steps = 1;
    ktop = 1;
    for (int k = nzc - 1; k >= 0; --k) {
      const R vmax = vmaxf(k);
      if (vmax > R(1.e-3)) {
        ktop = max(ktop, k + 1);
        const R dz_by_vt = a.f[F_DZ][TEMPO_IDX_STATE(a, c, k)] / vmax;
        steps = max(steps, int(R(dt) / dz_by_vt + R(1.0)));
      }
    }
    if (ktop == nzc) ktop = nzc - 1;
  };

  // rain forward values needed by the snow adjoint: filled speed and mass after rain sedimentation
  int r_steps = 1, r_ktop = 1;
  if (any & PF_QR) {
    auto vfill = [&](int f, int k) { const int s0 = fill_src(SD_RR, k); return s0 < 0 ? 0.0 : raw(f, s0); };
    cfl([&](int k) { return mx(R(vfill(SR_VTRR, k)), R(vfill(SR_VTNR, k))); }, r_steps, r_ktop);
    for (int k = 0; k < nzc; ++k) {
      vr[S(k)] = vfill(SR_VTRR, k);
      rrpost[S(k)] = mass(SD_RR, k);
    }
    const double ost = 1.0 / double(r_steps);
    for (int n = 0; n < r_steps; ++n) {
      double above = 0.0;
      for (int k = nzc - 1; k >= 0; --k) {
        const size_t s = S(k);
        const double sed = vr[s] * rrpost[s];
        if (k == nzc - 1 || k <= r_ktop - 1) {
          const double flux = (k == nzc - 1) ? -sed : (above - sed);
          const double v = rrpost[s] + flux * odz(k) * dt * ost;
          rrpost[s] = v > kR1 ? v : kR1;
        }
        above = sed;
      }
    }
  } else {
    for (int k = 0; k < nzc; ++k) {
      vr[S(k)] = 0.0;
      rrpost[S(k)] = mass(SD_RR, k);
    }
  }
  zero(vrbar);
  zero(rrbar);
  zero(rhob);

  // ---- cloud ----
  if (any & PF_QC) {
    int kt = 1;
    R hgt = R(0.0);
    for (int k = 0; k < nzc - 1; ++k) {
      if (mass(SD_RC, k) > kR2) kt = k + 1;
      hgt = hgt + a.f[F_DZ][TEMPO_IDX_STATE(a, c, k)];
      if (hgt > R(500.0)) break;
    }
    for (int f2 = 0; f2 < 2; ++f2) {
      const int fm = f2 ? SD_NC : SD_RC, fr = f2 ? SR_VTNC : SR_VTRC;
      zero(xb);
      zero(vb1);
      sed_adjoint(a, q, g, [&](int k) { return mass(fm, k); },
                  [&](int k) { return k < kt ? raw(fr, k) : 0.0; }, odz, orho, ubp(f2 ? 9 : 8), f2 ? kR2 : kR1, 1, 1,
                  kt, f2 ? 0.0 : 0.0, vb1, rhob);
      for (int k = 0; k < nzc; ++k) {
        zb(f2 ? AZ_NC : AZ_RC, k) += xb[S(k)];
        if (k < kt) zb(f2 ? AZ_VTNC : AZ_VTRC, k) += vb1[S(k)];
      }
    }
  }
  // ---- ice: one application with steps = substeps ----
  if (any & PF_QI) {
    int st, kt;
    auto vfill = [&](int f, int k) { const int s0 = fill_src(SD_RI, k); return s0 < 0 ? 0.0 : raw(f, s0); };
    cfl([&](int k) { return R(vfill(SR_VTRI, k)); }, st, kt);
    for (int f2 = 0; f2 < 2; ++f2) {
      const int fm = f2 ? SD_NI : SD_RI, fr = f2 ? SR_VTNI : SR_VTRI;
      zero(xb);
      zero(vb1);
      sed_adjoint(a, q, g, [&](int k) { return mass(fm, k); }, [&](int k) { return vfill(fr, k); }, odz, orho,
                  ubp(f2 ? 7 : 6), f2 ? kR2 : kR1, 1, st, kt, f2 ? 0.0 : pb_ice, vb1, rhob);
      for (int k = 0; k < nzc; ++k) {
        zb(f2 ? AZ_NI : AZ_RI, k) += xb[S(k)];
        const int s0 = fill_src(SD_RI, k);
        if (s0 >= 0) zb(f2 ? AZ_VTNI : AZ_VTRI, s0) += vb1[S(k)];
      }
    }
  }
  // ---- snow: speed mixed with rain where snow melts ----
  if (any & PF_QS) {
    auto vraw = [&](int k) {
      R v = R(raw(SR_VTS, k));
      if (a.pf[S(k)] & PF_SNOWMIX) {
        const double rs = mass(SD_RS, k);
        const R sr_ = R(rs / (rs + rrpost[S(k)]));
        v = v * sr_ + (R(1.0) - sr_) * R(vr[S(k)]);
      }
      return double(value(v));
    };
    auto vfill = [&](int k) { const int s0 = fill_src(SD_RS, k); return s0 < 0 ? 0.0 : vraw(s0); };
    int st, kt;
    cfl([&](int k) { return R(vfill(k)); }, st, kt);
    zero(xb);
    zero(vb1);
    sed_adjoint(a, q, g, [&](int k) { return mass(SD_RS, k); }, vfill, odz, orho, ubp(5), kR1, st, st, kt, pb_snow,
                vb1, rhob);
    zero(vb2);  // cotangent of the mixed raw speed per source level
    for (int k = 0; k < nzc; ++k) {
      zb(AZ_RS, k) += xb[S(k)];
      const int s0 = fill_src(SD_RS, k);
      if (s0 >= 0) vb2[S(s0)] += vb1[S(k)];
    }
    for (int k = 0; k < nzc; ++k) {
      const double vbk = vb2[S(k)];
      if (vbk == 0.0) continue;
      if (a.pf[S(k)] & PF_SNOWMIX) {
        const double rs = mass(SD_RS, k), rr = rrpost[S(k)], den = rs + rr;
        const double sr_ = rs / den, rawv = raw(SR_VTS, k), vrv = vr[S(k)];
        zb(AZ_VTS, k) += vbk * sr_;
        vrbar[S(k)] += vbk * (1.0 - sr_);
        const double srb = vbk * (rawv - vrv);
        zb(AZ_RS, k) += srb * rr / (den * den);
        rrbar[S(k)] += -srb * rs / (den * den);
      } else {
        zb(AZ_VTS, k) += vbk;
      }
    }
  }
  // ---- graupel ----
  if (any & PF_QG) {
    int st, kt;
    auto vfill = [&](int f, int k) { const int s0 = fill_src(SD_RG, k); return s0 < 0 ? 0.0 : raw(f, s0); };
    cfl([&](int k) { return R(vfill(SR_VTRG, k)); }, st, kt);
    const double rb_lim = kM3ToL * kR1 / 800.0;
    // the three fields are independent given the speeds
    const int fms[3] = {SD_RG, SD_NG, SD_RB}, frs[3] = {SR_VTRG, SR_VTNG, SR_VTRG}, azm[3] = {AZ_RG, AZ_NG, AZ_RB},
              azv[3] = {AZ_VTRG, AZ_VTNG, AZ_VTRG}, ten[3] = {2, 3, 4};
    const double lims[3] = {kR1, kR2, rb_lim};
    for (int f3 = 0; f3 < 3; ++f3) {
      zero(xb);
      zero(vb1);
      sed_adjoint(a, q, g, [&](int k) { return mass(fms[f3], k); }, [&](int k) { return vfill(frs[f3], k); }, odz,
                  orho, ubp(ten[f3]), lims[f3], st, st, kt, f3 == 0 ? pb_graupel : 0.0, vb1, rhob);
      for (int k = 0; k < nzc; ++k) {
        zb(azm[f3], k) += xb[S(k)];
        const int s0 = fill_src(SD_RG, k);
        if (s0 >= 0) zb(azv[f3], s0) += vb1[S(k)];
      }
    }
  }
  // ---- rain: the final masses and filled speed carry the snow cotangents ----
  if (any & PF_QR) {
    auto vfill = [&](int f, int k) { const int s0 = fill_src(SD_RR, k); return s0 < 0 ? 0.0 : raw(f, s0); };
    for (int f2 = 0; f2 < 2; ++f2) {
      const int fm = f2 ? SD_NR : SD_RR, fr = f2 ? SR_VTNR : SR_VTRR;
      for (int k = 0; k < nzc; ++k) {
        xb[S(k)] = f2 ? 0.0 : rrbar[S(k)];
        vb1[S(k)] = f2 ? 0.0 : vrbar[S(k)];
      }
      sed_adjoint(a, q, g, [&](int k) { return mass(fm, k); }, [&](int k) { return vfill(fr, k); }, odz, orho,
                  ubp(f2 ? 1 : 0), f2 ? kR2 : kR1, r_steps, r_steps, r_ktop, f2 ? 0.0 : pb_rain, vb1, rhob);
      for (int k = 0; k < nzc; ++k) {
        zb(f2 ? AZ_NR : AZ_RR, k) += xb[S(k)];
        const int s0 = fill_src(SD_RR, k);
        if (s0 >= 0) zb(f2 ? AZ_VTNR : AZ_VTRR, s0) += vb1[S(k)];
      }
    }
  } else {
    for (int k = 0; k < nzc; ++k) zb(AZ_RR, k) += rrbar[S(k)];
  }
  for (int k = 0; k < nzc; ++k) zb(AZ_RHO, k) += rhob[S(k)];
}

}  // namespace tempo

namespace tempo {

// caller array (layout g, precision T, may be NULL = zeros) <-> compact FP64 plane
template <class T>
__global__ void __launch_bounds__(256) k_to_compact(const T* __restrict__ src, DualGeom g, double* __restrict__ dst) {
  const size_t ncol = size_t(g.ni) * g.nj;
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ncol * size_t(g.nk)) return;
  const size_t c = e % ncol, k = e / ncol;
  dst[e] = src ? double(src[(c % g.ni) + k * g.ldk + (c / g.ni) * g.ldj]) : 0.0;
}

template <class T>
__global__ void __launch_bounds__(256) k_from_compact(const double* __restrict__ src, DualGeom g, T* __restrict__ dst) {
  const size_t ncol = size_t(g.ni) * g.nj;
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ncol * size_t(g.nk)) return;
  const size_t c = e % ncol, k = e / ncol;
  dst[(c % g.ni) + k * g.ldk + (c / g.ni) * g.ldj] = T(src[e]);
}


// several arrays of one layout in one launch (blockIdx.y: the array)
constexpr int kBatchMax = 24;
template <class S, class D>
struct PlaneBatch {
  const S* src[kBatchMax];
  D* dst[kBatchMax];
};
template <class T>
__global__ void __launch_bounds__(256) k_to_compact_n(PlaneBatch<T, double> b, DualGeom g) {
  const size_t ncol = size_t(g.ni) * g.nj;
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ncol * size_t(g.nk)) return;
  const T* src = b.src[blockIdx.y];
  const size_t c = e % ncol, k = e / ncol;
  b.dst[blockIdx.y][e] = src ? double(src[(c % g.ni) + k * g.ldk + (c / g.ni) * g.ldj]) : 0.0;
}
template <class T>
__global__ void __launch_bounds__(256) k_from_compact_n(PlaneBatch<double, T> b, DualGeom g) {
  const size_t ncol = size_t(g.ni) * g.nj;
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ncol * size_t(g.nk)) return;
  const size_t c = e % ncol, k = e / ncol;
  b.dst[blockIdx.y][(c % g.ni) + k * g.ldk + (c / g.ni) * g.ldj] = T(b.src[blockIdx.y][e]);
}
// values into compact duals with zero tangents
template <class T>
__global__ void __launch_bounds__(256) k_dual_pack_n(PlaneBatch<T, Dual<T>> b, DualGeom g) {
  const size_t ncol = size_t(g.ni) * g.nj;
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ncol * size_t(g.nk)) return;
  const size_t c = e % ncol, k = e / ncol;
  b.dst[blockIdx.y][e] = Dual<T>(b.src[blockIdx.y][(c % g.ni) + k * g.ldk + (c / g.ni) * g.ldj], T(0));
}
template <class K, class B>
hipError_t launch_batch(K kernel, const B& b, int n, const DualGeom& g, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (n > kBatchMax) return hipErrorInvalidValue;
  const size_t m = size_t(g.ni) * g.nj * g.nk;
  hipLaunchKernelGGL(kernel, dim3(unsigned((m + 255) / 256), unsigned(n)), dim3(256), 0, s, b, g);
  return hipGetLastError();
}
template <class T>
hipError_t launch_to_compact_n(const PlaneBatch<T, double>& b, int n, const DualGeom& g, hipStream_t s) {
  return launch_batch(k_to_compact_n<T>, b, n, g, s);
}
template <class T>
hipError_t launch_from_compact_n(const PlaneBatch<double, T>& b, int n, const DualGeom& g, hipStream_t s) {
  return launch_batch(k_from_compact_n<T>, b, n, g, s);
}
template <class T>
hipError_t launch_dual_pack_n(const PlaneBatch<T, Dual<T>>& b, int n, const DualGeom& g, hipStream_t s) {
  return launch_batch(k_dual_pack_n<T>, b, n, g, s);
}

template <class T>
hipError_t launch_to_compact(const T* src, const DualGeom& g, double* dst, hipStream_t s) {
  const size_t n = size_t(g.ni) * g.nj * g.nk;
  hipLaunchKernelGGL(k_to_compact<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, src, g, dst);
  return hipGetLastError();
}

template <class T>
hipError_t launch_from_compact(const double* src, const DualGeom& g, T* dst, hipStream_t s) {
  const size_t n = size_t(g.ni) * g.nj * g.nk;
  hipLaunchKernelGGL(k_from_compact<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, src, g, dst);
  return hipGetLastError();
}

}  // namespace tempo

// S^T with one wavefront per column, levels across lanes (nz <= 128): the
// serial column adjoint of tempo_adjoint.hpp recast on the wavefront
// primitives of the sedimentation kernel.  Every substep's adjoint is a
// two-point stencil in the level index (the flux of level k reads the
// sedimentation rate of k+1), so all levels update at once: the rate
// cotangent of level k collects its own flux cotangent and that of level
// k-1 by a one-lane shift.  Fill sources come from a ballot of the levels
// holding mass; the cotangents of the filled speeds are gathered back to
// their source levels through a per-wavefront LDS row, summed in level
// order.
#pragma once

#include "tempo_adjoint.hpp"
#include "tempo_sedi.hpp"

namespace tempo {

namespace wv {
// value of x at level k-1 for every lane's level k in chunk j (0 below level 0)
template <class T>
__device__ __forceinline__ T below(const T (&x)[kSedChunks], int j) {
  const T dn = __shfl_up(x[j], 1, 64);
  T prv = T(0);
  if (j > 0) prv = __shfl(x[j > 0 ? j - 1 : 0], 63, 64);
  return lane() == 0 ? prv : dn;
}
}  // namespace wv

struct WaveCol {
  int nzc, top;
  double dt;
  double odz[kSedChunks], orho[kSedChunks];
};

// fill sources: lowest level >= k whose mass exceeds r1 (-1: none)
__device__ __forceinline__ void wave_fill_src(const double (&mass)[kSedChunks], int nzc, int (&src)[kSedChunks],
                                              unsigned long long (&m)[kSedChunks]) {
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) m[j] = __ballot(j * 64 + wv::lane() < nzc && mass[j] > kR1);
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) {
    const int l = wv::lane();
    int s = -1;
    const unsigned long long mm = m[j] & (~0ull << l);
    if (mm) {
      s = j * 64 + __ffsll((long long)mm) - 1;
    } else {
#pragma unroll
      for (int j2 = j + 1; j2 < kSedChunks; ++j2)
        if (s < 0 && m[j2]) s = j2 * 64 + __ffsll((long long)m[j2]) - 1;
    }
    src[j] = s;
  }
}

// out(k) = raw(src(k)) or 0
__device__ __forceinline__ void wave_gather(const double (&raw)[kSedChunks], const int (&src)[kSedChunks],
                                            double (&out)[kSedChunks]) {
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) {
    const double r = wv::at_level(raw, src[j] < 0 ? 0 : src[j]);
    out[j] = src[j] < 0 ? 0.0 : r;
  }
}

// acc(k') += sum over levels k with src(k) = k' of vb(k), summed in level order
__device__ __forceinline__ void wave_scatter(const double (&vb)[kSedChunks], const int (&src)[kSedChunks],
                                             const unsigned long long (&m)[kSedChunks], double* row, int nzc,
                                             double (&acc)[kSedChunks]) {
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) row[j * 64 + wv::lane()] = vb[j];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) {
    const int k = j * 64 + wv::lane();
    if (k >= nzc || !((m[j] >> wv::lane()) & 1ull)) continue;
    // previous source level below k (its segment ends there)
    int lo = -1;
    const unsigned long long below_mask = wv::lane() ? (m[j] & ((1ull << wv::lane()) - 1ull)) : 0ull;
    if (below_mask) {
      lo = j * 64 + 63 - __clzll((long long)below_mask);
    } else {
#pragma unroll
      for (int j2 = j - 1; j2 >= 0; --j2)
        if (lo < 0 && m[j2]) lo = j2 * 64 + 63 - __clzll((long long)m[j2]);
    }
    double s = 0.0;
    for (int kk = lo + 1; kk <= k; ++kk) s += row[kk];
    acc[j] += s;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// one field's substeps backwards (see sed_adjoint in tempo_adjoint.hpp)
__device__ __forceinline__ void wave_sed_adjoint(const WaveCol& q, const double (&xr0)[kSedChunks],
                                                 const double (&vt)[kSedChunks], const double (&tb)[kSedChunks],
                                                 double limit, int nsub, int steps, int ktop, double pbar,
                                                 double (&xb)[kSedChunks], double (&vb)[kSedChunks],
                                                 double (&rhob)[kSedChunks]) {
  // every operation rounded on its own, as in the forward sedimentation it
  // recomputes (so that the recomputed masses and clamps are the forward's)
#pragma clang fp contract(off)
  const double ost = 1.0 / double(steps), dt = q.dt;
  auto updated = [&](int k) { return k < q.nzc && (k == q.top || k <= ktop - 1); };
  for (int n = nsub - 1; n >= 0; --n) {
    double xr[kSedChunks];
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) xr[j] = xr0[j];
    for (int r = 0; r < n; ++r) {
      double sed[kSedChunks];
#pragma unroll
      for (int j = 0; j < kSedChunks; ++j) sed[j] = vt[j] * xr[j];
#pragma unroll
      for (int j = 0; j < kSedChunks; ++j) {
        const double up = wv::above(sed, j);
        const int k = j * 64 + wv::lane();
        if (updated(k)) {
          const double flux = (k == q.top) ? -sed[j] : (up - sed[j]);
          const double v = xr[j] + flux * q.odz[j] * dt * ost;
          xr[j] = v > limit ? v : limit;
        }
      }
    }
    double sed[kSedChunks], fb[kSedChunks], post[kSedChunks];
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) sed[j] = vt[j] * xr[j];
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const double up = wv::above(sed, j);
      const int k = j * 64 + wv::lane();
      fb[j] = 0.0;
      post[j] = xr[j];
      if (updated(k)) {
        const double flux = (k == q.top) ? -sed[j] : (up - sed[j]);
        const double v = xr[j] + flux * q.odz[j] * dt * ost;
        const double vbar = v > limit ? xb[j] : 0.0;
        fb[j] = vbar * q.odz[j] * dt * ost + tb[j] * q.odz[j] * ost * q.orho[j];
        if (tb[j] != 0.0) rhob[j] += -q.orho[j] * q.orho[j] * tb[j] * flux * q.odz[j] * ost;
        xb[j] = vbar;
        post[j] = v > limit ? v : limit;
      }
    }
    const double x0post = __shfl(post[0], 0, 64);
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const double fbl = wv::below(fb, j);
      const int k = j * 64 + wv::lane();
      if (k >= q.nzc) continue;
      double sb = updated(k) ? -fb[j] : 0.0;
      if (k >= 1 && updated(k - 1) && k - 1 != q.top) sb += fbl;
      if (k == 0 && pbar != 0.0 && x0post > kLowLimitPrecip) sb += pbar * dt * ost;
      xb[j] += sb * vt[j];
      vb[j] += sb * xr[j];
    }
  }
}

// (126 VGPRs, four waves per SIMD; held to five it was measured slower)
template <class D>
__global__ void __launch_bounds__(64 * kScrW) k_adj_sed_wave(StepArgs<D> a, AdjArgs g) {
  // every operation rounded on its own, as in the forward sedimentation it
  // recomputes (so that the recomputed masses and clamps are the forward's)
#pragma clang fp contract(off)
  __shared__ double rows[kScrW][kSedLevels];
  const int w = int(threadIdx.x >> 6);
  const int c = int(blockIdx.x) * kScrW + w;
  if (c >= a.ncol) return;  // whole wavefront; no block-wide barrier below
  double* row = rows[w];
  const size_t plane = a.plane, n3 = size_t(a.ncol) * a.nz;
  const int l = wv::lane();
  WaveCol q;
  q.nzc = a.nlev ? min(max(a.nlev[c], 1), a.nz) : a.nz;
  q.top = q.nzc - 1;
  q.dt = a.dt;
  auto S = [&](int k) { return TEMPO_IDX_SCR(a, c, k); };
  auto E = [&](int k) { return size_t(c) + size_t(k) * a.ncol; };
  // this column's cotangents: a plane's first contribution is stored (as
  // 0 + x, the sum into a zeroed plane), later ones added; the planes and
  // levels no contribution reaches are zeroed at the end
  auto zero_z = [&](int z, int k0) {
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const int k = j * 64 + l;
      if (k >= k0 && k < a.nz) g.zbar[size_t(z) * n3 + E(k)] = 0.0;
    }
  };
  if (a.cf[c] & CF_EXIT) {
    for (int z = 0; z < AZ_N; ++z) zero_z(z, 0);
    return;
  }
  unsigned wrote = 0;  // planes with a stored contribution (wavefront-uniform)
  auto load_sd = [&](int f, double (&x)[kSedChunks]) {
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const int k = j * 64 + l;
      x[j] = k < q.nzc ? double(a.sd[size_t(f) * plane + S(k)].v) : 0.0;
    }
  };
  auto load_sr = [&](int f, double (&x)[kSedChunks]) {
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const int k = j * 64 + l;
      x[j] = k < q.nzc ? double(a.sr[size_t(f) * plane + S(k)].v) : 0.0;
    }
  };
  auto load_ub = [&](int m, double (&x)[kSedChunks]) {
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const int k = j * 64 + l;
      x[j] = k < q.nzc ? g.ubar[size_t(m) * n3 + E(k)] : 0.0;
    }
  };
  auto add_z = [&](int z, const double (&x)[kSedChunks]) {
    const bool first = !(wrote & (1u << z));
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const int k = j * 64 + l;
      if (k < q.nzc) {
        double* p = g.zbar + size_t(z) * n3 + E(k);
        *p = first ? 0.0 + x[j] : *p + x[j];
      }
    }
    wrote |= 1u << z;
  };
  auto zero = [&](double (&x)[kSedChunks]) {
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) x[j] = 0.0;
  };
  unsigned any = 0;
  using R = D;
  R dzr[kSedChunks];
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) {
    const int k = j * 64 + l;
    if (k < q.nzc) any |= a.pf[S(k)];
    dzr[j] = k < q.nzc ? a.f[F_DZ][TEMPO_IDX_STATE(a, c, k)] : R(1.0);
    q.odz[j] = 1.0 / double(value(dzr[j]));
    q.orho[j] = k < q.nzc ? 1.0 / double(a.sr[size_t(SR_RHO) * plane + S(k)].v) : 0.0;
  }
  any = wv::uor(any);
  const double dt = a.dt;
  const double pb_rain = g.pbar[c], pb_graupel = g.pbar[a.ncol + c], pb_snow = g.pbar[2 * size_t(a.ncol) + c],
               pb_ice = g.pbar[3 * size_t(a.ncol) + c];
  // CFL substeps and ktop from filled speeds (working precision)
  auto cfl = [&](auto vmaxf, int& steps, int& ktop) {
    int st = 1, kt = 1;
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const int k = j * 64 + l;
      if (k >= q.nzc) continue;
      const R vmax = vmaxf(j);
      if (vmax > R(1.e-3)) {
        kt = max(kt, k + 1);
        const R dz_by_vt = dzr[j] / vmax;
        st = max(st, int(R(dt) / dz_by_vt + R(1.0)));
      }
    }
    steps = wv::imax(st);
    ktop = wv::imax(kt);
    if (ktop == q.nzc) ktop = q.nzc - 1;
  };
  double rhob[kSedChunks];
  zero(rhob);

  // rain forward values for the snow adjoint
  double rr0[kSedChunks], vr[kSedChunks], rrpost[kSedChunks], vrbar[kSedChunks], rrbar[kSedChunks];
  int rsrc[kSedChunks];
  unsigned long long rm[kSedChunks];
  int r_steps = 1, r_ktop = 1;
  load_sd(SD_RR, rr0);
  zero(vrbar);
  zero(rrbar);
  zero(vr);
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) rrpost[j] = rr0[j];
  if (any & PF_QR) {
    double raw1[kSedChunks], raw2[kSedChunks], v2[kSedChunks];
    load_sr(SR_VTRR, raw1);
    load_sr(SR_VTNR, raw2);
    wave_fill_src(rr0, q.nzc, rsrc, rm);
    wave_gather(raw1, rsrc, vr);
    wave_gather(raw2, rsrc, v2);
    cfl([&](int j) { return mx(R(vr[j]), R(v2[j])); }, r_steps, r_ktop);
    const double ost = 1.0 / double(r_steps);
    for (int n = 0; n < r_steps; ++n) {
      double sed[kSedChunks];
#pragma unroll
      for (int j = 0; j < kSedChunks; ++j) sed[j] = vr[j] * rrpost[j];
#pragma unroll
      for (int j = 0; j < kSedChunks; ++j) {
        const double up = wv::above(sed, j);
        const int k = j * 64 + l;
        if (k < q.nzc && (k == q.top || k <= r_ktop - 1)) {
          const double flux = (k == q.top) ? -sed[j] : (up - sed[j]);
          const double v = rrpost[j] + flux * q.odz[j] * dt * ost;
          rrpost[j] = v > kR1 ? v : kR1;
        }
      }
    }
  }

  // ---- cloud ----
  if (any & PF_QC) {
    double xc[kSedChunks];
    load_sd(SD_RC, xc);
    int kt = 1;
    R hgt = R(0.0);
    for (int k = 0; k < q.nzc - 1; ++k) {
      if (wv::at_level(xc, k) > kR2) kt = k + 1;
      hgt = hgt + wv::at_level(dzr, k);
      if (hgt > R(500.0)) break;
    }
    for (int f2 = 0; f2 < 2; ++f2) {
      double x0[kSedChunks], raw[kSedChunks], vt[kSedChunks], tb[kSedChunks], xb[kSedChunks], vb[kSedChunks];
      load_sd(f2 ? SD_NC : SD_RC, x0);
      load_sr(f2 ? SR_VTNC : SR_VTRC, raw);
      load_ub(f2 ? 9 : 8, tb);
#pragma unroll
      for (int j = 0; j < kSedChunks; ++j) vt[j] = (j * 64 + l < kt) ? raw[j] : 0.0;
      zero(xb);
      zero(vb);
      wave_sed_adjoint(q, x0, vt, tb, f2 ? kR2 : kR1, 1, 1, kt, 0.0, xb, vb, rhob);
#pragma unroll
      for (int j = 0; j < kSedChunks; ++j)
        if (j * 64 + l >= kt) vb[j] = 0.0;
      add_z(f2 ? AZ_NC : AZ_RC, xb);
      add_z(f2 ? AZ_VTNC : AZ_VTRC, vb);
    }
  }
  // ---- ice: one application with steps = substeps ----
  if (any & PF_QI) {
    double xi[kSedChunks], raw1[kSedChunks], raw2[kSedChunks], v1[kSedChunks], v2[kSedChunks];
    int src[kSedChunks];
    unsigned long long m[kSedChunks];
    load_sd(SD_RI, xi);
    load_sr(SR_VTRI, raw1);
    load_sr(SR_VTNI, raw2);
    wave_fill_src(xi, q.nzc, src, m);
    wave_gather(raw1, src, v1);
    wave_gather(raw2, src, v2);
    int st, kt;
    cfl([&](int j) { return R(v1[j]); }, st, kt);
    for (int f2 = 0; f2 < 2; ++f2) {
      double x0[kSedChunks], tb[kSedChunks], xb[kSedChunks], vb[kSedChunks], acc[kSedChunks];
      load_sd(f2 ? SD_NI : SD_RI, x0);
      load_ub(f2 ? 7 : 6, tb);
      zero(xb);
      zero(vb);
      wave_sed_adjoint(q, x0, f2 ? v2 : v1, tb, f2 ? kR2 : kR1, 1, st, kt, f2 ? 0.0 : pb_ice, xb, vb, rhob);
      add_z(f2 ? AZ_NI : AZ_RI, xb);
      zero(acc);
      wave_scatter(vb, src, m, row, q.nzc, acc);
      add_z(f2 ? AZ_VTNI : AZ_VTRI, acc);
    }
  }
  // ---- snow: speed mixed with rain where snow melts ----
  if (any & PF_QS) {
    double xs[kSedChunks], raw[kSedChunks], vraw[kSedChunks], v1[kSedChunks], tb[kSedChunks];
    unsigned mix[kSedChunks];
    int src[kSedChunks];
    unsigned long long m[kSedChunks];
    load_sd(SD_RS, xs);
    load_sr(SR_VTS, raw);
    load_ub(5, tb);
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const int k = j * 64 + l;
      mix[j] = k < q.nzc ? (a.pf[S(k)] & PF_SNOWMIX) : 0u;
      R v = R(raw[j]);
      if (mix[j]) {
        const R sr_ = R(xs[j] / (xs[j] + rrpost[j]));
        v = v * sr_ + (R(1.0) - sr_) * R(vr[j]);
      }
      vraw[j] = double(value(v));
    }
    wave_fill_src(xs, q.nzc, src, m);
    wave_gather(vraw, src, v1);
    int st, kt;
    cfl([&](int j) { return R(v1[j]); }, st, kt);
    double xb[kSedChunks], vb[kSedChunks], acc[kSedChunks], zs[kSedChunks], zv[kSedChunks];
    zero(xb);
    zero(vb);
    wave_sed_adjoint(q, xs, v1, tb, kR1, st, st, kt, pb_snow, xb, vb, rhob);
    zero(acc);
    wave_scatter(vb, src, m, row, q.nzc, acc);
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      zs[j] = xb[j];
      zv[j] = 0.0;
      const double vbk = acc[j];
      if (vbk == 0.0) continue;
      if (mix[j]) {
        const double rs = xs[j], rr = rrpost[j], den = rs + rr;
        const double sr_ = rs / den;
        zv[j] = vbk * sr_;
        vrbar[j] += vbk * (1.0 - sr_);
        const double srb = vbk * (raw[j] - vr[j]);
        zs[j] += srb * rr / (den * den);
        rrbar[j] += -srb * rs / (den * den);
      } else {
        zv[j] = vbk;
      }
    }
    add_z(AZ_RS, zs);
    add_z(AZ_VTS, zv);
  }
  // ---- graupel ----
  if (any & PF_QG) {
    double xg[kSedChunks], raw1[kSedChunks], raw2[kSedChunks], v1[kSedChunks], v2[kSedChunks];
    int src[kSedChunks];
    unsigned long long m[kSedChunks];
    load_sd(SD_RG, xg);
    load_sr(SR_VTRG, raw1);
    load_sr(SR_VTNG, raw2);
    wave_fill_src(xg, q.nzc, src, m);
    wave_gather(raw1, src, v1);
    wave_gather(raw2, src, v2);
    int st, kt;
    cfl([&](int j) { return R(v1[j]); }, st, kt);
    const double rb_lim = kM3ToL * kR1 / 800.0;
    const int fms[3] = {SD_RG, SD_NG, SD_RB}, azm[3] = {AZ_RG, AZ_NG, AZ_RB}, azv[3] = {AZ_VTRG, AZ_VTNG, AZ_VTRG},
              ten[3] = {2, 3, 4};
    const double lims[3] = {kR1, kR2, rb_lim};
    for (int f3 = 0; f3 < 3; ++f3) {
      double x0[kSedChunks], tb[kSedChunks], xb[kSedChunks], vb[kSedChunks], acc[kSedChunks];
      load_sd(fms[f3], x0);
      load_ub(ten[f3], tb);
      zero(xb);
      zero(vb);
      wave_sed_adjoint(q, x0, f3 == 1 ? v2 : v1, tb, lims[f3], st, st, kt, f3 == 0 ? pb_graupel : 0.0, xb, vb, rhob);
      add_z(azm[f3], xb);
      zero(acc);
      wave_scatter(vb, src, m, row, q.nzc, acc);
      add_z(azv[f3], acc);
    }
  }
  // ---- rain: final masses and filled speed carry the snow cotangents ----
  if (any & PF_QR) {
    for (int f2 = 0; f2 < 2; ++f2) {
      double x0[kSedChunks], raw[kSedChunks], vt[kSedChunks], tb[kSedChunks], xb[kSedChunks], vb[kSedChunks],
          acc[kSedChunks];
      load_sd(f2 ? SD_NR : SD_RR, x0);
      load_sr(f2 ? SR_VTNR : SR_VTRR, raw);
      wave_gather(raw, rsrc, vt);
      load_ub(f2 ? 1 : 0, tb);
#pragma unroll
      for (int j = 0; j < kSedChunks; ++j) {
        xb[j] = f2 ? 0.0 : rrbar[j];
        vb[j] = f2 ? 0.0 : vrbar[j];
      }
      wave_sed_adjoint(q, x0, vt, tb, f2 ? kR2 : kR1, r_steps, r_steps, r_ktop, f2 ? 0.0 : pb_rain, xb, vb, rhob);
      add_z(f2 ? AZ_NR : AZ_RR, xb);
      zero(acc);
      wave_scatter(vb, rsrc, rm, row, q.nzc, acc);
      add_z(f2 ? AZ_VTNR : AZ_VTRR, acc);
    }
  } else {
    add_z(AZ_RR, rrbar);
  }
  add_z(AZ_RHO, rhob);
  for (int z = 0; z < AZ_N; ++z) zero_z(z, (wrote & (1u << z)) ? q.nzc : 0);
}

// All tangent passes and S^T of one adjoint step.  a: dual step arguments
// over the working state copy (whose fields point into `work_state`);
// `inputs` is the packed input state.  snap_a / snap_b receive the working
// state and scratch after phase A and after phase B of a tangent-free run,
// so that the later passes restart from them instead of recomputing.
// scratch snapshots of the adjoint: working-precision and FP64 planes after
// phase A (used for nz > 128 only), and the post-sedimentation tendency
// values
struct AdjSnap {
  void* sr_a;  // sr_bytes
  void* sd_a;  // sd_bytes
  void* sd_b;  // sd_bytes (holds 10 FP64 planes of the post-sedimentation tendencies)
};

// The passes of one adjoint.  The state is out of place: a.f holds the
// packed inputs (duals, tangents zero except the seeded field), a.fa the
// state after the initial checks (phase A's output), a.fo the new state
// (the final update's output), so that no pass has to restore the state.
template <class D>
hipError_t adjoint_passes(const StepArgs<D>& a, const AdjArgs& g, Dual<double>* vtbuf, size_t sr_bytes,
                          size_t sd_bytes, AdjSnap snap, const bool* have_x, hipStream_t s) {
  const unsigned gc = unsigned((a.ncol + 255) / 256);
  const dim3 g2(gc, unsigned(a.nz));
  const size_t n3 = size_t(a.ncol) * a.nz;
  hipError_t e;
  auto cp = [&](void* dst, const void* src, size_t n) { return hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, s); };
  auto phase_pre_a = [&]() {
    hipLaunchKernelGGL(k_prepass<D>, prepass_grid(a), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_phase_a<D>, level_grid(a, 256), dim3(256), 0, s, a);
  };
  // the final update's passes and contractions skip the levels without
  // output cotangents
  StepArgs<D> ac = a;
  ac.lvl_on = g.lvl_on;
  hipLaunchKernelGGL(k_adj_level_zero<D>, dim3(unsigned((a.nz + 255) / 256)), dim3(256), 0, s, a, g);
  hipLaunchKernelGGL(k_adj_level_flags<D>, g2, dim3(256), 0, s, a, g);
  auto phase_c = [&]() { hipLaunchKernelGGL(k_phase_c<D>, level_grid(a, 256), dim3(256), 0, s, ac); };

  // tangent-free forward.  Its sedimentation on the values, one wavefront
  // per column (the tangents are zero in this pass), writes only the ten
  // tendency planes, which S^T does not read and every later reader finds
  // rewritten by a local pass's phase A (it writes all planes of each live
  // point it runs), so the scratch after A needs no snapshot.  The
  // lane-per-column form (nz > 128) uses state planes as work space: the
  // scratch after A is snapshotted and restored before S^T.
  const bool wave_b = a.nz <= kSedLevels;
  phase_pre_a();
  if (!wave_b) {
    if ((e = cp(snap.sr_a, a.sr, sr_bytes)) != hipSuccess) return e;
    if ((e = cp(snap.sd_a, a.sd, sd_bytes)) != hipSuccess) return e;
  }
  if (wave_b)
    hipLaunchKernelGGL(k_phase_b_wave<D>, dim3(unsigned((a.ncol + kSedCols - 1) / kSedCols)), dim3(64 * kSedCols),
                       sed_lds_bytes(a.nz), s, a);
  else
    hipLaunchKernelGGL(k_phase_b<D>, dim3(unsigned((a.ncol + 127) / 128)), dim3(128), 0, s, a, vtbuf);
  double* post = static_cast<double*>(snap.sd_b);
  hipLaunchKernelGGL(k_adj_keep_post<D>, g2, dim3(256), 0, s, a, post);
  // passes seeded on the post-sedimentation tendencies (the final update
  // reads the inputs, a.fa and the scratch, and writes only a.fo).  Without
  // prognostic graupel number and volume the final update does not read
  // their tendencies, so those cotangents are zero and need no pass.
  // Each pass's contraction moves the seed on to the next pass.
  const bool ngqb = a.f[F_NG] && a.f[F_QB];
  int ms[10], nm = 0;
  for (int m = 0; m < 10; ++m) {
    if (!ngqb && (SD_QRTEN + m == SD_NGTEN || SD_QRTEN + m == SD_QBTEN)) {
      double* ub = g.ubar + size_t(m) * n3;
      hipLaunchKernelGGL(k_adj_zero<double>, dim3(unsigned((n3 + 255) / 256)), dim3(256), 0, s, ub, n3);
      continue;
    }
    ms[nm++] = m;
  }
  if (nm > 0) hipLaunchKernelGGL(k_adj_seed_ten<D>, g2, dim3(256), 0, s, a, ms[0], -1);
  for (int j = 0; j < nm; ++j) {
    phase_c();
    const AdjSeed nx{2, j + 1 < nm ? ms[j + 1] : -1, j + 1 < nm ? ms[j] : -1};
    hipLaunchKernelGGL(k_adj_contract<D>, g2, dim3(256), 0, s, ac, g, g.ubar + size_t(ms[j]) * n3, 0, nx);
  }
  // S^T on the values after phase A
  if (!wave_b) {
    if ((e = cp(a.sr, snap.sr_a, sr_bytes)) != hipSuccess) return e;
    if ((e = cp(a.sd, snap.sd_a, sd_bytes)) != hipSuccess) return e;
  }
  if (wave_b)
    hipLaunchKernelGGL(k_adj_sed_wave<D>, dim3(unsigned((a.ncol + kScrW - 1) / kScrW)), dim3(64 * kScrW), 0, s, a, g);
  else
    hipLaunchKernelGGL(k_adj_sedimentation<D>, dim3(unsigned((a.ncol + 63) / 64)), dim3(64), 0, s, a, g);
  // the local passes' phase A runs at the points where they contribute,
  // their final update on the post-sedimentation tendency values
  StepArgs<D> al = a, acl = ac;
  al.pt_on = g.pt_loc;
  acl.post = post;
  hipLaunchKernelGGL(k_adj_point_flags<D>, g2, dim3(256), 0, s, a, g);
  // local passes, one per requested input field: xbar_i = (dz/dx_i) . zbar
  // + the local part through A and C, phase B replaced by the forward's
  // post-sedimentation tendency values.  The pre-pass results (column flags,
  // heights) depend on values only and stay from the first pass.
  int xs[AX_N], nxs = 0;
  for (int i = 0; i < AX_N; ++i)
    if (have_x[i]) xs[nxs++] = i;
  if (nxs > 0) hipLaunchKernelGGL(k_adj_seed_input<D>, g2, dim3(256), 0, s, a, adj_x_field(xs[0]), -1);
  for (int j = 0; j < nxs; ++j) {
    double* xb = g.xbar + size_t(xs[j]) * n3;
    hipLaunchKernelGGL(k_phase_a<D>, level_grid(a, 256), dim3(256), 0, s, al);
    hipLaunchKernelGGL(k_adj_capture<D>, g2, dim3(256), 0, s, a, g, xb);
    hipLaunchKernelGGL(k_phase_c<D>, level_grid(a, 256), dim3(256), 0, s, acl);
    const AdjSeed nx{1, j + 1 < nxs ? adj_x_field(xs[j + 1]) : -1, j + 1 < nxs ? adj_x_field(xs[j]) : -1};
    hipLaunchKernelGGL(k_adj_contract<D>, g2, dim3(256), 0, s, ac, g, xb, 1, nx);
  }
  return hipGetLastError();
}

}  // namespace tempo

// FP64 lookup-table construction on the GPU, plus packing into the gather
// layouts used by the column step.  Compiled with -ffp-contract=off: every
// accumulation is a separate IEEE multiply and add, in the reference loop
// order, so each entry rounds exactly as the reference builder's.
//
// The O(N^3) factors (size distributions, kernel products, freezing
// probabilities) come from the host in FP64 (tempo_host.cpp); these kernels
// do the O(N^5) sums: 1.7e11 multiply-adds for the rain-graupel table, which
// the reference builds in 3000 s on one core (tools/build_tables.F90:10).
//
// Memory layouts (all double):
//   n_r     [m][k][n2]          rain distribution for r_r(m), n0r_exp(k)
//   qg_a    [q][n3][n2][n]      q = t1 z1 y1 t2 y2 kernel products
//   qg_ng   [n3][j][i][n]       graupel distribution for n0g_exp(i), r_g(j)
//   qs_a    [q][n2][n]          q = t1 z1 y1 t3 z3 y3 t2 y2 z2 t4 y4 z4
//   qs_ns   [j][i][n]           snow distribution for r_s(i), tc(j)
//   fz_prob [m][k][n]           Bigg freezing probability for nt_in(m), T index k
//   fz_nc   [j][i][n]           cloud distribution for t_nc(j), r_c(i)
// Outputs are the reference table arrays, Fortran order, file order.
#include <hip/hip_runtime.h>

#include "tempo_const.hpp"

namespace tempo {

constexpr int kNb = kNbins;

// rain-graupel collection (qr_acr_qg, src/module_mp_tempo_tables.F90:624-714)
// grid: x = entries (k,m) of one (i,j,n3); y = i + 37*(j + 37*n3)
__global__ void __launch_bounds__(256) k_table_qg(const double* __restrict__ qg_a,
                                                  const double* __restrict__ qg_ng,
                                                  const double* __restrict__ n_r,
                                                  const double* __restrict__ r_r,
                                                  double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kNtbR1 * kNtbR) return;
  const int k = e % kNtbR1, m = e / kNtbR1;
  const int b = blockIdx.y;
  const int i = b % kNtbG1, j = (b / kNtbG1) % kNtbG, n3 = b / (kNtbG1 * kNtbG);
  const size_t qstride = size_t(kNrhg) * kNb * kNb;
  const double* A = qg_a + size_t(n3) * kNb * kNb;
  const double* g = qg_ng + ((size_t(n3) * kNtbG + j) * kNtbG1 + i) * kNb;
  const double* r = n_r + (size_t(m) * kNtbR1 + k) * kNb;
  double t1 = 0.0, z1 = 0.0, y1 = 0.0, t2 = 0.0, y2 = 0.0;
  for (int n2 = 0; n2 < kNb; ++n2) {
    const double rr = r[n2];
    const double* a = A + n2 * kNb;
    for (int n = 0; n < kNb; ++n) {
      const double gg = g[n];
      t1 = t1 + (a[0 * qstride + n] * gg) * rr;
      z1 = z1 + (a[1 * qstride + n] * gg) * rr;
      y1 = y1 + (a[2 * qstride + n] * gg) * rr;
      t2 = t2 + (a[3 * qstride + n] * gg) * rr;
      y2 = y2 + (a[4 * qstride + n] * gg) * rr;
    }
  }
  const size_t idx = size_t(i) + kNtbG1 * (size_t(j) + kNtbG * (size_t(n3) + kNrhg * (size_t(k) + kNtbR1 * m)));
  const double rm = r_r[m];
  out[0 * kNQg1 + idx] = t1;
  out[1 * kNQg1 + idx] = z1 < rm ? z1 : rm;
  out[2 * kNQg1 + idx] = t2;
  out[3 * kNQg1 + idx] = y1;
  out[4 * kNQg1 + idx] = y2;
}

// rain-snow collection (qr_acr_qs, :493-621); y = i + 37*j (snow mass, T)
__global__ void __launch_bounds__(256) k_table_qs(const double* __restrict__ qs_a,
                                                  const double* __restrict__ qs_ns,
                                                  const double* __restrict__ n_r,
                                                  const double* __restrict__ r_r,
                                                  double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kNtbR1 * kNtbR) return;
  const int k = e % kNtbR1, m = e / kNtbR1;
  const int i = blockIdx.y % kNtbS, j = blockIdx.y / kNtbS;
  const size_t qs = size_t(kNb) * kNb;
  const double* s = qs_ns + (size_t(j) * kNtbS + i) * kNb;
  const double* r = n_r + (size_t(m) * kNtbR1 + k) * kNb;
  double acc[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) acc[q] = 0.0;
  for (int n2 = 0; n2 < kNb; ++n2) {
    const double rr = r[n2];
    const double* a = qs_a + size_t(n2) * kNb;
    for (int n = 0; n < kNb; ++n) {
      const double ss = s[n];
#pragma unroll
      for (int q = 0; q < 12; ++q) acc[q] = acc[q] + (a[q * qs + n] * ss) * rr;
    }
  }
  const size_t idx = size_t(i) + kNtbS * (size_t(j) + kNtbT * (size_t(k) + kNtbR1 * m));
  const double rm = r_r[m];
  // accumulators t1 z1 y1 t3 z3 y3 t2 y2 z2 t4 y4 z4 -> file order
  // tcs_racs1 tmr_racs1 tcs_racs2 tmr_racs2 tcr_sacr1 tms_sacr1 tcr_sacr2
  // tms_sacr2 tnr_racs1 tnr_racs2 tnr_sacr1 tnr_sacr2 (:605-616)
  const double v[12] = {acc[0], acc[1] < rm ? acc[1] : rm, acc[3], acc[4], acc[6], acc[8],
                        acc[9], acc[11], acc[2], acc[5], acc[7], acc[10]};
#pragma unroll
  for (int q = 0; q < 12; ++q) out[q * kNQs1 + idx] = v[q];
}

// Bigg freezing of rain (freezewater :747-782): scan from the largest drop
// down with early exit once the frozen mass reaches r_r(i).
// one thread per (i, j, k, m), i fastest
__global__ void __launch_bounds__(256) k_table_fz_rain(const double* __restrict__ n_r,
                                                       const double* __restrict__ prob,
                                                       const double* __restrict__ massr,
                                                       const double* __restrict__ r_r, double xm0g,
                                                       double* __restrict__ out) {
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= kNFzR1) return;
  const int i = int(e % kNtbR), j = int((e / kNtbR) % kNtbR1);
  const int k = int((e / (kNtbR * kNtbR1)) % kNtbT1), m = int(e / (size_t(kNtbR) * kNtbR1 * kNtbT1));
  const double* nr = n_r + (size_t(i) * kNtbR1 + j) * kNb;  // r_r(i), n0r_exp(j)
  const double* pr = prob + (size_t(m) * kNtbT1 + k) * kNb;
  const double lim = r_r[i];
  double sum1 = 0.0, sum2 = 0.0, sumn1 = 0.0, sumn2 = 0.0;
  for (int n2 = kNb - 1; n2 >= 0; --n2) {
    const double pn = pr[n2] * nr[n2];
    if (massr[n2] < xm0g) {
      sumn1 = sumn1 + pn;
      sum1 = sum1 + pn * massr[n2];
    } else {
      sumn2 = sumn2 + pn;
      sum2 = sum2 + pn * massr[n2];
    }
    if ((sum1 + sum2) >= lim) break;
  }
  out[0 * kNFzR1 + e] = sum1;   // tpi_qrfz
  out[1 * kNFzR1 + e] = sumn1;  // tni_qrfz
  out[2 * kNFzR1 + e] = sum2;   // tpg_qrfz
  out[3 * kNFzR1 + e] = sumn2;  // tnr_qrfz
}

// Bigg freezing of cloud drops (freezewater :784-802)
__global__ void __launch_bounds__(256) k_table_fz_cloud(const double* __restrict__ n_c,
                                                        const double* __restrict__ prob,
                                                        const double* __restrict__ massc,
                                                        const double* __restrict__ t_nc,
                                                        const double* __restrict__ r_c,
                                                        double* __restrict__ out) {
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= kNFzC1) return;
  const int i = int(e % kNtbC), j = int((e / kNtbC) % kNbins);
  const int k = int((e / (size_t(kNtbC) * kNbins)) % kNtbT1), m = int(e / (size_t(kNtbC) * kNbins * kNtbT1));
  const double* nc = n_c + (size_t(j) * kNtbC + i) * kNb;
  const double* pr = prob + (size_t(m) * kNtbT1 + k) * kNb;
  const double lim = r_c[i], cap = t_nc[j];
  double sum1 = 0.0, sumn2 = 0.0;
  for (int n = kNb - 1; n >= 0; --n) {
    const double pn = pr[n] * nc[n];
    const double s = sumn2 + pn;
    sumn2 = cap < s ? cap : s;
    sum1 = sum1 + pn * massc[n];
    if (sum1 >= lim) break;
  }
  out[0 * kNFzC1 + e] = sum1;   // tpi_qcfz
  out[1 * kNFzC1 + e] = sumn2;  // tni_qcfz
}

// Pack narr reference-order arrays (each n elements) into an interleaved
// [idx][narr] gather layout in precision T, so that one lookup touches one
// cache line instead of narr separate ones.
template <class T>
__global__ void __launch_bounds__(256) k_pack(const double* __restrict__ src, size_t n, int narr,
                                              T* __restrict__ dst) {
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  for (int q = 0; q < narr; ++q) dst[e * narr + q] = T(src[q * n + e]);
}

template __global__ void k_pack<float>(const double*, size_t, int, float*);
template __global__ void k_pack<double>(const double*, size_t, int, double*);

}  // namespace tempo

// ---- launch wrappers (C++ linkage, used by tempo_capi.cpp) ----
namespace tempo {

hipError_t launch_table_qg(const double* qg_a, const double* qg_ng, const double* n_r, const double* r_r,
                           double* out, hipStream_t s) {
  dim3 grid((kNtbR1 * kNtbR + 255) / 256, kNtbG1 * kNtbG * kNrhg);
  hipLaunchKernelGGL(k_table_qg, grid, dim3(256), 0, s, qg_a, qg_ng, n_r, r_r, out);
  return hipGetLastError();
}

hipError_t launch_table_qs(const double* qs_a, const double* qs_ns, const double* n_r, const double* r_r,
                           double* out, hipStream_t s) {
  dim3 grid((kNtbR1 * kNtbR + 255) / 256, kNtbS * kNtbT);
  hipLaunchKernelGGL(k_table_qs, grid, dim3(256), 0, s, qs_a, qs_ns, n_r, r_r, out);
  return hipGetLastError();
}

hipError_t launch_table_fz(const double* n_r, const double* prob_r, const double* massr, const double* r_r,
                           double xm0g, const double* n_c, const double* prob_c, const double* massc,
                           const double* t_nc, const double* r_c, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_table_fz_rain, dim3((kNFzR1 + 255) / 256), dim3(256), 0, s, n_r, prob_r, massr, r_r, xm0g,
                     out);
  hipLaunchKernelGGL(k_table_fz_cloud, dim3((kNFzC1 + 255) / 256), dim3(256), 0, s, n_c, prob_c, massc, t_nc, r_c,
                     out + 4 * kNFzR1);
  return hipGetLastError();
}

template <class T>
hipError_t launch_pack(const double* src, size_t n, int narr, T* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_pack<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, src, n, narr, dst);
  return hipGetLastError();
}
template hipError_t launch_pack<float>(const double*, size_t, int, float*, hipStream_t);
template hipError_t launch_pack<double>(const double*, size_t, int, double*, hipStream_t);

}  // namespace tempo

// Column step instantiated in double.  FP64 twin: compiled with -ffp-contract=off so
// every operation rounds as in the reference FP64 build.
#include "tempo_kernels.hpp"
#include "tempo_sedi.hpp"
#include "tempo_adjoint_wave.hpp"

namespace tempo {

// ev (optional): 5 events recorded around the 4 launches for per-phase timing
hipError_t launch_step(const StepArgs<double>& a, const LaunchCfg& lc, double* vtbuf, hipStream_t s, hipEvent_t* ev) {
  const unsigned gc = unsigned((a.ncol + 255) / 256);
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_prepass<double>, dim3(gc), dim3(256), 0, s, a);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_phase_a<double>, level_grid(a, lc.block_a), dim3(unsigned(lc.block_a)), 0, s, a);
  if (ev) (void)hipEventRecord(ev[2], s);
  if (lc.sed_wave && a.nz <= kSedLevels && !a.fl.semi_sedi)
    hipLaunchKernelGGL(k_phase_b_wave<double>, dim3(unsigned((a.ncol + kSedCols - 1) / kSedCols)), dim3(64 * kSedCols),
                       sed_lds_bytes(a.nz), s, a);
  else
    hipLaunchKernelGGL(k_phase_b<double>, dim3(unsigned((a.ncol + 127) / 128)), dim3(128), 0, s, a, vtbuf);
  if (a.fl.refl_from_melting && a.fl.refl10cm && a.d3[D3_REFL]) {
    hipLaunchKernelGGL(k_melt_level<double>, dim3(unsigned((a.ncol + 255) / 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_melt_refl<double>, level_grid(a, 256), dim3(256), 0, s, a);
  }
  if (ev) (void)hipEventRecord(ev[3], s);
  hipLaunchKernelGGL(k_phase_c<double>, level_grid(a, lc.block_c), dim3(unsigned(lc.block_c)), 0, s, a);
  if (ev) (void)hipEventRecord(ev[4], s);
  return hipGetLastError();
}

// tangent-linear step in this precision (same contraction flags as the step)
hipError_t launch_step_tlm(const StepArgs<Dual<double>>& a, Dual<double>* vtbuf, hipStream_t s) {
  return launch_step_tlm_impl(a, vtbuf, s);
}
template hipError_t launch_check_update<double>(const CheckArgs<double>&, int, hipStream_t);
template hipError_t launch_surface_emissions<double>(double*, const double*, int, int, long long, long long, long long, int,
                                                     double, hipStream_t);
template hipError_t launch_dual_pack<double>(const double*, const double*, const DualGeom&, Dual<double>*, hipStream_t);
template hipError_t launch_dual_unpack<double>(const Dual<double>*, const DualGeom&, double*, double*, hipStream_t);

hipError_t launch_adjoint(const StepArgs<Dual<double>>& a, const AdjArgs& g, Dual<double>* vtbuf, size_t sr_bytes,
                          size_t sd_bytes, AdjSnap snap, const bool* have_x, hipStream_t s) {
  return adjoint_passes(a, g, vtbuf, sr_bytes, sd_bytes, snap, have_x, s);
}
template hipError_t launch_to_compact<double>(const double*, const DualGeom&, double*, hipStream_t);
template hipError_t launch_from_compact<double>(const double*, const DualGeom&, double*, hipStream_t);

}  // namespace tempo

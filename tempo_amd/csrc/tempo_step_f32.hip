// Column step instantiated in float.  Production FP32 path (sedimentation state and
// fluxes carried in FP64 inside phase B).
#include "tempo_kernels.hpp"
#include "tempo_sedi.hpp"
#include "tempo_adjoint_wave.hpp"

namespace tempo {

// ev (optional): 5 events recorded around the 4 launches for per-phase timing
hipError_t launch_step(const StepArgs<float>& a, const LaunchCfg& lc, double* vtbuf, hipStream_t s, hipEvent_t* ev) {
  const unsigned gc = unsigned((a.ncol + 255) / 256);
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_prepass<float>, dim3(gc), dim3(256), 0, s, a);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_phase_a<float>, level_grid(a, lc.block_a), dim3(unsigned(lc.block_a)), 0, s, a);
  if (ev) (void)hipEventRecord(ev[2], s);
  if (lc.sed_wave && a.nz <= kSedLevels && !a.fl.semi_sedi)
    hipLaunchKernelGGL(k_phase_b_wave<float>, dim3(unsigned((a.ncol + kSedCols - 1) / kSedCols)), dim3(64 * kSedCols),
                       sed_lds_bytes(a.nz), s, a);
  else
    hipLaunchKernelGGL(k_phase_b<float>, dim3(unsigned((a.ncol + 127) / 128)), dim3(128), 0, s, a, vtbuf);
  if (a.fl.refl_from_melting && a.fl.refl10cm && a.d3[D3_REFL]) {
    hipLaunchKernelGGL(k_melt_level<float>, dim3(unsigned((a.ncol + 255) / 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_melt_refl<float>, level_grid(a, 256), dim3(256), 0, s, a);
  }
  if (ev) (void)hipEventRecord(ev[3], s);
  hipLaunchKernelGGL(k_phase_c<float>, level_grid(a, lc.block_c), dim3(unsigned(lc.block_c)), 0, s, a);
  if (ev) (void)hipEventRecord(ev[4], s);
  return hipGetLastError();
}

// tangent-linear step in this precision (same contraction flags as the step)
hipError_t launch_step_tlm(const StepArgs<Dual<float>>& a, Dual<double>* vtbuf, hipStream_t s) {
  return launch_step_tlm_impl(a, vtbuf, s);
}
template hipError_t launch_check_update<float>(const CheckArgs<float>&, int, hipStream_t);
template hipError_t launch_surface_emissions<float>(float*, const float*, int, int, long long, long long, long long, int,
                                                     double, hipStream_t);
template hipError_t launch_dual_pack<float>(const float*, const float*, const DualGeom&, Dual<float>*, hipStream_t);
template hipError_t launch_dual_unpack<float>(const Dual<float>*, const DualGeom&, float*, float*, hipStream_t);

hipError_t launch_adjoint(const StepArgs<Dual<float>>& a, const AdjArgs& g, Dual<double>* vtbuf, size_t sr_bytes,
                          size_t sd_bytes, AdjSnap snap, const bool* have_x, hipStream_t s) {
  return adjoint_passes(a, g, vtbuf, sr_bytes, sd_bytes, snap, have_x, s);
}
template hipError_t launch_to_compact<float>(const float*, const DualGeom&, double*, hipStream_t);
template hipError_t launch_from_compact<float>(const double*, const DualGeom&, float*, hipStream_t);

}  // namespace tempo

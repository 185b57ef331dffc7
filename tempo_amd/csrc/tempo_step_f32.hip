// Column step instantiated in float.  Production FP32 path (sedimentation state and
// fluxes carried in FP64 inside phase B).
#include "tempo_kernels.hpp"
#include "tempo_sedi.hpp"
#include "tempo_fused.hpp"
#include "tempo_adjoint_wave.hpp"

namespace tempo {

// ev (optional): 5 events recorded around the 4 launches for per-phase timing
hipError_t launch_step(const StepArgs<float>& a, const LaunchCfg& lc, double* vtbuf, hipStream_t s, hipEvent_t* ev) {
  return launch_step_impl(a, a, lc, vtbuf, s, ev);
}

// the step in one persistent launch of `grid` blocks (fused_grid)
hipError_t launch_step_fused(const StepArgs<float>& a, int grid, hipStream_t s, hipEvent_t* ev) {
  return launch_step_fused_impl(a, grid, s, ev);
}
int fused_grid_size(const StepArgs<float>& a) { return fused_grid<float>(a.ncol, a.nz); }

// the same step with the level-local kernels recording branch signatures
// into a.sig
hipError_t launch_step_sig(const StepArgs<float>& a, const LaunchCfg& lc, double* vtbuf, hipStream_t s) {
  return launch_step_impl(a, sig_args(a), lc, vtbuf, s, nullptr);
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
template hipError_t launch_to_compact_n<float>(const PlaneBatch<float, double>&, int, const DualGeom&, hipStream_t);
template hipError_t launch_from_compact_n<float>(const PlaneBatch<double, float>&, int, const DualGeom&, hipStream_t);
template hipError_t launch_dual_pack_n<float>(const PlaneBatch<float, Dual<float>>&, int, const DualGeom&, hipStream_t);

}  // namespace tempo

// Column step instantiated in double.  FP64 twin: compiled with -ffp-contract=off so
// every operation rounds as in the reference FP64 build.
#include "tempo_kernels.hpp"
#include "tempo_sedi.hpp"
#include "tempo_fused.hpp"
#include "tempo_adjoint_wave.hpp"

namespace tempo {

// ev (optional): 5 events recorded around the 4 launches for per-phase timing
hipError_t launch_step(const StepArgs<double>& a, const LaunchCfg& lc, double* vtbuf, hipStream_t s, hipEvent_t* ev) {
  return launch_step_impl(a, a, lc, vtbuf, s, ev);
}

// the step in one persistent launch of `grid` blocks (fused_grid)
hipError_t launch_step_fused(const StepArgs<double>& a, int grid, hipStream_t s, hipEvent_t* ev) {
  return launch_step_fused_impl(a, grid, s, ev);
}
int fused_grid_size(const StepArgs<double>& a) { return fused_grid<double>(a.ncol, a.nz); }

// the same step with the level-local kernels recording branch signatures
// into a.sig
hipError_t launch_step_sig(const StepArgs<double>& a, const LaunchCfg& lc, double* vtbuf, hipStream_t s) {
  return launch_step_impl(a, sig_args(a), lc, vtbuf, s, nullptr);
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
template hipError_t launch_to_compact_n<double>(const PlaneBatch<double, double>&, int, const DualGeom&, hipStream_t);
template hipError_t launch_from_compact_n<double>(const PlaneBatch<double, double>&, int, const DualGeom&, hipStream_t);
template hipError_t launch_dual_pack_n<double>(const PlaneBatch<double, Dual<double>>&, int, const DualGeom&, hipStream_t);

}  // namespace tempo

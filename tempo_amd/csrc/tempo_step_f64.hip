// Column step instantiated in double.  FP64 twin: compiled with -ffp-contract=off so
// every operation rounds as in the reference FP64 build.
#include "tempo_kernels.hpp"

namespace tempo {

// ev (optional): 5 events recorded around the 4 launches for per-phase timing
hipError_t launch_step(const StepArgs<double>& a, double* vtbuf, hipStream_t s, hipEvent_t* ev) {
  const unsigned gc = unsigned((a.ncol + 255) / 256);
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_prepass<double>, dim3(gc), dim3(256), 0, s, a);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_phase_a<double>, dim3(gc, unsigned(a.nz)), dim3(256), 0, s, a);
  if (ev) (void)hipEventRecord(ev[2], s);
  hipLaunchKernelGGL(k_phase_b<double>, dim3(unsigned((a.ncol + 127) / 128)), dim3(128), 0, s, a, vtbuf);
  if (ev) (void)hipEventRecord(ev[3], s);
  hipLaunchKernelGGL(k_phase_c<double>, dim3(gc, unsigned(a.nz)), dim3(256), 0, s, a);
  if (ev) (void)hipEventRecord(ev[4], s);
  return hipGetLastError();
}

}  // namespace tempo

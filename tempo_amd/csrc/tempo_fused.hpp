// The whole column step in one persistent, work-queued launch.
//
// The four-launch step (k_prepass, k_phase_a, the sedimentation, k_phase_c)
// passes 37 planes of per-point scratch through HBM between its kernels and
// runs its level-local kernels over every (column, level) slot of the batch,
// inactive levels of ragged columns included.  k_step_fused instead keeps a
// fixed grid of blocks resident (as many as fit on the chip) that pull
// column groups -- kSedCols consecutive columns, one scratch column tile --
// from a global queue until it is empty.  A block runs, for its group:
//
//   P  one wavefront per column: the early-exit and first-step graupel tests
//      and the height prefix of the default aerosol profile, over the
//      column's active levels (lanes over levels, wavefront reductions);
//   A  the level-local physics of every active point of the group (lanes over
//      8 columns x 8 levels, so that the scratch writes are contiguous);
//   B  the sedimentation, one wavefront per column (phase_b_group);
//   C  the final update of every active point.
//
// The per-point scratch of a group lives in a slot private to the block and
// reused for every group it takes, so the 37 planes stay in the caches
// instead of crossing HBM; only the state, the diagnostics and the column
// results (flags, substeps, precipitation) are global.  Each point's
// arithmetic is the four-launch step's, so the results are identical.
#pragma once

#include "tempo_sedi.hpp"

namespace tempo {

constexpr int kFusedThreads = 64 * kSedCols;

// stage P for column c (wavefront w of the block; all 64 lanes call it):
// k_prepass's column tests with the levels across the lanes; the sums and
// the height prefix keep the reference's sequential order (lane 0)
template <class R>
__device__ __forceinline__ void prepass_wave(const StepArgs<R>& a, const Scr<R>& scr, int c, int* nzc_s, int w) {
  const int lane = int(threadIdx.x & 63u);
  if (c >= a.ncol) {
    if (lane == 0) nzc_s[w] = 0;
    return;
  }
  const int nzc = a.nlev ? min(max(a.nlev[c], 1), a.nz) : a.nz;
  unsigned cf = 0;
  bool q = false;
#pragma unroll
  for (int j = 0; j < kSedChunks; ++j) {
    const int k = j * 64 + lane;
    if (k < nzc) {
      const size_t o = TEMPO_IDX_STATE(a, c, k);
      q = q || a.f[F_QC][o] > R(kR1) || a.f[F_QR][o] > R(kR1) || a.f[F_QI][o] > R(kR1) ||
          a.f[F_QS][o] > R(kR1) || a.f[F_QG][o] > R(kR1);
    }
  }
  const bool any_q = __ballot(q) != 0ull;
  if (!any_q) {
    bool s = false;
#pragma unroll
    for (int j = 0; j < kSedChunks; ++j) {
      const int k = j * 64 + lane;
      if (k < nzc) {
        const size_t o = TEMPO_IDX_STATE(a, c, k);
        const R temp = load_t(a, o), pres = a.f[F_P][o], qv = mx(R(kMinQv), a.f[F_QV][o]);
        const R qvs = rslf(pres, temp);
        const R qvsi = (temp - R(kT0) <= R(0.0)) ? rsif(pres, temp) : qvs;
        R ssati = qv / qvsi - R(1.0);
        if (mabs(ssati) < R(kEps)) ssati = R(0.0);
        s = s || ssati > R(0.0);
      }
    }
    if (__ballot(s) == 0ull) cf |= CF_EXIT;
  }
  if (lane == 0) {
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
    if (!a.f[F_NWFA] || !a.f[F_NIFA]) {
      R h = R(0);
      for (int k = 0; k < nzc; ++k) {
        if (k > 0) h = h + a.f[F_DZ][TEMPO_IDX_STATE(a, c, k)];
        scr.sr[size_t(SR_HGT) * scr.plane + scr.at(c, k)] = h;
      }
    }
    a.cf[c] = cf;
    nzc_s[w] = nzc;
  }
}

// the stages (inlined)
template <class R>
__device__ __forceinline__ void fused_stage_a(const StepArgs<R>& a, const Scr<R>& scr, int c, int k) {
  phase_a_point(a, scr, c, k);
}
template <class R>
__device__ __forceinline__ void fused_stage_b(const StepArgs<R>& a, const Scr<R>& scr, int c0, double* tile_mem) {
  phase_b_group(a, scr, c0, tile_mem);
}
template <class R>
__device__ __forceinline__ void fused_stage_c(const StepArgs<R>& a, const Scr<R>& scr, int c, int k) {
  phase_c_point(a, scr, c, k);
}

// register budget of the fused kernel: as phase A's for FP32 (four waves
// per SIMD), the compiler's choice otherwise
template <class R>
struct FusedWaves {
  static constexpr int n = 1;
};
template <>
struct FusedWaves<float> {
  static constexpr int n = 4;
};

// a.sr, a.sd and a.pf hold gridDim.x slots of fused_slot_plane(nz)-element
// planes (SR_N, SD_N and 1 of them per slot); a.queue is zero at launch
// A fresh view of the kernel's arguments for each stage: loads through it
// are not merged with another stage's, so the uniform operands one stage
// loads (field pointers, constants) are not kept live through the others.
template <class R>
__device__ __forceinline__ const StepArgs<R>& fresh(const StepArgs<R>& a) {
  auto k = __builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(k));
  return *(const StepArgs<R>*)(k);
}

template <class R>
__global__ void __launch_bounds__(kFusedThreads) __attribute__((amdgpu_waves_per_eu(FusedWaves<R>::n)))
k_step_fused(StepArgs<R> a) {
  extern __shared__ double tile_mem[];  // sed_lds_bytes(nz)
  __shared__ int item_s;
  __shared__ int nzc_s[kSedCols];
  const int tid = int(threadIdx.x), w = tid >> 6;
  const size_t sp = a.plane;
  const size_t b = blockIdx.x;
  Scr<R> scr{a.sr + b * SR_N * sp, a.sd + b * SD_N * sp, a.pf + b * sp, sp, 0, a.nzs};
  const int ngroups = (a.ncol + kSedCols - 1) / kSedCols;
  for (;;) {
    // the barrier also ends the previous group's final update before this
    // group's writes reuse the slot
    if (tid == 0) item_s = atomicAdd(fresh(a).queue, 1);
    __syncthreads();
    const int g = item_s;
    if (g >= ngroups) break;  // block-uniform: every wavefront leaves
    const int c0 = g * kSedCols;
    scr.c0 = c0;
    prepass_wave(fresh(a), scr, c0 + w, nzc_s, w);
    __syncthreads();
    // lanes over 8 columns x 8 levels: thread t takes column c0 + t % 8 and
    // level r*64 + t / 8
#pragma unroll 1
    for (int r = 0; r < kSedChunks; ++r) {
      const int cc = tid & (kSedCols - 1), k = r * (kFusedThreads / kSedCols) + (tid / kSedCols);
      if (k < nzc_s[cc]) fused_stage_a(fresh(a), scr, c0 + cc, k);
    }
    __syncthreads();
    fused_stage_b(fresh(a), scr, c0, tile_mem);
    __syncthreads();
#pragma unroll 1
    for (int r = 0; r < kSedChunks; ++r) {
      const int cc = tid & (kSedCols - 1), k = r * (kFusedThreads / kSedCols) + (tid / kSedCols);
      if (k < nzc_s[cc]) fused_stage_c(fresh(a), scr, c0 + cc, k);
    }
  }
}

template <class R>
__global__ void __launch_bounds__(64) k_fused_reset(int* queue) {
  if (threadIdx.x == 0) *queue = 0;
}

// blocks of one fused launch: as many as are resident on the device at once
// (never more than the column groups)
template <class R>
int fused_grid(int ncol, int nz) {
  static int per_cu[2][kSedLevels + 1] = {};
  int& n = per_cu[sizeof(S_t<R>) == 8 ? 1 : 0][nz];
  if (n == 0) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_step_fused<R>, kFusedThreads, sed_lds_bytes(nz)) !=
            hipSuccess ||
        blocks < 1)
      blocks = 1;
    n = blocks;
  }
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
  const int groups = (ncol + kSedCols - 1) / kSedCols;
  return max(1, min(groups, n * cus));
}

template <class R>
hipError_t launch_step_fused_impl(const StepArgs<R>& a, int grid, hipStream_t s, hipEvent_t* ev) {
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_fused_reset<R>, dim3(1), dim3(64), 0, s, a.queue);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_step_fused<R>, dim3(unsigned(grid)), dim3(kFusedThreads), sed_lds_bytes(a.nz), s, a);
  if (ev) {
    (void)hipEventRecord(ev[2], s);
    (void)hipEventRecord(ev[3], s);
    (void)hipEventRecord(ev[4], s);
  }
  return hipGetLastError();
}

}  // namespace tempo

/* tempo_hip.h -- C ABI of the MI355X-native TEMPO column microphysics step.
 *
 * Drop-in for the reference driver's public procedures
 * (src/module_mp_tempo_driver.F90:14): tempo_init, tempo_run and the
 * lookup-table files they read.  All entry points are plain C: pointers,
 * sizes and status codes, no C++ or framework types.  A NULL pointer for an
 * optional field reproduces an absent optional argument of the reference
 * (present() tests, src/module_mp_tempo_main.F90:328,364,388,446,...).
 *
 * Thread safety: one context per device; calls on one context are
 * stream-ordered and must not overlap from several host threads.
 */
#ifndef TEMPO_HIP_H
#define TEMPO_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference has none: it aborts with `error stop`) */
#define TEMPO_HIP_OK 0
#define TEMPO_HIP_E_ARG -1        /* invalid argument or missing required field */
#define TEMPO_HIP_E_UNSUPPORTED -2 /* option not implemented by this library */
#define TEMPO_HIP_E_TABLE -3      /* table file missing or wrong size */
#define TEMPO_HIP_E_HIP -4        /* HIP runtime error */
#define TEMPO_HIP_E_NOMEM -5      /* device allocation failed */

/* working precision of the state arrays passed to tempo_hip_run */
#define TEMPO_HIP_FP32 4
#define TEMPO_HIP_FP64 8

/* table identifiers for tempo_hip_table_export */
#define TEMPO_HIP_TABLE_QR_ACR_QG 1   /* 5 x (37,37,9,37,37) real64, file order */
#define TEMPO_HIP_TABLE_QR_ACR_QS 2   /* 12 x (37,9,37,37) real64 */
#define TEMPO_HIP_TABLE_FREEZEWATER 3 /* 4 x (37,37,45,55) + 2 x (37,100,45,55) real64 */

typedef struct tempo_hip_ctx tempo_hip_ctx;

/* Run configuration: the fields of ty_tempo_cfgs
 * (src/module_mp_tempo_cfgs.F90:14-38) plus the tempo_init flag overrides
 * (src/module_mp_tempo_driver.F90:42-79).  Fill with tempo_hip_cfg_default. */
typedef struct {
  int aerosolaware_flag;          /* selects the aerosol-aware constants; physics follows field presence */
  int hailaware_flag;             /* hail-aware graupel fall-speed constants */
  int ml_for_bl_nc_flag;          /* cloud-number network loaded (for qc_bl/qcfrac_bl) */
  int ml_for_nc_flag;             /* one-moment cloud number from the network (nc absent) */
  int semi_sedi_flag;             /* semi-Lagrangian rain and graupel sedimentation (tempo_hip_run only) */
  int refl10cm_from_melting_flag; /* reflectivity of melting snow and graupel below the melting level */
  int turn_off_micro_flag;
  int cloud_condensation_flag;
  int verbose;
  int check_tables;               /* verify table file sizes before reading */
  int nc_option;                  /* single_moment_nc_opt: 0 'land', 1 'ocean', 2 numeric */
  double nc_value;                /* used when nc_option == 2 [m^-3] */
  int refl10cm_flag;
  int re_cloud_flag;
  int re_ice_flag;
  int re_snow_flag;
  int max_hail_diameter_flag;
  int rain_med_vol_diam_flag;
  int graupel_med_vol_diam_flag;
  int cloud_number_mixing_ratio_flag;
} tempo_hip_cfg;

/* State of tempo_run (src/module_mp_tempo_driver.F90:207-241), device
 * pointers in the precision given to tempo_hip_run.  Either t, or th and pii,
 * must be given.  p, w, dz, qv, qc, qr, qi, qs, qg, ni, nr are required;
 * nc, nwfa, nifa, ng, qb and land (int per column) are optional (NULL).
 * qc_bl and qcfrac_bl (inputs, laid out like the state) are the boundary
 * layer's cloud water and cloud fraction: given together, with the
 * cloud-number network on, they enter the cloud effective radius
 * (src/module_mp_tempo_main.F90:839-857). */
typedef struct {
  void *t, *th, *pii, *p, *w, *dz;
  void *qv, *qc, *qr, *qi, *qs, *qg, *ni, *nr;
  void *nc, *nwfa, *nifa, *ng, *qb;
  const int *land;
  const void *qc_bl, *qcfrac_bl;
} tempo_hip_fields;

/* ty_tempo_driver_diags (src/module_mp_tempo_driver.F90:18-34); any entry may
 * be NULL.  2-D entries have one value per column, 3-D one per level. */
typedef struct {
  void *rain_precip, *ice_liquid_equiv_precip, *snow_liquid_equiv_precip, *graupel_liquid_equiv_precip;
  void *frozen_fraction, *frz_rain_precip, *max_hail_diameter_sfc, *max_hail_diameter_column;
  void *refl10cm, *re_cloud, *re_ice, *re_snow, *rain_med_vol_diam, *graupel_med_vol_diam;
  void *cloud_number_mixing_ratio;
} tempo_hip_diags;

/* Array geometry.  Element (i, k, j) of a state array is at
 * i + k*ldk + j*ldj (Fortran (ims:ime, kms:kme, jms:jme) with the base
 * pointer at the tile origin); of a 3-D diagnostic at i + k*dldk + j*dldj;
 * of a 2-D diagnostic at i + j*d2ldj.  Column index c = i + j*ni.
 * nlev (device, ni*nj ints, may be NULL) gives each column's active level
 * count (levels k >= nlev[c] are not touched).  An empty batch (ni or nj
 * zero) is a no-op: the step entry points and the surface emissions return
 * TEMPO_HIP_OK without reading the arrays, as the reference's loops over an
 * empty tile do nothing. */
typedef struct {
  int ni, nj, nz;
  long long ldk, ldj;
  long long dldk, dldj;
  long long d2ldj;
  const int *nlev;
} tempo_hip_dims;

/* Fills the defaults of ty_tempo_cfgs (src/module_mp_tempo_cfgs.F90:15-34). */
void tempo_hip_cfg_default(tempo_hip_cfg *cfg);

/* The run-time flags of later tempo_hip_run / tempo_hip_run_host calls, as
 * a host changes its ty_tempo_cfgs between tempo_run calls (tempo_run reads
 * tempo_cfgs on every call, src/module_mp_tempo_driver.F90:187-205).
 * aerosolaware_flag and hailaware_flag select constants at tempo_hip_init
 * and must keep their values (E_ARG otherwise). */
int tempo_hip_set_run_cfg(tempo_hip_ctx *ctx, const tempo_hip_cfg *cfg);

/* Replaces tempo_init (src/module_mp_tempo_driver.F90:42-182): computes the
 * constants and init-time tables, and provides the three collection/freezing
 * tables -- read from the reference files in table_dir
 * (read_table_*, :532-622), or, when table_dir is NULL, built on the GPU in
 * FP64 (replacing tools/build_tables.F90).  ccn_path names the CCN
 * activation table (read_table_ccn, :625-649).  Uses the current device. */
int tempo_hip_init(const tempo_hip_cfg *cfg, const char *table_dir, const char *ccn_path,
                   tempo_hip_ctx **ctx);

/* Pre-allocates the per-point scratch for batches of up to ncol columns of
 * nz levels, so that tempo_hip_run allocates nothing (capture-safe). */
int tempo_hip_reserve(tempo_hip_ctx *ctx, int ncol, int nz);

/* Replaces tempo_run (src/module_mp_tempo_driver.F90:187-513): one
 * microphysics step of every column, state updated in place, diagnostics
 * written; stream-ordered on `stream` (a hipStream_t, NULL = default). */
int tempo_hip_run(tempo_hip_ctx *ctx, const tempo_hip_fields *f, const tempo_hip_diags *d,
                  const tempo_hip_dims *dims, double dt, int itimestep, int precision, void *stream);

/* tempo_hip_run without allocation (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The step's per-point
 * scratch comes from a static device array of 256 MiB per GPU instead of
 * the context's buffers; batches needing more are rejected with
 * TEMPO_HIP_E_ARG (per point, levels rounded up to 16 and columns to 8:
 * 188 bytes in FP32 -- 348 where the lane-per-column sedimentation runs --
 * and 348 in FP64, plus 88 per column; about 14,800 FP32 columns of 90
 * levels).  Results equal tempo_hip_run's bit for bit.
 * Every call on one GPU shares that array: calls that overlap on a GPU
 * (several streams or host threads) race, and must be serialised by the
 * caller.  No scratch needs clearing: every value the step reads from it was
 * written earlier in the same step.  Call with the context's device
 * current. */
int tempo_hip_run_noalloc(tempo_hip_ctx *ctx, const tempo_hip_fields *f, const tempo_hip_diags *d,
                          const tempo_hip_dims *dims, double dt, int itimestep, int precision, void *stream);

/* Branch accounting (no counterpart in the reference): the step of
 * tempo_hip_run, bit for bit, that also records per point 32-bit hashes of
 * the threshold decisions its level-local physics took, in two classes:
 * q = 0 every comparison of an `if` condition on a value; q = 1 the discrete
 * outcomes -- every table index, nint and bin count, and the checks'
 * presence flags (l_q*), number clamps and graupel density index.  sig
 * (device, 6*nz*ni*nj words): sig[(2p + q)*nz*ncol + k*ncol + c] for point
 * (column c = i + j*ni, level k), p = 0 the physics up to sedimentation
 * (src/module_mp_tempo_main.F90:320-588), p = 1 the final update
 * (:729-863), p = 2 the semi-Lagrangian sedimentation (:1562-1749; its
 * interface-speed replacements and limiters, monotone limiter and clamps
 * as comparisons, each level's remap source cells, the surface-flux cells
 * and the checks between substeps as outcomes; 0 in flux form, whose
 * decisions are per column); levels k >= nlev[c] are not written.  Two
 * runs took the same decisions of a class at a point where its words of
 * that class agree; the sedimentation's substep counts and tops follow
 * from tempo_hip_substeps. */
int tempo_hip_branch_signature(tempo_hip_ctx *ctx, const tempo_hip_fields *f, const tempo_hip_diags *d,
                               const tempo_hip_dims *dims, double dt, int itimestep, int precision, unsigned *sig,
                               void *stream);

/* As tempo_hip_run with host arrays (FP64, Fortran layout of `dims`): copies
 * to the device, runs and copies back; for hosts that keep state on the CPU
 * (tempo_run, src/module_mp_tempo_driver.F90:187-279, on host memory).
 * land and nlev are host pointers here.  The step runs as a pipeline over
 * slabs of columns (ranges of j, or of i when nj = 1; TEMPO_HIP_TUNE_HOST_*):
 * host threads copy a slab's inputs into pinned memory of the context (not
 * for arrays already in page-locked memory, tempo_hip_host_alloc), a copy
 * stream moves them up, the slab is stepped, a second copy stream moves its
 * results down while later slabs go up, and host threads copy them into the
 * caller's arrays.  Only the fields the step updates come back, and the
 * diagnostics go in only where the step leaves points of their spans
 * unwritten (padded rows, levels k >= nlev[c]).  Diagnostic strides must not
 * overlap (dldk >= ni, dldj >= dldk*nz, d2ldj >= ni where nj > 1; E_ARG).
 * After a step in more than one slab, tempo_hip_substeps has no record.
 * Returns when the caller's arrays hold the results. */
int tempo_hip_run_host(tempo_hip_ctx *ctx, const tempo_hip_fields *f, const tempo_hip_diags *d,
                       const tempo_hip_dims *dims, double dt, int itimestep);
/* The same, stepped in `precision` on the device (TEMPO_HIP_FP32: on FP32
 * copies of the caller's FP64 arrays, converted back at the points the step
 * writes; halos and padding keep their FP64 values). */
int tempo_hip_run_host_prec(tempo_hip_ctx *ctx, const tempo_hip_fields *f, const tempo_hip_diags *d,
                            const tempo_hip_dims *dims, double dt, int itimestep, int precision);

/* Page-locked host memory (hipHostMalloc), and its release; no counterpart
 * in the reference.  A host that keeps its state and diagnostic arrays in
 * such memory lets tempo_hip_run_host(_prec) move them between host and
 * device directly, without the staging copies through the context's pinned
 * buffers (each array is checked, so pageable and page-locked arrays can be
 * mixed).  *ptr is NULL after an error. */
int tempo_hip_host_alloc(size_t bytes, void **ptr);
int tempo_hip_host_free(void *ptr);

/* Replaces tempo_aerosol_surface_emissions
 * (src/module_mp_tempo_driver.F90:516-529): nwfa(i, k, j) += nwfa2d(i, j) * dt
 * at level k (0-based: kts - kms) of every column described by dims (ni, nj,
 * ldk, ldj; nz bounds k), the reference looping over the whole memory extent
 * ims:ime, jms:jme.  nwfa2d element (i, j) at i + j*d2ldj.  Device arrays in
 * the given precision, stream-ordered. */
int tempo_hip_aerosol_surface_emissions(tempo_hip_ctx *ctx, double dt, void *nwfa, const void *nwfa2d,
                                        const tempo_hip_dims *dims, int k, int precision, void *stream);

/* As tempo_hip_aerosol_surface_emissions with FP64 host arrays. */
int tempo_hip_aerosol_surface_emissions_host(tempo_hip_ctx *ctx, double dt, double *nwfa, const double *nwfa2d,
                                             const tempo_hip_dims *dims, int k);

/* The public check-and-update routines a host calls outside the step
 * (src/module_mp_tempo_main.F90:18-20; FV3 recomputes the effective-radius
 * inputs after advection with them), over n independent points of device
 * arrays in the given precision.  Arguments follow the reference:
 *   cloud_check_and_update (:891-973): rho in; qc1d, qcten, ncten inout;
 *     nc1d inout or NULL (one-moment cloud), ncsave in or NULL (one-moment
 *     numbers; NULL: the land value); l_qc (0/1), rc, nc, ilamc, mvd_c out;
 *   ice_check_and_update (:1045-1107): rho in; qi1d, ni1d, qiten, niten
 *     inout; l_qi, ri, ni, ilami out;
 *   snow_check_and_update (:1110-1134): rho in; qs1d, qsten inout; l_qs,
 *     rs out.
 * ilamc and ilami are in the working precision (FP64 in the reference). */
int tempo_hip_cloud_check_and_update(tempo_hip_ctx *ctx, int n, const void *rho, int *l_qc, void *qc1d, void *nc1d,
                                     const void *ncsave, void *rc, void *nc, void *qcten, void *ncten, void *ilamc,
                                     void *mvd_c, double dt, int precision, void *stream);
int tempo_hip_ice_check_and_update(tempo_hip_ctx *ctx, int n, const void *rho, int *l_qi, void *qi1d, void *ni1d,
                                   void *ri, void *ni, void *qiten, void *niten, void *ilami, double dt, int precision,
                                   void *stream);
int tempo_hip_snow_check_and_update(tempo_hip_ctx *ctx, int n, const void *rho, int *l_qs, void *qs1d, void *rs,
                                    void *qsten, double dt, int precision, void *stream);
/* The same on FP64 host arrays. */
int tempo_hip_cloud_check_and_update_host(tempo_hip_ctx *ctx, int n, const double *rho, int *l_qc, double *qc1d,
                                          double *nc1d, const double *ncsave, double *rc, double *nc, double *qcten,
                                          double *ncten, double *ilamc, double *mvd_c, double dt);
int tempo_hip_ice_check_and_update_host(tempo_hip_ctx *ctx, int n, const double *rho, int *l_qi, double *qi1d,
                                        double *ni1d, double *ri, double *ni, double *qiten, double *niten,
                                        double *ilami, double dt);
int tempo_hip_snow_check_and_update_host(tempo_hip_ctx *ctx, int n, const double *rho, int *l_qs, double *qs1d,
                                         double *rs, double *qsten, double dt);

/* Tangent-linear step (no counterpart in the reference): the step of
 * tempo_hip_run on f (updated in place) together with its derivative in the
 * direction f_tl -- device arrays laid out like f, holding the tangents of
 * the inputs (NULL entries: zero tangent) and overwritten with the tangents
 * of the outputs; d_tl (may be NULL) receives the tangents of the requested
 * diagnostics.  Thresholds, clamps, table indices and substep counts take
 * their forward branches (branch-frozen derivatives; with semi_sedi_flag
 * also the semi-Lagrangian remap's source cells, limiters and clamps); in
 * FP64 the values are those of tempo_hip_run bit for bit.  qc_bl and
 * qcfrac_bl of f, when given, are constants of the step.  Returns
 * TEMPO_HIP_E_UNSUPPORTED for a context with refl10cm_from_melting_flag. */
int tempo_hip_tlm(tempo_hip_ctx *ctx, const tempo_hip_fields *f, const tempo_hip_fields *f_tl,
                  const tempo_hip_diags *d, const tempo_hip_diags *d_tl, const tempo_hip_dims *dims, double dt,
                  int itimestep, int precision, void *stream);

/* Adjoint of one step (no counterpart in the reference).  f: the input state
 * (device, not modified; t must be given).  ybar: cotangents of the outputs
 * t, qv, qc, qr, qi, qs, qg, ni, nr, nc, nwfa, nifa, ng, qb (NULL entries:
 * zero); dbar: cotangents of rain_precip, ice_liquid_equiv_precip,
 * snow_liquid_equiv_precip, graupel_liquid_equiv_precip and refl10cm (other
 * entries ignored).  Writes dL/d(input) of t, qv, qc, qr, qi, qs, qg, nc,
 * nr, ni, nwfa, nifa, ng and qb into the matching entries of xbar (NULL
 * entries are skipped and cost nothing), with the branches frozen as in
 * tempo_hip_tlm; p, w, dz enter as constants.  Working memory (about 2.5 kB
 * per column-level) is allocated on first use.  Returns
 * TEMPO_HIP_E_UNSUPPORTED for a context with semi_sedi_flag or
 * refl10cm_from_melting_flag. */
int tempo_hip_adjoint(tempo_hip_ctx *ctx, const tempo_hip_fields *f, const tempo_hip_fields *ybar,
                      const tempo_hip_diags *dbar, const tempo_hip_fields *xbar, const tempo_hip_dims *dims, double dt,
                      int itimestep, int precision, void *stream);

/* Copies a lookup table (FP64, reference file order, native endian) to host
 * memory; n must equal the table's element count. */
int tempo_hip_table_export(tempo_hip_ctx *ctx, int table, double *host_out, size_t n);

/* Writes the three tables in the reference file format (big-endian stream,
 * src/module_mp_tempo_tables.F90:96-444) into dir. */
int tempo_hip_table_write(tempo_hip_ctx *ctx, const char *dir);

/* Milliseconds the GPU table build took at init (0 when read from files). */
double tempo_hip_table_build_ms(tempo_hip_ctx *ctx);

/* Per-column substep counts and sedimentation tops of the last run:
 * out[q*ncol + c], q = rain steps, rain ktop, graupel steps, graupel ktop,
 * snow steps, snow ktop, ice steps, ice ktop, cloud steps, cloud ktop;
 * n = 10*ncol, or 12*ncol for two more rows: 32-bit hashes of the decisions
 * of the semi-Lagrangian rain and graupel sedimentation (semi_sedi_flag;
 * 0 in flux form), equal in two runs that took the same decisions.  E_ARG
 * when there is no record: no step yet, or a host-array step in slabs. */
int tempo_hip_substeps(tempo_hip_ctx *ctx, int *host_out, size_t n);

/* Launch-shape tuning of subsequent runs in one precision (no counterpart in
 * the reference; set by tempo_amd/autotune.py).  Keys and values:
 *   TEMPO_HIP_TUNE_SED_WAVE  1: sedimentation with one wavefront per column
 *                               (default; used when nz <= 128), 0: one lane
 *                               per column, 2: the wavefront kernel issuing
 *                               every load pass of its LDS staging, also
 *                               those past the column group's end (a check
 *                               form of the staging)
 *   TEMPO_HIP_TUNE_BLOCK_A   threads per block of the level-local physics
 *   TEMPO_HIP_TUNE_BLOCK_C   ... and of the final update: 64, 128 or 256
 *   TEMPO_HIP_TUNE_FUSED     1: the whole step as one persistent launch that
 *                               pulls column groups from a work queue and
 *                               keeps their scratch in the caches (for
 *                               flux-form sedimentation on at most 128 levels
 *                               without the melting-layer reflectivity) --
 *                               currently switched off: 1 returns
 *                               TEMPO_HIP_E_UNSUPPORTED;
 *                               0: four launches over the batch (default)
 *   TEMPO_HIP_TUNE_ORDER     point grouping of the level-local kernels, a
 *                               bit mask: bit 0 (level-local physics), bit 1
 *                               (final update) set = each block hands its
 *                               points with microphysics to do to its first
 *                               wavefronts, so that they do not share
 *                               wavefronts with idle points; clear = lane i
 *                               takes column i (default 0)
 *   TEMPO_HIP_TUNE_WAVES_A   1: the level-local physics compiled for one
 *                               wave per SIMD fewer (more registers, fewer
 *                               spills: 3 instead of 4 in FP32, 2 instead
 *                               of 3 in FP64), 0: the default
 *   TEMPO_HIP_TUNE_CCN_LDS   1: each block of the level-local physics
 *                               copies the CCN activation table's slice
 *                               (441 floats) into LDS and reads it there,
 *                               0: read from global memory (default)
 *   TEMPO_HIP_TUNE_HOST_SLABS   host-array steps (tempo_hip_run_host,
 *                               _run_host_prec): slabs of columns of the
 *                               transfer pipeline, 1..4096 (fewer where the
 *                               tile has fewer j rows, or fewer than 256
 *                               columns per slab when nj = 1); 0 = about
 *                               128 MB of uploads per slab, at most 32
 *                               (default)
 *   TEMPO_HIP_TUNE_HOST_THREADS ... and host threads copying between the
 *                               caller's arrays and pinned memory, 1..64
 *                               (default 8)
 * Every setting gives identical results. */
#define TEMPO_HIP_TUNE_SED_WAVE 1
#define TEMPO_HIP_TUNE_BLOCK_A 2
#define TEMPO_HIP_TUNE_BLOCK_C 3
#define TEMPO_HIP_TUNE_ORDER 4
#define TEMPO_HIP_TUNE_FUSED 5
#define TEMPO_HIP_TUNE_HOST_SLABS 6
#define TEMPO_HIP_TUNE_HOST_THREADS 7
#define TEMPO_HIP_TUNE_WAVES_A 8
#define TEMPO_HIP_TUNE_CCN_LDS 9
int tempo_hip_tune(tempo_hip_ctx *ctx, int precision, int key, int value);

/* Per-phase timing of subsequent runs (no counterpart in the reference):
 * enable != 0 records HIP events around the four launches of each run
 * (pre-pass, level-local physics, sedimentation, final update); read
 * returns the accumulated milliseconds per phase (ms[0..3]) and the number
 * of runs, synchronising on the recorded events. */
int tempo_hip_profile(tempo_hip_ctx *ctx, int enable);
int tempo_hip_profile_read(tempo_hip_ctx *ctx, double *ms, int n, int *runs);

/* Message of the last failing call on this thread. */
const char *tempo_hip_last_error(void);

/* Releases the context and its device memory (replaces nothing: the
 * reference never frees its tables). */
void tempo_hip_finalize(tempo_hip_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* TEMPO_HIP_H */

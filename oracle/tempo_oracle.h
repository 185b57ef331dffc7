/* Test oracle: plain-C FP64 restatement of the TEMPO column step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product library never links or calls it.
 *
 * Every function follows the reference file:line cited at its definition,
 * in the reference's operation order (left-to-right products, libm pow()
 * for real exponents as in the -O0 reference build, repeated squaring for
 * integer exponents), so that it can be compared bit-for-bit with the
 * reference FP64 build in oracle/_ref.
 *
 * Field layout: (nz, ncol) row-major, element (k, i) at k*ncol + i, i.e. the
 * Fortran (i, k) layout of tempo_run with k = 0 at the surface.
 */
#ifndef TEMPO_ORACLE_H
#define TEMPO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* flags[0..17]: aerosolaware, hailaware, semi_sedi, refl_from_melting,
 * turn_off_micro, cloud_condensation, refl10cm, re_cloud, re_ice, re_snow,
 * max_hail, mvd_r, mvd_g, nc_mixing_ratio, ml_for_bl_nc, ml_for_nc,
 * nc_opt (0 land, 1 ocean), unused.
 * Tables are flat, in reference file order: ccn float32 (8820),
 * fz/qs/qg float64.  Returns 0, or a negative code for an unsupported
 * option (melting-layer reflectivity). */
int oracle_init(const int *flags, const float *ccn, const double *fz, const double *qs,
                const double *qg);

/* state[0..18]: t, p, w, dz, qv, qc, qr, qi, qs, qg, ni, nr, nc, nwfa, nifa,
 * ng, qb, th, pii (NULL = absent); updated in place.
 * land: per-column int or NULL.  nlev: per-column active level count or NULL.
 * diag2[0..7]: rain, ice, snow, graupel precip, frozen fraction, freezing
 * rain, max hail sfc, max hail column (per column, may be NULL).
 * diag3[0..6]: refl10cm, re_cloud, re_ice, re_snow, mvd_r, mvd_g,
 * nc mixing ratio ((nz, ncol), may be NULL). */
int oracle_run(int ncol, int nz, double dt, int itimestep, double *const *state, const int *land,
               const int *nlev, double *const *diag2, double *const *diag3);
/* Boundary-layer cloud water and fraction ((nz, ncol), NULL = absent) for the
 * following oracle_run calls (tempo_run's qc_bl, qcfrac_bl). */
void oracle_set_bl(const double *qc_bl, const double *qcfrac_bl);

double oracle_gamma_p(double a, double x);
double oracle_rslf(double p, double t);
double oracle_rsif(double p, double t);
void oracle_snow_moments(double rs, double tc, double *out10);
void oracle_surface_emissions(double dt, double *nwfa, const double *nwfa2d, int ni, int nk, int nj, int kts);
int oracle_get_array(int which, double *out, int n);

/* Single lookup-table entries, 1-based reference indices.  Built with the
 * table builder's hail-aware constants (src/module_mp_tempo_tables.F90:50). */
void oracle_qg_entry(int i, int j, int n3, int k, int m, double *out5);
void oracle_qs_entry(int i, int j, int k, int m, double *out12);
void oracle_fz_rain_entry(int i, int j, int k, int m, double *out4);
void oracle_fz_cloud_entry(int i, int j, int k, int m, double *out2);

/* Trace big-table reads into buf (NULL stops tracing); returns the number of
 * entries recorded since the previous call. */
long long oracle_trace(long long *buf, long long cap);
/* sensitivity probe: one in `rate` real powers moved up by one ulp (0: off) */
void oracle_perturb_pow(long long rate);

/* Table-index helpers for the known-answer tests; see tempo_oracle.c. */
int oracle_index(int kind, double a, double b, int *i1, int *i2);

#ifdef __cplusplus
}
#endif
#endif

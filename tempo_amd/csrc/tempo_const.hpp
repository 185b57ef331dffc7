// Constant block of the TEMPO column step, shared by host and device code.
//
// Fixed physical parameters follow src/module_mp_tempo_params.F90:39-167;
// the derived block (gamma-function constants, bins, rate prefactors) is
// computed once on the host in FP64 (tempo_host.cpp, following
// initialize_parameters / initialize_bins_for_tables,
// src/module_mp_tempo_params.F90:411-722) and uploaded to __constant__
// memory in the working precision R, where every lane reads it through the
// scalar cache (all accesses are wave-uniform).
#pragma once

#include <cstddef>
#include <cstdint>

namespace tempo {

// ---- fixed parameters (src/module_mp_tempo_params.F90:39-167) ----
constexpr double kAvR = 4854.0, kBvR = 1.0, kAvS = 40.0, kBvS = 0.55, kFvS = 100.0;
constexpr double kBvC = 2.0, kBvI = 1.0, kAvGOld = 442.0, kBvGOld = 0.89;
constexpr double kAvGNew = 161.794724, kBvGNew = 0.640961647, kFvR = 195.0;
constexpr double kAvC = 0.316946e8, kACoeff = 0.47244157, kBCoeff = 0.54698726, kAvI = 1493.9;
constexpr double kAmS = 0.069, kBmS = 2.0, kBmG = 3.0, kBmI = 3.0, kBmR = 3.0;
constexpr double kRhoI = 890.0, kXm0I = 1.e-12, kD0C = 1.e-6, kD0R = 50.e-6, kD0S = 300.e-6;
constexpr double kD0G = 350.e-6, kD0RMax = 2.5e-3, kCCube = 0.5, kCSqrd = 0.15;
constexpr double kMuR = 0.0, kMuS = 0.6357, kMuG = 0.0, kMuI = 0.0;
constexpr double kNuCScale = 1000.e6;
constexpr double kNaccn0 = 300.0e6, kNaccn1 = 50.0e6, kNain0 = 1.5e6, kNain1 = 0.5e6;
constexpr double kNwfaDefault = 11.1e6, kAeroMax = 9999.e6, kHgFrz = 235.16;
constexpr double kNifaDefault = 0.5e6 * 0.01;
constexpr double kNtCO = 50.e6, kNtCL = 100.e6, kNtCMax = 1999.e6, kNtCMin = 2.0;
constexpr double kTno = 5.0, kAto = 0.304, kRhoS = 100.0;
constexpr double kDemottSsati = 0.25, kDemottTempc = -20.0, kMaxNi = 4999.e3, kIcenucMax = 1000.e3;
constexpr double kRimeThreshold = 2.0, kRimeConversion = 0.95;
constexpr double kEfSI = 0.05, kEfRS = 0.95, kEfRG = 0.75, kEfRI = 0.95;
constexpr double kAutoconNrFactor = 10.0, kDtRimeToRain = 120.0;
constexpr double kMinQv = 1.e-10, kR1 = 1.e-12, kR2 = 1.e-6, kLowLimitPrecip = 1.e-9, kEps = 1.e-15;
constexpr double kM3ToL = 1000.0, kGonvMin = 1.e2, kGonvMax = 1.e6;
constexpr double kSc = 0.632, kGravity = 9.8;
constexpr double kPi = 3.1415926536, kT0 = 273.15, kRhoW = 1000.0, kLsub = 2.834e6, kLvap0 = 2.5e6;
constexpr double kRv = 461.5, kRdry = 287.04, kCp = 1004.0, kRUni = 8.314;
constexpr double kKap0 = 490.6, kKap1 = 17.46, kLam0 = 20.78, kLam1 = 3.29;
constexpr int kIdxBg1 = 6;  // graupel density index when hail-aware is off (500 kg/m3)
constexpr int kNrhg = 9;

// lookup-table dimensions (src/module_mp_tempo_params.F90:170-191)
constexpr int kNbins = 100, kNtbI = 64, kNtbI1 = 55, kNtbC = 37, kNtbT = 9, kNtbR = 37, kNtbR1 = 37;
constexpr int kNtbS = 37, kNtbG = 37, kNtbG1 = 37, kNtbT1 = 45, kNtbIn = 55;
constexpr int kNtbArc = 7, kNtbArw = 9, kNtbArt = 7, kNtbArr = 5, kNtbArk = 4;
constexpr int kNhbins = 50;

// element counts of the reference table files (Fortran order, per array)
constexpr size_t kNQs1 = size_t(kNtbS) * kNtbT * kNtbR1 * kNtbR;             // 455,877
constexpr size_t kNQg1 = size_t(kNtbG1) * kNtbG * kNrhg * kNtbR1 * kNtbR;     // 16,867,449
constexpr size_t kNFzR1 = size_t(kNtbR) * kNtbR1 * kNtbT1 * kNtbIn;           // 3,388,275
constexpr size_t kNFzC1 = size_t(kNtbC) * kNbins * kNtbT1 * kNtbIn;           // 9,157,500
constexpr size_t kNCcn = size_t(kNtbArc) * kNtbArw * kNtbArt * kNtbArr * kNtbArk;
constexpr int kNQsArr = 12, kNQgArr = 5, kNFzRArr = 4, kNFzCArr = 2;

// Derived constants (gamma-function values, prefactors, bins) in precision R.
template <class R>
struct Const {
  R av_g[kNrhg], bv_g[kNrhg], am_g[kNrhg];
  R am_i, am_r, ar_volume, lfus, olfus, roverrv, orv, rho_not, rho_not0, sc3;
  R d0i, obmi, obmr, obmg, oams;
  R cce[5][16], ccg[5][16], ocg1[16], ocg2[16];  // [a-1][n], n = 1..15 (index 0 unused)
  R cie[7], cig[7], oig1, oig2;
  R cre[13], crg[13], ore1, org1, org2, org3;
  R cse[17], csg[17];
  R cge[12][kNrhg], cgg[12][kNrhg], oge1, ogg1, ogg2, ogg3;
  R t1_qr_qc, t1_qr_qi, t2_qr_qi, t1_qs_qc, t1_qs_qi, t1_qr_ev, t2_qr_ev;
  R t1_qs_sd, t2_qs_sd, t1_qs_me, t2_qs_me, t1_qg_sd, t1_qg_me;
  // snow moment regression, with the per-moment polynomial in the second
  // index pre-collected in the reference's evaluation order
  R sa[10], sb[10];
  // bins used by the step
  R dr1, log_dr_ratio, ds1, log_ds_ratio, t_nc1, nic1;
  R hbins[kNhbins], dhbins[kNhbins];
  // decade-index bases: nint(log10(first bin))
  int nic2, nii2, nir2, nis2, nig2, nii3, nir3, nig3, niin2;
  int hail;  // hail-aware graupel constants in use
};

constexpr int kRadarBins = 50;

// Constants of the melting reflectivity, built on the host in FP64
// (tempo_host.cpp, build_radar); used by tempo_radar.hpp.
struct Radar {
  double sbins[kRadarBins], dsbins[kRadarBins];  // snow diameters 100 um .. 2 cm and widths
  double gbins[kRadarBins], dgbins[kRadarBins];  // graupel diameters 100 um .. 5 cm and widths
  double simpson[kRadarBins];                    // Simpson weights of the bins (:226-233)
  double ocmg[kNrhg], ocms, obms;                // oamg^obmg, oams^obms, 1/bm_s
  // the first (ice-water) Maxwell-Garnett mixing's operands: m_w^2, m_i^2
  // and its two shape factors, fixed by the 10-cm indices of water and ice at
  // 0 C (complex_water_ray, complex_ice_maetzler, :144-201)
  double m1_re, m1_im, m3_re, m3_im, b2_re, b2_im, b3_re, b3_im;
  double k_w;  // |(m_w^2 - 1)/(m_w^2 + 2)|^2
};

// Per-run configuration flags (ty_tempo_cfgs, src/module_mp_tempo_cfgs.F90:14-38).
struct RunFlags {
  int aerosolaware, hailaware, semi_sedi, refl_from_melting, turn_off_micro, cloud_condensation;
  int refl10cm, re_cloud, re_ice, re_snow, max_hail, mvd_r, mvd_g, nc_mixing_ratio;
  int ml_for_bl_nc, ml_for_nc;
  double nc_const;  // one-moment cloud number when nc is absent and no land mask is given
};

}  // namespace tempo

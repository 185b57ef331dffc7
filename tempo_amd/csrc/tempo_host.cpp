// Host-side FP64 set-up for the TEMPO column step (compiled with
// -ffp-contract=off so every expression rounds as in the reference FP64
// build): the derived constant block, the small init-time tables, and the
// O(N^3) factor arrays that the GPU table builder consumes.
//
// Libm calls (pow, exp, log, tgamma) are the host's, i.e. the same functions
// the reference build calls, in the reference's operation order.
#include "tempo_host.hpp"

#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <string>

namespace tempo {
namespace {

// Literal decade bins 1..9 x 10^e, parsed from the same decimal literals as
// the reference parameter arrays (src/module_mp_tempo_params.F90:194-263).
void decade_bins(double* out, int n, int e0) {
  for (int idx = 0; idx < n; ++idx) {
    int e = e0 + idx / 9, d = idx % 9 + 1;
    std::string lit = std::to_string(d) + ".e" + std::to_string(e);
    out[idx] = std::strtod(lit.c_str(), nullptr);
  }
}

// create_bins (src/module_mp_tempo_params.F90:695-722)
void log_bins(int numbins, double lowbin, double highbin, double* bins, double* delta) {
  std::vector<double> x(numbins + 1);
  x[0] = lowbin;
  x[numbins] = highbin;
  for (int n = 2; n <= numbins; ++n)
    x[n - 1] = std::exp(double(n - 1) / double(numbins) * std::log(x[numbins] / x[0]) + std::log(x[0]));
  for (int n = 0; n < numbins; ++n) bins[n] = std::sqrt(x[n] * x[n + 1]);
  if (delta)
    for (int n = 0; n < numbins; ++n) delta[n] = x[n + 1] - x[n];
}

int nint(double v) { return int(std::lround(v)); }

// ---- melting-layer reflectivity constants (src/module_mp_tempo_diags.F90) ----
// complex products and quotients formed as the reference's build forms them
// ((ac - bd, ad + bc), the quotient over |b|^2); square root and logarithm
// from the C library, as there
using cxd = std::complex<double>;
static cxd cx_mul(cxd a, cxd b) {
  return cxd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
static cxd cx_div(cxd a, cxd b) {
  const double d = b.real() * b.real() + b.imag() * b.imag();
  return cxd((a.real() * b.real() + a.imag() * b.imag()) / d, (a.imag() * b.real() - a.real() * b.imag()) / d);
}

// complex_water_ray (:144-172) and complex_ice_maetzler (:175-201)
static cxd water_ray(double lambda, double t) {
  const double pi = kPi;
  const double epsinf = 5.27137 + 0.02164740 * t - 0.00131198 * t * t;
  const double epss = 78.54 * (1.0 - 4.579e-3 * (t - 25.0) + 1.190e-5 * (t - 25.0) * (t - 25.0) -
                               2.800e-8 * (t - 25.0) * (t - 25.0) * (t - 25.0));
  const double alpha = -16.8129 / (t + 273.16) + 0.0609265;
  const double lambdas = 0.00033836 * std::exp(2513.98 / (t + 273.16)) * 1e-2;
  const double p1 = std::pow(lambdas / lambda, 1.0 - alpha);
  const double nenner = 1.0 + 2.0 * p1 * std::sin(alpha * pi * 0.5) + std::pow(lambdas / lambda, 2.0 - 2.0 * alpha);
  const double epsr = epsinf + ((epss - epsinf) * (p1 * std::sin(alpha * pi * 0.5) + 1.0)) / nenner;
  const double epsi = ((epss - epsinf) * (p1 * std::cos(alpha * pi * 0.5) + 0.0)) / nenner + lambda * 1.25664 / 1.88496;
  return std::sqrt(cxd(epsr, -epsi));
}
static cxd ice_maetzler(double lambda, double t) {
  const double c = 2.99e8, tk = t + 273.16, f = c / lambda * 1e-9;
  const double b1 = 0.0207, b2 = 1.16e-11, b = 335.0;
  const double deltabeta = std::exp(-10.02 + 0.0364 * (tk - 273.16));
  const double em1 = std::exp(b / tk) - 1.0;
  const double betam = (b1 / tk) * (std::exp(b / tk) / (em1 * em1)) + b2 * f * f;
  const double beta = betam + deltabeta;
  const double theta = 300.0 / tk - 1.0;
  const double alfa = (0.00504 + 0.0062 * theta) * std::exp(-22.1 * theta);
  cxd ri(3.1884 + 9.1e-4 * (tk - 273.16), 0.0);
  ri = ri + cxd(0.0, alfa / f + beta * f);
  return std::sqrt(std::conj(ri));
}


}  // namespace

void build_radar(const Const<double>& c, Radar& r) {
  std::memset(&r, 0, sizeof(r));
  // initialize_bins_for_radar (src/module_mp_tempo_params.F90:674-692)
  log_bins(kRadarBins, 100.e-6, 0.02, r.sbins, r.dsbins);
  log_bins(kRadarBins, 100.e-6, 0.05, r.gbins, r.dgbins);
  // Simpson weights (:226-233): 1/3, 4/3, 1/3 added over consecutive pairs
  double w[kRadarBins + 1] = {};
  for (int n = 0; n < kRadarBins - 1; n += 2) {
    w[n] += 1.0 / 3.0;
    w[n + 1] += 4.0 / 3.0;
    w[n + 2] += 1.0 / 3.0;
  }
  for (int n = 0; n < kRadarBins; ++n) r.simpson[n] = w[n];
  // obms, ocms = oams**obms, ocmg = oamg**obmg (src/module_mp_tempo_params.F90:468-473)
  r.obms = 1.0 / kBmS;
  r.ocms = std::pow(c.oams, r.obms);
  for (int m = 0; m < kNrhg; ++m) r.ocmg[m] = std::pow(1.0 / c.am_g[m], c.obmg);
  // the first Maxwell-Garnett mixing: m1t = m_w^2, m2t = m_air^2 = 1, m3t = m_i^2
  const cxd m_w = water_ray(0.10, 0.0), m_i = ice_maetzler(0.10, 0.0), m_air(1.0, 0.0);
  const cxd m1t = cx_mul(m_w, m_w), m2t = cx_mul(m_air, m_air), m3t = cx_mul(m_i, m_i);
  const cxd two(2.0, 0.0), one(1.0, 0.0);
  const cxd beta2 = cx_mul(cx_div(cx_mul(two, m1t), m2t - m1t), cx_mul(cx_div(m2t, m2t - m1t), std::log(cx_div(m2t, m1t))) - one);
  const cxd beta3 = cx_mul(cx_div(cx_mul(two, m1t), m3t - m1t), cx_mul(cx_div(m3t, m3t - m1t), std::log(cx_div(m3t, m1t))) - one);
  r.m1_re = m1t.real(); r.m1_im = m1t.imag();
  r.m3_re = m3t.real(); r.m3_im = m3t.imag();
  r.b2_re = beta2.real(); r.b2_im = beta2.imag();
  r.b3_re = beta3.real(); r.b3_im = beta3.imag();
  const double kw = std::abs(cx_div(m1t - one, m1t + cxd(2.0, 0.0)));
  r.k_w = kw * kw;
}

Bins make_bins() {
  Bins b;
  decade_bins(b.r_c, 37, -6);
  decade_bins(b.r_i, 64, -10);
  decade_bins(b.r_r, 37, -6);
  decade_bins(b.r_s, 37, -6);
  decade_bins(b.r_g, 37, -6);
  decade_bins(b.n0r_exp, 37, 6);
  decade_bins(b.n0g_exp, 37, 2);
  decade_bins(b.nt_i, 55, 0);
  decade_bins(b.nt_in, 55, 0);
  return b;
}

// initialize_graupel_vars + initialize_parameters + initialize_bins_for_tables
// (src/module_mp_tempo_params.F90:411-658) and the hail-size bins (:661-671).
void build_const(Const<double>& c, HostBins& hb, bool hail, bool cube_as_product) {
  std::memset(&c, 0, sizeof(c));
  const Bins lb = make_bins();
  hb.lit = lb;
  c.hail = hail;
  static const double av_g0[kNrhg] = {45.9173813, 67.0867386, 98.0158463, 122.353378, 143.204224,
                                      161.794724, 178.762115, 194.488785, 209.225876};
  for (int m = 0; m < kNrhg; ++m) {
    c.av_g[m] = av_g0[m];
    c.bv_g[m] = 0.640961647;
  }
  c.av_g[kIdxBg1 - 1] = hail ? kAvGNew : kAvGOld;
  c.bv_g[kIdxBg1 - 1] = hail ? kBvGNew : kBvGOld;

  const double rho_g[kNrhg] = {50., 100., 200., 300., 400., 500., 600., 700., 800.};
  c.am_i = kPi * kRhoI / 6.0;
  c.am_r = kPi * kRhoW / 6.0;
  for (int m = 0; m < kNrhg; ++m) c.am_g[m] = kPi * rho_g[m] / 6.0;
  {
    const double a = 2.5e-6;  // (2.5e-6)**3, folded by the compiler as a product
    c.ar_volume = 4.0 / 3.0 * kPi * (a * a * a);
  }
  c.lfus = kLsub - kLvap0;
  c.olfus = 1.0 / c.lfus;
  c.roverrv = kRdry / kRv;
  c.orv = 1.0 / kRv;
  c.rho_not = 101325.0 / (kRdry * 298.0);
  c.rho_not0 = 101325.0 / (kRdry * kT0);
  c.sc3 = std::pow(kSc, 1.0 / 3.0);
  c.d0i = std::pow(kXm0I / c.am_i, 1.0 / kBmI);
  c.obmi = 1.0 / kBmI;
  c.obmr = 1.0 / kBmR;
  c.oams = 1.0 / kAmS;
  c.obmg = 1.0 / kBmG;
  hb.xm0g = cube_as_product ? c.am_g[kNrhg - 1] * (kD0G * (kD0G * kD0G))
                            : c.am_g[kNrhg - 1] * std::pow(kD0G, kBmG);

  for (int n = 1; n <= 15; ++n) {
    c.cce[0][n] = n + 1.0;
    c.cce[1][n] = kBmR + n + 1.0;
    c.cce[2][n] = kBmR + n + 4.0;
    c.cce[3][n] = n + kBvC + 1.0;
    c.cce[4][n] = kBmR + n + kBvC + 1.0;
    for (int a = 0; a < 5; ++a) c.ccg[a][n] = std::tgamma(c.cce[a][n]);
    c.ocg1[n] = 1.0 / c.ccg[0][n];
    c.ocg2[n] = 1.0 / c.ccg[1][n];
  }
  const double cie[7] = {kMuI + 1.0, kBmI + kMuI + 1.0, kBmI + kMuI + kBvI + 1.0, kMuI + kBvI + 1.0,
                         kMuI + 2.0, kBmI * 0.5 + kMuI + kBvI + 1.0, kBmI * 0.5 + kMuI + 1.0};
  for (int n = 0; n < 7; ++n) {
    c.cie[n] = cie[n];
    c.cig[n] = std::tgamma(cie[n]);
  }
  c.oig1 = 1.0 / c.cig[0];
  c.oig2 = 1.0 / c.cig[1];
  const double cre[13] = {kBmR + 1.0,
                          kMuR + 1.0,
                          kBmR + kMuR + 1.0,
                          kBmR * 2.0 + kMuR + 1.0,
                          kMuR + kBvR + 1.0,
                          kBmR + kMuR + kBvR + 1.0,
                          kBmR * 0.5 + kMuR + kBvR + 1.0,
                          kBmR + kMuR + kBvR + 3.0,
                          kMuR + kBvR + 3.0,
                          kMuR + 2.0,
                          0.5 * (kBvR + 5.0 + 2.0 * kMuR),
                          kBmR * 0.5 + kMuR + 1.0,
                          kBmR * 2.0 + kMuR + kBvR + 1.0};
  for (int n = 0; n < 13; ++n) {
    c.cre[n] = cre[n];
    c.crg[n] = std::tgamma(cre[n]);
  }
  c.ore1 = 1.0 / c.cre[0];
  c.org1 = 1.0 / c.crg[0];
  c.org2 = 1.0 / c.crg[1];
  c.org3 = 1.0 / c.crg[2];
  const double cse[17] = {kBmS + 1.0,
                          kBmS + 2.0,
                          kBmS * 2.0,
                          kBmS + kBvS + 1.0,
                          kBmS * 2.0 + kBvS + 1.0,
                          kBmS * 2.0 + 1.0,
                          kBmS + kMuS + 1.0,
                          kBmS + kMuS + 2.0,
                          kBmS + kMuS + 3.0,
                          kBmS + kMuS + kBvS + 1.0,
                          kBmS * 2.0 + kMuS + kBvS + 1.0,
                          kBmS * 2.0 + kMuS + 1.0,
                          kBvS + 2.0,
                          kBmS + kBvS,
                          kMuS + 1.0,
                          1.0 + (1.0 + kBvS) / 2.0,
                          kBmS + kBvS + 2.0};
  for (int n = 0; n < 17; ++n) {
    c.cse[n] = cse[n];
    c.csg[n] = std::tgamma(cse[n]);
  }
  for (int m = 0; m < kNrhg; ++m) {
    double bv = c.bv_g[m];
    double cge[12] = {kBmG + 1.0,
                      kMuG + 1.0,
                      kBmG + kMuG + 1.0,
                      kBmG * 2. + kMuG + 1.0,
                      kBmG * 2.0 + kMuG + bv + 1.0,
                      kBmG + kMuG + bv + 1.0,
                      kBmG * 0.5 + kMuG + bv + 1.0,
                      kMuG + bv + 1.0,
                      kMuG + bv + 3.0,
                      kMuG + 2.0,
                      0.5 * (bv + 5.0 + 2.0 * kMuG),
                      kBmG * 0.5 + kMuG + 1.0};
    for (int n = 0; n < 12; ++n) {
      c.cge[n][m] = cge[n];
      c.cgg[n][m] = std::tgamma(cge[n]);
    }
  }
  c.oge1 = 1.0 / c.cge[0][0];
  c.ogg1 = 1.0 / c.cgg[0][0];
  c.ogg2 = 1.0 / c.cgg[1][0];
  c.ogg3 = 1.0 / c.cgg[2][0];
  c.t1_qr_qc = kPi * 0.25 * kAvR * c.crg[8];
  c.t1_qr_qi = kPi * 0.25 * kAvR * c.crg[8];
  c.t2_qr_qi = kPi * 0.25 * c.am_r * kAvR * c.crg[7];
  c.t1_qs_qc = kPi * 0.25 * kAvS;
  c.t1_qs_qi = kPi * 0.25 * kAvS;
  c.t1_qr_ev = 0.78 * c.crg[9];
  c.t2_qr_ev = 0.308 * c.sc3 * std::sqrt(kAvR) * c.crg[10];
  c.t1_qs_sd = 0.86;
  c.t2_qs_sd = 0.28 * c.sc3 * std::sqrt(kAvS);
  c.t1_qs_me = kPi * 4.0 * kCSqrd * c.olfus * 0.86;
  c.t2_qs_me = kPi * 4.0 * kCSqrd * c.olfus * 0.28 * c.sc3 * std::sqrt(kAvS);
  c.t1_qg_sd = 0.86 * c.cgg[9][0];
  c.t1_qg_me = kPi * 4.0 * kCCube * c.olfus * 0.86 * c.cgg[9][0];

  static const double sa[10] = {5.065339, -0.062659, -3.032362, 0.029469, -0.000285,
                                0.31255,  0.000204,  0.003199,  0.0,      -0.015952};
  static const double sb[10] = {0.476221, -0.015896, 0.165977, 0.007468, -0.000141,
                                0.060366, 0.000079,  0.000594, 0.0,      -0.003577};
  for (int i = 0; i < 10; ++i) {
    c.sa[i] = sa[i];
    c.sb[i] = sb[i];
  }

  // bins (initialize_bins_for_tables)
  c.nic2 = nint(std::log10(lb.r_c[0]));
  c.nii2 = nint(std::log10(lb.r_i[0]));
  c.nir2 = nint(std::log10(lb.r_r[0]));
  c.nis2 = nint(std::log10(lb.r_s[0]));
  c.nig2 = nint(std::log10(lb.r_g[0]));
  c.nii3 = nint(std::log10(lb.nt_i[0]));
  c.nir3 = nint(std::log10(lb.n0r_exp[0]));
  c.nig3 = nint(std::log10(lb.n0g_exp[0]));
  c.niin2 = nint(std::log10(lb.nt_in[0]));
  hb.dc[0] = kD0C;
  hb.dtc[0] = kD0C;
  for (int n = 1; n < kNbins; ++n) {
    hb.dc[n] = hb.dc[n - 1] + 1.0e-6;
    hb.dtc[n] = hb.dc[n] - hb.dc[n - 1];
  }
  log_bins(kNbins, c.d0i, 2.0 * kD0S, hb.di, hb.dti);
  log_bins(kNbins, kD0R, 0.005, hb.dr, hb.dtr);
  log_bins(kNbins, kD0S, 0.02, hb.ds, hb.dts);
  log_bins(kNbins, kD0G, 0.05, hb.dg, hb.dtg);
  log_bins(kNbins, 1.0, 3000.0, hb.t_nc, nullptr);
  for (int n = 0; n < kNbins; ++n) hb.t_nc[n] = hb.t_nc[n] * 1.0e6;
  c.nic1 = std::log(hb.t_nc[kNbins - 1] / hb.t_nc[0]);
  c.dr1 = hb.dr[0];
  c.log_dr_ratio = std::log(hb.dr[kNbins - 1] / hb.dr[0]);
  c.ds1 = hb.ds[0];
  c.log_ds_ratio = std::log(hb.ds[kNbins - 1] / hb.ds[0]);
  c.t_nc1 = hb.t_nc[0];
  log_bins(kNhbins, 500.e-6, 0.075, c.hbins, c.dhbins);
}

static int get_nuc(double nc) {  // src/module_mp_tempo_utils.F90:334-346
  if ((kNuCScale / nc) >= 12.5) return 15;
  int v = nint(kNuCScale / nc) + 2;
  return v < 15 ? v : 15;
}

static double gamma_p(double a, double x) {  // src/module_mp_tempo_utils.F90:27-145
  if (x < (a + 1.0)) {
    double aj = a, term = 1.0 / std::tgamma(aj + 1.0), sum = term;
    for (int k = 1; k <= 100; ++k) {
      aj = aj + 1.0;
      term = term * x / aj;
      sum = sum + term;
      if (std::fabs(term) < (std::fabs(sum) * 1.e-7)) break;
    }
    return sum * std::pow(x, a) * std::exp(-x);
  }
  const double tiny = 1.e-30;
  double b = 1.0 - a + x, d = 1.0 / b, cc = b + (1.0 / tiny), delta = cc * d, f = tiny * delta;
  for (int k = 1; k <= 100; ++k) {
    double aj = k * (a - k);
    b = b + 2.0;
    d = b + aj * d;
    if (std::fabs(d) < tiny) d = tiny;
    cc = b + aj / cc;
    if (std::fabs(cc) < tiny) cc = tiny;
    d = 1.0 / d;
    delta = cc * d;
    f = f * delta;
    if (std::fabs(delta - 1.0) < 1.e-7) break;
  }
  return 1.0 - std::exp(-x + a * std::log(x)) * f / std::tgamma(a);
}

// Collision efficiency fit shared by rain/snow collecting cloud water
// (Beard & Grover, src/module_mp_tempo_utils.F90:383-394, 425-434).
static double beard_grover(double dcj, double stokes, double p) {
  double reynolds = 9.0 * stokes / (p * p * kRhoW);
  double f = std::log(reynolds);
  double g = -0.1007 - 0.358 * f + 0.0261 * f * f;
  double k0 = std::exp(g);
  double z = std::log(stokes / (k0 + 1.e-15));
  double h = 0.1465 + 1.302 * z - 0.607 * z * z + 0.293 * z * z * z;
  double yc0 = 2.0 / kPi * std::atan(h);
  (void)dcj;
  return (yc0 + p) * (yc0 + p) / ((1. + p) * (1. + p));
}

void build_small_tables(const Const<double>& c, const HostBins& hb, SmallTables& t) {
  const Bins& lb = hb.lit;
  // compute_efrw (src/module_mp_tempo_utils.F90:349-400), stored (i=rain bin, j=cloud bin)
  t.efrw.assign(kNbins * kNbins, 0.0);
  for (int j = 0; j < kNbins; ++j)
    for (int i = 0; i < kNbins; ++i) {
      double dcj = hb.dc[j], dri = hb.dr[i], p = dcj / dri, ef = 0.0;
      if (dri < 50.e-6 || dcj < 3.e-6) {
        ef = 0.0;
      } else if (p > 0.25) {
        double x = dcj * 1.e6;
        if (dri < 75.e-6)
          ef = 0.026794 * x - 0.20604;
        else if (dri < 125.e-6)
          ef = -0.00066842 * x * x + 0.061542 * x - 0.37089;
        else if (dri < 175.e-6)
          ef = 4.091e-06 * x * x * x * x - 0.00030908 * x * x * x + 0.0066237 * x * x - 0.0013687 * x -
               0.073022;
        else if (dri < 250.e-6)
          ef = 9.6719e-5 * x * x * x - 0.0068901 * x * x + 0.17305 * x - 0.65988;
        else if (dri < 350.e-6)
          ef = 9.0488e-5 * x * x * x - 0.006585 * x * x + 0.16606 * x - 0.56125;
        else
          ef = 0.00010721 * x * x * x - 0.0072962 * x * x + 0.1704 * x - 0.46929;
      } else {
        double vtr = -0.1021 + 4.932e3 * dri - 0.9551e6 * dri * dri + 0.07934e9 * dri * dri * dri -
                     0.002362e12 * dri * dri * dri * dri;
        double stokes = dcj * dcj * vtr * kRhoW / (9.0 * 1.718e-5 * dri);
        ef = beard_grover(dcj, stokes, p);
      }
      t.efrw[i + kNbins * j] = std::fmax(0.0, std::fmin(ef, 0.95));
    }
  // compute_efsw (:403-440), stored (i=snow bin, j=cloud bin)
  t.efsw.assign(kNbins * kNbins, 0.0);
  for (int j = 0; j < kNbins; ++j) {
    double dcj = hb.dc[j];
    double vtc = 1.19e4 * (1.0e4 * dcj * dcj * 0.25);
    for (int i = 0; i < kNbins; ++i) {
      double dsi = hb.ds[i];
      double vts = kAvS * std::pow(dsi, kBvS) * std::exp(-kFvS * dsi) - vtc;
      double ds_m = std::pow(kAmS * std::pow(dsi, kBmS) / c.am_r, c.obmr);
      double p = dcj / ds_m;
      if (p > 0.25 || dsi < kD0S || dcj < 6.e-6 || vts < 1.e-3) continue;
      double stokes = dcj * dcj * vts * kRhoW / (9. * 1.718e-5 * ds_m);
      t.efsw[i + kNbins * j] = std::fmax(0.0, std::fmin(beard_grover(dcj, stokes, p), 0.95));
    }
  }
  // compute_drop_evap, tnc_wev(i=diameter bin, j=mass bin, k=number bin) (:493-527)
  t.tnc_wev.assign(size_t(kNbins) * kNtbC * kNbins, 0.0);
  for (int k = 0; k < kNbins; ++k) {
    int nu_c = get_nuc(hb.t_nc[k]);
    for (int j = 0; j < kNtbC; ++j) {
      double lamc = std::pow(hb.t_nc[k] * c.am_r * c.ccg[1][nu_c] * c.ocg1[nu_c] / lb.r_c[j], c.obmr);
      double n0_c = hb.t_nc[k] * c.ocg1[nu_c] * std::pow(lamc, c.cce[0][nu_c]);
      double summ2 = 0.0;
      for (int i = 0; i < kNbins; ++i) {
        double r = 1.0, a = hb.dc[i];
        for (int b = nu_c;;) {  // dc**nu_c by repeated squaring (integer power)
          if (b & 1) r *= a;
          b /= 2;
          if (b == 0) break;
          a *= a;
        }
        summ2 = summ2 + n0_c * r * std::exp(-lamc * hb.dc[i]) * hb.dtc[i];
        t.tnc_wev[i + kNbins * (j + kNtbC * size_t(k))] = summ2;
      }
    }
  }
  // qi_aut_qs (:443-490), tables (i=mass bin 64, j=number bin 55)
  t.tps_iaus.assign(kNtbI * kNtbI1, 0.0);
  t.tni_iaus.assign(kNtbI * kNtbI1, 0.0);
  t.tpi_ide.assign(kNtbI * kNtbI1, 0.0);
  for (int j = 0; j < kNtbI1; ++j)
    for (int i = 0; i < kNtbI; ++i) {
      double lami = std::pow(c.am_i * c.cig[1] * c.oig1 * lb.nt_i[j] / lb.r_i[i], c.obmi);
      double di_mean = (kBmI + kMuI + 1.) / lami;
      double n0_i = lb.nt_i[j] * c.oig1 * std::pow(lami, c.cie[0]);
      double t1 = 0., t2 = 0., ide;
      if (di_mean > 5. * kD0S) {
        t1 = lb.r_i[i];
        t2 = lb.nt_i[j];
        ide = 0.;
      } else if (di_mean < c.d0i) {
        ide = 1.;
      } else {
        ide = gamma_p(kMuI + 2.0, lami * kD0S);
        for (int n = 0; n < kNbins; ++n) {
          double ni = n0_i * std::pow(hb.di[n], kMuI) * std::exp(-lami * hb.di[n]) * hb.dti[n];
          if (hb.di[n] >= kD0S) {
            t1 = t1 + ni * c.am_i * std::pow(hb.di[n], kBmI);
            t2 = t2 + ni;
          }
        }
      }
      t.tps_iaus[i + kNtbI * j] = t1;
      t.tni_iaus[i + kNtbI * j] = t2;
      t.tpi_ide[i + kNtbI * j] = ide;
    }
}

double gamma_p_host(double a, double x) { return gamma_p(a, x); }

// ---- factors for the GPU table builder (hail-aware constants, as the
// reference table build forces, src/module_mp_tempo_tables.F90:31,50) ----

// Rain size distribution for n0r_exp(k), r_r(m): qr_acr_qs :531-536,
// qr_acr_qg :658-663, freezewater :753-761 (same expressions).
static void rain_dist(const Const<double>& c, const HostBins& hb, int k, int m, double* n_r) {
  const Bins& lb = hb.lit;
  double lam_exp = std::pow(lb.n0r_exp[k] * c.am_r * c.crg[0] / lb.r_r[m], c.ore1);
  double lamr = lam_exp * std::pow(c.crg[2] * c.org2 * c.org1, c.obmr);
  double n0_r = lb.n0r_exp[k] / (c.crg[1] * lam_exp) * std::pow(lamr, c.cre[1]);
  for (int n2 = 0; n2 < kNbins; ++n2)
    n_r[n2] = n0_r * std::pow(hb.dr[n2], kMuR) * std::exp(-lamr * hb.dr[n2]) * hb.dtr[n2];
}

static double rain_vt(double d) {
  return -0.1021 + 4.932e3 * d - 0.9551e6 * d * d + 0.07934e9 * d * d * d - 0.002362e12 * d * d * d * d;
}

// host snow_moments for the 2nd and (bm_s+1)th moments only
// (src/module_mp_tempo_utils.F90:148-200)
void snow_m2_m3(const Const<double>& c, double rs, double tc, double& smob, double& smoc) {
  const double* sa = c.sa;
  const double* sb = c.sb;
  smob = rs * c.oams;
  double loga = sa[0] + sa[1] * tc + sa[2] * kBmS + sa[3] * tc * kBmS + sa[4] * tc * tc + sa[5] * kBmS * kBmS +
                sa[6] * tc * tc * kBmS + sa[7] * tc * kBmS * kBmS + sa[8] * tc * tc * tc + sa[9] * kBmS * kBmS * kBmS;
  double a = std::pow(10.0, loga);
  double b = sb[0] + sb[1] * tc + sb[2] * kBmS + sb[3] * tc * kBmS + sb[4] * tc * tc + sb[5] * kBmS * kBmS +
             sb[6] * tc * tc * kBmS + sb[7] * tc * kBmS * kBmS + sb[8] * tc * tc * tc + sb[9] * kBmS * kBmS * kBmS;
  double smo2 = std::pow(smob / a, 1.0 / b);
  double e = c.cse[0];
  loga = sa[0] + sa[1] * tc + sa[2] * e + sa[3] * tc * e + sa[4] * tc * tc + sa[5] * e * e + sa[6] * tc * tc * e +
         sa[7] * tc * e * e + sa[8] * tc * tc * tc + sa[9] * e * e * e;
  a = std::pow(10.0, loga);
  b = sb[0] + sb[1] * tc + sb[2] * e + sb[3] * tc * e + sb[4] * tc * tc + sb[5] * e * e + sb[6] * tc * tc * e +
      sb[7] * tc * e * e + sb[8] * tc * tc * tc + sb[9] * e * e * e;
  smoc = a * std::pow(smo2, b);
}

void build_table_factors(TableFactors& f) {
  Const<double> c;
  HostBins hb;
  // the reference table file is produced by its -O2 build, where bin cubes
  // d**3 are evaluated as d*(d*d); mirror that so the GPU tables match it
  build_const(c, hb, /*hail=*/true, /*cube_as_product=*/true);
  const Bins& lb = hb.lit;
  auto cube = [](double d) { return d * (d * d); };

  // rain distributions n_r[(m*37 + k)*100 + n2] for (n0r_exp(k), r_r(m))
  f.n_r.assign(size_t(kNtbR) * kNtbR1 * kNbins, 0.0);
  for (int m = 0; m < kNtbR; ++m)
    for (int k = 0; k < kNtbR1; ++k) rain_dist(c, hb, k, m, &f.n_r[(size_t(m) * kNtbR1 + k) * kNbins]);

  std::vector<double> vr(kNbins), massr(kNbins);
  for (int n2 = 0; n2 < kNbins; ++n2) {
    vr[n2] = rain_vt(hb.dr[n2]);
    massr[n2] = c.am_r * cube(hb.dr[n2]);
  }

  // ---- rain-graupel (qr_acr_qg, src/module_mp_tempo_tables.F90:624-714) ----
  // kernel factors A[q][n3][n2][n], q = t1,z1,y1,t2,y2; term = (A*n_g(n))*n_r(n2)
  f.qg_a.assign(size_t(5) * kNrhg * kNbins * kNbins, 0.0);
  f.qg_ng.assign(size_t(kNrhg) * kNtbG * kNtbG1 * kNbins, 0.0);
  const double cg = kPi * .25 * kEfRG;
  for (int n3 = 0; n3 < kNrhg; ++n3) {
    for (int n2 = 0; n2 < kNbins; ++n2)
      for (int n = 0; n < kNbins; ++n) {
        double vg = c.av_g[n3] * std::pow(hb.dg[n], c.bv_g[n3]);
        double massg = c.am_g[n3] * cube(hb.dg[n]);
        double dvg = 0.5 * ((vr[n2] - vg) + std::fabs(vr[n2] - vg));
        double dvr = 0.5 * ((vg - vr[n2]) + std::fabs(vg - vr[n2]));
        double s = hb.dg[n] + hb.dr[n2];
        double base = cg * s * s;
        size_t o = (size_t(n3) * kNbins + n2) * kNbins + n;
        size_t stride = size_t(kNrhg) * kNbins * kNbins;
        f.qg_a[0 * stride + o] = base * dvg * massg;
        f.qg_a[1 * stride + o] = base * dvg * massr[n2];
        f.qg_a[2 * stride + o] = base * dvg;
        f.qg_a[3 * stride + o] = base * dvr * massr[n2];
        f.qg_a[4 * stride + o] = base * dvr;
      }
    for (int j = 0; j < kNtbG; ++j)
      for (int i = 0; i < kNtbG1; ++i) {
        double lam_exp = std::pow(lb.n0g_exp[i] * c.am_g[n3] * c.cgg[0][0] / lb.r_g[j], c.oge1);
        double lamg = lam_exp * std::pow(c.cgg[2][0] * c.ogg2 * c.ogg1, c.obmg);
        double n0_g = lb.n0g_exp[i] / (c.cgg[1][0] * lam_exp) * std::pow(lamg, c.cge[1][0]);
        double* ng = &f.qg_ng[((size_t(n3) * kNtbG + j) * kNtbG1 + i) * kNbins];
        for (int n = 0; n < kNbins; ++n)
          ng[n] = n0_g * std::pow(hb.dg[n], kMuG) * std::exp(-lamg * hb.dg[n]) * hb.dtg[n];
      }
  }

  // ---- rain-snow (qr_acr_qs, :493-621) ----
  // A[q][n2][n], q = t1,z1,y1 (massr>1.5 masss) | t3,z3,y3 (else) with dvs;
  // t2,y2,z2 | t4,y4,z4 with dvr.  The branch is a property of (n2,n), so
  // the factor of the branch not taken is stored as zero: adding +0 leaves
  // the accumulator bit-identical to the reference's branchy loop.
  const double cs = kPi * .25 * kEfRS;
  f.qs_a.assign(size_t(12) * kNbins * kNbins, 0.0);
  std::vector<double> vs(kNbins);
  for (int n = 0; n < kNbins; ++n) vs[n] = 1.5 * kAvS * std::pow(hb.ds[n], kBvS) * std::exp(-kFvS * hb.ds[n]);
  for (int n2 = 0; n2 < kNbins; ++n2)
    for (int n = 0; n < kNbins; ++n) {
      double masss = kAmS * (hb.ds[n] * hb.ds[n]);
      double dvs = 0.5 * ((vr[n2] - vs[n]) + std::fabs(vr[n2] - vs[n]));
      double dvr = 0.5 * ((vs[n] - vr[n2]) + std::fabs(vs[n] - vr[n2]));
      double s = hb.ds[n] + hb.dr[n2];
      double base = cs * s * s;
      bool big = massr[n2] > 1.5 * masss;
      size_t o = size_t(n2) * kNbins + n, st = size_t(kNbins) * kNbins;
      // accumulator order: t1 z1 y1 t3 z3 y3 t2 y2 z2 t4 y4 z4
      double v[12] = {base * dvs * masss, base * dvs * massr[n2], base * dvs,
                      base * dvs * masss, base * dvs * massr[n2], base * dvs,
                      base * dvr * massr[n2], base * dvr, base * dvr * masss,
                      base * dvr * massr[n2], base * dvr, base * dvr * masss};
      for (int q = 0; q < 12; ++q) {
        bool first = (q < 3) || (q >= 6 && q < 9);
        f.qs_a[q * st + o] = (first == big) ? v[q] : 0.0;
      }
    }
  f.qs_ns.assign(size_t(kNtbT) * kNtbS * kNbins, 0.0);
  static const double tc_tab[kNtbT] = {-0.01, -5., -10., -15., -20., -25., -30., -35., -40.};
  for (int j = 0; j < kNtbT; ++j)
    for (int i = 0; i < kNtbS; ++i) {
      double m2, m3;
      snow_m2_m3(c, lb.r_s[i], tc_tab[j], m2, m3);
      double om3 = 1.0 / m3;
      double mrat = m2 * (m2 * om3) * (m2 * om3) * (m2 * om3);
      double m0 = std::pow(m2 * om3, kMuS);
      double slam1 = m2 * om3 * kLam0, slam2 = m2 * om3 * kLam1;
      double* ns = &f.qs_ns[(size_t(j) * kNtbS + i) * kNbins];
      for (int n = 0; n < kNbins; ++n)
        ns[n] = mrat * (kKap0 * std::exp(-slam1 * hb.ds[n]) + kKap1 * m0 * std::pow(hb.ds[n], kMuS) *
                                                                  std::exp(-slam2 * hb.ds[n])) *
                hb.dts[n];
    }

  // ---- freezing (freezewater, :717-805) ----
  const double orho_w = 1.0 / kRhoW;
  f.fz_massr = massr;
  f.fz_xm0g = hb.xm0g;
  f.fz_prob_r.assign(size_t(kNtbIn) * kNtbT1 * kNbins, 0.0);
  f.fz_prob_c.assign(size_t(kNtbIn) * kNtbT1 * kNbins, 0.0);
  f.fz_massc.assign(kNbins, 0.0);
  for (int n = 0; n < kNbins; ++n) f.fz_massc[n] = c.am_r * cube(hb.dc[n]);
  for (int m = 0; m < kNtbIn; ++m) {
    double t_adjust = std::fmax(-3.0, std::fmin(3.0 - std::log10(lb.nt_in[m]), 3.0));
    for (int k = 0; k < kNtbT1; ++k) {
      double texp = std::exp(double(k + 1) - t_adjust) - 1.0;
      for (int n = 0; n < kNbins; ++n) {
        double volr = massr[n] * orho_w, volc = f.fz_massc[n] * orho_w;
        f.fz_prob_r[(size_t(m) * kNtbT1 + k) * kNbins + n] =
            std::fmax(0.0, 1.0 - std::exp(-120.0 * volr * 5.2e-4 * texp));
        f.fz_prob_c[(size_t(m) * kNtbT1 + k) * kNbins + n] =
            std::fmax(0.0, 1.0 - std::exp(-120.0 * volc * 5.2e-4 * texp));
      }
    }
  }
  // cloud drop distributions n_c[(j*37 + i)*100 + n] for t_nc(j), r_c(i)
  f.fz_nc.assign(size_t(kNbins) * kNtbC * kNbins, 0.0);
  for (int j = 0; j < kNbins; ++j) {
    int nu_c = nint(kNuCScale / hb.t_nc[j]) + 2;
    if (nu_c > 15) nu_c = 15;
    for (int i = 0; i < kNtbC; ++i) {
      double lamc = std::pow(hb.t_nc[j] * c.am_r * c.ccg[1][nu_c] * c.ocg1[nu_c] / lb.r_c[i], c.obmr);
      double n0_c = hb.t_nc[j] * c.ocg1[nu_c] * std::pow(lamc, c.cce[0][nu_c]);
      double* nc = &f.fz_nc[(size_t(j) * kNtbC + i) * kNbins];
      for (int n = 0; n < kNbins; ++n) {
        double r = 1.0, a = hb.dc[n];
        for (int b = nu_c;;) {
          if (b & 1) r *= a;
          b /= 2;
          if (b == 0) break;
          a *= a;
        }
        nc[n] = n0_c * r * std::exp(-lamc * hb.dc[n]) * hb.dtc[n];
      }
    }
  }
  f.fz_t_nc.assign(hb.t_nc, hb.t_nc + kNbins);
  f.r_r.assign(lb.r_r, lb.r_r + kNtbR);
  f.r_c.assign(lb.r_c, lb.r_c + kNtbC);
}

}  // namespace tempo

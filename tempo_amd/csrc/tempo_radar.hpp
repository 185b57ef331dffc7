// Reflectivity of melting snow and graupel (refl10cm_from_melting_flag):
// the binned Rayleigh backscatter of wet particles with Maxwell-Garnett
// mixing of ice, water and air (src/module_mp_tempo_diags.F90:124-395).
// Computed in FP64 in either working precision, as the reference declares
// these routines real(dp).  Only the levels below the column's melting level
// that hold snow or graupel reach it.
#pragma once

#include "tempo_step.hpp"

namespace tempo {
namespace radar {

struct Cx {
  double re, im;
};
__device__ __forceinline__ Cx cx(double r) { return Cx{r, 0.0}; }
__device__ __forceinline__ Cx operator+(Cx a, Cx b) { return Cx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cx operator-(Cx a, Cx b) { return Cx{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cx operator+(Cx a, double b) { return Cx{a.re + b, a.im}; }
__device__ __forceinline__ Cx operator-(Cx a, double b) { return Cx{a.re - b, a.im}; }
// products and quotients as the reference's build forms them: (ac - bd,
// ad + bc) and the quotient over |b|^2
__device__ __forceinline__ Cx mul(Cx a, Cx b) { return Cx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cx div(Cx a, Cx b) {
  const double d = b.re * b.re + b.im * b.im;
  return Cx{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// real functions: correctly rounded in the FP64 step (as the host libm of
// the reference), the device library in the FP32 step
template <bool kCR>
struct Fn {
  __device__ static double pow(double x, double y) { return crm::pow_cr(x, y); }
  __device__ static double exp(double x) { return crm::exp_cr(x); }
  __device__ static double log(double x) { return crm::log_cr(x); }
};
template <>
struct Fn<false> {
  __device__ static double pow(double x, double y) { return ::pow(x, y); }
  __device__ static double exp(double x) { return ::exp(x); }
  __device__ static double log(double x) { return ::log(x); }
};

// principal square root and logarithm
template <bool kCR>
__device__ Cx csqrt(Cx z) {
  if (z.re == 0.0 && z.im == 0.0) return Cx{0.0, z.im};
  const double t = ::sqrt(0.5 * (::fabs(z.re) + ::hypot(z.re, z.im)));
  if (z.re >= 0.0) return Cx{t, z.im / (2.0 * t)};
  return Cx{::fabs(z.im) / (2.0 * t), ::copysign(t, z.im)};
}
template <bool kCR>
__device__ Cx clog(Cx z) {
  return Cx{Fn<kCR>::log(::hypot(z.re, z.im)), ::atan2(z.im, z.re)};
}

// one Maxwell-Garnett mixing with operands m1t, m2t, m3t and volume
// fractions vol2, vol3 (:355-363, 372-380); b2, b3 given when precomputed
template <bool kCR>
__device__ Cx mg_mix(Cx m1t, Cx m2t, Cx m3t, double vol2, double vol3, const Cx* b2 = nullptr, const Cx* b3 = nullptr) {
  const Cx beta2 = b2 ? *b2 : mul(div(mul(cx(2.0), m1t), m2t - m1t), mul(div(m2t, m2t - m1t), clog<kCR>(div(m2t, m1t))) - 1.0);
  const Cx beta3 = b3 ? *b3 : mul(div(mul(cx(2.0), m1t), m3t - m1t), mul(div(m3t, m3t - m1t), clog<kCR>(div(m3t, m1t))) - 1.0);
  const Cx num = mul(cx(1.0 - vol2 - vol3), m1t) + mul(mul(cx(vol2), beta2), m2t) + mul(mul(cx(vol3), beta3), m3t);
  const Cx den = cx(1.0 - vol2 - vol3) + mul(cx(vol2), beta2) + mul(cx(vol3), beta3);
  return csqrt<kCR>(div(num, den));
}

// rayleigh_soak_wetgraupel (:309-395) at 10 cm with melt water 0.9 outside
template <bool kCR>
__device__ double backscatter(const Radar& Q, double x_g, double a_geo, double b_geo, double fmelt) {
  using F = Fn<kCR>;
  constexpr double lambda_radar = 0.10, pi = kPi;
  const double fm = ::fmax(::fmin(fmelt, 1.0), 0.0);
  double mra = 0.9;
  mra = mra + (1.0 - mra) * fm;
  const double x_w = x_g * fm;
  const double d_g = a_geo * F::pow(x_g, b_geo);
  if (!(d_g >= kR1)) return 0.0;
  double vg = pi / 6.0 * (d_g * d_g * d_g);  // d**3 by repeated squaring: d * d^2
  const double rhog = ::fmax(::fmin(x_g / vg, 900.0), 10.0);
  vg = x_g / rhog;
  double volg;
  if (mra <= 1.0 - rhog / 1000.0) {
    volg = vg * (1.0 - mra * fm);
  } else {
    const double fmgrenz = (900.0 - rhog) / (mra * 900.0 - rhog + 900.0 * rhog / 1000.0);
    volg = fm <= fmgrenz ? (1.0 - mra * fm) * vg : (x_g - x_w) / 900.0 + x_w / 1000.0;
  }
  const double d_large = F::pow(6.0 / pi * volg, 1.0 / 3.0);
  const double volice = (x_g - x_w) / (volg * 900.0);
  const double volwater = x_w / (1000.0 * volg);
  const double volair = 1.0 - volice - volwater;
  const double volmix1 = volice / ::fmax(volice + volwater, 1e-10);
  const Cx one = Cx{1.0, 0.0};
  const Cx b2{Q.b2_re, Q.b2_im}, b3{Q.b3_re, Q.b3_im};
  const Cx mt = mg_mix<kCR>(Cx{Q.m1_re, Q.m1_im}, one, Cx{Q.m3_re, Q.m3_im}, 0.0, volmix1, &b2, &b3);
  const Cx core = mg_mix<kCR>(mul(mt, mt), one, Cx{4.0, 0.0}, volair, 0.0);
  const Cx c2 = mul(core, core);
  const Cx q = div(c2 - 1.0, c2 + 2.0);
  const double a = ::hypot(q.re, q.im);
  const double d2 = d_large * d_large, d4 = d2 * d2;  // d**6 by repeated squaring: d^4 * d^2
  return a * a * pi * pi * pi * pi * pi * (d4 * d2) / (lambda_radar * lambda_radar * lambda_radar * lambda_radar);
}

// reflectivity_from_melting_graupel (:204-245); am_g, ogg2, cge21 from the
// step's constants
template <bool kCR>
__device__ double melting_graupel(const Radar& Q, double am_g, double ogg2, double cge21, double rg, double ng,
                                  double ilamg, int idx, double rr) {
  using F = Fn<kCR>;
  constexpr double lambda_radar = 0.10, pi = kPi;
  const double sr = ::fmax(0.01, ::fmin(1.0 - rg / ::fmax(rg + rr, kR1), 0.99));
  const double fmelt = sr * sr;
  double eta = 0.0;
  const double lamg = 1.0 / ilamg;
  const double n0_g = ng * ogg2 * F::pow(lamg, cge21);
  for (int n = 0; n < kRadarBins; ++n) {
    const double mass = am_g * F::pow(Q.gbins[n], kBmG);
    const double bsc = backscatter<kCR>(Q, mass, Q.ocmg[idx - 1], 1.0 / kBmG, fmelt);
    const double f_d = n0_g * F::pow(Q.gbins[n], kMuG) * F::exp(-lamg * Q.gbins[n]);
    eta = eta + f_d * bsc * Q.simpson[n] * Q.dgbins[n];
  }
  return lambda_radar * lambda_radar * lambda_radar * lambda_radar / (pi * pi * pi * pi * pi * Q.k_w) * eta;
}

// reflectivity_from_melting_snow (:248-300)
template <bool kCR>
__device__ double melting_snow(const Radar& Q, double rs, double smob, double smoc, double rr) {
  using F = Fn<kCR>;
  constexpr double lambda_radar = 0.10, pi = kPi;
  const double sr = ::fmax(0.01, ::fmin(1.0 - rs / (rs + rr), 0.99));
  const double fmelt = sr * sr;
  double eta = 0.0;
  const double om3 = 1.0 / smoc, m0 = smob * om3, mrat = smob * m0 * m0 * m0;
  const double slam1 = m0 * kLam0, slam2 = m0 * kLam1;
  for (int n = 0; n < kRadarBins; ++n) {
    const double mass = kAmS * F::pow(Q.sbins[n], kBmS);
    const double bsc = backscatter<kCR>(Q, mass, Q.ocms, Q.obms, fmelt);
    const double f_d = mrat * (kKap0 * F::exp(-slam1 * Q.sbins[n]) +
                               kKap1 * F::pow(m0 * Q.sbins[n], kMuS) * F::exp(-slam2 * Q.sbins[n]));
    eta = eta + f_d * bsc * Q.simpson[n] * Q.dsbins[n];
  }
  return lambda_radar * lambda_radar * lambda_radar * lambda_radar / (pi * pi * pi * pi * pi * Q.k_w) * eta;
}

}  // namespace radar
}  // namespace tempo

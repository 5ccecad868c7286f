// Per-lane two-stream physics shared by the kernels (device) and the host
// self-checks. Written for one (column, k-term, sub-column) at a time; the
// kernels call these in their layer loops.
#pragma once

#include <float.h>
#include <math.h>

#include "sg_types.h"

#define SG_HD __host__ __device__ __forceinline__

namespace sg {

constexpr double kElsasser = 1.66;       // diffusivity_factor.F90: elsasser_factor
constexpr double kDiffMinor = 1.66;      // diffusivity_factor_minor
constexpr double kEps = DBL_EPSILON;
constexpr double kTiny = DBL_MIN;

struct Coef {
  double T, R, T0, s1, s2;  // trans, reflect, trans_0, source coeff 1, 2
};

// Two-stream coefficients of one homogeneous layer: two_coeff_basic
// (two_coeff_basic.F90:108-161; PIFM80 for solar, Elsasser for infra-red),
// lambda = sqrt(sum*diff) (two_coeff.F90:158-163), the solar resonance guard
// and direct-beam coupling (solar_coefficient_basic.F90:124-219), and
// transmission, reflection and source coefficients (trans_source_coeff.F90:
// 142-263, quadratic infra-red source).
SG_HD void two_coeff(bool solar, double tau, double om, double g, double sec0, Coef &c) {
  double sum, diff;
  if (solar) {
    sum = 2.0 - 1.5 * om * g - 0.5 * om;
    diff = 2.0 * (1.0 - om);
  } else {
    sum = kElsasser - 1.5 * om * g;
    diff = kElsasser * (1.0 - om);
  }
  double lambda = sqrt(sum * diff);
  double gup = 0.0, gdn = 0.0;
  if (solar) {
    const double tolp = 32.0 * kEps;
    if (fabs(lambda - sec0) < tolp) {
      sum *= 1.0 + tolp;
      diff *= 1.0 + tolp;
      lambda *= 1.0 + tolp;
    }
    double ksi0 = 1.5 * g / sec0;
    double factor = 0.5 * om * sec0 / ((lambda - sec0) * (lambda + sec0));
    gup = factor * (sum - sec0 - ksi0 * (diff - sec0));
    gdn = factor * (sum + sec0 + ksi0 * (diff + sec0));
  }
  const double sqe = 1.4901161193847656e-08;  // sqrt(DBL_EPSILON)
  const double tol = 1.220703125e-04;         // sqrt(sqrt(DBL_EPSILON))
  double e = exp(-lambda * tau);
  double e2 = e * e;
  double gam = (sum - lambda) / (sum + lambda);
  double g2 = gam * gam;
  double inv = 1.0 / (1.0 - e2 * g2);
  c.T = e * (1.0 - g2) * inv;
  c.R = gam * (1.0 - e2) * inv;
  if (solar) {
    c.T0 = exp(-tau * sec0);
    c.s1 = (gup - c.R * (1.0 + gdn)) - gup * c.T * c.T0;
    c.s2 = c.T0 * (1.0 + gdn - gup * c.R) - (1.0 + gdn) * c.T;
  } else {
    c.T0 = 0.0;
    c.s1 = (1.0 - c.T + c.R + sqe) / (sqe + tau * sum);
    double q = tau > tol ? -2.0 * (1.0 - c.T - c.R + sqe) / (diff * tau + sqe)
                         : -2.0 + diff * tau;
    c.s2 = -(1.0 + c.R + c.T + q) / (sum * tau + sqe);
  }
}

// Horner evaluation of the band Planck polynomial (planck_flux_band_mod.F90:33-49)
SG_HD double planck_band(const SgBand &b, int n_deg, double t_ref, double temp) {
  double x = temp / t_ref;
  double v = b.planck[n_deg];
  for (int j = n_deg - 1; j >= 0; j--) v = v * x + b.planck[j];
  return v;
}

// k_abs of one term at one layer: k * max(0, (p/p_ref)^a (T/T_ref)^b * mmr)
// (radiance_calc.F90:1457-1491 with scale_absorb.F90:137-152)
SG_HD double k_term(const SgTerm &t, double lnp, double lnt, double mmr) {
  double s = exp(t.a * (lnp - t.lnpref) + t.b * (lnt - t.lntref));
  return t.k * fmax(0.0, s * mmr);
}

// Pade-2 droplet optics (opt_prop_pade_2_mod.F90:100-143): extinction,
// scattering, first phase-function moment and forward scattering, all
// weighted by the in-cloud mixing ratio.
SG_HD void pade2(const double *cp, double mmr, double re, double &kext, double &kscat,
                 double &phf, double &fwd) {
  kext = mmr * (cp[0] + re * (cp[1] + re * cp[2])) /
         (1.0 + re * (cp[3] + re * (cp[4] + re * cp[5])));
  kscat = kext * (1.0 - (cp[6] + re * (cp[7] + re * cp[8])) / (1.0 + re * (cp[9] + re * cp[10])));
  double g = (cp[11] + re * (cp[12] + re * cp[13])) / (1.0 + re * (cp[14] + re * cp[15]));
  phf = kscat * g;
  fwd = phf * g;
}

// Slingo-Schrecker droplet optics (opt_prop_water_cloud.F90:231-247)
SG_HD void slingo(const double *cp, double mmr, double re, double &kext, double &kscat,
                  double &phf, double &fwd) {
  kext = mmr * (cp[0] + cp[1] / re);
  kscat = kext * (1.0 - cp[2] - cp[3] * re);
  double g = cp[4] + cp[5] * re;
  phf = kscat * g;
  fwd = kscat * g * g;
}

// McICA cloud-generator linear congruential step (socrates_cloud_gen.F90:48-50
// constants, :196-205 step with the FP64 multiply-truncate of the reference)
constexpr int kLcgM = 86436, kLcgA = 1093, kLcgC = 18257;
SG_HD int lcg_step(int x) {
  const double rm = 1.0 / (double)kLcgM;
  int t = x * kLcgA + kLcgC;
  double q = (double)t * rm;
  return t - (int)q * kLcgM;
}

}  // namespace sg

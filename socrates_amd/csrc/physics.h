// Per-lane two-stream physics shared by the kernels (device) and the host
// self-checks. Written for one (column, k-term, sub-column) at a time; the
// kernels call these in their layer loops.
//
// Contraction of a*b+c into fused multiply-adds is fixed per function: off in
// the FP64 forms of the reference's formulas (as the reference is built), on
// within one expression in the mixed form's. Either way the code gives the
// same bits wherever it is inlined, so a layer recomputed by the recomputing
// solve tile, or solved in another kernel variant of a batch, equals its
// first computation.
#pragma once

#include <float.h>
#include <math.h>

#include "sg_types.h"

#define SG_HD __host__ __device__ __forceinline__

namespace sg {

// The library's device pointers address global memory (HBM, or host memory
// mapped for the device), never LDS or private memory. A pointer read from a
// descriptor or table in memory is generic to the compiler: loads through it
// become FLAT instructions, and a FLAT load can only be waited for with
// lgkmcnt(0), which also waits for every other load in flight (a prefetch of
// the next layer included). Seen through G, the pointer is a global one, so
// its loads are global loads, waited for one by one with vmcnt.
template <class T>
SG_HD T *G(T *p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (T *)(__attribute__((address_space(1))) T *)(unsigned long long)p;
#else
  return p;
#endif
}

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
// 142-263, quadratic infra-red source). Evaluated without fused multiply-add
// contraction: the reference's source coefficients cancel (solar near
// lambda = sec(theta_0), infra-red quadratic source for tau near 2^-13), and
// a contracted product there moves fluxes by up to 1e-4 W m-2; unfused, the
// device differs from the reference only by its exp.
SG_HD void two_coeff(bool solar, double tau, double om, double g, double sec0, Coef &c) {
#pragma clang fp contract(off)
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
#pragma clang fp contract(off)
  double x = temp / t_ref;
  double v = b.planck[n_deg];
  for (int j = n_deg - 1; j >= 0; j--) v = v * x + b.planck[j];
  return v;
}

// k_abs of one term at one layer: k * max(0, (p/p_ref)^a (T/T_ref)^b * mmr)
// (radiance_calc.F90:1457-1491 with scale_absorb.F90:137-152), the two
// powers as one exponential of the layer's logarithms. (The reference's two
// pow calls, rtor_v in vectlib_mod.F90:80-103, were measured on the device
// too: the FP64 fluxes moved by under 1e-11 relative, and the solve took 26 %
// longer.) lnp, lnt: logarithms of the layer pressure and temperature.
SG_HD double k_term(const SgTerm &t, double lnp, double lnt, double mmr) {
#pragma clang fp contract(off)
  double s = exp(t.a * (lnp - t.lnpref) + t.b * (lnt - t.lntref));
  return t.k * fmax(0.0, s * mmr);
}

// k_abs of one term at one layer for the other scalings on the path:
//  - power quadratic, k * max(0, (p/p_ref)^a (1 + b x + c x^2) * mmr) with
//    x = T/T_ref - 1 (scale_absorb.F90:176-193);
//  - Wenyi's CO2 tables in band 4 (scale_absorb.F90:238-259): gk4 bilinear
//    in ln(p/100 Pa) (step 0.5) and T (step 20 K), over gk250b; table block
//    w = [gk4 (10 x 26 x 14) | gk250b (14) | plg (26) | ttb (10)];
//  - P/T look-up table, max(0, max(0, sum of the four weighted entries) *
//    mmr) (radiance_calc.F90:1136-1152, 1432-1442) with the layer's
//    factors and indices g (inter_pt_lookup.F90:147-191).
// p, tk: layer pressure and temperature.
SG_HD double k_scaled(const SgTerm &t, const SgSpec *spec, double lnp, double p, double tk,
                      const double *g, double mmr) {
#pragma clang fp contract(off)
  if (t.fnc == SG_FNC_LOOKUP) {
    const int idx = (int)g[SG_G_IDX];
    const int jp = idx & 255, jt = (idx >> 8) & 255, jtt = idx >> 16, nt = spec->n_tmp;
    const double *kl = G(spec->klut) + (size_t)t.lut * spec->n_pre * nt;
    const double ke = fmax(0.0, g[SG_G_F00] * kl[jp * nt + jt] + g[SG_G_F10] * kl[(jp + 1) * nt + jtt] +
                                    g[SG_G_F01] * kl[jp * nt + jt + 1] +
                                    g[SG_G_F11] * kl[(jp + 1) * nt + jtt + 1]);
    return fmax(0.0, ke * mmr);
  }
  double s;
  if (t.fnc == SG_FNC_POWER_QUAD) {
    const double x = tk * t.tinv - 1.0;
    s = exp(t.a * (lnp - t.lnpref)) * (1.0 + x * t.b + t.c * x * x);
  } else {  // SG_FNC_WENYI
    const double *gk4 = G(spec->wenyi), *gk250b = gk4 + 3640, *plg = gk250b + 14, *ttb = plg + 26;
    const double cgp = fmax(-5.5, log(p / 100.0));
    int jp = (int)((5.5 + cgp) * 2.0) + 1;
    jp = jp < 1 ? 1 : (jp > 26 ? 26 : jp);
    const int jp1 = jp + 1 < 26 ? jp + 1 : 26;
    int jt = (int)((tk - 240.0) / 20.0 + 7.0);
    jt = jt < 1 ? 1 : (jt > 10 ? 10 : jt);
    const int jt1 = jt + 1 < 10 ? jt + 1 : 10;
    const int ix = t.lut;
#define SG_GK4(a, b) gk4[((a) - 1) * 364 + ((b) - 1) * 14 + ix]
    const double dl = (cgp - plg[jp - 1]) * 2.0;
    const double gkpb = SG_GK4(jt, jp) + (SG_GK4(jt, jp1) - SG_GK4(jt, jp)) * dl;
    const double gkpc = SG_GK4(jt1, jp) + (SG_GK4(jt1, jp1) - SG_GK4(jt1, jp)) * dl;
#undef SG_GK4
    s = (gkpb + (gkpc - gkpb) * (tk - ttb[jt - 1]) / 20.0) / gk250b[ix];
  }
  return t.k * fmax(0.0, s * mmr);
}

// Water-vapour continuum absorption of band B at one layer: the sum over the
// band's continua of k_cont * amount (grey_opt_prop.F90:801-816), amount by
// rescale_continuum.F90:107-179 (self-broadened: times the molar density of
// water and the water mixing ratio; foreign-broadened: times rho q (1 - q) /
// M_air, l_mixing_ratio false), with the reference's constants
// (rad_ccf.F90:42,58). h2o: the layer's water mixing ratio (clipped at 0).
SG_HD double continuum_k(const SgBand &B, double p, double tk, double rho, double h2o) {
#pragma clang fp contract(off)
  const double mol_weight_air = 28.966e-03, repsilon = 18.0153 / 28.966;
  double kg = 0.0;
  for (int j = 0; j < B.n_cont; j++) {
    const double *sc = B.cont_s[j];
    const double pwk = pow(p / B.cont_pref[j], sc[0]);
    double amount;
    if (B.cont_fnc[j] == SG_FNC_POWER_LAW) {
      amount = pwk * pow(tk / B.cont_tref[j], sc[1]);
    } else {
      const double x = tk / B.cont_tref[j] - 1.0;
      amount = pwk * (1.0 + sc[1] * x + sc[2] * (x * x));
    }
    if (B.cont_type[j] == 1) amount = amount * (rho * h2o / (repsilon * mol_weight_air)) * h2o;
    else amount = amount * rho * h2o * (1.0 - h2o) / mol_weight_air;
    kg = kg + B.cont_k[j] * amount;
  }
  return kg;
}

// P/T look-up interpolation of one layer (inter_pt_lookup.F90:147-191):
// ln p clamped inside the table, jp the last table pressure at or below it
// (MINLOC over the non-negative differences), T clamped per pressure row,
// jt / jtt the rows' lower neighbours; writes the four factors and the
// packed indices into g.
SG_HD void lookup_factors(const SgSpec *spec, double p, double tk, double *g) {
#pragma clang fp contract(off)
  const double eps = kEps;
  const int np = spec->n_pre, nt = spec->n_tmp;
  const double *lp = spec->lut_lnp;
  const double pmin = lp[0] + fmax(fabs(lp[0]) * eps, eps);
  const double pmax = lp[np - 1] - fmax(fabs(lp[np - 1]) * eps, eps);
  const double pl = fmin(fmax(log(p), pmin), pmax);
  int jp = 0;
  for (int k = 1; k < np; k++)
    if (pl - lp[k] >= 0.0) jp = k;
  const double fp = (pl - lp[jp]) / (lp[jp + 1] - lp[jp]);
  int jt[2];
  double ft[2];
  for (int r = 0; r < 2; r++) {
    const double *tl = spec->lut_t[jp + r];
    const double tc = fmin(fmax(tk, tl[0] * (1.0 + eps)), tl[nt - 1] * (1.0 - eps));
    int j = 0;
    for (int k = 1; k < nt; k++)
      if (tc - tl[k] >= 0.0) j = k;
    jt[r] = j;
    ft[r] = (tc - tl[j]) / (tl[j + 1] - tl[j]);
  }
  const double compfp = 1.0 - fp;
  g[SG_G_F00] = compfp * (1.0 - ft[0]);
  g[SG_G_F10] = fp * (1.0 - ft[1]);
  g[SG_G_F01] = compfp * ft[0];
  g[SG_G_F11] = fp * ft[1];
  g[SG_G_IDX] = (double)(jp + 256 * jt[0] + 65536 * jt[1]);
}

// Pade-2 droplet optics (opt_prop_pade_2_mod.F90:100-143): extinction,
// scattering, first phase-function moment and forward scattering, all
// weighted by the in-cloud mixing ratio.
SG_HD void pade2(const double *cp, double mmr, double re, double &kext, double &kscat,
                 double &phf, double &fwd) {
#pragma clang fp contract(off)
  kext = mmr * (cp[0] + re * (cp[1] + re * cp[2])) /
         (1.0 + re * (cp[3] + re * (cp[4] + re * cp[5])));
  kscat = kext * (1.0 - (cp[6] + re * (cp[7] + re * cp[8])) / (1.0 + re * (cp[9] + re * cp[10])));
  double g = (cp[11] + re * (cp[12] + re * cp[13])) / (1.0 + re * (cp[14] + re * cp[15]));
  phf = kscat * g;
  fwd = phf * g;
}

// Slingo-Schrecker droplet optics (opt_prop_water_cloud.F90:231-247); the
// forward scattering k_scat g**2 forms the power first (:292-298)
SG_HD void slingo(const double *cp, double mmr, double re, double &kext, double &kscat,
                  double &phf, double &fwd) {
#pragma clang fp contract(off)
  kext = mmr * (cp[0] + cp[1] / re);
  kscat = kext * (1.0 - cp[2] - cp[3] * re);
  double g = cp[4] + cp[5] * re;
  phf = kscat * g;
  fwd = kscat * (g * g);
}

// Droplet optics by the type's scheme (5 Pade-2, 1 Slingo-Schrecker).
SG_HD void drop_optics(const SgDrop &dr, int b, double mmr, double re, double &kext, double &kscat,
                       double &phf, double &fwd) {
  if (dr.scheme == 5) pade2(dr.parm[b], mmr, re, kext, kscat, phf, fwd);
  else slingo(dr.parm[b], mmr, re, kext, kscat, phf, fwd);
}

// Ice optics of the two-stream case of opt_prop_ice_cloud.F90:227-256 (one
// phase-function moment, rescaled, forward scattering of order 2): Fu's
// phase-function fit (scheme 12, opt_prop_fu_phf_mod.F90:93-130), Baran's
// temperature fit (scheme 19, opt_prop_baran_mod.F90:93-125, t**4 formed
// left to right) or Pade-2 (scheme 20, the droplet form). t: the layer
// temperature.
SG_HD void ice_optics(const SgDrop &ic, int b, double iwc, double de, double t, double &kext,
                      double &kscat, double &phf, double &fwd) {
#pragma clang fp contract(off)
  const double *cp = ic.parm[b];
  if (ic.scheme == 12) {
    double x = cp[3] / de;
    kext = iwc * ((cp[2] * x + cp[1]) * x + cp[0]);
    x = de / cp[8];
    kscat = kext * (1.0 - (cp[4] + x * (cp[5] + x * (cp[6] + x * cp[7]))));
    x = de / cp[13];
    const double g = cp[9] + x * (cp[10] + x * (cp[11] + x * cp[12]));
    phf = kscat * g;
    fwd = phf * g;
  } else if (ic.scheme == 19) {
    double x = fmin(iwc + 1.0e-12, 1.0e-01);
    kext = cp[0] * x / (t * t * t * t);
    x = fmin(iwc + 1.0e-12, 1.0e-03 - kEps);
    const double om = fmin(cp[1] + (cp[2] * x * t), 1.0 - kEps);
    kscat = kext * om;
    const double g = fmin(cp[3] + (cp[4] * x * t), 1.0 - kEps);
    phf = kscat * g;
    fwd = phf * g;
  } else {
    pade2(cp, iwc, de, kext, kscat, phf, fwd);
  }
}

// McICA cloud-generator linear congruential step (socrates_cloud_gen.F90:48-50
// constants, :196-205 step with the FP64 multiply-truncate of the reference)
constexpr int kLcgM = 86436, kLcgA = 1093, kLcgC = 18257;
// k steps of the generator from x in one: F^k(x) = (a_k x + c_k) mod m with
// a_k = a^k mod m and c_k = (a c_(k-1) + c) mod m. lcg_step's FP64
// multiply-truncate is the exact residue for every int32 argument (checked
// over the whole range), so F^k is exactly k applications of lcg_step.
constexpr int kLcgSkipMax = 2 * 64 + 64;
struct LcgSkip {
  int a[kLcgSkipMax + 1], c[kLcgSkipMax + 1];
};
constexpr LcgSkip make_lcg_skip() {
  LcgSkip s{};
  long long A = 1, C = 0;
  s.a[0] = 1;
  s.c[0] = 0;
  for (int k = 1; k <= kLcgSkipMax; k++) {
    A = (A * kLcgA) % kLcgM;
    C = (C * kLcgA + kLcgC) % kLcgM;
    s.a[k] = (int)A;
    s.c[k] = (int)C;
  }
  return s;
}
SG_HD int lcg_step(int x) {
  const double rm = 1.0 / (double)kLcgM;
  int t = x * kLcgA + kLcgC;
  double q = (double)t * rm;
  return t - (int)q * kLcgM;
}

}  // namespace sg

// The mixed form's functions fuse a*b+c into FMAs within one expression
// only (contract(on)): the backend then makes the same choice wherever the
// code is inlined, so a layer recomputed by the recomputing solve tile gives
// the bits of its first computation.
//
// FP32 transcendentals of the mixed form: single hardware instructions on
// the device (v_exp_f32 after the scaling, v_rcp_f32 then a multiply, about
// 1.5 ulp, v_sqrt_f32); the host build used by the CPU tests has the
// correctly rounded library forms.
#ifdef __HIP_DEVICE_COMPILE__
#define SG_EXPF(x) __expf(x)
#define SG_FDIV(a, b) ((a) * __builtin_amdgcn_rcpf(b))
#define SG_SQRTF(x) __builtin_amdgcn_sqrtf(x)
#else
#define SG_EXPF(x) expf(x)
#define SG_FDIV(a, b) ((a) / (b))
#define SG_SQRTF(x) sqrtf(x)
#endif

namespace sg {

struct CoefF {
  float T, R, T0, s1, s2;
  float u;    // 1 - T - R (layer diffuse absorption), to relative accuracy
  float em0;  // 1 - T0 (direct-beam extinction), to relative accuracy
};

// 1 - exp(-x) = -expm1(-x) in FP32 given e = exp(-x): Taylor series below
// x = 0.2 (terms to x^7, relative error < 1e-9 before rounding), 1 - e above,
// where the subtraction loses under 2 bits. Both sides are evaluated and one
// selected, so lanes of a wave do not diverge.
SG_HD float one_minus_exp(float x, float e) {
#pragma clang fp contract(on)
  float p = x * (1.0f - x * (0.5f - x * (0.16666667f - x * (0.041666668f - x * (0.008333334f -
            x * (0.0013888889f - x * 1.984127e-4f))))));
  return x < 0.2f ? p : 1.0f - e;
}

// two_coeff_fast_lw (two_coeff_fast_lw.F90:62-113): the infra-red layer
// without scattering (diffusivity 1.66): transmission, no reflection, and the
// linear and quadratic source coefficients, in the reference's operation
// order.
SG_HD void two_coeff_fast_lw(double tau, Coef &o) {
#pragma clang fp contract(off)
  const double sq_eps_r = 1.4901161193847656e-08;  // sqrt(epsilon)
  const double tol = 1.220703125e-04;               // sqrt(sq_eps_r)
  o.T = exp(-1.66 * tau);
  o.R = 0.0;
  o.T0 = 0.0;
  o.s1 = (1.0 - o.T + sq_eps_r) / (1.66 * tau + sq_eps_r);
  o.s2 = tau > tol ? -(1.0 + o.T - 2.0 * o.s1) / (1.66 * tau + sq_eps_r)
                   : -(1.0 + o.T - 2.0 + 1.66 * tau) / (1.66 * tau + sq_eps_r);
}

// Mixed-precision form of two_coeff (see two_coeff above for the reference
// routines): the same two-stream solution evaluated in FP32 through forms
// free of cancellation. With e = exp(-lambda tau), em = 1 - e,
// gc = 1 - gamma = 2 lambda / (S + lambda), d1 = 1 - e gamma = em + gc e:
//   T = e gc (1 + gamma) / (d1 (1 + e gamma)),  R = gamma em (1 + e) / (d1 (1 + e gamma)),
//   1 - T + R = em (1 + gamma) / d1,            1 - T - R = em gc / (1 + e gamma),
// and gamma = S (S - D) / (S + lambda)^2 with S - D taken from omega.
// omc = 1 - omega, computed by the caller from the absorption, keeps the
// near-conservative limit exact (diff = 2 omc or 1.66 omc). The
// regularisation constants keep their FP64 values (sqrt(eps) = 2^-26,
// tol = 2^-13).
//
// Solar sources without the resonance. The reference forms
//   s1 = f [A (1 - T T0) - B R] - R,   s2 = f [B (T0 - T) - A R T0] + T0 - T,
// f = omega mu / (2 (lambda - mu)(lambda + mu)), A, B = S -+ mu -+ ksi0 (D -+ mu),
// ksi0 = 1.5 g / mu, mu = sec0, have brackets that vanish at mu = lambda
// (there A = gamma B, since lambda^2 = S D). Dividing each bracket by
// (mu - lambda) as a divided difference in mu gives, for every mu,
//   s1 = -omega mu Q / (2 (lambda + mu)) - R,
//   Q = (c - 1)(1 - T T0) + A_l T tphi - (1 - c) R,
//   s2 = -omega mu M / (2 (lambda + mu)) + T0 - T,
//   M = (1 - c)(T0 - T) - B_l tphi - (c - 1) R T0 + A_l R tphi,
// with c = 1.5 g D / (mu lambda), A_l = S - lambda - 1.5 g (D / lambda - 1),
// B_l = S + lambda + 1.5 g (D / lambda + 1) (A, B at mu = lambda) and
// tphi = (T0 - e) / (lambda - mu) = tau e (1 - exp(-x)) / x, x = (mu - lambda) tau.
// This is the same function as the reference's (which perturbs lambda by
// 32 eps at exact resonance and loses digits near it), evaluated without a
// 1 / (lambda - mu) amplification, so FP32 needs no FP64 fallback.
SG_HD void two_coeff_f32(bool solar, float tau, float om, float omc, float g, float sec0,
                         CoefF &c) {
#pragma clang fp contract(on)
  float S, D, smd;  // smd = S - D, formed from omega without cancellation
  if (solar) {
    S = 2.0f - 1.5f * om * g - 0.5f * om;
    D = 2.0f * omc;
    smd = 1.5f * om * (1.0f - g);
  } else {
    S = 1.66f - 1.5f * om * g;
    D = 1.66f * omc;
    smd = om * (1.66f - 1.5f * g);
  }
  float lambda = SG_SQRTF(S * D);
  const float sqe = 1.4901161193847656e-08f;
  const float tol = 1.220703125e-04f;
  float x = lambda * tau;
  float e = SG_EXPF(-x);
  float em = one_minus_exp(x, e);
  float inv_sl = SG_FDIV(1.0f, S + lambda);
  float gc = 2.0f * lambda * inv_sl;
  float gam = S * smd * inv_sl * inv_sl;             // (S - lambda) / (S + lambda)
  float d1 = em + gc * e;
  float ope = 1.0f + e * gam;
  float rden = SG_FDIV(1.0f, d1 * ope);
  c.T = e * gc * (1.0f + gam) * rden;
  c.R = gam * em * (1.0f + e) * rden;
  const float u = em * gc * SG_FDIV(1.0f, ope);      // 1 - T - R
  c.u = u;
  if (solar) {
    const float mu = sec0;
    float xs = tau * mu;
    c.T0 = SG_EXPF(-xs);
    float em0 = one_minus_exp(xs, c.T0);
    c.em0 = em0;
    float omt = em * (1.0f + gam * gam * e) * rden;   // 1 - T
    float omtt0 = omt + c.T * em0;                    // 1 - T T0
    float t0mt = omt - em0;                           // T0 - T
    float r = SG_FDIV(D, lambda);
    float cc = SG_FDIV(1.5f * g * r, mu);
    float al = S - lambda + 1.5f * g * (1.0f - r);
    float bl = S + lambda + 1.5f * g * (r + 1.0f);
    float xd = (mu - lambda) * tau;
    // (1 - exp(-x)) / x, Taylor terms to x^7, near the resonance or for thin
    // layers; the direct quotient elsewhere
    float e1 = 1.0f - xd * (0.5f - xd * (0.16666667f - xd * (0.041666668f - xd * (0.008333334f -
               xd * (0.0013888889f - xd * (1.984127e-4f - xd * 2.4801587e-5f))))));
    float tphi = fabsf(xd) < 0.2f ? tau * e * e1 : SG_FDIV(c.T0 - e, lambda - mu);
    float fac = SG_FDIV(-0.5f * om * mu, lambda + mu);
    float q = (cc - 1.0f) * omtt0 + al * c.T * tphi - (1.0f - cc) * c.R;
    float m = (1.0f - cc) * t0mt - bl * tphi - (cc - 1.0f) * c.R * c.T0 + al * c.R * tphi;
    c.s1 = fac * q - c.R;
    c.s2 = fac * m + t0mt;
  } else {
    c.T0 = 0.0f;
    c.em0 = 1.0f;
    float omtr = em * (1.0f + gam) * SG_FDIV(1.0f, d1);   // 1 - T + R
    float den = sqe + tau * S;
    c.s1 = SG_FDIV(omtr + sqe, den);
    // w = D tau - u, exact to O(x^3) terms for thin layers
    float f = SG_FDIV(em, fmaxf(x, 1e-30f));             // (1 - e) / x
    float g1 = 0.5f - x * (0.16666667f - x * (0.041666668f - x * (0.008333334f - x * 0.0013888889f)));
    float ws = D * tau * x * SG_FDIV(2.0f * S * g1 - (S - lambda) * f, 2.0f * S - (S - lambda) * x * f);
    float w = x < 0.2f ? (x > 0.0f ? ws : 0.0f) : D * tau - u;
    // s2 = -(1 + R + T + q) / (S tau + sqe), q the reference's two branches;
    // tau > tol: 1 + R + T + q = 2 - u + q = (2 (D tau - u) - u dts) / dts, dts = D tau + sqe
    // (the sqe terms of the reference's 2 (u + sqe) / dts cancel against 2 dts / dts);
    // tau <= tol: 1 + R + T - 2 + D tau = D tau - u. The numerator's
    // difference is formed in FP64 from the FP32 pieces (exact products),
    // then rounded once.
    double dts = (double)D * (double)tau + 1.4901161193847656e-08;
    float nf = (float)(2.0 * (double)w - (double)u * dts);
    float qn = tau > tol ? SG_FDIV(nf, (float)dts) : w;
    c.s2 = -SG_FDIV(qn, den);
  }
}

}  // namespace sg

namespace sg {

// FP32 k-term (the mixed form's k-term exponent sum and exponential)
SG_HD float k_term_f32(const SgTerm &t, float lnp, float lnt, float mmr) {
#pragma clang fp contract(on)
  float s = SG_EXPF((float)t.a * (lnp - (float)t.lnpref) + (float)t.b * (lnt - (float)t.lntref));
  return (float)t.k * fmaxf(0.0f, s * mmr);
}

}  // namespace sg

namespace sg {

// FP64 coefficients from the mixed form's FP32 ones. The smaller of T and R
// is taken as computed (FP32 relative accuracy) and the larger closes the
// balance T + R + (1 - T - R) = 1 in FP64 with the absorption u = 1 - T - R,
// which is itself accurate to FP32 relative precision; T0 = 1 - (1 - T0)
// likewise. The layer's absorption and the small one of T and R then keep
// their relative accuracy in the FP64 elimination: thin layers (T ~ 1) keep
// their flux divergence, and so their heating rates, and thick
// near-conservative layers (R ~ 1) keep their transmission.
SG_HD void coef_from_f32(const CoefF &f, Coef &o) {
#pragma clang fp contract(on)
  if (f.T < f.R) {
    o.T = f.T;
    o.R = (1.0 - (double)f.T) - (double)f.u;
  } else {
    o.R = f.R;
    o.T = (1.0 - (double)f.R) - (double)f.u;
  }
  o.T0 = 1.0 - (double)f.em0;
  o.s1 = f.s1;
  o.s2 = f.s2;
}

}  // namespace sg

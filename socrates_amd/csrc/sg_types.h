// Device-resident tables of one spectral file (plus its McICA sampling
// tables) and the per-call column records, shared by host and device code.
//
// The spectral content is the subset of the reference's StrSpecData
// (def_spectrum.F90:33-472) that the two-stream flux path reads: per band the
// major-gas k-terms (ESFT terms), minor-gas k-terms for equivalent
// extinction, the solar flux fraction, the Rayleigh coefficient, the Planck
// polynomial and the droplet parametrisations.
#pragma once

#include <stdint.h>

#define SG_MAXB 16          // bands per file
#define SG_MAXG 96          // major k-terms (g-points) per file
#define SG_MAXMINOR_GAS 11  // minor gases per band
#define SG_MAXMINOR 256     // minor k-terms per file
#define SG_MAXSUB 192       // McICA sub-column items per file (sum of subcol_k)
#define SG_NPL 8            // Planck polynomial coefficients
#define SG_NDROP 8          // droplet types kept
#define SG_NICE 16          // ice types kept
#define SG_NDPARM 16        // droplet fit parameters
#define SG_MAXL 192         // layers per column
#define SG_MAXSUBCOL 64     // generated McICA sub-columns per column
#define SG_MAXCONT 2        // continua per band
#define SG_MAXPRE 64        // pressures of the k-term look-up table
#define SG_MAXTMP 8         // temperatures per pressure of the look-up table

// gas slots of the column record
#define SG_GAS_H2O 0
#define SG_GAS_CO2 1
#define SG_GAS_O3 2
#define SG_GAS_ZERO 3       // absorber the call supplies no amount for
#define SG_GAS_WM0 4        // first of the well-mixed further gases (slot
                            // SG_GAS_WM0 + k for the call's wm_mix_ratio[k])
#define SG_NWM 10           // well-mixed further gases: n2o, ch4, o2, so2, ocs,
                            // cfc11, cfc12, cfc113, hcfc22, hfc134a

// scaling functions of the k-terms (rad_pcf.F90:509-532)
#define SG_FNC_POWER_LAW 1
#define SG_FNC_POWER_QUAD 2
#define SG_FNC_WENYI 4      // ip_scale_wenyi, CO2 in band 4 only
#define SG_FNC_LOOKUP 9

struct SgTerm {             // one k-term (ESFT term)
  double k, w;              // absorption coefficient (m2/kg), weight
  double a, b, c;           // scaling parameters: power law (a, b), power
                            // quadratic (a, b, c)
  double lnpref, lntref;    // log of the reference pressure and temperature
  double tinv;              // 1 / reference temperature
  int band, gas;            // band, gas slot
  int n_sub, sub0;          // McICA: sampled sub-columns, first sub item
  int fnc;                  // SG_FNC_*
  int lut;                  // look-up table of the term (SG_FNC_LOOKUP), or
                            // the term's index in the Wenyi tables
  int iscat;                // the term's treatment of scattering in the file
                            // (i_scat: 0 or 1 full, 2 none, 3 extinction as
                            // absorption), used when the call asks for hybrid
};

struct SgBand {
  int n_major, g0;                     // major terms [g0, g0+n_major)
  int n_minor_gas, minor0;             // minor terms start at minor0
  int minor_nk[SG_MAXMINOR_GAS];       // terms per minor gas
  double solar_flux;                   // normalised solar flux in the band
  double rayleigh;                     // Rayleigh mass scattering coefficient
  double planck[SG_NPL];               // thermal_coeff(0:n_deg)
  // continua, in the band's order (radiance_calc.F90:897-923): continuum
  // type (1 self-broadened water vapour, 2 foreign-broadened), scaling
  // function, coefficient, reference pressure and temperature, parameters
  int n_cont;
  int cont_type[SG_MAXCONT], cont_fnc[SG_MAXCONT];
  double cont_k[SG_MAXCONT], cont_pref[SG_MAXCONT], cont_tref[SG_MAXCONT];
  double cont_s[SG_MAXCONT][3];
};

struct SgDrop {
  int present, scheme, n_parm, pad;
  double dim_min, dim_max;
  double parm[SG_MAXB][SG_NDPARM];
};

#define SG_MAXABS 32

struct SgSpec {
  int n_band, n_g, n_minor, n_deg;
  int n_mband;                   // bands with minor gases, and which
  int mband[SG_MAXB];
  int n_absorb;                  // gaseous absorbers of the file and their
  int type_absorb[SG_MAXABS];    // identifiers (rad_pcf.F90 ip_h2o = 1, ...)
  double t_ref_planck;
  // look-up-table k-terms (read_spectrum.F90:1404-1507): the P/T grid (ln p,
  // and t_lookup(it, ip) as [ip][it]) and the device tables, term t.lut at
  // klut + t.lut * n_pre * n_tmp as [ip][it]; the Wenyi CO2 tables
  // (10 x 26 x 14 gk4, then gk250b, plg, ttb) when a term uses them
  int n_pre, n_tmp, has_lut, has_cont;
  double lut_lnp[SG_MAXPRE];
  double lut_t[SG_MAXPRE][SG_MAXTMP];
  const double *klut;
  const double *wenyi;
  SgBand band[SG_MAXB];
  SgTerm g[SG_MAXG];
  SgTerm minor[SG_MAXMINOR];
  SgDrop drop[SG_NDROP];
  SgDrop ice[SG_NICE];           // ice crystals (block 12): the same record,
                                 // scheme = the ice parametrisation
  // McICA sampling (set_cld_mcica, socrates_set_cld_mcica.F90:103-151) for
  // [0] single and [1] optimal sampling
  int has_mcica, n_gen, ipph, n1, n2;
  int n_req[2];
  int n_sub_items[2];
  int sub_g[2][SG_MAXSUB];      // g-point of each sub item
  int sub_isub[2][SG_MAXSUB];   // generated sub-column (1-based) it samples
  double sub_inv_s[2][SG_MAXSUB];
  int g_n_sub[2][SG_MAXG];      // sub items per g-point
  int g_sub0[2][SG_MAXG];       // first sub item of a g-point
};

// per-layer column record, [col][layer][SG_NQ]
#define SG_Q_LNP 0
#define SG_Q_LNT 1
#define SG_Q_MASS 2
#define SG_Q_MMR 3     // 4 slots: h2o, co2, o3, zero
#define SG_Q_WCLD 7    // cloud fraction w_cloud (0 if below threshold)
#define SG_Q_LWC 8     // in-cloud liquid mass mixing ratio
#define SG_Q_RE 9      // effective radius clamped to the fit range
#define SG_Q_FRAC1 10  // frac_cloud of the homogeneous type (1 in cloud)
#define SG_Q_P 11      // layer pressure
#define SG_Q_T 12      // layer temperature
#define SG_Q_RSD 13    // condensate relative standard deviation (McICA)
// triple-column representation: slots WCLD, LWC, RE and FRAC1 carry the
// total cloud fraction and the stratiform type; these the convective one
#define SG_Q_LWC2 14   // in-cloud convective liquid mass mixing ratio
#define SG_Q_RE2 15    // convective effective radius clamped to the fit range
#define SG_Q_FRAC2 16  // frac_cloud of the convective type
// McICA representation: the cloud generator's exponential-random overlap
// parameters of the layer and the one above (in the triple-column slots)
#define SG_Q_ALPHA 14
#define SG_Q_RCORR 15
// ice of the stratiform (or homogeneous) and the convective type: in-cloud
// mass mixing ratio and the dimension clamped to the ice type's range
#define SG_Q_IWC 17
#define SG_Q_DE 18
#define SG_Q_IWC2 19
#define SG_Q_DE2 20
#define SG_NQ 21

// per-layer gas record of calls whose spectrum has look-up-table terms or
// continua, [col][layer][SG_NGQ]: the look-up interpolation factors
// (inter_pt_lookup.F90:184-189), the indices jp, jt, jtt (0-based, packed
// as jp + 256 jt + 65536 jtt) and the layer density
#define SG_G_F00 0
#define SG_G_F10 1
#define SG_G_F01 2
#define SG_G_F11 3
#define SG_G_IDX 4
#define SG_G_RHO 5
#define SG_NGQ 6

// per-column scalars, [col][SG_NS]
#define SG_S_SEC0 0    // zen_0 (secant of the solar zenith angle, or 1)
#define SG_S_SOL 1     // solar irradiance (0 when the sun is down)
#define SG_S_TGROUND 2
#define SG_S_FC 3      // frac_cloudy
#define SG_S_NCLD 4    // number of cloudy sub-columns (as a real)
#define SG_S_CT 5      // first layer (0-based) with cloud, n_layer if none (McICA)
#define SG_S_LFG 6     // 1: the ground source of each band is given (band scalar 1)
#define SG_NS 7

// per (col, band, layer) record from the preparation kernel, [col][band][layer][4]
//  LW: k_grey_tot, planck diff, planck diff_2, planck at the bottom level
//  SW: k_grey_tot, adjust_solar_ke, 0, 0
// per (col, band) scalars [col][band][4]
//  LW: planck at level 0, planck at the ground (or the given surface emission
//      over 1 - albedo diff), albedo diff, albedo dir
//  SW: band solar irradiance, 0, albedo diff, albedo dir
#define SG_NBQ 4

/* CPU restatement of the SOCRATES two-stream flux path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the parity tests and the CPU
 * baseline of the benchmark. Nothing in the product links or calls it.
 * Build: oracle/Makefile (gcc -O2 -fno-fast-math -ffp-contract=off).
 *
 * Each routine below restates one reference routine (cited file:line, paths
 * relative to the reference tree) keeping its operation order, so that the
 * results agree with the compiled reference to rounding.
 */
#include "soc_oracle.h"
#include "../socrates_amd/csrc/wenyi_tables.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef SOC_ORACLE_EXP_JITTER
/* Conditioning probe (tests only): every exp() result is moved by -1, 0 or
 * +1 ulp with a fixed pseudo-random sequence, to measure how strongly the
 * reference formulas amplify last-bit differences of exp. */
static unsigned long long jit_state = 88172645463325252ULL;
static double jit_exp(double x) {
  double v = exp(x);
  jit_state ^= jit_state << 13; jit_state ^= jit_state >> 7; jit_state ^= jit_state << 17;
  int r = (int)(jit_state % 3ULL) - 1;
  return r == 0 ? v : nextafter(v, r > 0 ? INFINITY : -INFINITY);
}
#define exp(x) jit_exp(x)
#endif

#define NB 16      /* bands */
#define NG 16      /* gases per file */
#define NK 32      /* k-terms per gas and band */
#define NCONT 4    /* continua per band */
#define NPRE 64    /* pressures of the k-term look-up table */
#define NTMP 8     /* temperatures of the k-term look-up table */
#define NDROP 16   /* droplet types */
#define NPARM 20   /* parameters per droplet fit */
#define NICE 16    /* ice types */

/* option codes, rad_pcf.F90 */
enum { IP_SOLAR = 1, IP_INFRA_RED = 2 };
/* rad_pcf.F90:539-547 */
enum { IP_SCATTER_FULL = 1, IP_NO_SCATTER_ABS = 2, IP_NO_SCATTER_EXT = 3, IP_SCATTER_APPROX = 4,
       IP_SCATTER_HYBRID = 5 };
/* gas identifiers (gas_list_pcf.F90) */
enum { IP_H2O = 1, IP_CO2 = 2, IP_O3 = 3, IP_N2O = 4, IP_CH4 = 6, IP_O2 = 7, IP_SO2 = 9,
       IP_CFC11 = 14, IP_CFC12 = 15, IP_CFC113 = 16, IP_HCFC22 = 17, IP_HFC134A = 19,
       IP_OCS = 25 };
enum { IP_SCALE_TERM = 2 };
/* scaling functions (rad_pcf.F90:509-532) */
enum { IP_SCALE_POWER_LAW = 1, IP_SCALE_POWER_QUAD = 2, IP_SCALE_WENYI = 4, IP_SCALE_LOOKUP = 9 };
/* continua (rad_pcf.F90) */
enum { IP_SELF_CONTINUUM = 1, IP_FRN_CONTINUUM = 2 };
/* rad_ccf.F90:42,58 */
static const double MOL_WEIGHT_AIR = 28.966e-03;
static const double REPSILON = 18.0153 / 28.966;
enum { IP_DROP_PADE_2 = 5, IP_SLINGO_SCHRECKER = 1 };
enum { IP_ICE_FU_PHF = 12, IP_ICE_BARAN = 19, IP_ICE_PADE_2_PHF = 20 };

static const double ELSASSER = 1.66;  /* diffusivity_factor.F90 */
static const double DIFF_MINOR = 1.66;

/* ------------------------------------------------------------------------ */
/* Spectral data: the subset of StrSpecData (def_spectrum.F90:33-472) that   */
/* the two-stream flux path reads.                                           */
typedef struct {
  int n_band, n_absorb;
  int type_absorb[NG];
  double solar_flux_band[NB];
  double rayleigh_coeff[NB];
  int n_band_absorb[NB];
  int index_absorb[NB][NG]; /* 1-based gas index */
  int i_band_k[NB][NG], i_scale_k[NB][NG], i_scale_fnc[NB][NG];
  double p_ref[NB][NG], t_ref[NB][NG];
  double k[NB][NG][NK], w[NB][NG][NK], scale[NB][NG][NK][3];
  int i_scat[NB][NG][NK]; /* the term's treatment of scattering (hybrid) */
  /* P/T look-up tables of ip_scale_lookup terms (read_spectrum.F90:1404-1507):
   * ln p_lookup, t_lookup(it, ip) as [ip][it], k_lookup per (band, gas) as
   * [term][ip][it] */
  int n_pre, n_tmp;
  double p_lookup[NPRE], t_lookup[NPRE][NTMP];
  double *k_lookup[NB][NG];
  /* continua (blocks 8 and 9, read_spectrum.F90:1881-1997): stored by the
   * continuum's index number, as the reader stores them */
  int n_band_continuum[NB], index_continuum[NB][NCONT], index_water;
  int i_scale_fnc_cont[NB][NCONT + 1];
  double k_cont[NB][NCONT + 1], scale_cont[NB][NCONT + 1][3];
  double p_ref_cont[NCONT + 1][NB], t_ref_cont[NCONT + 1][NB];
  int n_deg_fit;
  double t_ref_planck;
  double thermal_coeff[NB][8];
  int l_drop_type[NDROP], i_drop_parm[NDROP], n_drop_parm[NDROP];
  double drop_min[NDROP], drop_max[NDROP];
  double drop_parm[NDROP][NB][NPARM];
  /* ice crystals (block 12, read_spectrum.F90:1175-1262) */
  int l_ice_type[NICE], i_ice_parm[NICE], n_ice_parm[NICE];
  double ice_min[NICE], ice_max[NICE];
  double ice_parm[NICE][NB][NPARM];
} spec_t;

typedef struct {
  int n_subcol_gen, n_req_single, n_req_optimal, ipph;
  int reorder_single[512], reorder_optimal[512];
  int sw_subcol_k[NB][NK], lw_subcol_k[NB][NK];
  int n1, n2;
  double *xcw; /* n1 x n2, column-major */
} mcica_t;

static int block_header(const char *ln, int *type, int *sub, int *ver) {
  /* '(a15, i4, 12x, i4, 12x, i4)' read_spectrum.F90:218 */
  char buf[8];
  if (strncmp(ln, "*BLOCK", 6) != 0 || strlen(ln) < 51) return 0;
  memcpy(buf, ln + 15, 4); buf[4] = 0; *type = atoi(buf);
  memcpy(buf, ln + 31, 4); buf[4] = 0; *sub = atoi(buf);
  memcpy(buf, ln + 47, 4); buf[4] = 0; *ver = atoi(buf);
  return 1;
}

static double fld(const char *ln, int c0, int w) {
  char buf[64];
  int n = (int)strlen(ln);
  if (c0 >= n) return 0.0;
  if (c0 + w > n) w = n - c0;
  memcpy(buf, ln + c0, (size_t)w); buf[w] = 0;
  return strtod(buf, NULL);
}

static char *next_line(FILE *f, char *ln, int n) {
  char *r = fgets(ln, n, f);
  if (r) { size_t m = strlen(ln); while (m && (ln[m-1] == '\n' || ln[m-1] == '\r')) ln[--m] = 0; }
  return r;
}

/* Numbers of a '(6(1PE13.6))' record set: n values, 6 to a line. */
static int read_e13(FILE *f, char *ln, int lnsz, double *out, int n) {
  int got = 0;
  while (got < n) {
    if (!next_line(f, ln, lnsz)) return -1;
    for (int q = 0; q < 6 && got < n; q++) out[got++] = fld(ln, 13 * q, 13);
  }
  return 0;
}

/* The k-term look-up tables of the '_k' file (read_spectrum.F90:1404-1507):
 * the block whose header names 'k-table', two header records, the P/T grid
 * (ln p stored), then per band and per gas with ip_scale_lookup (in the
 * order block 5 lists them) two header records and per term and pressure
 * one record set of n_tmp values. */
static int read_k_table(const char *path, spec_t *s, int order[NB][NG], const int *n_order) {
  char kp[1100], ln[512];
  snprintf(kp, sizeof kp, "%s_k", path);
  FILE *f = fopen(kp, "r");
  if (!f) return -20;
  int found = 0;
  while (next_line(f, ln, sizeof ln))
    if (strncmp(ln, "*BLOCK", 6) == 0 && strncmp(ln + 8, "k-table", 7) == 0) { found = 1; break; }
  if (!found) { fclose(f); return -21; }
  next_line(f, ln, sizeof ln); next_line(f, ln, sizeof ln);
  double row[NTMP + 1];
  for (int ip = 0; ip < s->n_pre; ip++) {
    if (read_e13(f, ln, sizeof ln, row, s->n_tmp + 1)) { fclose(f); return -22; }
    s->p_lookup[ip] = log(row[0]);
    for (int it = 0; it < s->n_tmp; it++) s->t_lookup[ip][it] = row[it + 1];
  }
  for (int b = 0; b < s->n_band; b++)
    for (int j = 0; j < n_order[b]; j++) {
      int g = order[b][j];
      if (s->i_scale_fnc[b][g] != IP_SCALE_LOOKUP) continue;
      next_line(f, ln, sizeof ln); next_line(f, ln, sizeof ln);
      int nt = s->i_band_k[b][g];
      double *kl = (double *)malloc(sizeof(double) * (size_t)nt * s->n_pre * s->n_tmp);
      s->k_lookup[b][g] = kl;
      for (int k = 0; k < nt; k++)
        for (int ip = 0; ip < s->n_pre; ip++)
          if (read_e13(f, ln, sizeof ln, kl + ((size_t)k * s->n_pre + ip) * s->n_tmp, s->n_tmp)) {
            fclose(f);
            return -22;
          }
    }
  fclose(f);
  return 0;
}

/* Text readers for the blocks on the flux path: a restatement of
 * read_spectrum.F90 block readers 0_0_1/0_0_2 (557-728), 2_0_0 (764-789),
 * 3_0_0 (872-900), 4_0_0/4_0_1 (951-1036), 5_0_1 (1218-1641: power-law,
 * power-quadratic, Wenyi and look-up-table terms), 6_0_0 (1746-1777),
 * 8_0_0 (1881-1935), 9_0_0 (1939-1997), 10_0_2 (2245-2331). Other blocks
 * are skipped. */
static int read_spectrum(const char *path, spec_t *s) {
  FILE *f = fopen(path, "r");
  char ln[512];
  int t, st, v;
  int order[NB][NG], n_order[NB];
  int any_lookup = 0;
  if (!f) return -1;
  for (int b = 0; b < NB; b++)
    for (int g = 0; g < NG; g++) free(s->k_lookup[b][g]);
  memset(s, 0, sizeof(*s));
  memset(n_order, 0, sizeof n_order);
  while (next_line(f, ln, sizeof ln)) {
    if (!block_header(ln, &t, &st, &v)) continue;
    if (t == 0) {
      /* summary lines are labelled; version 2 adds optional ones */
      while (next_line(f, ln, sizeof ln)) {
        if (strncmp(ln, "Number of spectral bands", 24) == 0) s->n_band = atoi(strchr(ln, '=') + 1);
        else if (strncmp(ln, "Total number of gaseous absorbers", 33) == 0)
          s->n_absorb = atoi(strchr(ln, '=') + 1);
        else if (strncmp(ln, "List of indexing numbers and absorbers.", 39) == 0) break;
      }
      next_line(f, ln, sizeof ln);
      for (int i = 0; i < s->n_absorb; i++) {
        next_line(f, ln, sizeof ln);
        s->type_absorb[i] = (int)fld(ln, 12, 5);
      }
    } else if (t == 2) {
      next_line(f, ln, sizeof ln); next_line(f, ln, sizeof ln);
      for (int b = 0; b < s->n_band; b++) {
        next_line(f, ln, sizeof ln); s->solar_flux_band[b] = fld(ln, 12, 16);
      }
    } else if (t == 3) {
      for (int j = 0; j < 3; j++) next_line(f, ln, sizeof ln);
      for (int b = 0; b < s->n_band; b++) {
        next_line(f, ln, sizeof ln); s->rayleigh_coeff[b] = fld(ln, 12, 16);
      }
    } else if (t == 4) {
      for (int j = 0; j < 5; j++) next_line(f, ln, sizeof ln);
      for (int b = 0; b < s->n_band; b++) {
        next_line(f, ln, sizeof ln);
        s->n_band_absorb[b] = (int)fld(ln, 12, 5);
        if (v == 1) {
          /* '(5x, 12(2x, i3))' (:1022) */
          if (s->n_band_absorb[b] > 0) {
            next_line(f, ln, sizeof ln);
            for (int q = 0; q < s->n_band_absorb[b]; q++)
              s->index_absorb[b][q] = (int)fld(ln, 5 + 5 * q + 2, 3);
          }
          continue;
        }
        int got = 0, first = 1;
        while (got < s->n_band_absorb[b]) {
          next_line(f, ln, sizeof ln);
          int c0 = first ? 5 : 0;
          first = 0;
          for (int q = 0; q < 4 && got < s->n_band_absorb[b]; q++)
            s->index_absorb[b][got++] = (int)fld(ln, c0 + 5 * q + 2, 3);
        }
      }
    } else if (t == 5) {
      while (next_line(f, ln, sizeof ln))
        if (strncmp(ln, "Band        Gas, Number of k-terms", 34) == 0) break;
      next_line(f, ln, sizeof ln); next_line(f, ln, sizeof ln);
      for (int b = 0; b < s->n_band; b++)
        for (int j = 0; j < s->n_band_absorb[b]; j++) {
          next_line(f, ln, sizeof ln);
          int band = (int)fld(ln, 0, 5) - 1, gas = (int)fld(ln, 12, 5) - 1;
          int nt = (int)fld(ln, 24, 5);
          int fnc = (int)fld(ln, 48, 5);
          s->i_band_k[band][gas] = nt;
          s->i_scale_k[band][gas] = (int)fld(ln, 36, 5);
          s->i_scale_fnc[band][gas] = fnc;
          order[band][n_order[band]++] = gas;
          next_line(f, ln, sizeof ln);
          if (fnc == IP_SCALE_LOOKUP) {
            /* '(2(7x, i5))' num_ref_p, num_ref_t (:1333-1336) */
            s->n_pre = (int)fld(ln, 7, 5);
            s->n_tmp = (int)fld(ln, 19, 5);
            any_lookup = 1;
            if (s->n_pre > NPRE || s->n_tmp > NTMP) { fclose(f); return -23; }
          } else {
            s->p_ref[band][gas] = fld(ln, 6, 16);
            s->t_ref[band][gas] = fld(ln, 28, 16);
          }
          /* n_scale_variable(fnc) parameters (rad_pcf.F90:505-506): the
           * first two on the k-term's record, a third on the next one */
          const int nsv = fnc == IP_SCALE_POWER_LAW ? 2 : fnc == IP_SCALE_POWER_QUAD ? 3 : 0;
          for (int k = 0; k < nt; k++) {
            next_line(f, ln, sizeof ln);
            s->k[band][gas][k] = fld(ln, 3, 16);
            s->w[band][gas][k] = fld(ln, 22, 16);
            s->i_scat[band][gas][k] = (int)fld(ln, 38, 3);
            if (nsv >= 1) s->scale[band][gas][k][0] = fld(ln, 41, 16);
            if (nsv >= 2) s->scale[band][gas][k][1] = fld(ln, 60, 16);
            if (nsv >= 3) {
              next_line(f, ln, sizeof ln);
              s->scale[band][gas][k][2] = fld(ln, 41, 16);
            }
          }
        }
    } else if (t == 8) {
      for (int j = 0; j < 5; j++) next_line(f, ln, sizeof ln);
      for (int b = 0; b < s->n_band; b++) {
        next_line(f, ln, sizeof ln);
        s->n_band_continuum[b] = (int)fld(ln, 12, 5);
        if (s->n_band_continuum[b] > NCONT) { fclose(f); return -24; }
        if (s->n_band_continuum[b] > 0) {
          next_line(f, ln, sizeof ln);
          for (int q = 0; q < s->n_band_continuum[b]; q++)
            s->index_continuum[b][q] = (int)fld(ln, 5 + 5 * q + 2, 3);
        }
      }
      next_line(f, ln, sizeof ln); next_line(f, ln, sizeof ln);
      s->index_water = (int)fld(ln, 22, 5);
    } else if (t == 9) {
      next_line(f, ln, sizeof ln); next_line(f, ln, sizeof ln); next_line(f, ln, sizeof ln);
      for (int b = 0; b < s->n_band; b++)
        for (int j = 0; j < s->n_band_continuum[b]; j++) {
          next_line(f, ln, sizeof ln);
          int band = (int)fld(ln, 0, 5) - 1, ic = (int)fld(ln, 12, 5), fnc = (int)fld(ln, 24, 5);
          if (ic < 1 || ic > NCONT) { fclose(f); return -24; }
          s->i_scale_fnc_cont[band][ic] = fnc;
          next_line(f, ln, sizeof ln);
          s->p_ref_cont[ic][band] = fld(ln, 6, 16);
          s->t_ref_cont[ic][band] = fld(ln, 28, 16);
          /* '(6x, 1pe16.9, (t23, 2(6x, 1pe16.9)))' */
          next_line(f, ln, sizeof ln);
          s->k_cont[band][ic] = fld(ln, 6, 16);
          const int nsv = fnc == IP_SCALE_POWER_LAW ? 2 : fnc == IP_SCALE_POWER_QUAD ? 3 : 0;
          if (nsv >= 1) s->scale_cont[band][ic][0] = fld(ln, 28, 16);
          if (nsv >= 2) s->scale_cont[band][ic][1] = fld(ln, 50, 16);
          if (nsv >= 3) {
            next_line(f, ln, sizeof ln);
            s->scale_cont[band][ic][2] = fld(ln, 28, 16);
          }
        }
    } else if (t == 6) {
      /* version 1 (read_block_6_0_1, :1781-1877) adds a 'Type of data' line */
      next_line(f, ln, sizeof ln);
      if (v == 1) next_line(f, ln, sizeof ln);
      next_line(f, ln, sizeof ln);
      /* '(23x, i5, 26x, 1pe16.9)' (:1757), '(24x, i5, 26x, 1pe16.9)' (:1802) */
      s->n_deg_fit = (int)fld(ln, v == 1 ? 24 : 23, 5);
      s->t_ref_planck = fld(ln, v == 1 ? 55 : 54, 16);
      next_line(f, ln, sizeof ln); next_line(f, ln, sizeof ln);
      for (int b = 0; b < s->n_band; b++) {
        int got = 0;
        double vals[16];
        while (got < s->n_deg_fit + 2) {
          next_line(f, ln, sizeof ln);
          char *p = ln;
          char *e;
          if (got == 0) { strtol(p, &e, 10); p = e; vals[got++] = 0; }
          for (;;) {
            double x = strtod(p, &e);
            if (e == p) break;
            vals[got++] = x;
            p = e;
          }
        }
        for (int k = 0; k <= s->n_deg_fit; k++) s->thermal_coeff[b][k] = vals[k + 1];
      }
    } else if (t == 12) {
      /* ice: the same layout as the droplet block, the type number after
       * its '=' */
      next_line(f, ln, sizeof ln);
      next_line(f, ln, sizeof ln); int iice = atoi(strchr(ln, '=') + 1);
      next_line(f, ln, sizeof ln);
      int iparm = (int)fld(ln, 34, 5), npar = (int)fld(ln, 66, 5);
      next_line(f, ln, sizeof ln);
      next_line(f, ln, sizeof ln);
      /* a type whose fit is larger than the restatement keeps is read past
       * and left unset (a call that names it is refused) */
      const int keep = iice >= 1 && iice < NICE && npar <= NPARM;
      if (keep) {
        s->ice_min[iice] = fld(ln, 39, 12);
        s->ice_max[iice] = fld(ln, 55, 12);
        s->l_ice_type[iice] = 1;
        s->i_ice_parm[iice] = iparm;
        s->n_ice_parm[iice] = npar;
      }
      for (int b = 0; b < s->n_band; b++) {
        next_line(f, ln, sizeof ln);
        int got = 0;
        while (got < npar) {
          next_line(f, ln, sizeof ln);
          char *p = ln, *e;
          for (;;) {
            double x = strtod(p, &e);
            if (e == p) break;
            if (keep) s->ice_parm[iice][b][got] = x;
            got++;
            p = e;
          }
        }
      }
    } else if (t == 10) {
      next_line(f, ln, sizeof ln);
      next_line(f, ln, sizeof ln); int idrop = (int)fld(ln, 27, 5);
      next_line(f, ln, sizeof ln);
      int iparm = (int)fld(ln, 34, 5), npar = (int)fld(ln, 66, 5);
      next_line(f, ln, sizeof ln);
      next_line(f, ln, sizeof ln);
      s->drop_min[idrop] = fld(ln, 39, 12);
      s->drop_max[idrop] = fld(ln, 55, 12);
      s->l_drop_type[idrop] = 1;
      s->i_drop_parm[idrop] = iparm;
      s->n_drop_parm[idrop] = npar;
      for (int b = 0; b < s->n_band; b++) {
        next_line(f, ln, sizeof ln);
        int got = 0;
        while (got < npar) {
          next_line(f, ln, sizeof ln);
          char *p = ln, *e;
          for (;;) {
            double x = strtod(p, &e);
            if (e == p) break;
            s->drop_parm[idrop][b][got++] = x;
            p = e;
          }
        }
      }
    }
  }
  fclose(f);
  if (s->n_band <= 0) return -2;
  if (any_lookup) return read_k_table(path, s, order, n_order);
  return 0;
}

/* read_mcica_data, def_mcica.F90:61-190 */
static int read_mcica(const char *path, mcica_t *m) {
  FILE *f = fopen(path, "r");
  char ln[512];
  if (!f) return -1;
  memset(m, 0, sizeof(*m));
  for (int b = 0; b < NB; b++)
    for (int k = 0; k < NK; k++) { m->sw_subcol_k[b][k] = 1; m->lw_subcol_k[b][k] = 1; }
  while (next_line(f, ln, sizeof ln))
    if (strncmp(ln, "*DATA", 5) == 0) break;
  while (next_line(f, ln, sizeof ln)) {
    if (strncmp(ln, "*END", 4) == 0) break;
    if (!strcmp(ln, "tot_subcol_gen") || !strcmp(ln, "n_subcol_gen")) {
      if (fscanf(f, "%d", &m->n_subcol_gen) != 1) return -3;
    } else if (!strcmp(ln, "subcol_need_single") || !strcmp(ln, "n_subcol_req_single")) {
      if (fscanf(f, "%d", &m->n_req_single) != 1) return -3;
    } else if (!strcmp(ln, "subcol_need_optimal") || !strcmp(ln, "n_subcol_req_optimal")) {
      if (fscanf(f, "%d", &m->n_req_optimal) != 1) return -3;
    } else if (!strcmp(ln, "ipph")) {
      if (fscanf(f, "%d", &m->ipph) != 1) return -3;
    } else if (!strcmp(ln, "lw_subcol_reorder_single")) {
      for (int i = 0; i < m->n_req_single; i++)
        if (fscanf(f, "%d", &m->reorder_single[i]) != 1) return -3;
    } else if (!strcmp(ln, "lw_subcol_reorder_optimal")) {
      for (int i = 0; i < m->n_req_optimal; i++)
        if (fscanf(f, "%d", &m->reorder_optimal[i]) != 1) return -3;
    } else if (!strcmp(ln, "sw_subcol_k") || !strcmp(ln, "lw_subcol_k")) {
      int sw = ln[0] == 's';
      for (;;) {
        int b, k, n;
        next_line(f, ln, sizeof ln);
        if (sscanf(ln, "%d %d %d", &b, &k, &n) != 3) return -3;
        if (b == -99) break;
        if (sw) m->sw_subcol_k[b - 1][k - 1] = n; else m->lw_subcol_k[b - 1][k - 1] = n;
      }
    } else if (!strcmp(ln, "n1")) {
      if (fscanf(f, "%d", &m->n1) != 1) return -3;
    } else if (!strcmp(ln, "n2")) {
      if (fscanf(f, "%d", &m->n2) != 1) return -3;
    } else if (!strcmp(ln, "xcw")) {
      m->xcw = (double *)malloc(sizeof(double) * (size_t)m->n1 * (size_t)m->n2);
      for (int i = 0; i < m->n1 * m->n2; i++)
        if (fscanf(f, "%lf", &m->xcw[i]) != 1) return -3;
    }
    /* list-directed READs leave the rest of the record: skip to a new line */
    if (ln[0] && strcmp(ln, "*END")) {
      int ch = fgetc(f);
      if (ch != '\n' && ch != EOF) ungetc(ch, f);
    }
  }
  fclose(f);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* planck_flux_band, planck_flux_band_mod.F90:18-49 (polynomial form)        */
static double planck_flux_band(const spec_t *s, int b, double temperature) {
  double t_ratio = temperature / s->t_ref_planck;
  double v = s->thermal_coeff[b][s->n_deg_fit];
  for (int j = s->n_deg_fit - 1; j >= 0; j--) v = v * t_ratio + s->thermal_coeff[b][j];
  return v;
}

/* two_coeff_basic (two_coeff_basic.F90:108-161), lambda (two_coeff.F90:
 * 158-163), solar_coefficient_basic (solar_coefficient_basic.F90:124-219),
 * trans_source_coeff (trans_source_coeff.F90:142-263), for one layer. */
typedef struct { double trans, reflect, trans_0, sc1, sc2; } coeff_t;

static void two_coeff_one(int isolir, double tau, double omega, double asymmetry,
                          double sec_0, coeff_t *o) {
  const double eps = DBL_EPSILON;
  double sum, diff, lambda;
  if (isolir == IP_SOLAR) {
    /* ip_pifm80 */
    sum = 2.0 - 1.5 * omega * asymmetry - 0.5 * omega;
    diff = 2.0 * (1.0 - omega);
  } else {
    /* ip_elsasser */
    sum = ELSASSER - 1.5 * omega * asymmetry;
    diff = ELSASSER * (1.0 - omega);
  }
  double temp = sum * diff;
  lambda = sqrt(temp);
  double gamma_up = 0.0, gamma_down = 0.0;
  if (isolir == IP_SOLAR) {
    double tol_perturb = 3.2e+01 * eps;
    if (fabs(lambda - sec_0) < tol_perturb) {
      sum = (1.0 + tol_perturb) * sum;
      diff = (1.0 + tol_perturb) * diff;
      lambda = (1.0 + tol_perturb) * lambda;
    }
    double ksi_0 = 1.5 * asymmetry / sec_0;
    double factor = 0.5 * omega * sec_0 / ((lambda - sec_0) * (lambda + sec_0));
    gamma_up = factor * (sum - sec_0 - ksi_0 * (diff - sec_0));
    gamma_down = factor * (sum + sec_0 + ksi_0 * (diff + sec_0));
  }
  double sq_eps_r = sqrt(eps);
  double tol = sqrt(sq_eps_r);
  double exponential = exp(-lambda * tau);
  double exponential2 = exponential * exponential;
  double gamma = (sum - lambda) / (sum + lambda);
  double gamma2 = gamma * gamma;
  double tmp_inv = 1.0 / (1.0 - exponential2 * gamma2);
  o->trans = exponential * (1.0 - gamma2) * tmp_inv;
  o->reflect = gamma * (1.0 - exponential2) * tmp_inv;
  if (isolir == IP_SOLAR) {
    o->trans_0 = exp(-tau * sec_0);
    o->sc1 = (gamma_up - o->reflect * (1.0 + gamma_down)) - gamma_up * o->trans * o->trans_0;
    o->sc2 = o->trans_0 * (1.0 + gamma_down - gamma_up * o->reflect) - (1.0 + gamma_down) * o->trans;
  } else {
    o->trans_0 = 0.0;
    o->sc1 = (1.0 - o->trans + o->reflect + sq_eps_r) / (sq_eps_r + tau * sum);
    double sc2;
    if (tau > tol)
      sc2 = -2.0 * (1.0 - o->trans - o->reflect + sq_eps_r) / (diff * tau + sq_eps_r);
    else
      sc2 = -2.0 + diff * tau;
    o->sc2 = -(1.0 + o->reflect + o->trans + sc2) / (sum * tau + sq_eps_r);
  }
}

void soc_oracle_two_coeff(int n, int isolir, const double *tau, const double *omega,
                          const double *asymmetry, const double *sec_0, double *trans,
                          double *reflect, double *trans_0, double *sc1, double *sc2) {
  for (int i = 0; i < n; i++) {
    coeff_t o;
    two_coeff_one(isolir, tau[i], omega[i], asymmetry[i], sec_0 ? sec_0[i] : 1.0, &o);
    trans[i] = o.trans; reflect[i] = o.reflect; trans_0[i] = o.trans_0;
    sc1[i] = o.sc1; sc2[i] = o.sc2;
  }
}

/* solver_homogen_direct.F90:93-127 for one column; flux_total 1-based
 * index k is stored at flux_total[k-1]. */
void soc_oracle_solver_homogen_direct(int n_layer, const double *trans,
                                      const double *reflect, const double *s_down,
                                      const double *s_up, double diffuse_albedo,
                                      double flux_inc_down, double source_ground,
                                      double *flux_total) {
  double *beta = (double *)malloc(sizeof(double) * 3 * (size_t)(n_layer + 1));
  double *gam = beta + (n_layer + 1), *h = gam + (n_layer + 1);
  double alpha = diffuse_albedo, s_up_prime = source_ground;
  for (int i = n_layer; i >= 1; i--) {
    beta[i] = 1.0 / (1.0 - alpha * reflect[i - 1]);
    gam[i] = alpha * trans[i - 1];
    h[i] = s_up_prime + alpha * s_down[i - 1];
    alpha = reflect[i - 1] + beta[i] * gam[i] * trans[i - 1];
    s_up_prime = s_up[i - 1] + beta[i] * trans[i - 1] * h[i];
  }
  flux_total[1] = flux_inc_down;
  flux_total[0] = alpha * flux_total[1] + s_up_prime;
  for (int i = 1; i <= n_layer; i++) {
    flux_total[2 * i] = beta[i] * (h[i] + gam[i] * flux_total[2 * i - 1]);
    flux_total[2 * i + 1] = s_down[i - 1] + trans[i - 1] * flux_total[2 * i - 1]
                            + reflect[i - 1] * flux_total[2 * i];
  }
  free(beta);
}

/* two_coeff_fast_lw.F90:62-113: non-scattering transmission and infra-red
 * source coefficients (the quadratic one when quad) */
static void two_coeff_fast_lw_one(double tau, int quad, coeff_t *o) {
  const double sq_eps_r = sqrt(DBL_EPSILON), tol = sqrt(sq_eps_r);
  double trans_in = -1.66e+00 * tau;
  o->trans = exp(trans_in);
  o->reflect = 0.0;
  o->trans_0 = 0.0;
  o->sc1 = (1.0 - o->trans + sq_eps_r) / (1.66e+00 * tau + sq_eps_r);
  o->sc2 = 0.0;
  if (quad) {
    if (tau > tol)
      o->sc2 = -(1.0 + o->trans - 2.0 * o->sc1) / (1.66e+00 * tau + sq_eps_r);
    else
      o->sc2 = -(1.0 + o->trans - 2.0 + 1.66e+00 * tau) / (1.66e+00 * tau + sq_eps_r);
  }
}

/* solver_no_scat.F90:60-97 for one column; flux_total as for
 * soc_oracle_solver_homogen_direct */
static void solver_no_scat(int n_layer, const double *trans, const double *s_down,
                           const double *s_up, double diffuse_albedo, double flux_inc_down,
                           double source_ground, double *flux_total) {
  flux_total[1] = flux_inc_down;
  for (int i = 1; i <= n_layer; i++)
    flux_total[2 * i + 1] = s_down[i - 1] + trans[i - 1] * flux_total[2 * i - 1];
  flux_total[2 * n_layer] = source_ground + diffuse_albedo * flux_total[2 * n_layer + 1];
  for (int i = n_layer; i >= 1; i--)
    flux_total[2 * i - 2] = s_up[i - 1] + trans[i - 1] * flux_total[2 * i];
}

/* two_stream.F90:160-205 with ir_source.F90:93-120, solar_source.F90:
 * 104-162 and column_solver.F90:160-185 (no spherical geometry, no
 * orography, direct tau unscaled: the default control). */
void soc_oracle_two_stream(int n_profile, int n_layer, int isolir, int l_scale_solar,
                           int l_ir_source_quad, const double *tau, const double *omega,
                           const double *asymmetry, const double *sec_0,
                           const double *flux_inc_down, const double *flux_inc_direct,
                           const double *diffuse_albedo, const double *direct_albedo,
                           const double *diff_planck, const double *diff_planck_2,
                           const double *d_planck_flux_surface,
                           const double *adjust_solar_ke, double *flux_direct,
                           double *flux_total) {
  const int P = n_profile, L = n_layer;
  double *tr = (double *)malloc(sizeof(double) * 4 * (size_t)L);
  double *rf = tr + L, *sd = rf + L, *su = sd + L;
  double *ft = (double *)malloc(sizeof(double) * (size_t)(2 * L + 2));
  for (int l = 0; l < P; l++) {
    double fd = isolir == IP_SOLAR ? flux_inc_direct[l] : 0.0;
    if (isolir == IP_SOLAR) flux_direct[l] = fd;
    for (int i = 0; i < L; i++) {
      const size_t k = (size_t)i * P + l;
      coeff_t o;
      two_coeff_one(isolir, tau[k], omega[k], asymmetry[k], isolir == IP_SOLAR ? sec_0[l] : 1.0, &o);
      tr[i] = o.trans;
      rf[i] = o.reflect;
      if (isolir == IP_SOLAR) {
        if (l_scale_solar) {
          double fnew = fd * o.trans_0 * adjust_solar_ke[k];
          su[i] = o.sc1 * fd;
          sd[i] = (o.sc2 - o.trans_0) * fd + fnew;
          fd = fnew;
        } else {
          double fnew = fd * o.trans_0;
          su[i] = o.sc1 * fd;
          sd[i] = o.sc2 * fd;
          fd = fnew;
        }
        flux_direct[(size_t)(i + 1) * P + l] = fd;
      } else if (l_ir_source_quad) {
        su[i] = o.sc1 * diff_planck[k] + o.sc2 * diff_planck_2[k];
        sd[i] = -o.sc1 * diff_planck[k] + o.sc2 * diff_planck_2[k];
      } else {
        su[i] = o.sc1 * diff_planck[k];
        sd[i] = -su[i];
      }
    }
    double source_ground = isolir == IP_SOLAR
        ? (direct_albedo[l] - diffuse_albedo[l]) * fd
        : (1.0 - diffuse_albedo[l]) * d_planck_flux_surface[l];
    soc_oracle_solver_homogen_direct(L, tr, rf, sd, su, diffuse_albedo[l], flux_inc_down[l],
                                     source_ground, ft);
    for (int k = 0; k < 2 * L + 2; k++) flux_total[(size_t)k * P + l] = ft[k];
  }
  free(ft);
  free(tr);
}

/* ------------------------------------------------------------------------ */
/* cloud_gen, socrates_cloud_gen.F90:54-478 (ioverlap 1-3 and 0)            */
#define LCG_M 86436
#define LCG_A 1093
#define LCG_C 18257

static inline int lcg(int x) {
  const double rm = 1.0 / (double)LCG_M;
  int temp = x * LCG_A + LCG_C;
  return temp - (int)((double)temp * rm) * LCG_M;
}

static double zcw_lookup(double y, double sigma, int n1, int n2, const double *xcw) {
  double rind1 = y * (double)(n1 - 1) + 1.0;
  int ind1 = (int)rind1;
  if (ind1 > n1 - 1) ind1 = n1 - 1;
  if (ind1 < 1) ind1 = 1;
  rind1 = rind1 - (double)ind1;
  rind1 = rind1 < 1.0 ? rind1 : 1.0;
  rind1 = rind1 > 0.0 ? rind1 : 0.0;
  double rind2 = 40.0 * sigma - 3.0;
  int ind2 = (int)rind2;
  if (ind2 > n2 - 1) ind2 = n2 - 1;
  if (ind2 < 1) ind2 = 1;
  rind2 = rind2 - (double)ind2;
  rind2 = rind2 < 1.0 ? rind2 : 1.0;
  rind2 = rind2 > 0.0 ? rind2 : 0.0;
#define XCW(a, b) xcw[((a) - 1) + (size_t)((b) - 1) * (size_t)n1]
  return (1.0 - rind1) * (1.0 - rind2) * XCW(ind1, ind2)
       + (1.0 - rind1) * rind2 * XCW(ind1, ind2 + 1)
       + rind1 * (1.0 - rind2) * XCW(ind1 + 1, ind2)
       + rind1 * rind2 * XCW(ind1 + 1, ind2 + 1);
#undef XCW
}

int soc_oracle_cloud_gen(int n_profile, int n_layer, int n_subcol, int n_req, int ipph,
                         int ioverlap, const int *rand_seed, const double *rlc_cf,
                         const double *rlc_cw, const double *sigma_qcw,
                         const double *avg_cf, const double *p_layer, int n1, int n2,
                         const double *xcw, double *c_sub, int *n_subcol_cld) {
  const int P = n_profile, L = n_layer;
  const int rand_seed_max = (2147483647 - LCG_C) / LCG_A;
  const double rm = 1.0 / (double)LCG_M;
  const double cut = DBL_EPSILON;
#define AT(a, i, l) (a)[(size_t)((i) - 1) * P + (l)]
#define CSUB(s, i, l) c_sub[((size_t)((s) - 1) * L + (size_t)((i) - 1)) * P + (l)]
  memset(c_sub, 0, sizeof(double) * (size_t)n_subcol * L * P);
  /* cloud top / base over the whole batch (cloud_gen:224-245) */
  int k_top = 0, k_base = 0;
  for (int k = 1; k <= L && !k_top; k++)
    for (int l = 0; l < P; l++) if (AT(avg_cf, k, l) > cut) { k_top = k; break; }
  if (k_top == 0) {
    for (int l = 0; l < P; l++) n_subcol_cld[l] = 0;
    return 0;
  }
  for (int k = L; k >= 1 && !k_base; k--)
    for (int l = 0; l < P; l++) if (AT(avg_cf, k, l) >= cut) { k_base = k; break; }
  double *alpha = (double *)calloc((size_t)(L + 1), sizeof(double));
  double *rcorr = (double *)calloc((size_t)(L + 1), sizeof(double));
  int *seed_x = (int *)malloc(sizeof(int) * (size_t)n_subcol);
  int *seed_y = (int *)malloc(sizeof(int) * (size_t)n_subcol);
  for (int l = 0; l < P; l++) {
    int rs = abs(rand_seed[l]) % rand_seed_max;
    for (int i = 0; i < n_subcol; i++) rs = lcg(rs);
    for (int i = 0; i < n_subcol; i++) { rs = lcg(rs); seed_x[i] = rs; }
    for (int i = 0; i < n_subcol; i++) { rs = lcg(rs); seed_y[i] = rs; }
    /* overlap factors (cloud_gen:339-376); cloud_top = 1 here */
    for (int k = k_top - 1; k <= k_base - 1; k++) {
      if (ioverlap == 1 || k < 1) { alpha[k] = 0.0; rcorr[k] = 0.0; continue; }
      if (ioverlap == 2 || AT(avg_cf, k, l) > 0.0) {
        alpha[k] = AT(rlc_cf, k, l) > 0.0
          ? exp((AT(p_layer, k, l) - AT(p_layer, k + 1, l)) / AT(rlc_cf, k, l)) : 0.0;
        rcorr[k] = AT(rlc_cw, k, l) > 0.0
          ? exp((AT(p_layer, k, l) - AT(p_layer, k + 1, l)) / AT(rlc_cw, k, l)) : 0.0;
      } else {
        alpha[k] = 0.0; rcorr[k] = 0.0;
      }
    }
    int i_loc = 1;
    double x = 0.0, y = 0.0;
    for (int i = 0; i < n_subcol; i++) {
      int int_x = seed_x[i], int_y = seed_y[i];
      int l_cld = 0;
      for (int k = k_top; k <= k_base; k++) {
        double cf = AT(avg_cf, k, l);
        if (!(cf > 0.0)) continue;
        if (ioverlap == 0) {
          int_x = lcg(int_x); x = (double)int_x * rm;
          int_y = lcg(int_y); y = (double)int_y * rm;
          if (k == 1) {
            int_x = lcg(int_x); x = (double)int_x * rm;
            int_y = lcg(int_y); y = (double)int_y * rm;
          } else if (x < (1.0 - AT(avg_cf, k - 1, l))) {
            int_x = lcg(int_x); double x1 = (double)int_x * rm;
            int_y = lcg(int_y); double y1 = (double)int_y * rm;
            x = x1 * (1.0 - AT(avg_cf, k - 1, l));
            y = y1;
          }
        } else {
          int_x = lcg(int_x); double x1 = (double)int_x * rm;
          int_y = lcg(int_y); double y1 = (double)int_y * rm;
          if (x1 >= alpha[k - 1]) { int_x = lcg(int_x); x = (double)int_x * rm; }
          if (y1 >= rcorr[k - 1]) { int_y = lcg(int_y); y = (double)int_y * rm; }
        }
        if (x < cf) {
          double zcw;
          if (ipph == 0) zcw = 1.0;
          else zcw = zcw_lookup(y, AT(sigma_qcw, k, l), n1, n2, xcw) * 1.0; /* ls_ratio = 1 */
          l_cld = 1;
          CSUB(i_loc, k, l) = zcw;
        }
      }
      if (l_cld) i_loc++;
    }
    n_subcol_cld[l] = i_loc - 1;
    /* fill of set_cld_mcica (socrates_set_cld_mcica.F90:202-220) */
    int n_fill = n_req < n_subcol ? n_req : n_subcol;
    int nc = n_subcol_cld[l];
    if (nc < n_fill && nc > 0)
      for (int j = nc + 1; j <= n_fill; j++)
        for (int k = 1; k <= L; k++) CSUB(j, k, l) = CSUB(j - nc, k, l);
  }
  free(alpha); free(rcorr); free(seed_x); free(seed_y);
#undef AT
#undef CSUB
  return 0;
}

/* ------------------------------------------------------------------------ */
/* runes(): socrates_runes.F90:47-1079 restricted to the flux path.          */

typedef struct {
  int L;
  double *trans, *reflect, *trans_0, *sc1, *sc2, *s_up, *s_down, *flux_direct,
      *flux_total;
} work_t;

/* column_solver for the homogeneous direct solver (column_solver.F90:160-243)
 * with solar_source (solar_source.F90:104-209, l_scale_solar) or ir_source
 * (ir_source.F90:93-120). Coefficients are in w; fluxes out in w. */
static void solve_column(int isolir, work_t *w, const double *adjust, const double *dplanck,
                         const double *dplanck2, double flux_inc_direct, double flux_inc_down,
                         double alb_diff, double alb_dir, double d_planck_flux_surface, int nosc) {
  const int L = w->L;
  double source_ground;
  if (isolir == IP_SOLAR) {
    w->flux_direct[0] = flux_inc_direct;
    for (int i = 1; i <= L; i++) {
      w->flux_direct[i] = w->flux_direct[i - 1] * w->trans_0[i - 1] * adjust[i - 1];
      w->s_up[i - 1] = w->sc1[i - 1] * w->flux_direct[i - 1];
      w->s_down[i - 1] = (w->sc2[i - 1] - w->trans_0[i - 1]) * w->flux_direct[i - 1]
                         + w->flux_direct[i];
    }
    source_ground = (alb_dir - alb_diff) * w->flux_direct[L];
  } else {
    for (int i = 0; i < L; i++) {
      w->s_up[i] = w->sc1[i] * dplanck[i] + w->sc2[i] * dplanck2[i];
      w->s_down[i] = -w->sc1[i] * dplanck[i] + w->sc2[i] * dplanck2[i];
    }
    source_ground = (1.0 - alb_diff) * d_planck_flux_surface;
  }
  /* column_solver.F90:192-199: the non-scattering solver when scattering is
   * ignored */
  if (nosc)
    solver_no_scat(L, w->trans, w->s_down, w->s_up, alb_diff, flux_inc_down, source_ground,
                   w->flux_total);
  else
    soc_oracle_solver_homogen_direct(L, w->trans, w->reflect, w->s_down, w->s_up, alb_diff,
                                     flux_inc_down, source_ground, w->flux_total);
}

/* Padé-2 droplet optics, opt_prop_pade_2_mod.F90:100-143 */
static void pade_2(const double *cp, double mmr, double re, double *k_ext_tot,
                   double *k_ext_scat, double *phf, double *fwd) {
  *k_ext_tot = mmr * (cp[0] + re * (cp[1] + re * cp[2]))
             / (1.0 + re * (cp[3] + re * (cp[4] + re * cp[5])));
  *k_ext_scat = *k_ext_tot * (1.0 - (cp[6] + re * (cp[7] + re * cp[8]))
                               / (1.0 + re * (cp[9] + re * cp[10])));
  double asym = (cp[11] + re * (cp[12] + re * cp[13])) / (1.0 + re * (cp[14] + re * cp[15]));
  *phf = *k_ext_scat * asym;
  *fwd = *phf * asym;
}

/* Droplet optics by the type's scheme: Pade-2 (above), or Slingo-Schrecker
 * (opt_prop_water_cloud.F90:231-247, forward scattering k_scat g^2 with the
 * power formed first, :292-298) */
static void drop_optics(const spec_t *sp, int type, int b, double mmr, double re, double *k_ext_tot,
                        double *k_ext_scat, double *phf, double *fwd) {
  const double *cp = sp->drop_parm[type][b];
  if (sp->i_drop_parm[type] == IP_SLINGO_SCHRECKER) {
    *k_ext_tot = mmr * (cp[0] + cp[1] / re);
    *k_ext_scat = *k_ext_tot * (1.0 - cp[2] - cp[3] * re);
    double asym = cp[4] + cp[5] * re;
    *phf = *k_ext_scat * asym;
    *fwd = *k_ext_scat * (asym * asym);
  } else {
    pade_2(cp, mmr, re, k_ext_tot, k_ext_scat, phf, fwd);
  }
}

/* Ice optics by the type's scheme, for the two-stream case of
 * opt_prop_ice_cloud.F90:227-256 (one phase-function moment, rescaled,
 * forward scattering of order 2): Fu's phase-function fit
 * (opt_prop_fu_phf_mod.F90:93-130), Baran's temperature fit
 * (opt_prop_baran_mod.F90:93-125, t**4 formed left to right as the compiled
 * reference forms it) or Pade-2 (the droplet form). t: the layer
 * temperature. */
static void ice_optics(const spec_t *sp, int type, int b, double iwc, double de, double t,
                       double *k_ext_tot, double *k_ext_scat, double *phf, double *fwd) {
  const double *cp = sp->ice_parm[type][b];
  switch (sp->i_ice_parm[type]) {
    case IP_ICE_FU_PHF: {
      double x = cp[3] / de;
      *k_ext_tot = iwc * ((cp[2] * x + cp[1]) * x + cp[0]);
      x = de / cp[8];
      *k_ext_scat = *k_ext_tot * (1.0 - (cp[4] + x * (cp[5] + x * (cp[6] + x * cp[7]))));
      x = de / cp[13];
      double asym = cp[9] + x * (cp[10] + x * (cp[11] + x * cp[12]));
      *phf = *k_ext_scat * asym;
      *fwd = *phf * asym;
      break;
    }
    case IP_ICE_BARAN: {
      double x = fmin(iwc + 1.0e-12, 1.0e-01);
      *k_ext_tot = cp[0] * x / (t * t * t * t);
      x = fmin(iwc + 1.0e-12, 1.0e-03 - DBL_EPSILON);
      double om = fmin(cp[1] + (cp[2] * x * t), 1.0 - DBL_EPSILON);
      *k_ext_scat = *k_ext_tot * om;
      double asym = fmin(cp[3] + (cp[4] * x * t), 1.0 - DBL_EPSILON);
      *phf = *k_ext_scat * asym;
      *fwd = *phf * asym;
      break;
    }
    default:
      pade_2(cp, iwc, de, k_ext_tot, k_ext_scat, phf, fwd);
      break;
  }
}

/* ------------------------------------------------------------------------ */
/* Triple-column cloud overlap: clear, stratiform and convective regions     */
/* (ip_cloud_conv_strat with maximum-random overlap, ip_cloud_triple, and    */
/* ip_solver_triple_hogan).                                                  */

/* overlap_coupled.F90:150-453 for one profile with ip_cloud_triple (the
 * correlation factor of every region is 1). Layers are 1-based as in the
 * reference; w_cloud, frac_st, frac_cv are indexed by layer-1. ov holds the
 * 18 coefficients of interfaces i = nct-1 .. L at ov[(i-(nct-1))*18 + idx-1]
 * (idx 1-9: v, 10-18: u, as the reference numbers them). */
void soc_oracle_overlap_triple(int L, int nct, const double *w_cloud, const double *frac_st,
                               const double *frac_cv, double *ov);
/* ioverlap: ip_max_rand (0, correlation 1 between like regions),
 * ip_rand (1, correlation 0) or ip_exponential_rand (3, correlation
 * exp((p(i) - p(i+1)) / dp) at interfaces 1 < i < L, with dp = 1 Pa for the
 * clear region and the stratiform / convective decorrelation pressures,
 * overlap_coupled.F90:190-231). p: the profile's layer pressures. */
static void overlap_triple_profile(int L, int nct, const double *w_cloud, const double *frac_st,
                                   const double *frac_cv, int ioverlap, const double *p, double dps,
                                   double dpc, double *ov) {
  const double tol_cloud = 1.0e+02 * DBL_EPSILON;
  double area_upper[4], area_lower[4], ao[4][4], aru[4], arl[4], corr[4];
  area_upper[1] = 1.0; area_upper[2] = 0.0; area_upper[3] = 0.0;
  for (int i = nct - 1; i <= L; i++) {
    double *co = ov + (size_t)(i - (nct - 1)) * 18 - 1; /* co[idx] for idx 1..18 */
    for (int k = 1; k <= 3; k++) {
      if (ioverlap == 0) corr[k] = 1.0;
      else if (ioverlap == 1) corr[k] = 0.0;
      else {
        const double dp_corr = k == 1 ? 1.0 : k == 2 ? dps : dpc;
        /* dp_corr = 0: exp(-inf) = 0 for layers of distinct pressure */
        corr[k] = (i < L && i > 1) ? (dp_corr > 0.0 ? exp((p[i - 1] - p[i]) / dp_corr) : 0.0) : 1.0;
      }
    }
    for (int k = 1; k <= 3; k++) {
      if (i < L) {
        if (k == 1) area_lower[k] = 1.0 - w_cloud[i];           /* w_free(i+1) */
        else area_lower[k] = w_cloud[i] * (k == 2 ? frac_st[i] : frac_cv[i]);
      } else {
        area_lower[k] = k == 1 ? 1.0 : 0.0;
      }
      ao[k][k] = corr[k] * fmin(area_lower[k], area_upper[k]);
    }
    for (int k = 1; k <= 3; k++)
      for (int j = 1; j < k; j++) { ao[k][j] = 0.0; ao[j][k] = 0.0; }
    double art = 1.0 - ao[1][1];
    for (int k = 2; k <= 3; k++) art = art - ao[k][k];
    for (int k = 1; k <= 3; k++) {
      aru[k] = area_upper[k] - ao[k][k];
      arl[k] = area_lower[k] - ao[k][k];
    }
    for (int k = 1; k <= 3; k++)
      for (int j = 1; j <= 3; j++)
        if (art > tol_cloud) ao[k][j] = ao[k][j] + aru[k] * arl[j] / art;
    for (int k = 1; k <= 3; k++) {
      co[3 * (k - 1) + k] = 1.0;
      for (int j = 1; j < k; j++) { co[3 * (j - 1) + k] = 0.0; co[3 * (k - 1) + j] = 0.0; }
    }
    for (int k = 1; k <= 3; k++)
      if (area_upper[k] > tol_cloud)
        for (int j = 1; j <= 3; j++) co[3 * (j - 1) + k] = ao[k][j] / area_upper[k];
    for (int k = 1; k <= 3; k++) {
      co[3 * (3 + k - 1) + k] = 1.0;
      for (int j = 1; j < k; j++) { co[3 * (3 + j - 1) + k] = 0.0; co[3 * (3 + k - 1) + j] = 0.0; }
    }
    for (int k = 1; k <= 3; k++)
      if (area_lower[k] > tol_cloud)
        for (int j = 1; j <= 3; j++) co[3 * (3 + j - 1) + k] = ao[j][k] / area_lower[k];
    if (i < L)
      for (int k = 1; k <= 3; k++) area_upper[k] = area_lower[k];
  }
}

/* exported for the tests */
void soc_oracle_overlap_triple(int L, int nct, const double *w_cloud, const double *frac_st,
                               const double *frac_cv, double *ov) {
  overlap_triple_profile(L, nct, w_cloud, frac_st, frac_cv, 0, NULL, 0.0, 0.0, ov);
}

typedef struct {
  int L, nct;
  double *tr[3], *rf[3], *t0[3], *sc1[3], *sc2[3], *su[3], *sd[3];
  double *binv[3], *gam[3], *h[3];
  double *flux_direct; /* L+1 */
  double *flux_total;  /* 2L+2, 1-based index k at [k-1] */
} tri_t;

/* one k-term of triple_column.F90:255-569 for one profile: two_coeff_region
 * (two_coeff_region.F90:200-445), ir_source per region with the area
 * weights (triple_column.F90), triple_solar_source (triple_solar_source.F90:
 * 180-412), the surface sources and solver_triple_hogan
 * (solver_triple_hogan.F90:207-411). Layer arrays are 0-based (layer i at
 * [i-1]); region 0 clear, 1 stratiform, 2 convective. */
static void solve_triple(int isolir, tri_t *t, const double *tau_c, const double *om_c,
                         const double *g_c, const double *tau_k[2], const double *om_k[2], const double *g_k[2],
                         const double *frac_k[2], const double *w_cloud, const double *adjust,
                         const double *dpl, const double *dpl2, double flux_inc_direct,
                         double flux_inc_down, double alb_diff, double alb_dir, double dps,
                         const double *ov, double zen_0) {
  const int L = t->L, nct = t->nct;
  const double tiny_frac = DBL_MIN;
#define OV(i, idx) ov[(size_t)((i) - (nct - 1)) * 18 + (idx) - 1]
  /* two_coeff_region: clear region, all layers */
  for (int i = 1; i <= L; i++) {
    coeff_t o;
    two_coeff_one(isolir, tau_c[i - 1], om_c[i - 1], g_c[i - 1], zen_0, &o);
    t->tr[0][i - 1] = o.trans; t->rf[0][i - 1] = o.reflect; t->t0[0][i - 1] = o.trans_0;
    t->sc1[0][i - 1] = o.sc1; t->sc2[0][i - 1] = o.sc2;
  }
  for (int r = 1; r <= 2; r++)
    for (int i = nct; i <= L; i++) {
      t->tr[r][i - 1] = 0.0; t->rf[r][i - 1] = 0.0; t->t0[r][i - 1] = 0.0;
      t->sc1[r][i - 1] = 0.0; t->sc2[r][i - 1] = 0.0;
    }
  for (int k = 0; k < 2; k++) {
    const int r = k + 1; /* stratiform type -> stratiform region, convective likewise */
    for (int i = nct; i <= L; i++) {
      const double f = frac_k[k][i - 1];
      if (!(f > tiny_frac)) continue;
      coeff_t o;
      two_coeff_one(isolir, tau_k[k][i - 1], om_k[k][i - 1], g_k[k][i - 1], zen_0, &o);
      t->tr[r][i - 1] = t->tr[r][i - 1] + f * o.trans;
      t->rf[r][i - 1] = t->rf[r][i - 1] + f * o.reflect;
      t->sc1[r][i - 1] = t->sc1[r][i - 1] + f * o.sc1;
      t->sc2[r][i - 1] = t->sc2[r][i - 1] + f * o.sc2;
      if (isolir == IP_SOLAR) t->t0[r][i - 1] = t->t0[r][i - 1] + f * o.trans_0;
    }
  }
  for (int r = 1; r <= 2; r++)
    for (int i = nct; i <= L; i++) {
      const double fr = frac_k[r - 1][i - 1]; /* frac_region = frac_cloud of the type */
      if (!(fr > tiny_frac)) continue;
      const double inv = 1.0 / fr;
      t->tr[r][i - 1] = t->tr[r][i - 1] * inv;
      t->rf[r][i - 1] = t->rf[r][i - 1] * inv;
      if (isolir == IP_SOLAR) t->t0[r][i - 1] = t->t0[r][i - 1] * inv;
      t->sc1[r][i - 1] = t->sc1[r][i - 1] * inv;
      t->sc2[r][i - 1] = t->sc2[r][i - 1] * inv;
    }
  double sfg[3];
  if (isolir == IP_INFRA_RED) {
    for (int r = 0; r < 3; r++) {
      const int n_top = r == 0 ? 1 : nct;
      for (int i = n_top; i <= L; i++) {
        t->su[r][i - 1] = t->sc1[r][i - 1] * dpl[i - 1] + t->sc2[r][i - 1] * dpl2[i - 1];
        t->sd[r][i - 1] = -t->sc1[r][i - 1] * dpl[i - 1] + t->sc2[r][i - 1] * dpl2[i - 1];
      }
    }
    for (int i = nct; i <= L; i++) {
      const double wf = 1.0 - w_cloud[i - 1], wc = w_cloud[i - 1];
      t->sd[0][i - 1] = wf * t->sd[0][i - 1];
      t->su[0][i - 1] = wf * t->su[0][i - 1];
      t->sd[1][i - 1] = wc * frac_k[0][i - 1] * t->sd[1][i - 1];
      t->su[1][i - 1] = wc * frac_k[0][i - 1] * t->su[1][i - 1];
      t->sd[2][i - 1] = wc * frac_k[1][i - 1] * t->sd[2][i - 1];
      t->su[2][i - 1] = wc * frac_k[1][i - 1] * t->su[2][i - 1];
    }
    sfg[0] = OV(L, 10) * (1.0 - alb_diff) * dps;
    sfg[1] = OV(L, 13) * (1.0 - alb_diff) * dps;
    sfg[2] = OV(L, 16) * (1.0 - alb_diff) * dps;
  } else {
    /* triple_solar_source (l_scale_solar) */
    double *fd = t->flux_direct;
    fd[0] = flux_inc_direct;
    for (int i = 1; i <= nct - 1; i++) {
      fd[i] = fd[i - 1] * t->t0[0][i - 1] * adjust[i - 1];
      t->su[0][i - 1] = t->sc1[0][i - 1] * fd[i - 1];
      t->sd[0][i - 1] = (t->sc2[0][i - 1] - t->t0[0][i - 1]) * fd[i - 1] + fd[i];
    }
    double base[3] = {fd[nct - 1], 0.0, 0.0}, top[3];
    for (int i = nct; i <= L; i++) {
      for (int a = 0; a < 3; a++)
        top[a] = OV(i - 1, 3 * a + 1) * base[0] + OV(i - 1, 3 * a + 2) * base[1] +
                 OV(i - 1, 3 * a + 3) * base[2];
      for (int r = 0; r < 3; r++) {
        base[r] = top[r] * t->t0[r][i - 1] * adjust[i - 1];
        t->su[r][i - 1] = t->sc1[r][i - 1] * top[r];
        t->sd[r][i - 1] = (t->sc2[r][i - 1] - t->t0[r][i - 1]) * top[r] + base[r];
      }
      fd[i] = base[0] + base[1] + base[2];
    }
    for (int r = 0; r < 3; r++) sfg[r] = (alb_dir - alb_diff) * base[r];
  }
  /* solver_triple_hogan */
  double a1 = 0, a2 = 0, a3 = 0, g1 = 0, g2 = 0, g3 = 0;
  double **T = t->tr, **R = t->rf, **SU = t->su, **SD = t->sd;
  for (int i = L + 1; i >= 1; i--) {
    if (i < L + 1 && i >= nct) {
      double th1 = a1 * OV(i, 1) + a2 * OV(i, 4) + a3 * OV(i, 7);
      double th2 = a1 * OV(i, 2) + a2 * OV(i, 5) + a3 * OV(i, 8);
      double th3 = a1 * OV(i, 3) + a2 * OV(i, 6) + a3 * OV(i, 9);
      const int j = i - 1;
      t->binv[0][j] = 1.0 / (1.0 - th1 * R[0][j]);
      t->gam[0][j] = th1 * T[0][j];
      t->h[0][j] = g1 + th1 * SD[0][j];
      t->binv[1][j] = 1.0 / (1.0 - th2 * R[1][j]);
      t->gam[1][j] = th2 * T[1][j];
      t->h[1][j] = g2 + th2 * SD[1][j];
      t->binv[2][j] = 1.0 / (1.0 - th3 * R[2][j]);
      t->gam[2][j] = th3 * T[2][j];
      t->h[2][j] = g3 + th3 * SD[2][j];
      double l1 = SU[0][j] + t->h[0][j] * T[0][j] * t->binv[0][j];
      double l2 = SU[1][j] + t->h[1][j] * T[1][j] * t->binv[1][j];
      double l3 = SU[2][j] + t->h[2][j] * T[2][j] * t->binv[2][j];
      a1 = R[0][j] + th1 * T[0][j] * T[0][j] * t->binv[0][j];
      g1 = OV(i - 1, 10) * l1 + OV(i - 1, 11) * l2 + OV(i - 1, 12) * l3;
      a2 = R[1][j] + th2 * T[1][j] * T[1][j] * t->binv[1][j];
      g2 = OV(i - 1, 13) * l1 + OV(i - 1, 14) * l2 + OV(i - 1, 15) * l3;
      a3 = R[2][j] + th3 * T[2][j] * T[2][j] * t->binv[2][j];
      g3 = OV(i - 1, 16) * l1 + OV(i - 1, 17) * l2 + OV(i - 1, 18) * l3;
    } else if (i < nct - 1) {
      const int j = i - 1;
      t->binv[0][j] = 1.0 / (1.0 - a1 * R[0][j]);
      t->gam[0][j] = a1 * T[0][j];
      t->h[0][j] = g1 + a1 * SD[0][j];
      double l1 = T[0][j] * t->binv[0][j];
      a1 = R[0][j] + l1 * t->gam[0][j];
      g1 = SU[0][j] + l1 * t->h[0][j];
    } else if (i == nct - 1) {
      const int j = i - 1;
      double th1;
      if (nct < L) th1 = a1 * OV(i, 1) + a2 * OV(i, 4) + a3 * OV(i, 7);
      else th1 = a1;
      t->binv[0][j] = 1.0 / (1.0 - th1 * R[0][j]);
      t->gam[0][j] = th1 * T[0][j];
      t->h[0][j] = g1 + th1 * SD[0][j];
      double lam = T[0][j] * t->binv[0][j];
      a1 = R[0][j] + lam * t->gam[0][j];
      g1 = SU[0][j] + lam * t->h[0][j];
    } else if (i == L + 1) {
      a1 = alb_diff; a2 = alb_diff; a3 = alb_diff;
      g1 = sfg[0]; g2 = sfg[1]; g3 = sfg[2];
    }
  }
  double *ft = t->flux_total - 1; /* ft[k] for k = 1 .. 2L+2 */
  ft[2] = flux_inc_down;
  if (nct > 1) ft[1] = a1 * ft[2] + g1;
  else ft[1] = g1 + flux_inc_down * (OV(0, 1) * a1 + OV(0, 4) * a2 + OV(0, 7) * a3);
  for (int i = 1; i <= nct - 1; i++) {
    const int j = i - 1;
    ft[2 * i + 1] = (t->gam[0][j] * ft[2 * i] + t->h[0][j]) * t->binv[0][j];
    ft[2 * i + 2] = T[0][j] * ft[2 * i] + R[0][j] * ft[2 * i + 1] + SD[0][j];
  }
  double fd1 = 0, fd2 = 0, fd3 = 0, fu1, fu2, fu3;
  for (int i = nct; i <= L; i++) {
    const int j = i - 1;
    if (i > nct) {
      double d1 = fd1, d2 = fd2, d3 = fd3;
      fd1 = OV(i - 1, 1) * d1 + OV(i - 1, 2) * d2 + OV(i - 1, 3) * d3;
      fd2 = OV(i - 1, 4) * d1 + OV(i - 1, 5) * d2 + OV(i - 1, 6) * d3;
      fd3 = OV(i - 1, 7) * d1 + OV(i - 1, 8) * d2 + OV(i - 1, 9) * d3;
    } else {
      fd1 = OV(i - 1, 1) * ft[2 * i];
      fd2 = OV(i - 1, 4) * ft[2 * i];
      fd3 = OV(i - 1, 7) * ft[2 * i];
    }
    fu1 = (t->gam[0][j] * fd1 + t->h[0][j]) * t->binv[0][j];
    fu2 = (t->gam[1][j] * fd2 + t->h[1][j]) * t->binv[1][j];
    fu3 = (t->gam[2][j] * fd3 + t->h[2][j]) * t->binv[2][j];
    fd1 = T[0][j] * fd1 + R[0][j] * fu1 + SD[0][j];
    fd2 = T[1][j] * fd2 + R[1][j] * fu2 + SD[1][j];
    fd3 = T[2][j] * fd3 + R[2][j] * fu3 + SD[2][j];
    ft[2 * i + 1] = fu1 + fu2 + fu3;
    ft[2 * i + 2] = fd1 + fd2 + fd3;
  }
#undef OV
}

/* atm%gas_mix_ratio of one absorber (socrates_set_atm.F90:227-415, 535-622):
 * full fields clipped at zero, well-mixed values as given, absent gases 0 */
/* gas amounts as set_gas_mix_ratio gives them (socrates_set_atm.F90:535-622):
 * a field clamped at 0, else the well-mixed ratio */
static double wm_amount(const soc_case *c, int slot, int i, int l) {
  const double *f = c->wm_field[slot];
  return f ? fmax(f[(size_t)i * c->n_profile + l], 0.0) : c->wm_mix_ratio[slot];
}

static double gas_mix_ratio(const soc_case *c, int type, int i, int l) {
  const int P = c->n_profile;
  switch (type) {
    case IP_H2O: return fmax(c->h2o[(size_t)i * P + l], 0.0);
    case IP_O3: return fmax(c->o3[(size_t)i * P + l], 0.0);
    case IP_CO2: return c->co2 ? fmax(c->co2[(size_t)i * P + l], 0.0) : c->co2_mix_ratio;
    case IP_N2O: return wm_amount(c, SOC_WM_N2O, i, l);
    case IP_CH4: return wm_amount(c, SOC_WM_CH4, i, l);
    case IP_O2: return wm_amount(c, SOC_WM_O2, i, l);
    case IP_SO2: return wm_amount(c, SOC_WM_SO2, i, l);
    case IP_OCS: return wm_amount(c, SOC_WM_OCS, i, l);
    case IP_CFC11: return wm_amount(c, SOC_WM_CFC11, i, l);
    case IP_CFC12: return wm_amount(c, SOC_WM_CFC12, i, l);
    case IP_CFC113: return wm_amount(c, SOC_WM_CFC113, i, l);
    case IP_HCFC22: return wm_amount(c, SOC_WM_HCFC22, i, l);
    case IP_HFC134A: return wm_amount(c, SOC_WM_HFC134A, i, l);
    default: return 0.0;
  }
}

/* MINLOC(x - tab, mask x - tab >= 0) over n ascending entries (0-based) */
static int minloc_below(double x, const double *tab, int n, int stride) {
  int best = -1;
  double db = 0.0;
  for (int k = 0; k < n; k++) {
    const double d = x - tab[(size_t)k * stride];
    if (d >= 0.0 && (best < 0 || d < db)) { best = k; db = d; }
  }
  return best;
}

/* inter_pt_lookup.F90:147-191 for one layer: indices (0-based) and factors
 * fac00, fac10, fac01, fac11 */
static void inter_pt_lookup(const spec_t *s, double p, double t, int *jp, int *jt, int *jtt,
                            double fac[4]) {
  const double eps = DBL_EPSILON;
  const int np = s->n_pre, nt = s->n_tmp;
  const double pmin = s->p_lookup[0] + fmax(fabs(s->p_lookup[0]) * eps, eps);
  const double pmax = s->p_lookup[np - 1] - fmax(fabs(s->p_lookup[np - 1]) * eps, eps);
  const double pl = fmin(fmax(log(p), pmin), pmax);
  const int ip = minloc_below(pl, s->p_lookup, np, 1);
  const double fp = (pl - s->p_lookup[ip]) / (s->p_lookup[ip + 1] - s->p_lookup[ip]);
  double tl = fmin(fmax(t, s->t_lookup[ip][0] * (1.0 + eps)), s->t_lookup[ip][nt - 1] * (1.0 - eps));
  const int it = minloc_below(tl, s->t_lookup[ip], nt, 1);
  const double ft = (tl - s->t_lookup[ip][it]) / (s->t_lookup[ip][it + 1] - s->t_lookup[ip][it]);
  tl = fmin(fmax(t, s->t_lookup[ip + 1][0] * (1.0 + eps)), s->t_lookup[ip + 1][nt - 1] * (1.0 - eps));
  const int itt = minloc_below(tl, s->t_lookup[ip + 1], nt, 1);
  const double ftt =
      (tl - s->t_lookup[ip + 1][itt]) / (s->t_lookup[ip + 1][itt + 1] - s->t_lookup[ip + 1][itt]);
  const double compfp = 1.0 - fp;
  fac[0] = compfp * (1.0 - ft);
  fac[1] = fp * (1.0 - ftt);
  fac[2] = compfp * ft;
  fac[3] = fp * ftt;
  *jp = ip; *jt = it; *jtt = itt;
}

int soc_oracle_runes(const char *spectral_file, const char *mcica_file, const soc_case *c,
                     double *flux_up, double *flux_down, double *flux_direct,
                     double *heating_rate) {
  static spec_t sp;
  static mcica_t mc;
  const int P = c->n_profile, L = c->n_layer;
  const int isolir = c->i_source;
  const double eps = DBL_EPSILON, tiny = DBL_MIN, huge = DBL_MAX;
  int rc = read_spectrum(spectral_file, &sp);
  if (rc) return rc;
  const int NBAND = sp.n_band;
  /* the scalings restated here: per-term power law, power-quadratic, look-up
   * table, and Wenyi's CO2 tables in band 4 */
  for (int b = 0; b < NBAND; b++)
    for (int j = 0; j < sp.n_band_absorb[b]; j++) {
      const int g = sp.index_absorb[b][j] - 1, fnc = sp.i_scale_fnc[b][g];
      if (sp.i_scale_k[b][g] != IP_SCALE_TERM) return -12;
      if (fnc != IP_SCALE_POWER_LAW && fnc != IP_SCALE_POWER_QUAD && fnc != IP_SCALE_LOOKUP &&
          !(fnc == IP_SCALE_WENYI && b == 3))
        return -12;
    }
  /* ip_cloud_conv_strat: triple-column overlap with maximum-random overlap;
   * McICA is switched off for it (socrates_set_control.F90:338-356) */
  /* ip_cloud_homogen without McICA is the mixed-column representation
   * (ip_cloud_mix_max / ip_cloud_mix_random / ip_cloud_part_corr with
   * solver_mix_direct_hogan, socrates_set_control.F90:319-334); its two
   * regions are the clear and stratiform regions of the triple-column
   * solve with an empty convective region, which the reference computes to
   * the same bits. ip_cloud_conv_strat with random overlap is
   * ip_cloud_mix_random, which the reference rejects (mix_column.F90). */
  int conv_strat = c->l_cloud && c->i_cloud_repr == 3;
  int mixcol = c->l_cloud && c->i_cloud_repr == 1 && c->i_inhom != 2;
  int triple = conv_strat || mixcol;
  int mcica = c->l_cloud && c->i_cloud_repr == 1 && c->i_inhom == 2;
  if (c->l_cloud && !mcica && !triple) return -10;
  if (triple && c->i_overlap != 0 && c->i_overlap != 1 && c->i_overlap != 3) return -10;
  if (conv_strat && c->i_overlap == 1) return -10;
  /* treatment of scattering: runes' default full; any other sets
   * l_grey_single (socrates_set_control.F90:272-277). Ignoring scattering
   * is an infra-red approximation (two_coeff_region_fast_lw.F90:116-120);
   * the triple and mixed columns keep full scattering here (their region
   * coefficients without scattering are not restated), and their
   * approximate-scattering solvers are not either. */
  const int scat = c->i_scatter_method > 0 ? c->i_scatter_method : IP_SCATTER_FULL;
  if (scat > IP_SCATTER_HYBRID) return -13;
  if (c->n_aer_mode > 0 && (!c->aer_mix_ratio || !c->aer_absorption || !c->aer_scattering ||
                            !c->aer_asymmetry))
    return -14;
  const int grey_single = scat != IP_SCATTER_FULL;
  for (int b = 0; b < sp.n_band; b++) {
    const int g = sp.index_absorb[b][0] - 1;
    for (int k = 0; k < sp.i_band_k[b][g]; k++) {
      int sm = scat == IP_SCATTER_HYBRID ? sp.i_scat[b][g][k] : scat;
      if (sm <= 0) sm = IP_SCATTER_FULL;
      if ((sm == IP_NO_SCATTER_ABS || sm == IP_NO_SCATTER_EXT) && (isolir == IP_SOLAR || triple))
        return -13;
      if (sm == IP_SCATTER_APPROX && triple) return -13;
    }
  }
  if (mcica) {
    rc = read_mcica(mcica_file, &mc);
    if (rc) return rc;
  }
#define LAY(a, i, l) (a)[(size_t)(i) * P + (l)]     /* i: 0-based layer */
  /* --- set_cld / set_cld_dim / set_cld_mcica inputs (McICA) ----------- */
  double *w_cloud = NULL, *cond_mmr = NULL, *re = NULL, *c_sub = NULL, *frac_cloudy = NULL;
  double *ice_in = NULL, *de = NULL;   /* in-cloud ice and its clamped dimension */
  const int ice = c->l_cloud && c->i_st_ice > 0;
  if (ice && (c->i_st_ice >= NICE || !sp.l_ice_type[c->i_st_ice])) return -12;
  int *n_subcol_cld = NULL, n_req = 1, n_gen = 1, *subcol_reorder = NULL;
  int first_subcol_k[NB][NK + 1];
  if (mcica) {
    const double min_cloud_fraction = 0.001;
    w_cloud = (double *)calloc((size_t)L * P, sizeof(double));
    cond_mmr = (double *)calloc((size_t)L * P, sizeof(double));
    re = (double *)calloc((size_t)L * P, sizeof(double));
    ice_in = (double *)calloc((size_t)L * P, sizeof(double));
    de = (double *)calloc((size_t)L * P, sizeof(double));
    double *rsd = (double *)calloc((size_t)L * P, sizeof(double));
    double *dp = (double *)calloc((size_t)L * P, sizeof(double));
    double *dpc = (double *)calloc((size_t)L * P, sizeof(double));
    const int it = c->i_st_water;
    for (int i = 0; i < L; i++)
      for (int l = 0; l < P; l++) {
        double frac = fmax(LAY(c->cloud_frac, i, l), 0.0);
        double mmr = fmax(LAY(c->liq_mmr, i, l), 0.0);
        /* in-cloud mixing ratio (socrates_set_cld.F90:817-826) */
        double m_in = frac > eps ? mmr / frac : 0.0;
        /* w_cloud and fraction normalisation (socrates_set_cld.F90:828-852) */
        double wc = frac, fr1 = frac;
        if (wc > min_cloud_fraction) fr1 = fr1 / wc;
        else { wc = 0.0; fr1 = 0.0; }
        if (wc > 1.0 + min_cloud_fraction) return -11;
        else if (wc > 1.0) wc = 1.0;
        LAY(w_cloud, i, l) = wc;
        LAY(cond_mmr, i, l) = m_in;
        /* effective radius clamp (socrates_set_cld_dim.F90:214-220) */
        double r = fmax(LAY(c->liq_dim, i, l), 0.0);
        LAY(re, i, l) = fmin(fmax(r, sp.drop_min[it]), sp.drop_max[it]);
        /* the ice component likewise (in-cloud mixing ratio over the same
         * type fraction, dimension clamped to the ice type's range) */
        double mi = (ice && c->ice_mmr) ? fmax(LAY(c->ice_mmr, i, l), 0.0) : 0.0;
        double i_in = frac > eps ? mi / frac : 0.0;
        LAY(ice_in, i, l) = i_in;
        if (ice) {
          double d = c->ice_dim ? fmax(LAY(c->ice_dim, i, l), 0.0) : 0.0;
          LAY(de, i, l) = fmin(fmax(d, sp.ice_min[c->i_st_ice]), sp.ice_max[c->i_st_ice]);
        }
        /* cond_rsd (socrates_set_cld_mcica.F90:168-180): liquid, then ice;
         * both carry the horizontal rsd */
        double wl = fmax(eps, m_in * fr1);
        double wi = fmax(eps, i_in * fr1);
        double sum_r = 0.0 + c->cloud_horizontal_rsd * wl;
        sum_r = sum_r + c->cloud_horizontal_rsd * wi;
        double sum_w = 0.0 + wl;
        sum_w = sum_w + wi;
        LAY(rsd, i, l) = sum_r / sum_w;
        LAY(dp, i, l) = c->cloud_vertical_decorr;
        LAY(dpc, i, l) = c->cloud_vertical_decorr * 0.5;
      }
    n_gen = c->n_subcol_gen > 0 ? c->n_subcol_gen : mc.n_subcol_gen;
    n_req = mc.n_req_optimal;
    /* sampling tables (socrates_set_cld_mcica.F90:103-151) */
    int (*subk)[NK] = isolir == IP_SOLAR ? mc.sw_subcol_k : mc.lw_subcol_k;
    first_subcol_k[0][0] = 1;
    for (int b = 0; b < NBAND; b++) {
      int maj = sp.index_absorb[b][0] - 1;
      int nk = sp.i_band_k[b][maj];
      for (int k = 0; k < nk; k++) first_subcol_k[b][k + 1] = first_subcol_k[b][k] + subk[b][k];
      if (b < NBAND - 1) first_subcol_k[b + 1][0] = first_subcol_k[b][nk];
    }
    subcol_reorder = (int *)malloc(sizeof(int) * (size_t)(n_req + 1));
    for (int l = 1; l <= n_req; l++) {
      int v = isolir == IP_SOLAR ? l % n_gen : mc.reorder_optimal[l - 1] % n_gen;
      subcol_reorder[l] = v == 0 ? n_gen : v;
    }
    c_sub = (double *)calloc((size_t)n_gen * L * P, sizeof(double));
    n_subcol_cld = (int *)calloc((size_t)P, sizeof(int));
    soc_oracle_cloud_gen(P, L, n_gen, n_req, mc.ipph, c->i_overlap, c->rand_seed, dp, dpc,
                         rsd, w_cloud, c->p_layer, mc.n1, mc.n2, mc.xcw, c_sub, n_subcol_cld);
    frac_cloudy = (double *)malloc(sizeof(double) * (size_t)P);
    for (int l = 0; l < P; l++) frac_cloudy[l] = (double)n_subcol_cld[l] / (double)n_gen;
    free(rsd); free(dp); free(dpc);
  }

  /* --- set_cld / set_cld_dim for the triple-column representation ----- */
  double *tw = NULL, *tfs = NULL, *tfc = NULL, *tms = NULL, *tmc = NULL, *trs = NULL, *trc = NULL;
  double *tis = NULL, *tic = NULL, *tds = NULL, *tdc = NULL;   /* ice of each type */
  double *ovl = NULL;
  int nct = L;
  if (triple) {
    const double min_cloud_fraction = 0.001;
    tw = (double *)calloc((size_t)7 * L * P, sizeof(double));
    tfs = tw + (size_t)L * P; tfc = tfs + (size_t)L * P; tms = tfc + (size_t)L * P;
    tmc = tms + (size_t)L * P; trs = tmc + (size_t)L * P; trc = trs + (size_t)L * P;
    tis = (double *)calloc((size_t)4 * L * P, sizeof(double));
    tic = tis + (size_t)L * P; tds = tic + (size_t)L * P; tdc = tds + (size_t)L * P;
    const int its = c->i_st_water, itc = c->i_st_water;  /* the case carries one droplet type for both */
    for (int i = 0; i < L; i++)
      for (int l = 0; l < P; l++) {
        /* type fractions and condensate (socrates_set_cld.F90:567-581, 284-330);
         * the mixed column (ip_cloud_homogen) reads no convective field
         * (socrates_set_cld.F90:477-495, l_combine false) */
        double fs = fmax(LAY(c->cloud_frac, i, l), 0.0);
        double fcv = conv_strat ? fmax(LAY(c->conv_frac, i, l), 0.0) : 0.0;
        double ms = fmax(LAY(c->liq_mmr, i, l), 0.0);
        double mcv = conv_strat ? fmax(LAY(c->liq_conv_mmr, i, l), 0.0) : 0.0;
        /* in-cloud mixing ratios (:817-826) */
        LAY(tms, i, l) = fs > eps ? ms / fs : 0.0;
        LAY(tmc, i, l) = fcv > eps ? mcv / fcv : 0.0;
        if (ice) {
          double is = c->ice_mmr ? fmax(LAY(c->ice_mmr, i, l), 0.0) : 0.0;
          double icv = (conv_strat && c->ice_conv_mmr) ? fmax(LAY(c->ice_conv_mmr, i, l), 0.0) : 0.0;
          LAY(tis, i, l) = fs > eps ? is / fs : 0.0;
          LAY(tic, i, l) = fcv > eps ? icv / fcv : 0.0;
          double ds = c->ice_dim ? fmax(LAY(c->ice_dim, i, l), 0.0) : 0.0;
          double dc = (conv_strat && c->ice_conv_dim) ? fmax(LAY(c->ice_conv_dim, i, l), 0.0) : 0.0;
          LAY(tds, i, l) = fmin(fmax(ds, sp.ice_min[c->i_st_ice]), sp.ice_max[c->i_st_ice]);
          LAY(tdc, i, l) = fmin(fmax(dc, sp.ice_min[c->i_st_ice]), sp.ice_max[c->i_st_ice]);
        }
        /* w_cloud and normalised type fractions (:828-852) */
        double wc = fs + fcv;
        if (wc > min_cloud_fraction) { fs = fs / wc; fcv = fcv / wc; }
        else { wc = 0.0; fs = 0.0; fcv = 0.0; }
        if (wc > 1.0 + min_cloud_fraction) return -11;
        else if (wc > 1.0) wc = 1.0;
        LAY(tw, i, l) = wc; LAY(tfs, i, l) = fs; LAY(tfc, i, l) = fcv;
        /* radius clamps (socrates_set_cld_dim.F90:214-220) */
        LAY(trs, i, l) = fmin(fmax(fmax(LAY(c->liq_dim, i, l), 0.0), sp.drop_min[its]), sp.drop_max[its]);
        LAY(trc, i, l) = fmin(fmax(conv_strat ? fmax(LAY(c->liq_conv_dim, i, l), 0.0) : 0.0,
                                   sp.drop_min[itc]), sp.drop_max[itc]);
      }
    /* cloud top over the whole call (set_cloud_geometry.F90:52-72) */
    nct = 1;
    for (;;) {
      int any = 0;
      for (int l = 0; l < P; l++) any |= LAY(tw, nct - 1, l) > 0.0;
      if (any || nct >= L) break;
      nct++;
    }
    ovl = (double *)malloc(sizeof(double) * (size_t)(L - nct + 2) * 18);
  }
  tri_t tri;
  double *tbuf = NULL;
  double *kcs[2][4], *tau_k[2], *om_k[2], *g_k[2], *frk[2], *wcol = NULL, *pcol = NULL;
  if (triple) {
    tbuf = (double *)malloc(sizeof(double) * (size_t)(61 * (L + 2) + 2 * L + 4));
    double *qq = tbuf;
    tri.L = L; tri.nct = nct;
    for (int r = 0; r < 3; r++) {
      tri.tr[r] = qq; qq += L; tri.rf[r] = qq; qq += L; tri.t0[r] = qq; qq += L;
      tri.sc1[r] = qq; qq += L; tri.sc2[r] = qq; qq += L; tri.su[r] = qq; qq += L;
      tri.sd[r] = qq; qq += L; tri.binv[r] = qq; qq += L; tri.gam[r] = qq; qq += L;
      tri.h[r] = qq; qq += L;
    }
    for (int k = 0; k < 2; k++) {
      for (int m = 0; m < 4; m++) { kcs[k][m] = qq; qq += L; }
      tau_k[k] = qq; qq += L; om_k[k] = qq; qq += L; g_k[k] = qq; qq += L; frk[k] = qq; qq += L;
    }
    wcol = qq; qq += L;
    pcol = qq; qq += L;
    tri.flux_direct = qq; qq += L + 1;
    tri.flux_total = qq; qq += 2 * L + 2;
  }

  /* --- per-column radiative transfer ------------------------------------ */
  work_t w, wc;
  double *buf = (double *)malloc(sizeof(double) * (size_t)(36 * (L + 2) + 16 * (2 * L + 2)));
  double *q = buf;
#define TAKE(n) (q += (n), q - (n))
  w.L = wc.L = L;
  w.trans = TAKE(L); w.reflect = TAKE(L); w.trans_0 = TAKE(L); w.sc1 = TAKE(L);
  w.sc2 = TAKE(L); w.s_up = TAKE(L); w.s_down = TAKE(L); w.flux_direct = TAKE(L + 1);
  w.flux_total = TAKE(2 * L + 2);
  wc.trans = TAKE(L); wc.reflect = TAKE(L); wc.trans_0 = TAKE(L); wc.sc1 = TAKE(L);
  wc.sc2 = TAKE(L); wc.s_up = TAKE(L); wc.s_down = TAKE(L); wc.flux_direct = TAKE(L + 1);
  wc.flux_total = TAKE(2 * L + 2);
  double *k_grey_tot = TAKE(L), *k_ext_scat = TAKE(L), *keqv = TAKE(L), *kgrey = TAKE(L);
  double *adjust = TAKE(L), *kmin = TAKE(L), *tau_clr = TAKE(L), *om_clr = TAKE(L);
  double *planck = TAKE(L + 1), *dpl = TAKE(L), *dpl2 = TAKE(L), *flux_gas = TAKE(L + 1);
  double *kc_tot = TAKE(L), *kc_scat = TAKE(L), *kc_phf = TAKE(L), *kc_fwd = TAKE(L);
  double *ki_tot = TAKE(L), *ki_scat = TAKE(L), *ki_phf = TAKE(L), *ki_fwd = TAKE(L);
  double *sum_flux = TAKE(2 * L + 3), *sum_k_flux = TAKE(2 * L + 3);
  double *ft = TAKE(2 * L + 2), *tg = TAKE(L), *src = TAKE(L);
  double *kg0 = TAKE(L);  /* clear-air grey extinction before k_eqv */
  /* clear-air products of scattering and the first phase-function moment
   * and forward scattering (aerosols), unnormalised; the normalised forward
   * scattering and delta-rescaled asymmetry of clear air */
  double *pf_clr = TAKE(L), *fw_clr = TAKE(L), *fw_n = TAKE(L), *g_clr = TAKE(L);
  double *acc_total = TAKE(2 * L + 2), *acc_direct = TAKE(L + 1);
  double *clr_total = TAKE(2 * L + 2), *clr_direct = TAKE(L + 1);
  double *up = TAKE(L + 1), *down = TAKE(L + 1), *dir = TAKE(L + 1);
  double *kabs = (double *)malloc(sizeof(double) * (size_t)NG * NK * L);
#define KABS(j, k, i) kabs[((size_t)(j) * NK + (k)) * L + (i)]
  const double temp_max = log(1.0 / eps);

  int *lu_jp = (int *)malloc(sizeof(int) * (size_t)3 * L), *lu_jt = lu_jp + L, *lu_jtt = lu_jt + L;
  double (*lu_fac)[4] = (double (*)[4])malloc(sizeof(double) * 4 * (size_t)L);
  for (int l = 0; l < P; l++) {
    if (sp.n_pre > 0)
      for (int i = 0; i < L; i++)
        inter_pt_lookup(&sp, LAY(c->p_layer, i, l), LAY(c->t_layer, i, l), &lu_jp[i], &lu_jt[i], &lu_jtt[i],
                        lu_fac[i]);
    if (triple) {
      double *fs1 = frk[0], *fc1 = frk[1];
      for (int i = 0; i < L; i++) {
        wcol[i] = LAY(tw, i, l); fs1[i] = LAY(tfs, i, l); fc1[i] = LAY(tfc, i, l);
      }
      for (int i = 0; i < L; i++) pcol[i] = LAY(c->p_layer, i, l);
      overlap_triple_profile(L, nct, wcol, fs1, fc1, c->i_overlap, pcol, c->cloud_vertical_decorr,
                             c->conv_vertical_decorr, ovl);
    }
    double zen_0 = 1.0, sol = 0.0;
    if (isolir == IP_SOLAR) {
      /* set_bound (socrates_set_bound.F90:611-618) */
      if (c->cos_zen[l] > 0.0) { sol = c->solar_irrad[l]; zen_0 = 1.0 / c->cos_zen[l]; }
      else { sol = 0.0; zen_0 = 1.0; }
    }
    int first = 1;
    for (int b = 0; b < NBAND; b++) {
      const double alb_diff = c->albedo_diff[(size_t)b * P + l];
      const double alb_dir = c->albedo_dir[(size_t)b * P + l];
      /* gas absorption per absorber and k-term (radiance_calc.F90:1457-1491,
       * scale_absorb.F90:137-152) */
      const int ngas = sp.n_band_absorb[b];
      for (int j = 0; j < ngas; j++) {
        int g = sp.index_absorb[b][j] - 1;
        int type = sp.type_absorb[g];
        const int fnc = sp.i_scale_fnc[b][g];
        for (int k = 0; k < sp.i_band_k[b][g]; k++) {
          double p_inv = 1.0 / (sp.p_ref[b][g] + 0.0), t_inv = 1.0 / sp.t_ref[b][g];
          const double *sc = sp.scale[b][g][k];
          for (int i = 0; i < L; i++) {
            const double p = LAY(c->p_layer, i, l), tt = LAY(c->t_layer, i, l);
            const double mr = gas_mix_ratio(c, type, i, l);
            if (fnc == IP_SCALE_LOOKUP) {
              /* radiance_calc.F90:1136-1152, then :1432-1442 */
              const double *kl = sp.k_lookup[b][g] + (size_t)k * sp.n_pre * sp.n_tmp;
              const int ip = lu_jp[i], jt = lu_jt[i], jtt = lu_jtt[i], nt = sp.n_tmp;
              double ke = fmax(0.0, lu_fac[i][0] * kl[ip * nt + jt] + lu_fac[i][1] * kl[(ip + 1) * nt + jtt] +
                                        lu_fac[i][2] * kl[ip * nt + jt + 1] +
                                        lu_fac[i][3] * kl[(ip + 1) * nt + jtt + 1]);
              KABS(j, k, i) = fmax(0.0, ke * mr);
              continue;
            }
            double frac;
            if (fnc == IP_SCALE_POWER_LAW) {
              /* scale_absorb.F90:137-152 */
              double pwk = pow((p + 0.0) * p_inv, sc[0]);
              double twk = pow(tt * t_inv, sc[1]);
              frac = pwk * twk;
            } else if (fnc == IP_SCALE_POWER_QUAD) {
              /* scale_absorb.F90:176-193 */
              double pwk = pow((p + 0.0) * p_inv, sc[0]);
              double tmp = tt * t_inv - 1.0;
              frac = pwk * (1.0 + tmp * sc[1] + sc[2] * tmp * tmp);
            } else {
              /* ip_scale_wenyi in band 4 (scale_absorb.F90:238-259) */
              double cgp = fmax(-5.5, log(p / 100.0));
              int jp = (int)((5.5 + cgp) * 2.0) + 1;
              jp = jp < 1 ? 1 : jp > 26 ? 26 : jp;
              int jp1 = jp + 1 < 26 ? jp + 1 : 26;
              int jw = (int)((tt - 240.0) / 20.0 + 7.0);
              jw = jw < 1 ? 1 : jw > 10 ? 10 : jw;
              int jw1 = jw + 1 < 10 ? jw + 1 : 10;
              const double gkpb = kWenyiGk4[jw - 1][jp - 1][k] +
                                  (kWenyiGk4[jw - 1][jp1 - 1][k] - kWenyiGk4[jw - 1][jp - 1][k]) *
                                      (cgp - kWenyiPlg[jp - 1]) * 2.0;
              const double gkpc = kWenyiGk4[jw1 - 1][jp - 1][k] +
                                  (kWenyiGk4[jw1 - 1][jp1 - 1][k] - kWenyiGk4[jw1 - 1][jp - 1][k]) *
                                      (cgp - kWenyiPlg[jp - 1]) * 2.0;
              frac = (gkpb + (gkpc - gkpb) * (tt - kWenyiTtb[jw - 1]) / 20.0) / kWenyiGk250b[k];
            }
            KABS(j, k, i) = sp.k[b][g][k] * fmax(0.0, frac * mr);
          }
        }
      }
      /* grey optics: continua (rescale_continuum.F90:107-179 via
       * radiance_calc.F90:897-923; grey_opt_prop.F90:801-816), then Rayleigh
       * (grey_opt_prop.F90:476-541, 816-835). The continuum data are taken
       * by the continuum's position in the band, as radiance_calc passes
       * them. */
      for (int i = 0; i < L; i++) {
        k_ext_scat[i] = (isolir == IP_SOLAR && c->l_rayleigh) ? sp.rayleigh_coeff[b] : 0.0;
        k_grey_tot[i] = 0.0;
        pf_clr[i] = 0.0;
        fw_clr[i] = 0.0;
      }
      /* modal aerosols (grey_opt_prop.F90:716-743 over every layer:
       * opt_prop_ukca_aerosol.F90:135-200, two-stream: one phase-function
       * moment, forward scattering of order 2), mode by mode */
      for (int j = 0; j < c->n_aer_mode; j++)
        for (int i = 0; i < L; i++) {
          const size_t im = ((size_t)j * L + i) * P + l;
          const size_t ib = (((size_t)b * c->n_aer_mode + j) * L + i) * P + l;
          const double mix = c->aer_mix_ratio[im];
          k_grey_tot[i] = k_grey_tot[i] + mix * c->aer_absorption[ib];
          const double ksc = mix * c->aer_scattering[ib];
          k_ext_scat[i] = k_ext_scat[i] + ksc;
          const double asym = c->aer_asymmetry[ib];
          pf_clr[i] = pf_clr[i] + ksc * asym;
          const double ks_phf = ksc * asym;
          fw_clr[i] = fw_clr[i] + ks_phf * asym;
        }
      for (int jc = 1; jc <= (c->l_no_continuum ? 0 : sp.n_band_continuum[b]); jc++) {
        const int icont = sp.index_continuum[b][jc - 1];
        const int fnc = sp.i_scale_fnc_cont[b][jc];
        const double *sc = sp.scale_cont[b][jc];
        const double pref = sp.p_ref_cont[jc][b], tref = sp.t_ref_cont[jc][b];
        for (int i = 0; i < L; i++) {
          const double p = LAY(c->p_layer, i, l), tt = LAY(c->t_layer, i, l);
          const double rho = LAY(c->density, i, l), wf = fmax(LAY(c->h2o, i, l), 0.0);
          double pwk = pow(p / pref, sc[0]), amount = 0.0;
          if (fnc == IP_SCALE_POWER_LAW) amount = pwk * pow(tt / tref, sc[1]);
          else if (fnc == IP_SCALE_POWER_QUAD)
            amount = pwk * (1.0 + sc[1] * (tt / tref - 1.0) + sc[2] * ((tt / tref - 1.0) * (tt / tref - 1.0)));
          if (icont == IP_SELF_CONTINUUM) {
            double mdw = rho * wf / (REPSILON * MOL_WEIGHT_AIR);
            amount = amount * mdw * wf;
          } else if (icont == IP_FRN_CONTINUUM) {
            amount = amount * rho * wf * (1.0 - wf) / MOL_WEIGHT_AIR;
          }
          k_grey_tot[i] = k_grey_tot[i] + sp.k_cont[b][jc] * amount;
        }
      }
      for (int i = 0; i < L; i++) {
        k_grey_tot[i] = k_grey_tot[i] + k_ext_scat[i];
        kg0[i] = k_grey_tot[i];
        /* the clear-air moments divided by the scattering where there is
         * any (grey_opt_prop.F90:1360-1440, tmp_inv), the phase function
         * delta-rescaled (rescale_phase_fnc.F90:116-124) */
        double fw = fw_clr[i], pf = pf_clr[i];
        if (k_ext_scat[i] > DBL_MIN) {
          const double inv = 1.0 / k_ext_scat[i];
          fw = fw * inv;
          pf = pf * inv;
        }
        fw_n[i] = fw;
        g_clr[i] = (pf - fw) / (1.0 - fw);
      }
      /* Planck terms (diff_planck_source_mod.F90:100-134) */
      double flux_ground = 0.0;
      if (isolir == IP_INFRA_RED) {
        planck[0] = planck_flux_band(&sp, b, LAY(c->t_level, 0, l));
        for (int i = 1; i <= L; i++) {
          planck[i] = planck_flux_band(&sp, b, LAY(c->t_level, i, l));
          dpl[i - 1] = planck[i] - planck[i - 1];
        }
        /* the quadratic term (diff_planck_source_mod.F90:112-121); a linear
         * source has none, which the sources below take as a zero term
         * (c2 * 0 added: ir_source.F90:40-59 gives the same) */
        for (int i = 1; i <= L; i++) {
          double bt = planck_flux_band(&sp, b, LAY(c->t_layer, i - 1, l));
          dpl2[i - 1] = c->l_ir_source_linear ? 0.0 : 2.0 * (planck[i] + planck[i - 1] - 2.0 * bt);
        }
        /* surface emission per band when given (socrates_set_bound.F90:
         * 251-290), else the Planck flux of the surface temperature
         * (diff_planck_source_mod.F90:123-136) */
        if (c->l_flux_ground)
          flux_ground = (c->flux_ground ? c->flux_ground[(size_t)b * P + l] : 0.0) / (1.0 - alb_diff);
        else
          flux_ground = planck_flux_band(&sp, b, c->t_ground ? c->t_ground[l] : 0.0);
      }
      /* equivalent extinction of minor gases (solve_band_k_eqv_scl.F90:
       * 572-1047) */
      for (int i = 0; i < L; i++) { keqv[i] = 0.0; kgrey[i] = 0.0; adjust[i] = 1.0; }
      for (int j = 1; j < ngas; j++) {
        int g = sp.index_absorb[b][j] - 1;
        int nk = sp.i_band_k[b][g];
        /* a single k-term, with l_grey_single: its absorption is added to
         * the grey extinction directly (solve_band_k_eqv_scl.F90:941-951) */
        if (nk == 1 && grey_single) {
          for (int i = 0; i < L; i++) kgrey[i] = kgrey[i] + KABS(j, 0, i);
          continue;
        }
        for (int i = 0; i < L; i++) kmin[i] = huge;
        if (isolir == IP_SOLAR) {
          double sflux = 0.0;
          flux_gas[0] = 1.0;
          for (int i = 1; i <= L; i++) { flux_gas[i] = 0.0; sum_k_flux[i] = 0.0; }
          for (int iex = 0; iex < nk; iex++) {
            double wk = sp.w[b][g][iex];
            double fdt = wk, fdt_prev;
            for (int i = 1; i <= L; i++) {
              double temp = -KABS(j, iex, i - 1) * LAY(c->mass, i - 1, l) * zen_0;
              fdt_prev = fdt;
              fdt = fdt_prev * exp(temp);
              flux_gas[i] = flux_gas[i] + fdt;
            }
            for (int i = 1; i <= L; i++) {
              sum_k_flux[i] = sum_k_flux[i] + KABS(j, iex, i - 1) * fdt;
              kmin[i - 1] = fmin(kmin[i - 1], KABS(j, iex, i - 1));
            }
            sflux = sflux + fdt;
          }
          for (int i = 1; i <= L; i++) {
            if (sflux > 0.0) keqv[i - 1] = keqv[i - 1] + sum_k_flux[i] / sflux;
            else keqv[i - 1] = keqv[i - 1] + kmin[i - 1];
            if (flux_gas[i - 1] > 0.0)
              adjust[i - 1] = adjust[i - 1] * flux_gas[i] / flux_gas[i - 1];
          }
        } else {
          for (int i = 1; i <= 2 * L + 2; i++) { sum_flux[i] = 0.0; sum_k_flux[i] = 0.0; }
          for (int iex = 0; iex < nk; iex++) {
            double wk = sp.w[b][g][iex];
            double flux_inc_down = -planck[0];
            double dps = flux_ground - planck[L];
            for (int i = 0; i < L; i++) {
              tg[i] = KABS(j, iex, i) * LAY(c->mass, i, l);
              kmin[i] = fmin(kmin[i], KABS(j, iex, i));
            }
            /* monochromatic_gas_flux.F90:115-196 (infra-red) */
            double eps_r = eps, sqe = sqrt(eps_r);
            for (int i = 0; i < L; i++) {
              double tr = exp(-DIFF_MINOR * tg[i]);
              src[i] = (1.0 - tr + sqe) * dpl[i] / (DIFF_MINOR * tg[i] + sqe);
              tg[i] = tr; /* reuse: transmission */
            }
            /* ft: 1-based index k at ft[k-1] */
            ft[1] = flux_inc_down;
            for (int i = 1; i <= L; i++) ft[2 * i + 1] = tg[i - 1] * ft[2 * i - 1] + (-src[i - 1]);
            ft[2 * L] = dps + alb_diff * ft[2 * L + 1];
            for (int i = L; i >= 1; i--) ft[2 * i - 2] = tg[i - 1] * ft[2 * i] + src[i - 1];
            for (int i = 2; i <= 2 * L + 1; i++) {
              sum_k_flux[i] = sum_k_flux[i] + KABS(j, iex, i / 2 - 1) * wk * fabs(ft[i - 1]);
              sum_flux[i] = sum_flux[i] + wk * fabs(ft[i - 1]);
            }
          }
          for (int i = 1; i <= L; i++) {
            double lkf = sum_k_flux[2 * i] + sum_k_flux[2 * i + 1];
            double lf = sum_flux[2 * i] + sum_flux[2 * i + 1];
            if (lf > 0.0) keqv[i - 1] = keqv[i - 1] + lkf / lf;
            else keqv[i - 1] = keqv[i - 1] + kmin[i - 1];
          }
        }
      }
      if (isolir == IP_SOLAR)
        for (int i = 0; i < L; i++)
          adjust[i] = adjust[i] * exp(fmin(keqv[i] * LAY(c->mass, i, l) * zen_0, temp_max));
      for (int i = 0; i < L; i++) k_grey_tot[i] = k_grey_tot[i] + keqv[i] + kgrey[i];

      /* cloud component optics (opt_prop_water_cloud.F90:212-300 with the
       * Padé-2 branch opt_prop_pade_2_mod.F90:100-143; opt_prop_ice_cloud.F90
       * :227-256 for the ice) */
      if (mcica) {
        const int it = c->i_st_water;
        for (int i = 0; i < L; i++) {
          if (LAY(w_cloud, i, l) > 0.0)
            drop_optics(&sp, it, b, LAY(cond_mmr, i, l), LAY(re, i, l), &kc_tot[i],
                        &kc_scat[i], &kc_phf[i], &kc_fwd[i]);
          else kc_tot[i] = kc_scat[i] = kc_phf[i] = kc_fwd[i] = 0.0;
          if (ice && LAY(w_cloud, i, l) > 0.0)
            ice_optics(&sp, c->i_st_ice, b, LAY(ice_in, i, l), LAY(de, i, l), LAY(c->t_layer, i, l),
                       &ki_tot[i], &ki_scat[i], &ki_phf[i], &ki_fwd[i]);
          else ki_tot[i] = ki_scat[i] = ki_phf[i] = ki_fwd[i] = 0.0;
        }
      }

      /* triple column: per cloud type, the grey totals with the droplet
       * optics (grey_opt_prop.F90:975-1000, 1185-1230), the equivalent
       * extinction (solve_band_k_eqv_scl.F90:1038-1045), normalised moments
       * (grey_opt_prop.F90:1474-1500) and the delta-rescaled phase function
       * (rescale_phase_fnc.F90, radiance_calc.F90:1027-1040) */
      if (triple) {
        for (int k = 0; k < 2; k++) {
          const int it = k == 0 ? c->i_st_water : c->i_st_water;
          const double *mm = k == 0 ? tms : tmc, *rr = k == 0 ? trs : trc;
          const double *im = k == 0 ? tis : tic, *id = k == 0 ? tds : tdc;
          for (int i = 0; i < L; i++) {
            /* every type starts from the clear-air totals (grey_opt_prop.F90:
             * 983-990): Rayleigh and the aerosols' scattering, and the
             * aerosols' unnormalised moment products */
            double kgt = kg0[i], ksc = k_ext_scat[i], phf = pf_clr[i], fwd = fw_clr[i];
            if (i + 1 >= nct && LAY(tw, i, l) > 0.0) {
              double ct, cs, cp, cf;
              drop_optics(&sp, it, b, LAY(mm, i, l), LAY(rr, i, l), &ct, &cs, &cp, &cf);
              kgt = kgt + ct;
              ksc = ksc + cs;
              phf = phf + cp;
              fwd = fwd + cf;
              /* the type's ice component */
              ct = cs = cp = cf = 0.0;
              if (ice)
                ice_optics(&sp, c->i_st_ice, b, LAY(im, i, l), LAY(id, i, l), LAY(c->t_layer, i, l),
                           &ct, &cs, &cp, &cf);
              kgt = kgt + ct;
              ksc = ksc + cs;
              phf = phf + cp;
              fwd = fwd + cf;
            }
            kgt = kgt + keqv[i] + kgrey[i];
            if (ksc > tiny) {
              double inv = 1.0 / ksc;
              fwd = fwd * inv;
              phf = phf * inv;
            }
            phf = (phf - fwd) / (1.0 - fwd);
            kcs[k][0][i] = kgt; kcs[k][1][i] = ksc; kcs[k][2][i] = phf; kcs[k][3][i] = fwd;
          }
        }
      }

      /* major k-term loop (solve_band_k_eqv_scl.F90:1051-1337) */
      const int gmaj = sp.index_absorb[b][0] - 1;
      const int nkmaj = sp.i_band_k[b][gmaj];
      for (int iex = 0; iex < nkmaj; iex++) {
        const double esft_weight = sp.w[b][gmaj][iex];
        double flux_inc_down, flux_inc_direct = 0.0, dps = 0.0;
        if (isolir == IP_SOLAR) {
          double sib = sol * sp.solar_flux_band[b];
          flux_inc_down = sib / zen_0;
          flux_inc_direct = sib / zen_0;
        } else {
          flux_inc_down = -planck[0];
          dps = flux_ground - planck[L];
        }
        /* the term's treatment of scattering (solve_band_k_eqv_scl.F90:
         * 1053-1058) */
        int sm = scat == IP_SCATTER_HYBRID ? sp.i_scat[b][gmaj][iex] : scat;
        if (sm <= 0) sm = IP_SCATTER_FULL;
        const int nosc = sm == IP_NO_SCATTER_ABS || sm == IP_NO_SCATTER_EXT;
        const int quad = !c->l_ir_source_linear;
        /* single scattering + delta-rescaling, clear (single_scattering.F90:
         * 114-128, rescale_tau_omega.F90:73-80 with zero forward scatter);
         * without scattering the optical depth of absorption (:132-143) or
         * of extinction (:147-158) and no albedo */
        for (int i = 0; i < L; i++) {
          if (nosc) {
            double tau = sm == IP_NO_SCATTER_ABS ? (k_grey_tot[i] + KABS(0, iex, i) - k_ext_scat[i])
                                                 : (k_grey_tot[i] + KABS(0, iex, i));
            tau_clr[i] = tau * LAY(c->mass, i, l);
            om_clr[i] = 0.0;
            continue;
          }
          double k_total = k_grey_tot[i] + KABS(0, iex, i);
          double tau = k_total * LAY(c->mass, i, l);
          double om = k_ext_scat[i] / (k_total + tiny);
          om = fmin(om, 1.0 - 3.2e+01 * eps);
          tau = tau * (1.0 - om * fw_n[i]);
          om = om * (1.0 - fw_n[i]) / (1.0 - om * fw_n[i]);
          tau_clr[i] = tau; om_clr[i] = om;
        }
        /* clear coefficients (two_stream / mcica_column.F90:266-298; without
         * scattering two_coeff_fast_lw, :287-298) */
        for (int i = 0; i < L; i++) {
          coeff_t o;
          if (nosc) two_coeff_fast_lw_one(tau_clr[i], quad, &o);
          else two_coeff_one(isolir, tau_clr[i], om_clr[i], g_clr[i], zen_0, &o);
          w.trans[i] = o.trans; w.reflect[i] = o.reflect; w.trans_0[i] = o.trans_0;
          w.sc1[i] = o.sc1; w.sc2[i] = o.sc2;
        }
        double *F, *FD;
        if (triple) {
          /* single_scattering_all.F90 and rescale_tau_omega.F90 per type */
          for (int k = 0; k < 2; k++)
            for (int i = 0; i < L; i++) {
              double k_total = kcs[k][0][i] + KABS(0, iex, i);
              double tau = k_total * LAY(c->mass, i, l);
              double om = kcs[k][1][i] / (k_total + tiny);
              om = fmin(om, 1.0 - 3.2e+01 * eps);
              const double fw = kcs[k][3][i];
              tau = tau * (1.0 - om * fw);
              om = om * (1.0 - fw) / (1.0 - om * fw);
              tau_k[k][i] = tau; om_k[k][i] = om; g_k[k][i] = kcs[k][2][i];
            }
          const double *tk[2] = {tau_k[0], tau_k[1]}, *ok[2] = {om_k[0], om_k[1]};
          const double *gk[2] = {g_k[0], g_k[1]}, *fk[2] = {frk[0], frk[1]};
          solve_triple(isolir, &tri, tau_clr, om_clr, g_clr, tk, ok, gk, fk, wcol, adjust, dpl, dpl2,
                       flux_inc_direct, flux_inc_down, alb_diff, alb_dir, dps, ovl, zen_0);
          F = tri.flux_total;
          FD = tri.flux_direct;
        } else {
          solve_column(isolir, &w, adjust, dpl, dpl2, flux_inc_direct, flux_inc_down, alb_diff,
                       alb_dir, dps, nosc);
          F = w.flux_total;
          FD = w.flux_direct;
        }
        if (mcica) {
          /* mcica_sample.F90:453-930 */
          memcpy(clr_total, w.flux_total, sizeof(double) * (2 * L + 2));
          memcpy(clr_direct, w.flux_direct, sizeof(double) * (L + 1));
          int m0 = first_subcol_k[b][iex], m1 = first_subcol_k[b][iex + 1];
          int nsk = m1 - m0;
          double subcol_k_inv = 1.0 / (double)nsk;
          for (int m = m0; m < m1; m++) {
            int isub = m % n_req;
            if (isub == 0) isub = n_req;
            isub = subcol_reorder[isub];
            for (int i = 0; i < L; i++) {
              wc.trans[i] = w.trans[i]; wc.reflect[i] = w.reflect[i];
              wc.trans_0[i] = w.trans_0[i]; wc.sc1[i] = w.sc1[i]; wc.sc2[i] = w.sc2[i];
              double cs = c_sub[((size_t)(isub - 1) * L + i) * P + l];
              double frac1 = LAY(w_cloud, i, l) > 0.0 ? 1.0 : 0.0;
              /* frac_cloud(l,i,1) = frac/w_cloud = 1 where cloudy, normalised
               * by the fraction before the clamp of w_cloud to 1
               * (socrates_set_cld.F90:828-853) */
              if (LAY(w_cloud, i, l) > 0.0) {
                double fr = fmax(LAY(c->cloud_frac, i, l), 0.0);
                frac1 = fr / fr;
              }
              if (!(cs * frac1 > 0.0)) continue;
              /* cloudy single-scattering (mcica_sample.F90:614-700) */
              double kgt = k_grey_tot[i] + kc_tot[i] * cs;
              kgt = kgt + ki_tot[i] * cs;
              double ksc = k_ext_scat[i] + kc_scat[i] * cs;
              ksc = ksc + ki_scat[i] * cs;
              double fwd = fw_clr[i] + kc_fwd[i] * cs;
              fwd = fwd + ki_fwd[i] * cs;
              double phf = pf_clr[i] + kc_phf[i] * cs;
              phf = phf + ki_phf[i] * cs;
              phf = (phf - fwd) / fmax(ksc - fwd, eps);
              fwd = fwd / fmax(ksc, eps);
              coeff_t o;
              if (nosc) {
                /* mcica_column.F90:509-522 */
                double tau = sm == IP_NO_SCATTER_ABS ? (kgt + KABS(0, iex, i) - ksc)
                                                     : (kgt + KABS(0, iex, i));
                two_coeff_fast_lw_one(tau * LAY(c->mass, i, l), quad, &o);
              } else {
                double k_total = kgt + KABS(0, iex, i);
                double tau = k_total * LAY(c->mass, i, l);
                double om = ksc / (k_total + tiny);
                om = fmin(om, 1.0 - 3.2e+01 * eps);
                tau = tau * (1.0 - om * fwd);
                om = om * (1.0 - fwd) / (1.0 - om * fwd);
                two_coeff_one(isolir, tau, om, phf, zen_0, &o);
              }
              /* mcica_column.F90:360-520: zero then frac-weighted sum */
              wc.trans[i] = 0.0 + frac1 * o.trans;
              wc.reflect[i] = 0.0 + frac1 * o.reflect;
              wc.sc1[i] = 0.0 + frac1 * o.sc1;
              wc.sc2[i] = 0.0 + frac1 * o.sc2;
              if (isolir == IP_SOLAR) wc.trans_0[i] = 0.0 + frac1 * o.trans_0;
            }
            solve_column(isolir, &wc, adjust, dpl, dpl2, flux_inc_direct, flux_inc_down,
                         alb_diff, alb_dir, dps, nosc);
            if (m == m0) {
              memcpy(acc_total, wc.flux_total, sizeof(double) * (2 * L + 2));
              memcpy(acc_direct, wc.flux_direct, sizeof(double) * (L + 1));
            } else {
              for (int k = 0; k < 2 * L + 2; k++) acc_total[k] = acc_total[k] + wc.flux_total[k];
              for (int k = 0; k <= L; k++) acc_direct[k] = acc_direct[k] + wc.flux_direct[k];
            }
          }
          double fc = frac_cloudy[l];
          for (int k = 0; k <= L; k++) {
            acc_direct[k] = acc_direct[k] * subcol_k_inv;
            acc_direct[k] = fc * acc_direct[k] + ((1.0 - fc) * clr_direct[k]);
          }
          for (int k = 0; k < 2 * L + 2; k++) {
            acc_total[k] = acc_total[k] * subcol_k_inv;
            acc_total[k] = (fc * acc_total[k]) + ((1.0 - fc) * clr_total[k]);
          }
          F = acc_total; FD = acc_direct;
        }
        /* augment_channel (augment_channel_mod.F90:159-233, 496-509) */
        double wincr = 1.0 * esft_weight;
        for (int i = 0; i <= L; i++) {
          if (first) {
            up[i] = wincr * F[2 * i];
            down[i] = wincr * F[2 * i + 1];
            if (isolir == IP_SOLAR) dir[i] = wincr * FD[i];
          } else {
            up[i] = up[i] + wincr * F[2 * i];
            down[i] = down[i] + wincr * F[2 * i + 1];
            if (isolir == IP_SOLAR) dir[i] = dir[i] + wincr * FD[i];
          }
        }
        first = 0;
      }
      /* adjust_ir_radiance (adjust_ir_radiance.F90:186-200) */
      if (isolir == IP_INFRA_RED)
        for (int i = 0; i <= L; i++) {
          up[i] = up[i] + planck[i] * 1.0;
          down[i] = down[i] + planck[i] * 1.0;
        }
    }
    /* set_diag (socrates_set_diag.F90:140-210) */
    for (int i = 0; i <= L; i++) {
      flux_up[(size_t)i * P + l] = up[i];
      flux_down[(size_t)i * P + l] = down[i];
      if (flux_direct) flux_direct[(size_t)i * P + l] = isolir == IP_SOLAR ? dir[i] : 0.0;
    }
    for (int i = 1; i <= L; i++) {
      double fdiv = down[i - 1] - down[i] + up[i] - up[i - 1];
      heating_rate[(size_t)(i - 1) * P + l] = fdiv / LAY(c->layer_heat_capacity, i - 1, l);
    }
  }
  free(kabs); free(buf); free(lu_jp); free(lu_fac);
  if (triple) { free(tw); free(tis); free(ovl); free(tbuf); }
  if (mcica) {
    free(w_cloud); free(cond_mmr); free(re); free(c_sub); free(n_subcol_cld);
    free(ice_in); free(de);
    free(frac_cloudy); free(subcol_reorder); free(mc.xcw); mc.xcw = NULL;
  }
#undef LAY
#undef TAKE
#undef KABS
  return 0;
}

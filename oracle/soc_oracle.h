/* CPU restatement of the SOCRATES two-stream flux path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the parity tests and the
 * benchmark's CPU-baseline leg use; the product (socrates_amd/, include/)
 * never links, loads or calls it.
 *
 * Each function cites the reference routine (file:line, relative to the
 * reference tree) whose arithmetic it restates, in the same operation order.
 *
 * Parity status (what pins each part; see tests/test_oracle.py):
 *   spectral / McICA file readers  - pinned by reading the committed files and
 *                                    by end-to-end agreement with oracle/_ref.
 *   gas optics, k_eqv, two_coeff,
 *   trans_source_coeff, solvers,
 *   accumulation, heating rates    - pinned against oracle/_ref (the reference
 *                                    compiled from its own sources) and the
 *                                    committed golden vectors; the solver also
 *                                    against a dense linear solve and the
 *                                    coefficients against closed-form limits.
 *   cloud_gen                      - pinned bit-exactly against oracle/_ref
 *                                    c_sub dumps (golden vectors).
 *   look-up-table, power-quadratic
 *   and Wenyi k-term scaling,
 *   continua, further gases        - pinned by the reference's own known-good
 *                                    output of its runes example on the GA9
 *                                    files (examples/runes) and by GA9 golden
 *                                    vectors from oracle/_ref (bit-exact).
 *   triple column (maximum-random
 *   and exponential-random) and
 *   mixed column (all overlaps)    - pinned bit-exactly against oracle/_ref
 *                                    (golden vectors); the mixed column is
 *                                    restated as the triple column with no
 *                                    convective cloud, which the reference
 *                                    computes to the same bits.
 */
#ifndef SOC_ORACLE_H
#define SOC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* One runes() call's inputs, arrays in the interface's full-field layout
 * (Fortran (n_profile, n_layer) = C [layer][profile]).  Level fields have
 * n_layer+1 rows; albedos are [band][profile]. */
typedef struct {
  int n_profile, n_layer, n_band;
  int i_source;            /* 1 solar, 2 infra-red */
  int l_cloud;             /* 0 clear, 1 cloudy */
  int i_cloud_repr;        /* ip_cloud_homogen (1) or ip_cloud_conv_strat (3) */
  int i_overlap;           /* ip_exponential_rand (3), ip_max_rand (0) */
  int i_inhom;             /* ip_mcica (2) */
  int i_mcica_sampling;    /* ip_mcica_optimal_sampling (2) */
  int l_rayleigh;
  int i_st_water;
  int n_subcol_gen;
  double co2_mix_ratio, cloud_vertical_decorr, cloud_horizontal_rsd,
      conv_vertical_decorr;
  const double *p_layer, *t_layer, *mass, *density, *layer_heat_capacity,
      *h2o, *o3, *p_level, *t_level, *t_ground, *cos_zen, *solar_irrad,
      *albedo_diff, *albedo_dir, *cloud_frac, *liq_mmr, *liq_dim, *conv_frac,
      *liq_conv_mmr, *liq_conv_dim;
  const int *rand_seed;
  /* well-mixed mass mixing ratios of the further gases (the *_mix_ratio
   * arguments of runes, socrates_runes.F90:69-83), by SOC_WM_* slot */
  double wm_mix_ratio[10];
  /* surface emission per band (runes' l_flux_ground / flux_ground,
   * [band][profile]) and CO2 as a full field ([layer][profile], NULL: the
   * well-mixed co2_mix_ratio) */
  int l_flux_ground;
  const double *flux_ground;
  const double *co2;
  /* fields of the further gases by SOC_WM_* slot ([layer][profile], NULL:
   * the well-mixed wm_mix_ratio) */
  const double *wm_field[10];
  /* ice (socrates_runes.F90:118-121): in-cloud-type stratiform and
   * convective ice mass mixing ratios and dimensions ([layer][profile];
   * NULL: zero) and the ice type of both (i_st_ice = i_cnv_ice); type 0:
   * no ice component */
  int i_st_ice;
  const double *ice_mmr, *ice_dim, *ice_conv_mmr, *ice_conv_dim;
  /* infra-red source linear in optical depth (control%l_ir_source_quad
   * false, ir_source.F90:52-59) in place of runes' quadratic one */
  int l_ir_source_linear;
  /* continuum absorption left out (control%l_continuum false; runes always
   * includes the file's continua) */
  int l_no_continuum;
  /* treatment of scattering (control%i_scatter_method, rad_pcf.F90:539-547):
   * 0 or 1 full, 2 scattering ignored (absorption only), 3 extinction taken
   * as absorption, 4 approximate (as full here), 5 per k-term from the
   * spectral file (hybrid) */
  int i_scatter_method;
  /* modal aerosols (runes' n_aer_mode, aer_mix_ratio, aer_absorption,
   * aer_scattering, aer_asymmetry; socrates_runes.F90:133-135, 609-621):
   * mix ratio [mode][layer][profile], optical properties
   * [band][mode][layer][profile]; n_aer_mode 0: none */
  int n_aer_mode;
  const double *aer_mix_ratio, *aer_absorption, *aer_scattering, *aer_asymmetry;
} soc_case;

enum { SOC_WM_N2O, SOC_WM_CH4, SOC_WM_O2, SOC_WM_SO2, SOC_WM_OCS, SOC_WM_CFC11, SOC_WM_CFC12,
       SOC_WM_CFC113, SOC_WM_HCFC22, SOC_WM_HFC134A };

/* Full runes() restatement. Outputs [level][profile] (n_layer+1 rows) and
 * heating_rate [layer][profile] in K/s. Returns 0 or a negative error. */
int soc_oracle_runes(const char *spectral_file, const char *mcica_file,
                     const soc_case *c, double *flux_up, double *flux_down,
                     double *flux_direct, double *heating_rate);

/* cloud_gen (socrates_cloud_gen.F90:54-478) for a batch of columns with the
 * inputs set_cld_mcica passes (socrates_set_cld_mcica.F90:160-192), followed
 * by the cyclic fill of set_cld_mcica:202-220. c_sub is [sub][layer][prof]
 * (n_subcol x n_layer x n_profile), n_subcol_cld per profile. */
int soc_oracle_cloud_gen(int n_profile, int n_layer, int n_subcol, int n_req,
                         int ipph, int ioverlap, const int *rand_seed,
                         const double *rlc_cf, const double *rlc_cw,
                         const double *sigma_qcw, const double *avg_cf,
                         const double *p_layer, int n1, int n2,
                         const double *xcw, double *c_sub, int *n_subcol_cld);

/* two_coeff + trans_source_coeff for one layer set (two_coeff.F90:148-191,
 * two_coeff_basic.F90:100-161, solar_coefficient_basic.F90:124-219,
 * trans_source_coeff.F90:142-263). isolir 1: solar (PIFM80) with sec_0,
 * 2: infra-red (Elsasser, quadratic source). n values in, outputs per value:
 * trans, reflect, trans_0 (solar), sc[2] (solar up/down or ir 1d/2d). */
void soc_oracle_two_coeff(int n, int isolir, const double *tau,
                          const double *omega, const double *asymmetry,
                          const double *sec_0, double *trans, double *reflect,
                          double *trans_0, double *sc1, double *sc2);

/* solver_homogen_direct (solver_homogen_direct.F90:93-127) for one column:
 * arrays of n_layer; flux_total of 2*n_layer+2 (1-based index k stored at
 * k-1: odd = up, even = down). */
void soc_oracle_solver_homogen_direct(int n_layer, const double *trans,
                                      const double *reflect,
                                      const double *s_down, const double *s_up,
                                      double diffuse_albedo,
                                      double flux_inc_down,
                                      double source_ground, double *flux_total);

/* two_stream (two_stream.F90:160-205) on prescribed optical properties:
 * two_coeff for every layer, ir_source (ir_source.F90:93-120) or
 * solar_source (solar_source.F90:104-162, column_solver.F90:160-185), the
 * surface source, and solver_homogen_direct. Arrays [layer][profile] or
 * [profile]; adjust_solar_ke is read only when l_scale_solar. Outputs:
 * flux_direct [level][profile] (solar), flux_total [k][profile] for the
 * 1-based interleaved index k = 1..2L+2 stored at k-1 (odd up, even down). */
void soc_oracle_two_stream(int n_profile, int n_layer, int isolir, int l_scale_solar,
                           int l_ir_source_quad, const double *tau, const double *omega,
                           const double *asymmetry, const double *sec_0,
                           const double *flux_inc_down, const double *flux_inc_direct,
                           const double *diffuse_albedo, const double *direct_albedo,
                           const double *diff_planck, const double *diff_planck_2,
                           const double *d_planck_flux_surface,
                           const double *adjust_solar_ke, double *flux_direct,
                           double *flux_total);

/* overlap_coupled (overlap_coupled.F90:150-453) for one profile with the
 * triple-column maximum-random overlap: the 18 transfer coefficients of
 * interfaces nct-1 .. L (1-based layers) at ov[(i-(nct-1))*18 + idx-1]. */
void soc_oracle_overlap_triple(int L, int nct, const double *w_cloud, const double *frac_st,
                               const double *frac_cv, double *ov);

#ifdef __cplusplus
}
#endif
#endif

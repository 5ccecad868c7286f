/* socgpu: MI355X (gfx950) two-stream radiative flux path for SOCRATES.
 *
 * C ABI: plain pointers and sizes only. Every entry point returns 0 on
 * success or a negative SOCGPU_E* code; socgpu_last_error() gives the message
 * (per host thread). Nothing here aborts the process: the Fortran shim
 * (socrates_amd/fortran/socrates_runes_gpu.f90) turns a non-zero return into
 * ereport(..., i_err_fatal), which is the reference's STOP-on-fatal-error
 * behaviour (ereport_mod.F90:27-40).
 *
 * Layouts follow the reference runes interface for full fields
 * (socrates_runes.F90:47-168 with l_invert = l_profile_last = .false.):
 * Fortran (n_profile, n_layer) arrays, i.e. C [layer][profile], layer 1 at the
 * top of the atmosphere; level fields (n_profile, 0:n_layer); per-band surface
 * albedos (n_profile, n_band). Reals are float64.
 */
#ifndef SOCGPU_H
#define SOCGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define SOCGPU_OK 0
#define SOCGPU_EINVAL (-1)     /* bad argument */
#define SOCGPU_ENOTFOUND (-2)  /* unknown spectrum / McICA name */
#define SOCGPU_EUNSUPPORTED (-3) /* option or file feature outside the path */
#define SOCGPU_EIO (-4)        /* file could not be read / parsed */
#define SOCGPU_EDEVICE (-5)    /* HIP runtime error */
#define SOCGPU_ENOMEM (-6)     /* device memory / admission limit */

/* precision forms */
#define SOCGPU_FP64 0          /* plain FP64: matches the reference to 1e-10 */
#define SOCGPU_MIXED 1         /* FP32 optical properties and two-stream
                                  coefficients, FP64 elimination and sums:
                                  within 0.05 W m-2 and 0.01 K/day of FP64 */

/* option codes, identical to the reference's rad_pcf.F90 values */
#define SOCGPU_SOURCE_SOLAR 1        /* ip_solar / ip_source_illuminate */
#define SOCGPU_SOURCE_THERMAL 2      /* ip_infra_red / ip_source_thermal */
#define SOCGPU_CLOUD_OFF 0           /* ip_cloud_off */
#define SOCGPU_CLOUD_HOMOGEN 1       /* ip_cloud_homogen */
#define SOCGPU_CLOUD_CONV_STRAT 3    /* ip_cloud_conv_strat: stratiform and
                                        convective cloud as triple-column
                                        regions (ip_cloud_triple with
                                        ip_solver_triple_hogan; maximum-random
                                        overlap only) */
#define SOCGPU_OVERLAP_MAX_RAND 0    /* ip_max_rand */
#define SOCGPU_OVERLAP_RANDOM 1      /* ip_rand */
#define SOCGPU_OVERLAP_EXP 2         /* ip_exponential (McICA) */
#define SOCGPU_OVERLAP_EXP_RAND 3    /* ip_exponential_rand */
#define SOCGPU_INHOM_HOMOGENEOUS 0   /* ip_homogeneous */
#define SOCGPU_INHOM_MCICA 2         /* ip_mcica */
#define SOCGPU_MCICA_SINGLE 1        /* ip_mcica_single_sampling */
#define SOCGPU_MCICA_OPTIMAL 2       /* ip_mcica_optimal_sampling */
/* treatment of scattering, runes' i_scatter_method (rad_pcf.F90:539-547) */
#define SOCGPU_SCATTER_FULL 1        /* ip_scatter_full (default) */
#define SOCGPU_SCATTER_NONE 2        /* ip_no_scatter_abs: absorption only */
#define SOCGPU_SCATTER_EXT 3         /* ip_no_scatter_ext: extinction as absorption */
#define SOCGPU_SCATTER_APPROX 4      /* ip_scatter_approx */
#define SOCGPU_SCATTER_HYBRID 5      /* ip_scatter_hybrid: per k-term, from the file */

/* Inputs of one radiation call: the subset of runes()'s optional arguments
 * on the two-stream flux path (socrates_runes.F90:47-168). A NULL pointer is
 * an absent optional argument. */
typedef struct SocgpuRunesArgs {
  int n_profile;              /* columns */
  int n_layer;                /* layers */
  int i_source;               /* SOCGPU_SOURCE_* */
  int l_rayleigh;             /* Rayleigh scattering (solar) */
  int i_cloud_representation; /* SOCGPU_CLOUD_* */
  int i_overlap;              /* SOCGPU_OVERLAP_* */
  int i_inhom;                /* SOCGPU_INHOM_* */
  int i_mcica_sampling;       /* SOCGPU_MCICA_* */
  int i_st_water;             /* droplet type of stratiform liquid */
  int i_cnv_water;            /* droplet type of convective liquid */
  int n_subcol_gen;           /* McICA sub-columns generated (<= 64) */
  int l_inputs_on_device;     /* pointers below are device pointers */
  double co2_mix_ratio;       /* well-mixed CO2 mass mixing ratio */
  double cloud_vertical_decorr; /* Pa */
  double conv_vertical_decorr;  /* Pa */
  double cloud_horizontal_rsd;
  int l_grey_albedo;          /* non-zero: grey_albedo for every band, both
                                 beams (socrates_set_bound.F90:138-153) */
  double grey_albedo;
  const double *p_layer, *t_layer, *mass, *density, *layer_heat_capacity;
  const double *h2o, *o3;                  /* [layer][profile] mmr */
  const double *p_level, *t_level;         /* [level][profile] */
  const double *t_ground, *cos_zenith_angle, *solar_irrad; /* [profile] */
  const double *albedo_diff, *albedo_dir;  /* [band][profile]; NULL = 0 */
  const double *cloud_frac, *liq_mmr, *liq_dim;           /* [layer][profile] */
  const double *conv_frac, *liq_conv_mmr, *liq_conv_dim;  /* [layer][profile] */
  const int *rand_seed;                    /* [profile] McICA seeds */
  /* Well-mixed mass mixing ratios of further gases, by SOCGPU_GAS_* index:
   * runes' n2o_mix_ratio, ch4_mix_ratio, o2_mix_ratio, so2_mix_ratio,
   * ocs_mix_ratio, cfc11_mix_ratio, cfc12_mix_ratio, cfc113_mix_ratio,
   * hcfc22_mix_ratio, hfc134a_mix_ratio (socrates_runes.F90:69-83). 0 is an
   * absent gas, as in set_atm (socrates_set_atm.F90:535-622). Read only for
   * gases the spectral file lists. */
  double gas_mix_ratio[10];
  /* Surface emission per band (runes' l_flux_ground and flux_ground,
   * socrates_runes.F90:103; socrates_set_bound.F90:251-290): when
   * l_flux_ground is set, a thermal call's ground source in band b is
   * flux_ground[b][profile] / (1 - albedo_diff) in place of the Planck flux
   * of t_ground (diff_planck_source_mod.F90:123-136); flux_ground NULL is
   * zero emission; t_ground is then not needed. */
  int l_flux_ground;
  /* CO2 as a full field (runes' co2, socrates_runes.F90:51): [layer][profile]
   * mass mixing ratio, negative values taken as 0; used unless NULL or
   * l_co2_well_mixed is set (runes' l_co2_well_mixed), when co2_mix_ratio
   * applies (set_gas_mix_ratio, socrates_set_atm.F90:535-622). */
  int l_co2_well_mixed;
  const double *flux_ground;   /* [band][profile] W m-2 */
  const double *co2;           /* [layer][profile] */
  /* Well-mixed H2O and O3 (runes' h2o_mix_ratio, o3_mix_ratio), used where
   * the h2o / o3 field is NULL (set_gas_mix_ratio, socrates_set_atm.F90:
   * 535-622: a field takes precedence unless l_*_well_mixed, which a
   * binding expresses by passing NULL). */
  double h2o_mix_ratio, o3_mix_ratio;
  /* Fields of the further gases, by SOCGPU_GAS_* index ([layer][profile]
   * mass mixing ratio, negative values taken as 0; runes' n2o, ch4, o2, ...
   * and their _1d forms, socrates_runes.F90:51-64); NULL: gas_mix_ratio. */
  const double *gas_field[10];
  /* Ice (socrates_runes.F90:118-121, set_cld and set_cld_dim): the ice types
   * of the stratiform (or homogeneous) and convective cloud (i_st_ice,
   * i_cnv_ice; the spectral file's types of Fu's phase-function fit, Baran's
   * fit or Pade-2; i_cnv_ice <= 0 takes i_st_ice) and their gridbox-mean
   * mass mixing ratios and dimensions ([layer][profile], NULL: zero).
   * i_st_ice <= 0: no ice component (the reference needs an ice type; with
   * one whose fit is zero at zero ice, Fu's or Pade-2, the results are the
   * same; Baran's fit adds a small extinction even at zero ice). */
  int i_st_ice, i_cnv_ice;
  const double *ice_mmr, *ice_dim, *ice_conv_mmr, *ice_conv_dim;
  /* Options of the reference's general driver (l_run_cdl.F90) that runes
   * fixes: an infra-red source linear in optical depth (l_ir_source_quad
   * false, ir_source.F90:52-59; runes takes it quadratic) and no continuum
   * absorption (l_continuum false; runes always includes the file's). */
  int l_ir_source_linear;
  int l_no_continuum;
  /* Treatment of scattering (socrates_runes.F90:589-591, 0 = full). Other
   * than full, minor gases with one k-term are added to the grey extinction
   * (l_grey_single, socrates_set_control.F90:272-277). SCATTER_NONE / EXT
   * (and hybrid terms the file marks so) solve without scattering
   * (two_coeff_fast_lw.F90, solver_no_scat.F90): infra-red only, clear sky
   * and McICA; APPROX is full scattering on those paths. The triple and
   * mixed columns take full scattering only (their approximate and
   * non-scattering region solvers are not built). */
  int i_scatter_method;
  /* Modal aerosols (runes' n_aer_mode, aer_mix_ratio, aer_absorption,
   * aer_scattering, aer_asymmetry, socrates_runes.F90:133-135, 609-621;
   * set_aer, socrates_set_aer.F90:299-309): per mode a mass mixing ratio
   * [mode][layer][profile] and, per band of the spectral file and mode,
   * specific absorption and scattering (m2 kg-1) and asymmetry
   * [band][mode][layer][profile], added to the clear-air grey optics of
   * every layer (grey_opt_prop.F90:716-743, opt_prop_ukca_aerosol.F90).
   * n_aer_mode 0: none (runes' l_aerosol_mode false is passed as 0). Every
   * cloud representation of the path takes them (clear sky, McICA, the
   * triple and mixed columns, whose cloudy regions start from the clear-air
   * totals). */
  int n_aer_mode;
  const double *aer_mix_ratio, *aer_absorption, *aer_scattering, *aer_asymmetry;
} SocgpuRunesArgs;

#define SOCGPU_GAS_N2O 0
#define SOCGPU_GAS_CH4 1
#define SOCGPU_GAS_O2 2
#define SOCGPU_GAS_SO2 3
#define SOCGPU_GAS_OCS 4
#define SOCGPU_GAS_CFC11 5
#define SOCGPU_GAS_CFC12 6
#define SOCGPU_GAS_CFC113 7
#define SOCGPU_GAS_HCFC22 8
#define SOCGPU_GAS_HFC134A 9

/* Outputs (StrDiag, socrates_def_diag.F90:31-76): caller-owned, NULL = not
 * requested. Fluxes [level][profile] in W m-2, heating rate [layer][profile]
 * in K s-1 (set_diag, socrates_set_diag.F90:140-210). The clear-sky fluxes
 * (flux_up_clear, flux_down_clear, flux_direct_clear,
 * socrates_def_diag.F90:46-52) are, as in the reference, the fluxes of the
 * same call with cloud off: asking for any of them runs a cloud-free
 * companion of the call in the same launch (socrates_set_control.F90:146-160
 * turn on the reference's clear-sky pass the same way). With
 * l_outputs_on_device, clear-sky output needs flux_up_clear and
 * flux_down_clear both. */
typedef struct SocgpuDiag {
  double *flux_up, *flux_down, *flux_direct, *heating_rate;
  int l_outputs_on_device;
  double *flux_up_clear, *flux_down_clear, *flux_direct_clear;
} SocgpuDiag;

/* Select the device and create the library's stream. Replaces nothing in the
 * reference (it has no device); called once before the other entry points. */
int socgpu_init(int device);

/* Read a spectral file and keep its tables resident on the device under
 * `name`. Replaces set_spectrum (socrates_set_spectrum.F90:29-149), which
 * reads into the module-saved spectrum_array (:20-23). */
int socgpu_spectrum_register(const char *name, const char *path);

/* socgpu_spectrum_register keeping only the gaseous absorbers whose
 * identifiers (rad_pcf.F90 ip_h2o = 1, ...) are listed: set_spectrum's
 * l_h2o, l_co2, ... flags without l_all_gases (compress_spectrum,
 * socrates_set_spectrum.F90:152-321: each band keeps the listed absorbers of
 * its list, in order, and the first one kept is its major gas). n_type < 0
 * keeps every absorber (l_all_gases). */
int socgpu_spectrum_register_gases(const char *name, const char *path, int n_type,
                                   const int *type_absorb);

/* Shape of a registered spectrum: band count and the identifiers of its
 * gaseous absorbers (type_absorb of the file's block 0, rad_pcf.F90 ip_h2o = 1,
 * ip_co2 = 2, ip_o3 = 3, ...; at most max_absorb written). Lets a binding
 * expand a grey albedo or check which gas amounts the file uses; the
 * reference keeps the same facts in StrSpecData%basic / %gas
 * (def_spectrum.F90:153-237). */
int socgpu_spectrum_query(const char *name, int *n_band, int *n_absorb,
                          int *type_absorb, int max_absorb);

/* Read McICA data and associate it with the SW/LW spectrum names (either may
 * be NULL). Replaces set_mcica (socrates_set_spectrum.F90:551-694) and
 * read_mcica_data (def_mcica.F90:61-190). */
int socgpu_mcica_register(const char *path, const char *sw_name,
                          const char *lw_name);

/* One radiation call on the library stream. Replaces runes
 * (socrates_runes.F90:47-1079): set_control defaults (socrates_set_control.F90
 * :236-398), set_atm/bound/cld/cld_dim/cld_mcica, radiance_calc on the
 * two-stream path, set_diag. precision: SOCGPU_FP64 or SOCGPU_MIXED.
 * Synchronous: returns when the outputs are written. */
int socgpu_runes(const char *spectrum_name, const SocgpuRunesArgs *args,
                 SocgpuDiag *diag, int precision);

/* Allocation-free form of socgpu_runes. No counterpart in the reference: it
 * is for callers that cannot allocate device memory while radiation runs
 * (code under HIP graph capture of their own, tight serving loops, harnesses
 * that forbid hipMalloc). The caller passes device inputs and outputs
 * (l_inputs_on_device and l_outputs_on_device set; flux_up and flux_down
 * required). The call's descriptor and workspace live in a fixed device
 * array of SOCGPU_NOALLOC_MAX_BYTES inside the library, mapped once per GPU
 * at socgpu_init and held until the process ends; a call whose workspace
 * does not fit (above about 500 all-sky McICA columns of 90 layers, or 4000
 * clear-sky ones) fails with SOCGPU_ENOMEM and nothing runs. Nothing here
 * calls hipMalloc or hipFree. Every call on a GPU reuses the same array,
 * which keeps the last call's contents: two calls using it at the same time
 * would race, so the entry point holds the library's lock for the whole call
 * and concurrent callers in one process run one after the other. Results
 * are those of socgpu_runes, bit for bit. Synchronous. */
#define SOCGPU_NOALLOC_MAX_BYTES ((size_t)64 << 20)
int socgpu_runes_noalloc(const char *spectrum_name, const SocgpuRunesArgs *args,
                         SocgpuDiag *diag, int precision);

/* Kernel-level entry: the two-stream solve on prescribed layer optical
 * properties, n_term independent solves per profile (for example the
 * g-points of a band). Replaces two_stream (two_stream.F90:160-205):
 * two_coeff, ir_source (ir_source.F90:93-120) or solar_source
 * (solar_source.F90:104-162), the surface source (column_solver.F90:
 * 160-185) and solver_homogen_direct, with the same argument meanings:
 * isolir SOCGPU_SOURCE_*; l_scale_solar multiplies the direct beam by
 * adjust_solar_ke; l_ir_source_quad selects the quadratic infra-red source;
 * infra-red fluxes are differential (Planck flux removed), as the solver
 * returns them. Layouts: tau, omega, asymmetry [term][layer][profile];
 * diff_planck, diff_planck_2, adjust_solar_ke [layer][profile]; sec_0,
 * flux_inc_down, flux_inc_direct, albedos, d_planck_flux_surface [profile];
 * outputs flux_direct (solar, may be NULL), flux_up, flux_down
 * [term][level][profile]. Pointers are device pointers when l_on_device.
 * Synchronous. */
int socgpu_two_stream(int n_profile, int n_layer, int n_term, int isolir,
                      int l_scale_solar, int l_ir_source_quad,
                      const double *tau, const double *omega,
                      const double *asymmetry, const double *sec_0,
                      const double *flux_inc_down,
                      const double *flux_inc_direct,
                      const double *diffuse_albedo,
                      const double *direct_albedo, const double *diff_planck,
                      const double *diff_planck_2,
                      const double *d_planck_flux_surface,
                      const double *adjust_solar_ke, double *flux_direct,
                      double *flux_up, double *flux_down, int l_on_device,
                      int precision);

/* McICA sub-column generator on the device (cloud_gen,
 * socrates_cloud_gen.F90:54-478, plus the fill of set_cld_mcica.F90:202-220)
 * for tests: inputs [layer][profile]; c_sub out [sub][layer][profile],
 * n_subcol_cld [profile]. Host pointers. */
int socgpu_cloud_gen(int n_profile, int n_layer, int n_subcol, int n_req,
                     int ipph, int ioverlap, const int *rand_seed,
                     const double *rlc_cf, const double *rlc_cw,
                     const double *sigma_qcw, const double *avg_cf,
                     const double *p_layer, int n1, int n2, const double *xcw,
                     double *c_sub, int *n_subcol_cld);

/* One radiation call of a batch: the arguments socgpu_runes takes. */
typedef struct SocgpuRequest {
  const char *spectrum_name;
  const SocgpuRunesArgs *args;
  SocgpuDiag *diag;
} SocgpuRequest;

/* Several radiation calls, each with its own spectrum, source, column count,
 * layer count and options (a ragged batch), packed into one launch
 * sequence: every kernel strides over the concatenated work of all calls.
 * Each call's results are those socgpu_runes gives for it alone (each call
 * keeps its own cloud top and base, as runes does per call,
 * socrates_cloud_gen.F90:224-245). No reference counterpart: the reference
 * runs one call at a time (socrates_runes.F90:47). Synchronous. */
int socgpu_runes_batch(int n_request, const SocgpuRequest *requests,
                       int precision);

/* Persistent form of socgpu_runes_batch for a host model that calls
 * radiation every time step: the batch object keeps capacity-sized device
 * buffers and a captured HIP graph of the kernel sequence, and replays the
 * graph for every batch that fits (capacity grows, and the graph is
 * re-captured, only when a batch does not fit or the solve tile changes). */
typedef struct SocgpuBatch SocgpuBatch;
int socgpu_batch_create(SocgpuBatch **batch, int precision);
int socgpu_batch_run(SocgpuBatch *batch, int n_request,
                     const SocgpuRequest *requests);
/* Device time of the last replay (ms), graph captures and replays so far. */
int socgpu_batch_stats(const SocgpuBatch *batch, double *last_ms,
                       int *captures, int *replays);
int socgpu_batch_destroy(SocgpuBatch *batch);

/* Serving: a request loop for callers that issue radiation calls
 * continuously (many model instances, or one model's column segments from
 * several threads). socgpu_serve_submit queues a call (its argument structs
 * are copied; the arrays they point to must stay valid until the call
 * completes); a server thread admits queued calls while the device
 * footprint of all admitted calls fits hbm_budget_bytes, packs everything
 * admitted into one batch per free slot (a stream with its own device
 * buffers) and launches it, so calls that arrive while other batches are on
 * the GPU are admitted as memory and slots free up and run alongside them.
 * Results are those of the same call made alone. No reference
 * counterpart. */
typedef struct SocgpuServeConfig {
  int precision;              /* SOCGPU_FP64 or SOCGPU_MIXED */
  int n_slots;                /* batches in flight (1-8) */
  int max_batch_requests;     /* calls per batch */
  int pad;
  size_t hbm_budget_bytes;    /* device bytes admitted calls may hold */
} SocgpuServeConfig;

typedef struct SocgpuServeStats {
  long long completed, batches, solves;  /* solves: column x g-point */
  double mean_batch;                     /* calls per batch */
  double p50_ms, p99_ms, max_ms;         /* submit-to-done latency */
  double elapsed_s;                      /* since start or reset */
} SocgpuServeStats;

typedef struct SocgpuServer SocgpuServer;
int socgpu_serve_start(SocgpuServer **server, const SocgpuServeConfig *cfg);
int socgpu_serve_submit(SocgpuServer *server, const SocgpuRequest *request,
                        long long *ticket);
/* Blocks until the call is done (outputs written) or timeout_s passes. */
int socgpu_serve_wait(SocgpuServer *server, long long ticket,
                      double timeout_s);
int socgpu_serve_stats(SocgpuServer *server, SocgpuServeStats *stats);
int socgpu_serve_reset_stats(SocgpuServer *server);
/* Drains the queue, stops the server thread and frees its buffers. */
int socgpu_serve_stop(SocgpuServer *server);
/* Load driver for measuring the serving loop: submits pool[pick[i]] at
 * arrival_s[i] seconds after the call starts (a schedule the caller draws,
 * for example Poisson arrivals), from the calling thread on a steady clock,
 * then waits for every call (timeout_s each); *wall_s gets the time from the
 * start to the last completion. A call picked several times writes the same
 * outputs each time. No reference counterpart. */
int socgpu_serve_replay(SocgpuServer *server, int n_pool, const SocgpuRequest *pool,
                        int n_arrival, const double *arrival_s, const int *pick,
                        double timeout_s, double *wall_s);

/* Launch shape of the column-solve kernel, per precision form (no reference
 * counterpart; socgpu_autotune picks it):
 *   recompute        0: keep each layer's elimination state in lane-private
 *                       memory between the two sweeps (64 B per lane and
 *                       layer of memory traffic); 1: keep a checkpoint every
 *                       4 layers and recompute each segment in registers on
 *                       the way up (twice the arithmetic, a sixth of the
 *                       traffic). Results are bit-identical either way.
 *   waves_per_block  1, 2 or 4 waves (64 items each: k-terms and McICA
 *                    samples of consecutive columns) per block.
 *   waves_per_simd   3 or 4: the register budget the kernel is compiled for
 *                    (168 or 128 registers per lane). */
typedef struct SocgpuTile {
  int recompute;
  int waves_per_block;
  int waves_per_simd;
} SocgpuTile;

int socgpu_set_tile(int precision, const SocgpuTile *tile);
/* Autotuner: times every tile (recompute 0/1 x waves_per_block 1/2/4 x
 * waves_per_simd 3/4) on the given batch with HIP events around the solve
 * kernel (best of `repeats` runs after a warm-up), keeps the fastest for the
 * precision form and returns it. */
int socgpu_autotune(int n_request, const SocgpuRequest *requests,
                    int precision, int repeats, SocgpuTile *best);
int socgpu_get_tile(int precision, SocgpuTile *tile);

/* Message of the last failure on this host thread. */
const char *socgpu_last_error(void);

/* Timing of the last socgpu_runes or socgpu_two_stream call on the device
 * (ms), measured with HIP events on the library stream around the whole
 * call and around the solve kernel alone. */
int socgpu_last_timing(double *total_ms, double *solve_ms);

/* Release device memory and the stream. */
int socgpu_finalize(void);

#ifdef __cplusplus
}
#endif
#endif

// socgpu: the SOCRATES two-stream flux path on MI355X (gfx950).
//
// One radiation call (socgpu_runes) runs five kernels on the library stream:
//   pack      (col,layer) lanes: caller arrays -> per-column layer records
//   cldrange  (col,layer) lanes: cloud top/base over the call (cloud_gen)
//   cloudgen  one wave per column, one lane per McICA sub-column
//   prep      (col,band) lanes: grey extinction, minor-gas equivalent
//             extinction, solar direct-beam adjustment, Planck terms
//   solve     one wave per column, one lane per (k-term, sub-column) item:
//             gas optics, single scattering, two-stream coefficients,
//             sources and the column solve; g-point-weighted band sums by
//             wavefront shuffles, one partial per wave
//   finish    (col,level) lanes: sum of partials, Planck, heating rates
// The spectral tables live on the device from registration on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <chrono>
#include <deque>
#include <condition_variable>
#include <thread>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/socgpu.h"
#include "physics.h"
#include "sg_types.h"
#include "spectrum.h"

namespace sg {

// ---------------------------------------------------------------- errors
static thread_local std::string g_err;
static int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return fail(SOCGPU_EDEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------- kernels
constexpr int kWave = 64;

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, kWave);
  return x;
}

// Cross-lane exchanges of doubles without the LDS: gfx950's
// v_permlane32_swap / v_permlane16_swap (swap the upper 32 lanes of one
// operand with the lower 32 of the other; the odd 16-lane rows of one with
// the even rows of the other) and DPP row_mirror (lane i <-> 15 - i in each
// row), row_half_mirror (i <-> 7 - i in each 8 lanes) and quad_perm. Each
// acts on the two 32-bit halves alike.
__device__ __forceinline__ unsigned lo32(double v) { return (unsigned)__double_as_longlong(v); }
__device__ __forceinline__ unsigned hi32(double v) { return (unsigned)(__double_as_longlong(v) >> 32); }
__device__ __forceinline__ double mk64(unsigned lo, unsigned hi) {
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Pairwise sum across the wave halves: lanes 0..31 get a + (a of lane i+32),
// lanes 32..63 get b + (b of lane i-32).
__device__ __forceinline__ double swap32_sum(double a, double b) {
  const auto l = __builtin_amdgcn_permlane32_swap(lo32(a), lo32(b), false, false);
  const auto h = __builtin_amdgcn_permlane32_swap(hi32(a), hi32(b), false, false);
  return mk64(l[0], h[0]) + mk64(l[1], h[1]);
}

// The same across 16-lane rows: even rows get a + (a of the row above),
// odd rows b + (b of the row below).
__device__ __forceinline__ double swap16_sum(double a, double b) {
  const auto l = __builtin_amdgcn_permlane16_swap(lo32(a), lo32(b), false, false);
  const auto h = __builtin_amdgcn_permlane16_swap(hi32(a), hi32(b), false, false);
  return mk64(l[0], h[0]) + mk64(l[1], h[1]);
}

template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)lo32(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)hi32(v), CTRL, 0xf, 0xf, false);
  return mk64((unsigned)lo, (unsigned)hi);
}

// The value of lane 63 in every lane, and the value of the lane below
// (wave_shr:1; lane 0 gets 0).
__device__ __forceinline__ double lane63(double v) {
  return mk64((unsigned)__builtin_amdgcn_readlane((int)lo32(v), 63),
              (unsigned)__builtin_amdgcn_readlane((int)hi32(v), 63));
}

// Sum over each group of 8 lanes, in every lane of the group.
__device__ __forceinline__ double sum8(double x) {
  x = x + dpp_move<0x141>(x);  // row_half_mirror
  x = x + dpp_move<0x4E>(x);   // quad_perm [2, 3, 0, 1]
  x = x + dpp_move<0xB1>(x);   // quad_perm [1, 0, 3, 2]
  return x;
}

__device__ __forceinline__ double lane_below(double v) { return dpp_move<0x138>(v); }

// Butterfly reduce-scatter of 16 values per lane over the 64 lanes of a
// wave: afterwards lanes 4j..4j+3 (j = 0..15) hold the wave total of value j
// in v[0]. Each step halves the values a lane carries, summing them with
// the matching values of its partner lane: 15 exchanges instead of the 96
// of sixteen separate all-reductions.
__device__ __forceinline__ void reduce16(double (&v)[16], int lane) {
#pragma unroll
  for (int k = 0; k < 8; k++) v[k] = swap32_sum(v[k], v[k + 8]);
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = swap16_sum(v[k], v[k + 4]);
  {
    const bool hi = lane & 8;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const double keep = hi ? v[k + 2] : v[k], give = hi ? v[k] : v[k + 2];
      v[k] = keep + dpp_move<0x140>(give);  // row_mirror: a partner in the other 8 lanes
    }
  }
  {
    const bool hi = lane & 4;
    const double keep = hi ? v[1] : v[0], give = hi ? v[0] : v[1];
    v[0] = keep + dpp_move<0x141>(give);    // row_half_mirror: the other quad
  }
  v[0] = v[0] + dpp_move<0x4E>(v[0]);
  v[0] = v[0] + dpp_move<0xB1>(v[0]);
}

// Same for 8 values per lane: lanes 8j..8j+7 (j = 0..7) hold the total of
// value j.
__device__ __forceinline__ void reduce8(double (&v)[8], int lane) {
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = swap32_sum(v[k], v[k + 4]);
#pragma unroll
  for (int k = 0; k < 2; k++) v[k] = swap16_sum(v[k], v[k + 2]);
  {
    const bool hi = lane & 8;
    const double keep = hi ? v[1] : v[0], give = hi ? v[0] : v[1];
    v[0] = keep + dpp_move<0x140>(give);
  }
  v[0] = sum8(v[0]);
}

struct CallDims {
  int P, L, NB, isolir;
  int mcica, mode;   // McICA on, sampling table (0 single, 1 optimal)
  int drop;          // droplet type index (0-based; the stratiform type)
  int nwpc;          // waves per column of the triple-column solve
  int n_items;       // lanes per column of the packed solve (its g-points, at least kMinItems)
  int n_waves;       // waves of the packed solve
  int n_gen, n_fill, ipph, ioverlap;
  int triple, drop2; // triple-column representation; convective droplet type
  int raw_range;     // cloud range on the fractions as given (kernel-level cloud_gen)
  int cg_pre;        // overlap parameters precomputed by the pack kernel
  int f32;           // mixed form: McICA c_sub and droplet optics kept in FP32
  int gx;            // per-layer gas record (look-up-table terms or continua)
  int gen;           // general gas optics (scalings other than the power law,
                     // continua, further gases)
  int conv;          // convective cloud fields are read (ip_cloud_conv_strat
                     // only: the reference ignores them for ip_cloud_homogen,
                     // socrates_set_cld.F90:477-495)
  int ice;           // an ice component (i_st_ice given); its types, 0-based:
  int ice_t, ice_t2; // stratiform (or homogeneous) and convective
  int co_stride;     // doubles per layer of the McICA cloud optics (4, or 8
                     // with the ice beside the droplets in the FP64 form)
  int ir_lin;        // infra-red source linear in optical depth
  int no_cont;       // continuum absorption left out
  int scat;          // treatment of scattering (SOCGPU_SCATTER_*)
  int grey1;         // minor gases with one k-term added to the grey extinction
  int n_aer;         // modal aerosols: modes (0: none; the general solve build)
  int ggen;          // the spectrum needs the general gas optics (band kernel build)
};

// Modal aerosol inputs of a call (device pointers: mixing ratio
// [mode][layer][profile], absorption, scattering and asymmetry
// [band][mode][layer][profile]) and the per-(column, band, layer) record the
// pack kernel writes from them: clear-air scattering (Rayleigh plus the
// modes'), the products of scattering with the asymmetry and with its
// square, unnormalised, and the modes' absorption
// (aerrec[col][band][layer][4]).
struct AerIn {
  const double *mix, *abs, *sca, *asy;
  double *rec;
};

// A major k-term's treatment of scattering for a call with method scat:
// SOCGPU_SCATTER_NONE or _EXT when the term is solved without scattering,
// else 0 (solve_band_k_eqv_scl.F90:1053-1058).
__host__ __device__ inline int term_nosc(int scat, int iscat) {
  const int m = scat == SOCGPU_SCATTER_HYBRID ? iscat : scat;
  return (m == SOCGPU_SCATTER_NONE || m == SOCGPU_SCATTER_EXT) ? m : 0;
}

// Caller arrays -> column records. set_atm (socrates_set_atm.F90:246-300),
// set_cld (socrates_set_cld.F90:817-852: in-cloud mixing ratio, w_cloud
// threshold 0.001, type-fraction normalisation), set_cld_dim
// (socrates_set_cld_dim.F90:214-220: radius clamp), cond_rsd of
// set_cld_mcica (socrates_set_cld_mcica.F90:168-180).
// Record of layer i of column col (caller index idx = i P + col) into r;
// dp_cf, dp_cw: the call's decorrelation pressure scales (McICA).
// Returns true where the layer's total cloud fraction exceeds 1 by more than
// the 0.001 threshold, which the reference treats as a fatal error
// (socrates_set_cld.F90:848-853); fractions up to that are clamped to 1.
__device__ __forceinline__ bool pack_layer(const CallDims &d, const SgSpec *__restrict__ spec,
                                           const SocgpuRunesArgs &a, int idx, double dp_cf, double dp_cw,
                                           double *r) {
  bool bad = false;
  {
    double p = a.p_layer[idx], t = a.t_layer[idx];
    r[SG_Q_LNP] = log(p);
    r[SG_Q_LNT] = log(t);
    r[SG_Q_MASS] = a.mass[idx];
    r[SG_Q_MMR + SG_GAS_H2O] = a.h2o ? fmax(a.h2o[idx], 0.0) : a.h2o_mix_ratio;
    // CO2: the field, clamped at 0, unless absent or declared well mixed
    // (set_gas_mix_ratio, socrates_set_atm.F90:560-612)
    r[SG_Q_MMR + SG_GAS_CO2] = (a.co2 && !a.l_co2_well_mixed) ? fmax(a.co2[idx], 0.0) : a.co2_mix_ratio;
    r[SG_Q_MMR + SG_GAS_O3] = a.o3 ? fmax(a.o3[idx], 0.0) : a.o3_mix_ratio;
    r[SG_Q_MMR + SG_GAS_ZERO] = 0.0;
    r[SG_Q_P] = p;
    r[SG_Q_T] = t;
    double w = 0.0, lwc = 0.0, re = 0.0, f1 = 0.0, rsd = 0.0;
    double lwc2 = 0.0, re2 = 0.0, f2 = 0.0;
    if (d.triple) {
      // stratiform and convective types (socrates_set_cld.F90:284-330,
      // 567-581): in-cloud mixing ratios over the raw type fractions,
      // w_cloud = sum of the types, fractions normalised by it above the
      // 0.001 threshold (:817-852)
      const double eps = kEps;
      double fs = fmax(a.cloud_frac[idx], 0.0);
      double fc = (d.conv && a.conv_frac) ? fmax(a.conv_frac[idx], 0.0) : 0.0;
      double ms = a.liq_mmr ? fmax(a.liq_mmr[idx], 0.0) : 0.0;
      double mc = (d.conv && a.liq_conv_mmr) ? fmax(a.liq_conv_mmr[idx], 0.0) : 0.0;
      lwc = fs > eps ? ms / fs : 0.0;
      lwc2 = fc > eps ? mc / fc : 0.0;
      w = fs + fc;
      if (w > 0.001) { f1 = fs / w; f2 = fc / w; }
      else { w = 0.0; f1 = 0.0; f2 = 0.0; }
      bad = w > 1.0 + 0.001;
      if (w > 1.0) w = 1.0;
      const SgDrop &ds = spec->drop[d.drop], &dc = spec->drop[d.drop2];
      re = fmin(fmax(a.liq_dim ? fmax(a.liq_dim[idx], 0.0) : 0.0, ds.dim_min), ds.dim_max);
      re2 = fmin(fmax((d.conv && a.liq_conv_dim) ? fmax(a.liq_conv_dim[idx], 0.0) : 0.0, dc.dim_min), dc.dim_max);
      if (d.ice) {
        // each type's ice over its own fraction, dimensions clamped to the
        // ice types' ranges (socrates_set_cld.F90:341-415, 817-826;
        // socrates_set_cld_dim.F90)
        const double is = a.ice_mmr ? fmax(a.ice_mmr[idx], 0.0) : 0.0;
        const double ic = (d.conv && a.ice_conv_mmr) ? fmax(a.ice_conv_mmr[idx], 0.0) : 0.0;
        const SgDrop &es = spec->ice[d.ice_t], &ec = spec->ice[d.ice_t2];
        r[SG_Q_IWC] = fs > eps ? is / fs : 0.0;
        r[SG_Q_IWC2] = fc > eps ? ic / fc : 0.0;
        r[SG_Q_DE] = fmin(fmax(a.ice_dim ? fmax(a.ice_dim[idx], 0.0) : 0.0, es.dim_min), es.dim_max);
        r[SG_Q_DE2] = fmin(fmax((d.conv && a.ice_conv_dim) ? fmax(a.ice_conv_dim[idx], 0.0) : 0.0, ec.dim_min),
                           ec.dim_max);
      }
    } else if (a.cloud_frac) {
      const double eps = kEps;
      double frac = fmax(a.cloud_frac[idx], 0.0);
      double mmr = a.liq_mmr ? fmax(a.liq_mmr[idx], 0.0) : 0.0;
      lwc = frac > eps ? mmr / frac : 0.0;
      w = frac;
      f1 = frac;
      if (w > 0.001) f1 = f1 / w;
      else { w = 0.0; f1 = 0.0; }
      bad = w > 1.0 + 0.001;
      if (w > 1.0) w = 1.0;
      const SgDrop &dr = spec->drop[d.drop];
      double r0 = a.liq_dim ? fmax(a.liq_dim[idx], 0.0) : 0.0;
      re = fmin(fmax(r0, dr.dim_min), dr.dim_max);
      // ice of the same type (socrates_set_cld.F90:341-381, 817-826)
      double iwc = 0.0;
      if (d.ice) {
        const double mi = a.ice_mmr ? fmax(a.ice_mmr[idx], 0.0) : 0.0;
        iwc = frac > eps ? mi / frac : 0.0;
        const SgDrop &e = spec->ice[d.ice_t];
        r[SG_Q_IWC] = iwc;
        r[SG_Q_DE] = fmin(fmax(a.ice_dim ? fmax(a.ice_dim[idx], 0.0) : 0.0, e.dim_min), e.dim_max);
      }
      // liquid and ice components, both with the call's horizontal relative
      // standard deviation (socrates_set_cld_mcica.F90:168-178)
      double wl = fmax(eps, lwc * f1), wi = fmax(eps, iwc * f1);
      double sr = 0.0 + a.cloud_horizontal_rsd * wl;
      sr = sr + a.cloud_horizontal_rsd * wi;
      double sw = 0.0 + wl;
      sw = sw + wi;
      rsd = sr / sw;
    }
    r[SG_Q_WCLD] = w;
    r[SG_Q_LWC] = lwc;
    r[SG_Q_RE] = re;
    r[SG_Q_FRAC1] = f1;
    r[SG_Q_RSD] = rsd;
    r[SG_Q_LWC2] = lwc2;
    r[SG_Q_RE2] = re2;
    r[SG_Q_FRAC2] = f2;
    if (d.mcica) {
      // exponential-random overlap parameters between this layer and the
      // one above (socrates_cloud_gen.F90:359-364), as the generator forms
      // them (same operations, no contraction)
#pragma clang fp contract(off)
      double alpha = 0.0, rcorr = 0.0;
      if (idx >= d.P) {
        const double dz = a.p_layer[idx - d.P] - p;
        alpha = dp_cf > 0.0 ? exp(dz / dp_cf) : 0.0;
        rcorr = dp_cw > 0.0 ? exp(dz / dp_cw) : 0.0;
      }
      r[SG_Q_ALPHA] = alpha;
      r[SG_Q_RCORR] = rcorr;
    }
  }
  return bad;
}

// Per-column scalars and per-(column, band) albedos.
__device__ __forceinline__ void pack_column(const CallDims &d, const SocgpuRunesArgs &a, int col,
                                            double *__restrict__ colscal, double *__restrict__ bandscal) {
  const int P = d.P;
  {
    double sec0 = 1.0, sol = 0.0;
    if (d.isolir == SOCGPU_SOURCE_SOLAR) {
      // set_bound (socrates_set_bound.F90:611-618)
      if (a.cos_zenith_angle[col] > 0.0) { sol = a.solar_irrad[col]; sec0 = 1.0 / a.cos_zenith_angle[col]; }
    }
    double *s = colscal + (size_t)col * SG_NS;
    s[SG_S_SEC0] = sec0;
    s[SG_S_SOL] = sol;
    s[SG_S_TGROUND] = a.t_ground ? a.t_ground[col] : 0.0;
    s[SG_S_FC] = 0.0;
    s[SG_S_NCLD] = 0.0;
    s[SG_S_CT] = (double)d.L;
    const bool lfg = a.l_flux_ground && d.isolir == SOCGPU_SOURCE_THERMAL;
    s[SG_S_LFG] = lfg ? 1.0 : 0.0;
    for (int b = 0; b < d.NB; b++) {
      double *bs = bandscal + ((size_t)col * d.NB + b) * SG_NBQ;
      // set_bound (socrates_set_bound.F90:138-233): grey albedo, else per band, else 0
      if (a.l_grey_albedo) {
        bs[2] = a.grey_albedo;
        bs[3] = a.grey_albedo;
      } else {
        bs[2] = a.albedo_diff ? a.albedo_diff[(size_t)b * P + col] : 0.0;
        bs[3] = a.albedo_dir ? a.albedo_dir[(size_t)b * P + col] : 0.0;
      }
      // given surface emission: the ground's Planck flux is the emission
      // over the diffuse co-albedo (diff_planck_source_mod.F90:123-129;
      // socrates_set_bound.F90:251-290, absent = 0)
      if (lfg) bs[1] = (a.flux_ground ? a.flux_ground[(size_t)b * P + col] : 0.0) / (1.0 - bs[2]);
    }
  }
}

// McICA sub-columns, one wave per column, one lane per generated
// sub-column: cloud_gen (socrates_cloud_gen.F90:54-478) followed by the
// cloudy-first compaction and the cyclic fill of set_cld_mcica
// (socrates_set_cld_mcica.F90:202-225). Bit-exact: the LCG, the overlap
// draws and the xcw interpolation use the reference's FP64 operations with
// no contraction.
__device__ double zcw_lookup(double y, double sigma, int n1, int n2, const double *xcw) {
#pragma clang fp contract(off)
  double r1 = y * (double)(n1 - 1) + 1.0;
  int i1 = (int)r1;
  i1 = i1 > n1 - 1 ? n1 - 1 : i1;
  i1 = i1 < 1 ? 1 : i1;
  r1 = fmax(0.0, fmin(r1 - (double)i1, 1.0));
  double r2 = 40.0 * sigma - 3.0;
  int i2 = (int)r2;
  i2 = i2 > n2 - 1 ? n2 - 1 : i2;
  i2 = i2 < 1 ? 1 : i2;
  r2 = fmax(0.0, fmin(r2 - (double)i2, 1.0));
  const double *c0 = xcw + (size_t)(i2 - 1) * n1, *c1 = xcw + (size_t)i2 * n1;
  double t1 = (1.0 - r1) * (1.0 - r2) * c0[i1 - 1];
  double t2 = (1.0 - r1) * r2 * c1[i1 - 1];
  double t3 = r1 * (1.0 - r2) * c0[i1];
  double t4 = r1 * r2 * c1[i1];
  return ((t1 + t2) + t3) + t4;
}

__constant__ LcgSkip kLcgSkip = make_lcg_skip();

// One layer k (1-based) of one sub-column's draws (cloud_gen:253-405):
// updates the x stream (cloud presence) and, with Y, the y stream
// (condensate rank). The x stream never depends on the y stream.
template <bool Y>
__device__ __forceinline__ void cg_layer(int ioverlap, bool pre, int k, const double *__restrict__ rec,
                                         double dp_cf, double dp_cw, int &ix, int &iy, double &x, double &y) {
#pragma clang fp contract(off)
  const double rm = 1.0 / (double)kLcgM;
  if (ioverlap == 0) {
    // maximum-random (cloud_gen:253-298)
    ix = lcg_step(ix); x = (double)ix * rm;
    if (Y) { iy = lcg_step(iy); y = (double)iy * rm; }
    if (k == 1) {
      ix = lcg_step(ix); x = (double)ix * rm;
      if (Y) { iy = lcg_step(iy); y = (double)iy * rm; }
    } else {
      double cfa = rec[(k - 2) * SG_NQ + SG_Q_WCLD];
      if (x < (1.0 - cfa)) {
        ix = lcg_step(ix); double x1 = (double)ix * rm;
        x = x1 * (1.0 - cfa);
        if (Y) { iy = lcg_step(iy); y = (double)iy * rm; }
      }
    }
  } else {
    // random / exponential / exponential-random (cloud_gen:339-405)
    double alpha = 0.0, rcorr = 0.0;
    if (ioverlap != 1 && k - 1 >= 1) {
      double cfa = rec[(k - 2) * SG_NQ + SG_Q_WCLD];
      // exponential overlap (2) correlates with the layer above even where
      // that layer is clear (cloud_gen:355-364)
      if (cfa > 0.0 || ioverlap == 2) {
        if (pre) {
          alpha = rec[(k - 1) * SG_NQ + SG_Q_ALPHA];
          if (Y) rcorr = rec[(k - 1) * SG_NQ + SG_Q_RCORR];
        } else {
          double dz = rec[(k - 2) * SG_NQ + SG_Q_P] - rec[(k - 1) * SG_NQ + SG_Q_P];
          alpha = dp_cf > 0.0 ? exp(dz / dp_cf) : 0.0;
          if (Y) rcorr = dp_cw > 0.0 ? exp(dz / dp_cw) : 0.0;
        }
      }
    }
    ix = lcg_step(ix); double x1 = (double)ix * rm;
    if (x1 >= alpha) { ix = lcg_step(ix); x = (double)ix * rm; }
    if (Y) {
      iy = lcg_step(iy); double y1 = (double)iy * rm;
      if (y1 >= rcorr) { iy = lcg_step(iy); y = (double)iy * rm; }
    }
  }
}

// Two passes over the layers: the x streams alone give each sub-column's
// cloud presence and so the cloudy-first order with the cyclic fill; then
// each lane replays, in full, the sub-column that lands in its slot and
// writes its condensate scalings in place (no second pass over c_sub).
__device__ __forceinline__ void cloudgen_body(const CallDims &d, const double *__restrict__ colrec,
                                              const int *__restrict__ seeds, double dp_cf,
                                              double dp_cw, const int *__restrict__ krange,
                                              const double *__restrict__ xcw, int n1, int n2,
                                              double *__restrict__ c_sub,
                                              double *__restrict__ colscal, int col, int lane) {
#pragma clang fp contract(off)
  const int L = d.L, n_gen = d.n_gen;
  const int seed_max = (2147483647 - kLcgC) / kLcgA;
  double *cs = c_sub + (size_t)col * L * SG_MAXSUBCOL;
  const double *rec = colrec + (size_t)col * L * SG_NQ;
  const int k_top = krange[0], k_base = krange[1];
  const bool valid = lane < n_gen;
  // seeds of this sub-column (cloud_gen:180-205): the stream from the
  // column's seed, warmed up by n_gen steps; sub-column j takes the values
  // after n_gen + j + 1 and 2 n_gen + j + 1 steps (taken in one skip each)
  const long long rs0 = abs(seeds[col]) % seed_max;
  int sx = 0, sy = 0;
  if (valid) {
    const int kx = n_gen + 1 + lane, ky = 2 * n_gen + 1 + lane;
    sx = (int)(((long long)kLcgSkip.a[kx] * rs0 + kLcgSkip.c[kx]) % kLcgM);
    sy = (int)(((long long)kLcgSkip.a[ky] * rs0 + kLcgSkip.c[ky]) % kLcgM);
  }
  const bool any = k_top <= L;
  const int k0 = any ? k_top : 1, k1 = any ? k_base : 0;
  // The draws x, y start a sub-column where the previous one left them
  // (cloud_gen keeps them across its sub-column loop). Only exponential
  // overlap can keep a value at a sub-column's first cloudy layer (the
  // others correlate only with a cloudy layer above); for it, the start of
  // every sub-column comes from running them in order, one lane at a time.
  double x0 = 0.0, y0 = 0.0;
  if (d.ioverlap == 2) {
    double xc = 0.0, yc = 0.0;
    for (int j = 0; j < n_gen; j++) {
      double xf = 0.0, yf = 0.0;
      if (lane == j) {
        x0 = xc; y0 = yc;
        int ix = sx, iy = sy;
        double x = xc, y = yc;
        for (int k = k0; k <= k1; k++)
          if (rec[(k - 1) * SG_NQ + SG_Q_WCLD] > 0.0)
            cg_layer<true>(d.ioverlap, d.cg_pre, k, rec, dp_cf, dp_cw, ix, iy, x, y);
        xf = x; yf = y;
      }
      xc = __shfl(xf, j, kWave);
      yc = __shfl(yf, j, kWave);
    }
  }
  // pass 1: cloud presence of this lane's sub-column
  bool cloudy = false;
  {
    int ix = sx, iy = 0;
    double x = x0, y = 0.0;
    for (int k = k0; k <= k1; k++) {
      const double cf = rec[(k - 1) * SG_NQ + SG_Q_WCLD];
      if (valid && cf > 0.0) {
        cg_layer<false>(d.ioverlap, d.cg_pre, k, rec, dp_cf, dp_cw, ix, iy, x, y);
        if (x < cf) cloudy = true;
      }
    }
  }
  // cloudy sub-columns first, then the cyclic fill up to n_fill
  unsigned long long mask = __ballot(cloudy);
  int ncld = __popcll(mask);
  int src = -1;
  if (ncld > 0 && (lane < ncld || lane < d.n_fill)) {
    int want = lane % ncld;
    unsigned long long m = mask;
    for (int j = 0; j < want; j++) m &= m - 1;
    src = __ffsll((long long)m) - 1;
  }
  // pass 2: the sub-column of slot `lane` (sub-column src), both streams
  int ix = __shfl(sx, src >= 0 ? src : 0, kWave), iy = __shfl(sy, src >= 0 ? src : 0, kWave);
  double x = __shfl(x0, src >= 0 ? src : 0, kWave), y = __shfl(y0, src >= 0 ? src : 0, kWave);
  for (int k = 1; k <= L; k++) {
    double val = 0.0;
    const double cf = rec[(k - 1) * SG_NQ + SG_Q_WCLD];
    if (src >= 0 && k >= k0 && k <= k1 && cf > 0.0) {
      cg_layer<true>(d.ioverlap, d.cg_pre, k, rec, dp_cf, dp_cw, ix, iy, x, y);
      if (x < cf)
        val = d.ipph == 0 ? 1.0 : zcw_lookup(y, rec[(k - 1) * SG_NQ + SG_Q_RSD], n1, n2, xcw) * 1.0;
    }
    if (d.f32) ((float *)c_sub + (size_t)col * L * SG_MAXSUBCOL)[(size_t)(k - 1) * SG_MAXSUBCOL + lane] = (float)val;
    else cs[(size_t)(k - 1) * SG_MAXSUBCOL + lane] = val;
  }
  // first layer of the column with cloud: above it every sub-column is clear
  int ct = L;
  for (int base = 0; base < L; base += kWave) {
    const int k = base + lane;
    const unsigned long long m = __ballot(k < L && rec[(size_t)k * SG_NQ + SG_Q_WCLD] > 0.0);
    if (m) { ct = base + __ffsll((long long)m) - 1; break; }
  }
  if (lane == 0) {
    colscal[(size_t)col * SG_NS + SG_S_NCLD] = (double)ncld;
    colscal[(size_t)col * SG_NS + SG_S_FC] = (double)ncld / (double)n_gen;
    colscal[(size_t)col * SG_NS + SG_S_CT] = (double)ct;
  }
}

// Wave-wide inclusive scans over the 64 lanes (lane order = scan order).
// A double moved across lanes by DPP (both 32-bit halves with the same
// control): row_shr:n (CTRL 0x110 + n) within rows of 16, row_bcast:15
// (0x142) and row_bcast:31 (0x143) across rows. Lanes with no source lane,
// or in rows ROWMASK leaves out, take `id`.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double v, double id) {
  const long long vb = __double_as_longlong(v), ib = __double_as_longlong(id);
  const int lo = __builtin_amdgcn_update_dpp((int)ib, (int)vb, CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(ib >> 32), (int)(vb >> 32), CTRL, ROWMASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Inclusive product scan over the 64 lanes (lane j ends with x_0 ... x_j):
// shifts by 1, 2, 4 and 8 within each row of 16, then the last lane of row 0
// (and of row 2) into rows 1 (and 3), then lane 31 into rows 2 and 3. DPP
// moves are register operations, unlike the LDS-routed shuffles.
__device__ __forceinline__ double scan_mul(double x, int lane) {
  (void)lane;
  x = x * dpp_f64<0x111, 0xf>(x, 1.0);
  x = x * dpp_f64<0x112, 0xf>(x, 1.0);
  x = x * dpp_f64<0x114, 0xf>(x, 1.0);
  x = x * dpp_f64<0x118, 0xf>(x, 1.0);
  x = x * dpp_f64<0x142, 0xa>(x, 1.0);
  x = x * dpp_f64<0x143, 0xc>(x, 1.0);
  return x;
}

// One step of the affine scan: the maps of lower lanes (A0, C0; identity
// (1, 0) where there is none) applied before the lane's own.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void affine_step(double &A, double &C) {
  const double A0 = dpp_f64<CTRL, ROWMASK>(A, 1.0), C0 = dpp_f64<CTRL, ROWMASK>(C, 0.0);
  C = A * C0 + C;
  A = A * A0;
}

// Composition of affine maps F -> A F + C in lane order: lane j ends with
// the map of lanes 0..j applied in turn (the steps of scan_mul).
__device__ __forceinline__ void scan_affine(double &A, double &C, int lane) {
  (void)lane;
  affine_step<0x111, 0xf>(A, C);
  affine_step<0x112, 0xf>(A, C);
  affine_step<0x114, 0xf>(A, C);
  affine_step<0x118, 0xf>(A, C);
  affine_step<0x142, 0xa>(A, C);
  affine_step<0x143, 0xc>(A, C);
}

// One wave per (column, band), lanes over layers (64 at a time, at most
// NCH chunks, with carries between them): the equivalent extinction of the
// band's minor gases (solve_band_k_eqv_scl.F90:572-1047), then the band's
// per-layer records (prep_layer). The band's gases and their k-terms in
// turn:
//  - solar: each term's direct-beam transmission down the column (a wave
//    product scan, :654-739); k_eqv = sum_k k T_k(surface) / sum_k
//    T_k(surface) per gas (:826-846), and adjust_solar_ke from the per-layer
//    ratios of the summed transmissions (:848-891);
//  - infra-red: each term's non-scattering fluxes (monochromatic_gas_flux.F90:
//    115-196, diffusivity 1.66) as two affine recurrences solved by wave
//    scans; k_eqv = sum k |F| / sum |F| over the layer's two levels
//    (:995-1019);
//  - where the weighting flux vanishes, the smallest k of the gas's terms.
// Per lane, the layer records and (infra-red) the band's Planck fluxes at
// the layer's levels are read or evaluated once for all terms and for the
// records. The scans associate the products differently from the
// reference's left-to-right sweeps (relative differences of order 1e-16).
// NCH: chunks of 64 layers the lane arrays hold (kernels for NCH = 2 and 3
// each take the calls with that many chunks, so a call of at most 128
// layers runs with the smaller register footprint).
constexpr int kMinorWaves = 4;
constexpr int kKeqMaxChunks = (SG_MAXL + kWave - 1) / kWave;

template <int kKeqChunks, bool GEN>
__device__ __forceinline__ void band_body(const CallDims &d, const SgSpec *__restrict__ spec,
                                          int l_rayleigh, const double *__restrict__ colrec,
                                          const double *__restrict__ grec, const double *__restrict__ wm,
                                          const double *const *gfield, const double *__restrict__ tlev,
                                          const double *__restrict__ colscal,
                                          double *__restrict__ bandrec,
                                          double *__restrict__ bandscal,
                                          double *__restrict__ cldopt, const int *__restrict__ krange,
                                          const AerIn &ae, int wave, int lane) {
  const int L = d.L, NB = d.NB;
  if (wave >= d.P * NB) return;
  const int col = wave / NB, b = wave % NB;
  const SgBand &B = spec->band[b];
  const double *rec = colrec + (size_t)col * L * SG_NQ;
  const int nch = (L + kWave - 1) / kWave;
  const bool solar = d.isolir == SOCGPU_SOURCE_SOLAR;
  double lnp[kKeqChunks], lnt[kKeqChunks], mass[kKeqChunks], b0[kKeqChunks], b1[kKeqChunks];
  double keq[kKeqChunks], adj[kKeqChunks], kgrey[kKeqChunks];
  const double *tl = tlev + (size_t)col * (L + 1);
#pragma unroll
  for (int c = 0; c < kKeqChunks; c++) {
    const int i = c * kWave + lane;
    lnp[c] = 0.0; lnt[c] = 0.0; mass[c] = 0.0; b0[c] = 0.0; b1[c] = 0.0;
    keq[c] = 0.0; adj[c] = 1.0; kgrey[c] = 0.0;
    if (c < nch && i < L) {
      const double *r = rec + (size_t)i * SG_NQ;
      lnp[c] = r[SG_Q_LNP]; lnt[c] = r[SG_Q_LNT]; mass[c] = r[SG_Q_MASS];
      if (!solar) {
        b0[c] = planck_band(B, spec->n_deg, spec->t_ref_planck, tl[i]);
        b1[c] = planck_band(B, spec->n_deg, spec->t_ref_planck, tl[i + 1]);
      }
    }
  }
  const double sec0 = colscal[(size_t)col * SG_NS + SG_S_SEC0];
  const double alb = bandscal[((size_t)col * d.NB + b) * SG_NBQ + 2];
  double b_top = 0.0, b_bot = 0.0, bg = 0.0;
  if (!solar) {
    b_top = planck_band(B, spec->n_deg, spec->t_ref_planck, tl[0]);
    b_bot = planck_band(B, spec->n_deg, spec->t_ref_planck, tl[L]);
    bg = colscal[(size_t)col * SG_NS + SG_S_LFG] != 0.0
             ? bandscal[((size_t)col * d.NB + b) * SG_NBQ + 1]
             : planck_band(B, spec->n_deg, spec->t_ref_planck, colscal[(size_t)col * SG_NS + SG_S_TGROUND]);
  }
  const double sqe = 1.4901161193847656e-08;
  int t0 = B.minor0;
  for (int j = 0; j < B.n_minor_gas; j++) {
    const int nk = B.minor_nk[j];
    const int gas = spec->minor[t0].gas;
    double mmr[kKeqChunks], num[kKeqChunks], den[kKeqChunks], kmin[kKeqChunks];
#pragma unroll
    for (int c = 0; c < kKeqChunks; c++) {
      const int i = c * kWave + lane;
      // further gases (general gas optics): the call's field, clamped at
      // 0, else its well-mixed ratio (socrates_set_atm.F90:535-622)
      const double *fld = (GEN && gas >= SG_GAS_WM0) ? G(gfield[gas - SG_GAS_WM0]) : nullptr;
      mmr[c] = (c < nch && i < L) ? ((!GEN || gas < SG_GAS_WM0) ? rec[(size_t)i * SG_NQ + SG_Q_MMR + gas]
                                     : fld ? fmax(fld[(size_t)i * d.P + col], 0.0)
                                           : wm[gas - SG_GAS_WM0])
                                  : 0.0;
      num[c] = 0.0; den[c] = 0.0; kmin[c] = DBL_MAX;
    }
    double sflux = 0.0;
    for (int q = 0; q < nk; q++) {
      const SgTerm &t = spec->minor[t0 + q];
      double kk[kKeqChunks];
#pragma unroll
      for (int c = 0; c < kKeqChunks; c++) {
        const int i = c * kWave + lane;
        if (c < nch && i < L) {
          if (!GEN || t.fnc == SG_FNC_POWER_LAW) kk[c] = k_term(t, lnp[c], lnt[c], mmr[c]);
          else {
            const double *r = rec + (size_t)i * SG_NQ;
            kk[c] = k_scaled(t, spec, lnp[c], r[SG_Q_P], r[SG_Q_T],
                             grec ? grec + ((size_t)col * L + i) * SG_NGQ : nullptr, mmr[c]);
          }
        } else {
          kk[c] = 0.0;
        }
        kmin[c] = fmin(kmin[c], kk[c]);
      }
      if (d.grey1 && nk == 1) {
        // a single k-term with l_grey_single: its absorption joins the grey
        // extinction as it is (solve_band_k_eqv_scl.F90:941-951)
#pragma unroll
        for (int c = 0; c < kKeqChunks; c++) kgrey[c] = kgrey[c] + kk[c];
        continue;
      }
      if (solar) {
        // direct-beam transmission: f = w prod exp(-k m sec0) down to the
        // bottom of each layer; den accumulates f (summed over terms, fg)
        double fdt = t.w, f[kKeqChunks];
#pragma unroll
        for (int c = 0; c < kKeqChunks; c++) {
          f[c] = 0.0;
          if (c < nch) {
            const int i = c * kWave + lane;
            const double x = i < L ? exp(-kk[c] * mass[c] * sec0) : 1.0;
            f[c] = fdt * scan_mul(x, lane);
            fdt = lane63(f[c]);
          }
        }
        // lanes below layer L-1 carry the identity, so fdt is now the
        // transmission to the surface (the value at layer L-1)
        const double fs = fdt;
#pragma unroll
        for (int c = 0; c < kKeqChunks; c++) {
          num[c] = num[c] + kk[c] * fs;
          den[c] = den[c] + f[c];
        }
        sflux = sflux + fs;
      } else {
        // downward differential flux F at the top of each layer
        double F = -b_top, tr[kKeqChunks], su[kKeqChunks], fdl[kKeqChunks];
#pragma unroll
        for (int c = 0; c < kKeqChunks; c++) {
          tr[c] = 1.0; su[c] = 0.0; fdl[c] = 0.0;
          if (c < nch) {
            const int i = c * kWave + lane;
            double A = 1.0, C = 0.0;
            if (i < L) {
              const double tau = kk[c] * mass[c];
              tr[c] = exp(-kDiffMinor * tau);
              su[c] = (1.0 - tr[c] + sqe) * (b1[c] - b0[c]) / (kDiffMinor * tau + sqe);
              A = tr[c];
              C = -su[c];
            }
            scan_affine(A, C, lane);
            const double below = A * F + C;
            double above = lane_below(below);
            if (lane == 0) above = F;
            fdl[c] = above;
            F = lane63(below);
          }
        }
        // lanes below layer L-1 carry the identity, so F is the downward
        // flux at the surface
        double U = (bg - b_bot) + alb * F;   // upward flux at the surface
#pragma unroll
        for (int cc = kKeqChunks - 1; cc >= 0; cc--) {
          if (cc < nch) {
            // bottom-up over this chunk: lane with layer i = cc*64 + lane,
            // scanned in reverse lane order
            const int rl = kWave - 1 - lane;     // reverse position
            const int i = cc * kWave + rl;
            double A = 1.0, C = 0.0;
            const double trr = __shfl(tr[cc], rl, kWave), sur = __shfl(su[cc], rl, kWave);
            if (i < L) { A = trr; C = sur; }
            scan_affine(A, C, lane);
            const double above = A * U + C;                  // F up at the top of layer i
            double belowu = lane_below(above);                // at its bottom
            if (lane == 0) belowu = U;
            // back to natural lane order: lane l holds layer cc*64 + l
            const double bl = __shfl(belowu, rl, kWave);
            const int il = cc * kWave + lane;
            if (il < L) {
              const double a = t.w * fabs(fdl[cc]), bb = t.w * fabs(bl);
              num[cc] = num[cc] + (kk[cc] * a + kk[cc] * bb);
              den[cc] = den[cc] + (a + bb);
            }
            U = lane63(above);
          }
        }
      }
    }
    // this gas's equivalent extinction and (solar) direct-beam adjustment
    if (!(d.grey1 && nk == 1)) {
#pragma unroll
    for (int c = 0; c < kKeqChunks; c++) {
      if (c < nch) {
        const int i = c * kWave + lane;
        double kq;
        if (solar) {
          kq = sflux > 0.0 ? num[c] / sflux : kmin[c];
          // flux_gas at the top of layer i: the summed transmission one layer
          // up (1 above the first layer)
          double fgp = lane_below(den[c]);
          const double carry = lane63(den[c > 0 ? c - 1 : 0]);
          if (lane == 0) fgp = c > 0 ? carry : 1.0;
          if (i == 0) fgp = 1.0;
          if (i < L && fgp > 0.0) adj[c] = adj[c] * den[c] / fgp;
        } else {
          kq = den[c] > 0.0 ? num[c] / den[c] : kmin[c];
        }
        if (i < L) keq[c] = keq[c] + kq;
      }
    }
    }
    t0 += nk;
  }
#pragma unroll
  for (int c = 0; c < kKeqChunks; c++) {
    const int i = c * kWave + lane;
    if (c < nch && i < L)
      prep_layer(d, spec, l_rayleigh, colrec, grec, colscal, bandrec, bandscal, cldopt, krange, ae, col, b, i,
                 keq[c], kgrey[c], adj[c], b0[c], b1[c], GEN);
  }
}

// Per (column, band, layer), given the layer's equivalent extinction and
// direct-beam adjustment product: the grey extinction (Rayleigh,
// grey_opt_prop.F90:476-541, 816-835, plus k_eqv), the solar direct-beam
// adjustment with the exponential of the equivalent optical depth
// (solve_band_k_eqv_scl.F90:848-891), the Planck source terms
// (diff_planck_source_mod.F90:100-134) and the cloud optics.
// bandrec[col][band][layer][4].
__device__ __forceinline__ void prep_layer(const CallDims &d, const SgSpec *__restrict__ spec,
                                           int l_rayleigh, const double *__restrict__ colrec,
                                           const double *__restrict__ grec,
                                           const double *__restrict__ colscal,
                                           double *__restrict__ bandrec,
                                           double *__restrict__ bandscal,
                                           double *__restrict__ cldopt, const int *__restrict__ krange,
                                           const AerIn &ae, int col, int b, int i, double keq, double kgrey,
                                           double adjust, double b0, double b1, bool gen) {
  const int NB = d.NB, L = d.L;
  const SgBand &B = spec->band[b];
  const double *rec = colrec + (size_t)col * L * SG_NQ;
  const double *r = rec + i * SG_NQ;
  if (d.mcica) {
    // droplet optics of the layer's in-cloud water at unit condensate
    // scaling (they depend on the band, not on the k-term; the solve scales
    // them by each sub-column's c_sub)
    // (written only where the layer has cloud: the solve reads them only
    // for cloudy layers of a sub-column)
    if (r[SG_Q_FRAC1] > 0.0) {
      double *co = cldopt + (((size_t)col * NB + b) * L + i) * d.co_stride;
      double kext, kscat, phf, fwd;
      drop_optics(spec->drop[d.drop], b, r[SG_Q_LWC], r[SG_Q_RE], kext, kscat, phf, fwd);
      double ie = 0.0, is = 0.0, ip = 0.0, iff = 0.0;
      if (d.ice) ice_optics(spec->ice[d.ice_t], b, r[SG_Q_IWC], r[SG_Q_DE], r[SG_Q_T], ie, is, ip, iff);
      if (d.f32) {
        // mixed form: droplets and ice together, (kext, kext - kscat, phf,
        // fwd) in FP32, the absorption formed here in FP64
        float4 *cf = (float4 *)((float *)cldopt + (((size_t)col * NB + b) * L + i) * 4);
        kext = kext + ie; kscat = kscat + is; phf = phf + ip; fwd = fwd + iff;
        *cf = make_float4((float)kext, (float)(kext - kscat), (float)phf, (float)fwd);
      } else {
        co[0] = kext; co[1] = kscat; co[2] = phf; co[3] = fwd;
        if (d.ice) { co[4] = ie; co[5] = is; co[6] = ip; co[7] = iff; }
      }
    }
  }
  double *br = bandrec + (((size_t)col * NB + b) * L + i) * SG_NBQ;
  double *bs = bandscal + ((size_t)col * NB + b) * SG_NBQ;
  const bool solar = d.isolir == SOCGPU_SOURCE_SOLAR;
  const double sec0 = colscal[(size_t)col * SG_NS + SG_S_SEC0];
  const double ray = (solar && l_rayleigh) ? B.rayleigh : 0.0;
  // clear-air grey extinction: continua, then Rayleigh scattering
  // (grey_opt_prop.F90:801-835), then the equivalent extinction
  const double kcont = (gen && B.n_cont > 0 && !d.no_cont) ? continuum_k(B, r[SG_Q_P], r[SG_Q_T],
                                                  grec[((size_t)col * L + i) * SG_NGQ + SG_G_RHO],
                                                  r[SG_Q_MMR + SG_GAS_H2O])
                                    : 0.0;
  // modal aerosols: their absorption precedes the continua in the clear-air
  // extinction, their scattering follows Rayleigh's (the sums over the
  // modes are formed by the pack kernel, aer_sums)
  double kaer = 0.0, ksc = ray, pfa = 0.0, fwa = 0.0;
  if (d.n_aer > 0) {
    const double *ar = ae.rec + (((size_t)col * NB + b) * L + i) * 4;
    ksc = ar[0]; pfa = ar[1]; fwa = ar[2]; kaer = ar[3];
  }
  br[0] = ((kaer + kcont) + ksc) + keq + kgrey;
  const double ksc_clr = ksc;
  if (d.triple) {
    // grey optics of each cloud type (grey_opt_prop.F90:975-1000,
    // 1185-1230: Rayleigh, then the type's droplets and its ice component
    // (zero without ice), at layers from the call's cloud top where the
    // column has cloud), the equivalent extinction (solve_band_k_eqv_scl.F90:
    // 1038-1045), moments normalised by the scattering (grey_opt_prop.F90:
    // 1474-1500) and the delta-rescaled phase function (rescale_phase_fnc.F90)
    const int nct = krange[0] < L ? krange[0] : L;
    double *co = cldopt + (((size_t)col * NB + b) * L + i) * 8;
    for (int k = 0; k < 2; k++) {
      // every type starts from the clear-air totals (aerosols included)
      double kgt = (kaer + kcont) + ksc_clr, ksc = ksc_clr, phf = pfa, fwd = fwa;
      if (i + 1 >= nct && r[SG_Q_WCLD] > 0.0) {
        const SgDrop &dr = spec->drop[k == 0 ? d.drop : d.drop2];
        const double mmr = r[k == 0 ? SG_Q_LWC : SG_Q_LWC2], re = r[k == 0 ? SG_Q_RE : SG_Q_RE2];
        double ct, cs, cp, cf;
        drop_optics(dr, b, mmr, re, ct, cs, cp, cf);
        kgt = kgt + ct;
        ksc = ksc + cs;
        phf = phf + cp;
        fwd = fwd + cf;
        // the type's ice component
        ct = 0.0; cs = 0.0; cp = 0.0; cf = 0.0;
        if (d.ice)
          ice_optics(spec->ice[k == 0 ? d.ice_t : d.ice_t2], b, r[k == 0 ? SG_Q_IWC : SG_Q_IWC2],
                     r[k == 0 ? SG_Q_DE : SG_Q_DE2], r[SG_Q_T], ct, cs, cp, cf);
        kgt = kgt + ct;
        ksc = ksc + cs;
        phf = phf + cp;
        fwd = fwd + cf;
      }
      kgt = kgt + keq + kgrey;
      if (ksc > kTiny) {
        const double inv = 1.0 / ksc;
        fwd = fwd * inv;
        phf = phf * inv;
      }
      phf = (phf - fwd) / (1.0 - fwd);
      co[4 * k + 0] = kgt;
      co[4 * k + 1] = ksc;
      co[4 * k + 2] = phf;
      co[4 * k + 3] = fwd;
    }
  }
  if (solar) {
    const double temp_max = 36.043653389117154;  // log(1/DBL_EPSILON)
    br[1] = adjust * exp(fmin(keq * r[SG_Q_MASS] * sec0, temp_max));
    br[2] = 0.0;
    br[3] = 0.0;
    if (i == 0) { bs[0] = colscal[(size_t)col * SG_NS + SG_S_SOL] * B.solar_flux; bs[1] = 0.0; }
  } else {
    // b0, b1: the band's Planck fluxes at the top and bottom of the layer
    double bt = planck_band(B, spec->n_deg, spec->t_ref_planck, r[SG_Q_T]);
    // diff_2 (diff_planck_source_mod.F90:112-121); zero for a linear source,
    // whose sources (ir_source.F90:52-59) are the quadratic ones without it
    br[1] = d.ir_lin ? 0.0 : 2.0 * (b1 + b0 - 2.0 * bt);
    br[2] = b1 - b0;                     // diff
    br[3] = b1;                          // planck at the bottom of the layer
    if (i == 0) {
      bs[0] = b0;
      // (the given surface emission was stored by the pack kernel)
      if (colscal[(size_t)col * SG_NS + SG_S_LFG] == 0.0)
        bs[1] = planck_band(B, spec->n_deg, spec->t_ref_planck,
                            colscal[(size_t)col * SG_NS + SG_S_TGROUND]);
    }
  }
}

// Optical properties of one layer for one lane: clear sky (gas + grey
// extinction, Rayleigh scattering; single_scattering.F90:114-128 and the
// delta-rescaling of rescale_tau_omega.F90:73-80, which is the identity for
// zero forward scattering) or, on a cloudy layer of the lane's McICA
// sub-column, the cloudy mix of mcica_sample.F90:470-720 (homogeneous
// representation, l_rescale) scaled by c_sub, then two_coeff. Cloudy
// coefficients carry the layer's type fraction as in mcica_column.F90:
// 360-520.
struct LaneCtx {
  bool solar;
  bool ice;                 // FP64 form: ice optics beside the droplets'
  int nosc;                 // the lane's term without scattering: SOCGPU_SCATTER_NONE
                            // (absorption only) or _EXT (extinction as absorption); 0
  double sec0, ksc_clear;
  bool aer;                 // modal aerosols: clear air scatters with a phase function
};

// One lane's inputs for one layer, loaded a layer ahead of their use so
// that the loads of layer i+1 are in flight while layer i is computed.
struct LayerIn {
  double lnp, lnt, mmr, mass, frac1;  // column record (wave-uniform but mmr)
  double kx;                          // the term's k_abs, for scalings other than the power law
  double kgt, q1, q2;                 // band record: grey extinction, adjust or Planck terms
  double ksa, pfa, fwa;               // clear-air scattering and its moment products
                                      // (Rayleigh only: ksc_clear, 0, 0)
  double cv;                          // the lane's sub-column condensate scaling
  double co[4];                       // cloud optics at unit condensate scaling:
                                      // FP64 form, the droplets' (k_ext, k_scat,
                                      // first moment, forward scattering); mixed
                                      // form, droplets and ice (k_ext, k_abs, ...)
  double ci[4];                       // FP64 form: the ice component's
};

struct LaneSrc {
  const double *rec, *br, *csub, *cop;
  const float *csubf, *copf;  // FP32 forms (mixed form)
  double ksc_clear;           // clear-air scattering (Rayleigh)
  const double *aer;          // the lane's aerosol record (column, band), or null
  int co_stride, ice;         // doubles per layer of cop; ice optics beside the droplets'
  int gas, isub;
  int fnc;                    // the term's scaling function
  const SgTerm *term;         // the term (device table), for fnc other than the power law
  const SgSpec *spec;
  const double *grec;         // the column's gas record, or null
};

// MIXED reads the FP32 c_sub and droplet optics (kext, kabs, first moment,
// forward scattering), the FP64 form (kext, kscat, ...) in FP64. CLR: a layer
// the lane knows to be clear (above the column's first cloudy layer, or its
// clear solve): no cloud fraction, scaling or cloud optics are read, so a
// prefetch of the next layer issues its loads and waits for none of them.
template <bool MIXED, bool GEN, bool CLR = false>
__device__ __forceinline__ void load_layer(const LaneSrc &s, int i, LayerIn &x) {
  const double *r = s.rec + i * SG_NQ;
  x.lnp = r[SG_Q_LNP];
  x.lnt = r[SG_Q_LNT];
  x.mmr = r[SG_Q_MMR + s.gas];
  x.mass = r[SG_Q_MASS];
  x.kx = (!GEN || s.fnc == SG_FNC_POWER_LAW)
             ? 0.0
             : k_scaled(*s.term, s.spec, x.lnp, r[SG_Q_P], r[SG_Q_T], s.grec ? s.grec + i * SG_NGQ : nullptr,
                        x.mmr);
  const double *q = s.br + i * SG_NBQ;
  x.kgt = q[0];
  x.q1 = q[1];
  x.q2 = q[2];
  if (GEN && s.aer) {
    const double *a = s.aer + i * 4;
    x.ksa = a[0]; x.pfa = a[1]; x.fwa = a[2];
  } else {
    x.ksa = s.ksc_clear; x.pfa = 0.0; x.fwa = 0.0;
  }
  if (CLR) {
    x.frac1 = 0.0;
    x.cv = 0.0;
    x.co[0] = 0.0; x.co[1] = 0.0; x.co[2] = 0.0; x.co[3] = 0.0;
    return;
  }
  x.frac1 = r[SG_Q_FRAC1];
  if (MIXED) x.cv = s.isub > 0 ? (double)s.csubf[(size_t)i * SG_MAXSUBCOL] : 0.0;
  else x.cv = s.isub > 0 ? s.csub[(size_t)i * SG_MAXSUBCOL] : 0.0;
  // A sampled sub-column's cloud optics are read whether or not the layer is
  // cloudy in it (the test is made where they are used, and the table spans
  // every layer), so that a prefetch waits for none of its loads.
  x.co[0] = 0.0; x.co[1] = 0.0; x.co[2] = 0.0; x.co[3] = 0.0;
  x.ci[0] = 0.0; x.ci[1] = 0.0; x.ci[2] = 0.0; x.ci[3] = 0.0;
  if (s.isub > 0) {
    if (MIXED) {
      const float4 co = *(const float4 *)(s.copf + i * 4);
      x.co[0] = co.x; x.co[1] = co.y; x.co[2] = co.z; x.co[3] = co.w;
    } else {
      const double *co = s.cop + i * s.co_stride;
      x.co[0] = co[0]; x.co[1] = co[1]; x.co[2] = co[2]; x.co[3] = co[3];
      if (s.ice) { x.ci[0] = co[4]; x.ci[1] = co[5]; x.ci[2] = co[6]; x.ci[3] = co[7]; }
    }
  }
}

// x.co: the layer's cloudy totals of the lane's sub-column (load_layer).
template <bool GEN>
__device__ __forceinline__ void layer_coef(const LaneCtx &c, const SgTerm &t, const LayerIn &x,
                                           Coef &o) {
#pragma clang fp contract(off)
  double kgas = (!GEN || t.fnc == SG_FNC_POWER_LAW) ? k_term(t, x.lnp, x.lnt, x.mmr) : x.kx;
  const double frac1 = x.frac1, kgt = x.kgt, csub = x.cv;
  if (csub * frac1 > 0.0) {
    // the sub-column's cloudy totals (mcica_sample.F90:614-640): each
    // component, droplets then ice, scaled by c_sub and added in turn
    const double cs = csub;
    double kg = kgt + x.co[0] * cs, ks = x.ksa + x.co[1] * cs;
    double pf = x.pfa + x.co[2] * cs, fw = x.fwa + x.co[3] * cs;
    if (c.ice) {
      kg = kg + x.ci[0] * cs;
      ks = ks + x.ci[1] * cs;
      pf = pf + x.ci[2] * cs;
      fw = fw + x.ci[3] * cs;
    }
    if (c.nosc) {
      // single_scattering.F90:132-158, mcica_column.F90:509-522
      const double kt = kg + kgas;
      two_coeff_fast_lw((c.nosc == SOCGPU_SCATTER_NONE ? kt - ks : kt) * x.mass, o);
      o.T *= frac1; o.s1 *= frac1; o.s2 *= frac1;
      return;
    }
    pf = (pf - fw) / fmax(ks - fw, kEps);
    fw = fw / fmax(ks, kEps);
    double ktot = kg + kgas;
    double tau = ktot * x.mass;
    double om = fmin(ks / (ktot + kTiny), 1.0 - 32.0 * kEps);
    tau = tau * (1.0 - om * fw);
    om = om * (1.0 - fw) / (1.0 - om * fw);
    two_coeff(c.solar, tau, om, pf, c.sec0, o);
    o.T *= frac1; o.R *= frac1; o.T0 *= frac1; o.s1 *= frac1; o.s2 *= frac1;
  } else {
    double ktot = kgt + kgas;
    if (c.nosc) {
      two_coeff_fast_lw((c.nosc == SOCGPU_SCATTER_NONE ? ktot - x.ksa : ktot) * x.mass, o);
      return;
    }
    double tau = ktot * x.mass;
    double om = fmin(x.ksa / (ktot + kTiny), 1.0 - 32.0 * kEps);
    double pf = 0.0;
    if (GEN && c.aer) {
      // clear air with aerosols: moments over the scattering
      // (grey_opt_prop.F90:1360-1440), the phase function delta-rescaled
      // (rescale_phase_fnc.F90:116-124), then rescale_tau_omega
      double fw = x.fwa;
      pf = x.pfa;
      if (x.ksa > kTiny) {
        const double inv = 1.0 / x.ksa;
        fw = fw * inv;
        pf = pf * inv;
      }
      pf = (pf - fw) / (1.0 - fw);
      tau = tau * (1.0 - om * fw);
      om = om * (1.0 - fw) / (1.0 - om * fw);
    }
    two_coeff(c.solar, tau, om, pf, c.sec0, o);
  }
}

// Mixed form of layer_coef: gas k-term, cloud optics, single scattering and
// two-stream coefficients in FP32 (two_coeff_f32), carrying the co-albedo
// 1 - omega from the absorption so the near-conservative limit keeps its
// accuracy; the reference's clamp omega <= 1 - 32 eps(FP64) is kept as
// 1 - omega >= 32 eps(FP64). No FP64 fallback: two_coeff_f32 has no
// resonance singularity.
template <bool GEN>
__device__ __forceinline__ void layer_coef_mixed(const LaneCtx &c, const SgTerm &t, const LayerIn &x,
                                                Coef &o) {
#pragma clang fp contract(on)
  const float kgas =
      (!GEN || t.fnc == SG_FNC_POWER_LAW) ? k_term_f32(t, (float)x.lnp, (float)x.lnt, (float)x.mmr)
                                          : (float)x.kx;
  const double csub = x.cv;
  const double *co = x.co;
  const float frac1 = (float)x.frac1;
  const float kag = (float)(x.kgt - x.ksa);  // grey absorption (equivalent extinction)
  const float mass = (float)x.mass;
  const float omc_min = (float)(32.0 * kEps);
  float tau, om, omc, g;
  bool cloudy = csub * frac1 > 0.0;
  if (cloudy) {
    // absorption kext - kscat formed in FP64 by the prep kernel (relative
    // accuracy ~1e-10 even for weakly absorbing droplets), kept in FP32
    const float kext = (float)co[0], kabs = (float)co[1];
    const float phf = (float)co[2], fwd = (float)co[3];
    const float cs = (float)csub;
    float ks = (float)x.ksa + (kext - kabs) * cs;
    float fw = (float)x.fwa + fwd * cs;
    float pf = (float)x.pfa + phf * cs;
    pf = SG_FDIV(pf - fw, fmaxf(ks - fw, (float)kEps));
    fw = SG_FDIV(fw, fmaxf(ks, (float)kEps));
    float ktot = (float)x.ksa + kag + kgas + kext * cs;
    float kab = kag + kgas + kabs * cs;
    if (c.nosc) {
      // without scattering: the optical depth of absorption or of
      // extinction, no albedo (single_scattering.F90:132-158)
      CoefF f;
      two_coeff_f32(false, (c.nosc == SOCGPU_SCATTER_NONE ? kab : ktot) * mass, 0.0f, 1.0f, 0.0f, 1.0f, f);
      coef_from_f32(f, o);
      o.T *= frac1; o.R *= frac1; o.s1 *= frac1; o.s2 *= frac1;
      return;
    }
    tau = ktot * mass;
    omc = fmaxf(SG_FDIV(kab, ktot), omc_min);
    om = 1.0f - omc;
    float sc = 1.0f - om * fw;
    tau = tau * sc;
    omc = SG_FDIV(omc, sc);
    om = 1.0f - omc;
    g = pf;
  } else {
    float ktot = (float)x.ksa + kag + kgas;
    if (c.nosc) {
      CoefF f;
      two_coeff_f32(false, (c.nosc == SOCGPU_SCATTER_NONE ? kag + kgas : ktot) * mass, 0.0f, 1.0f, 0.0f, 1.0f,
                    f);
      coef_from_f32(f, o);
      return;
    }
    tau = ktot * mass;
    omc = ktot > 0.0f ? fmaxf(SG_FDIV(kag + kgas, ktot), omc_min) : 1.0f;
    om = 1.0f - omc;
    g = 0.0f;
    if (GEN && c.aer) {
      // clear air with aerosols (layer_coef): moments over the scattering,
      // delta-rescaling
      float fw = (float)x.fwa, pf = (float)x.pfa;
      const float ksa = (float)x.ksa;
      if (ksa > 0.0f) {
        const float inv = SG_FDIV(1.0f, ksa);
        fw = fw * inv;
        pf = pf * inv;
      }
      g = SG_FDIV(pf - fw, 1.0f - fw);
      const float sc = 1.0f - om * fw;
      tau = tau * sc;
      omc = SG_FDIV(omc, sc);
      om = 1.0f - omc;
    }
  }
  CoefF f;
  two_coeff_f32(c.solar, tau, om, omc, g, (float)c.sec0, f);
  coef_from_f32(f, o);
  if (cloudy) { o.T *= frac1; o.R *= frac1; o.T0 *= frac1; o.s1 *= frac1; o.s2 *= frac1; }
}

// The elimination update of one layer: with the layer's transmission and
// reflection (o.T, o.R) and its upward and downward sources (su, sd), the
// reflectance of the atmosphere above (rho) and the diffuse flux it sends
// down (dn) move to the layer's bottom; returns the back-substitution pair
// fu(top) = a fu(bottom) + bb. The mixed form takes 1 / (1 - R rho) from the
// hardware reciprocal estimate and two Newton steps (a few ulp of FP64).
template <bool MIXED>
__device__ __forceinline__ void elim_step(const Coef &o, double su, double sd, double &rho, double &dn,
                                          double &a, double &bb) {
#pragma clang fp contract(on)
  double bp;
  if (MIXED) {
    const double x = 1.0 - o.R * rho;
    bp = __builtin_amdgcn_rcp(x);
    bp = fma(bp, fma(-x, bp, 1.0), bp);
    bp = fma(bp, fma(-x, bp, 1.0), bp);
  } else {
    bp = 1.0 / (1.0 - o.R * rho);
  }
  a = o.T * bp;
  bb = (o.R * dn + su) * bp;
  double rho_n = o.R + o.T * rho * a;
  dn = o.T * (rho * bb + dn) + sd;
  rho = rho_n;
}

// One layer of the top-down elimination for one lane: the layer's
// coefficients (FP64 layer_coef or the mixed layer_coef_mixed), its sources
// (solar_source.F90:104-162 with the k_eqv direct-beam adjustment, or
// ir_source.F90:93-120), and the update of the reflectance of the atmosphere
// above (rho) and the diffuse flux it sends down (dn). Returns the
// back-substitution pair fu(i) = a fu(i+1) + bb.
template <bool MIXED, bool GEN>
__device__ __forceinline__ void layer_step(const LaneCtx &c, const SgTerm &t, const LayerIn &x,
                                           double &fdir, double &rho, double &dn, double &a,
                                           double &bb) {
#pragma clang fp contract(on)
  Coef o;
  if (MIXED) layer_coef_mixed<GEN>(c, t, x, o);
  else layer_coef<GEN>(c, t, x, o);
  double su, sd;
  if (c.solar) {
    double fnew = fdir * o.T0 * x.q1;
    su = o.s1 * fdir;
    sd = (o.s2 - o.T0) * fdir + fnew;
    fdir = fnew;
  } else {
    su = o.s1 * x.q2 + o.s2 * x.q1;
    sd = -o.s1 * x.q2 + o.s2 * x.q1;
  }
  elim_step<MIXED>(o, su, sd, rho, dn, a, bb);
}

// The column solve, one lane per (column, g-point): the g-point's clear-sky
// solve (weight w_k (1 - frac_cloudy)) and its McICA sub-column samples
// (weight w_k frac_cloudy / subcol_k each) (mcica_sample.F90:453-930;
// monochromatic_radiance.F90:429-691; mcica_column.F90:255-622;
// solver_homogen_direct.F90:93-127). The lanes of a call are numbered column
// by column (n_items per column: its g-points) and cut into waves of 64, so a
// wave spans up to kMaxSeg columns ("segments") and no lane idles between
// columns. The two-stream system is eliminated top-down (reflectance of the
// atmosphere above and the downward diffuse flux it emits) and
// back-substituted bottom-up.
//
// Above the column's first cloudy layer (ic) every sample is clear, so its
// layers, its elimination and its direct beam there are those of the clear
// solve, to the bit. The lane therefore
//  1. eliminates the clear layers 0..ic-1 once;
//  2. for each of its solves, continues from that state through layers
//     ic..L-1, applies the surface condition and back-substitutes up to
//     level ic, adding the weighted fluxes of levels ic..L into acc;
//  3. back-substitutes the layers above ic once for all its solves together:
//     they share fu(i) = a(i) fu(i+1) + bb(i), so the weighted sum
//     X(i) = sum_j w_j fu_j(i) obeys X(i) = a(i) X(i+1) + bb(i) W with
//     W = sum_j w_j, and the downward flux sum is rho(i) X(i) + dn(i) W.
// This does the work of the reference's clear layers once per g-point
// instead of once per sample (mcica_column.F90:266-298 recomputes them for
// every sample); the sums associate differently from a sum of separate
// solves, by rounding only.
//
// Output: partial[wave][segment][3][L+1] (up, down, direct) of the segment's
// weighted sums, formed by butterflies over the wave with the other
// segments' lanes masked.
//
// RECOMP = false keeps the elimination state of every clear layer above ic
// (a, bb, rho, dn and the direct flux) in lane-private memory between the
// sweeps. RECOMP = true keeps only (fdir, rho, dn) at the top of every
// kSeg-th layer and recomputes each segment's layers (coefficients and
// elimination, bit-identical to the first sweep) into registers on the way
// up: a second pass of arithmetic in place of the memory traffic. The layers
// below ic are few (the cloud sits in the lowest of the log-spaced layers)
// and are kept in private memory for each solve in turn.
constexpr int kSeg = 4;
constexpr int kMaxSeg = 4;     // columns a wave can span (n_items >= kMinItems)
constexpr int kMinItems = 22;  // floor(63 / 22) + 2 = 4

// Sum over the lanes of segment s of each of v[0..7] (other lanes count
// zero); lanes 8j (j = 0..7) receive the sum of v[j].
__device__ __forceinline__ void seg_reduce8(const double (&v)[8], int seg, int s, int nseg, double (&o)[8],
                                            int lane) {
#pragma unroll
  for (int k = 0; k < 8; k++) o[k] = (nseg == 1 || seg == s) ? v[k] : 0.0;
  reduce8(o, lane);
}

template <bool MIXED, bool RECOMP, bool GEN>
__device__ __forceinline__ void solve_body(const CallDims &d, const SgSpec *__restrict__ spec,
                                           int l_rayleigh, const double *__restrict__ colrec,
                                           const double *__restrict__ grec,
                                           const double *__restrict__ colscal,
                                           const double *__restrict__ bandrec,
                                           const double *__restrict__ bandscal,
                                           const double *__restrict__ c_sub,
                                           const double *__restrict__ cldopt,
                                           const double *__restrict__ aerrec,
                                           double *__restrict__ partial, int wave, int lane) {
#pragma clang fp contract(on)
  static_assert(kSeg == 4, "checkpoints and output blocks share the 4-layer grid");
  const int L = d.L, NB = d.NB, IM = d.n_items;
  const long long g0 = (long long)wave * kWave, gmax = (long long)d.P * IM;
  if (g0 >= gmax) return;
  const int colA = (int)(g0 / IM);
  const long long glast = (g0 + kWave < gmax ? g0 + kWave : gmax) - 1;
  const int nseg = (int)(glast / IM) - colA + 1;  // wave-uniform, 1..kMaxSeg
  const long long g = g0 + lane;
  const bool inrange = g < gmax;
  const int col = inrange ? (int)(g / IM) : colA;
  const int item = inrange ? (int)(g - (long long)col * IM) : 0;
  const int seg = col - colA;
  const bool solar = d.isolir == SOCGPU_SOURCE_SOLAR;
  const double *cs = colscal + (size_t)col * SG_NS;
  const double fc = d.mcica ? cs[SG_S_FC] : 0.0;
  const bool act = inrange && item < spec->n_g;
  const int gi = act ? item : 0;
  const SgTerm t = spec->g[gi];
  // the lane's solves: the clear one unless every sub-column is cloudy, then
  // the g-point's samples if any sub-column is
  const bool clear_on = act && fc < 1.0;
  const int n_s = (act && d.mcica && fc > 0.0) ? spec->g_n_sub[d.mode][gi] : 0;
  const int s0 = n_s > 0 ? spec->g_sub0[d.mode][gi] : 0;
  const int n_solve = act ? (clear_on ? 1 : 0) + n_s : 0;
  // first layer of the column with cloud (every layer above is clear in
  // every sub-column); L where the lane has no sample
  const int ic = n_s > 0 ? min(max((int)cs[SG_S_CT], 0), L) : L;
  const int b = t.band;
  double *out = partial + (size_t)wave * kMaxSeg * 3 * (L + 1);
  const double *rec = colrec + (size_t)col * L * SG_NQ;
  const double *br = bandrec + ((size_t)col * NB + b) * L * SG_NBQ;
  const double *bs = bandscal + ((size_t)col * NB + b) * SG_NBQ;
  LaneCtx c;
  c.solar = solar;
  c.sec0 = cs[SG_S_SEC0];
  c.ksc_clear = (solar && l_rayleigh) ? spec->band[b].rayleigh : 0.0;
  c.ice = d.ice && !MIXED;
  c.nosc = term_nosc(d.scat, t.iscat);
  c.aer = GEN && aerrec != nullptr;
  LaneSrc src;
  src.aer = c.aer ? aerrec + ((size_t)col * NB + b) * L * 4 : nullptr;
  src.rec = rec;
  src.br = br;
  src.cop = d.mcica ? cldopt + ((size_t)col * NB + b) * L * d.co_stride : nullptr;
  src.co_stride = d.co_stride;
  src.ice = d.ice && !MIXED;
  src.ksc_clear = c.ksc_clear;
  src.copf = d.mcica ? (const float *)cldopt + ((size_t)col * NB + b) * L * 4 : nullptr;
  src.csub = c_sub + (size_t)col * L * SG_MAXSUBCOL;
  src.csubf = (const float *)c_sub + (size_t)col * L * SG_MAXSUBCOL;
  src.gas = t.gas;
  src.isub = 0;
  src.fnc = t.fnc;
  src.term = spec->g + gi;
  src.spec = spec;
  src.grec = grec ? grec + (size_t)col * L * SG_NGQ : nullptr;
  const double alb_diff = bs[2], alb_dir = bs[3];
  const double finc = solar ? bs[0] / c.sec0 : -bs[0];  // flux_inc_down, flux_inc_direct

  // 1. the clear layers above ic, once. Every load and store of the loop is
  // unconditional (the prefetch index clamped, the checkpoint of a layer
  // that starts no segment written to a spare row), so the compiler's wait
  // counts see one path and the prefetch of layer i+1 stays in flight while
  // layer i is computed.
  constexpr int kNst = RECOMP ? (SG_MAXL + kSeg - 1) / kSeg : SG_MAXL;
  double st[kNst + 1][RECOMP ? 3 : 5];
  double fdir = solar ? finc : 0.0, rho = 0.0, dn = finc;
  const int ic1 = act ? ic : 0;
  {
    LayerIn cur, nxt;
    load_layer<MIXED, GEN, true>(src, 0, cur);
    for (int i = 0; i < ic1; i++) {
      load_layer<MIXED, GEN, true>(src, min(i + 1, ic1 - 1), nxt);
      if (RECOMP) {
        const int row = i % kSeg == 0 ? i / kSeg : kNst;
        st[row][0] = fdir; st[row][1] = rho; st[row][2] = dn;
      } else {
        st[i][4] = fdir;
      }
      double a, bb;
      const double rho0 = rho, dn0 = dn;
      layer_step<MIXED, GEN>(c, t, cur, fdir, rho, dn, a, bb);
      if (!RECOMP) {
        st[i][0] = a; st[i][1] = bb; st[i][2] = rho0; st[i][3] = dn0;
      }
      cur = nxt;
    }
  }

  // 2. each solve from level ic down, back to level ic; weighted fluxes of
  // levels ic..L in acc[lev - ic]. The downward sweep's loop loads and
  // stores unconditionally (as step 1); the direct flux at each level is kept
  // with the layer's elimination pair and added to acc on the way back up.
  double acc[SG_MAXL + 1][3];
  const int nacc = act ? L - ic : -1;
  for (int k = 0; k <= nacc; k++) { acc[k][0] = 0.0; acc[k][1] = 0.0; acc[k][2] = 0.0; }
  double W = 0.0;
  {
    double st2[SG_MAXL][5];
    const double sg_lw = (1.0 - alb_diff) * (bs[1] - br[(L - 1) * SG_NBQ + 3]);
    for (int j = 0; j < n_solve; j++) {
      const bool clr = clear_on && j == 0;
      const int sI = s0 + j - (clear_on ? 1 : 0);
      const int isub = clr ? 0 : spec->sub_isub[d.mode][sI];
      const double wj = clr ? t.w * (1.0 - fc) : t.w * (fc * spec->sub_inv_s[d.mode][sI]);
      W = W + wj;
      LaneSrc sj = src;
      sj.isub = isub;
      sj.csub = src.csub + (isub > 0 ? isub - 1 : 0);
      sj.csubf = src.csubf + (isub > 0 ? isub - 1 : 0);
      double f2 = fdir, r2 = rho, d2 = dn;
      LayerIn cur, nxt;
      load_layer<MIXED, GEN>(sj, min(ic, L - 1), cur);
      for (int i = ic; i < L; i++) {
        load_layer<MIXED, GEN>(sj, min(i + 1, L - 1), nxt);
        double a, bb;
        const double r0 = r2, d0 = d2, f0 = f2;
        layer_step<MIXED, GEN>(c, t, cur, f2, r2, d2, a, bb);
        st2[i - ic][0] = a; st2[i - ic][1] = bb; st2[i - ic][2] = r0; st2[i - ic][3] = d0;
        st2[i - ic][4] = f0;
        cur = nxt;
      }
      if (L > ic) acc[L - ic][2] += wj * f2;
      const double sgs = solar ? (alb_dir - alb_diff) * f2 : sg_lw;
      double fu = (alb_diff * d2 + sgs) / (1.0 - alb_diff * r2);
      double fd = r2 * fu + d2;
      acc[L - ic][0] += wj * fu;
      acc[L - ic][1] += wj * fd;
      for (int lev = L - 1; lev >= ic; lev--) {
        fu = st2[lev - ic][0] * fu + st2[lev - ic][1];
        fd = st2[lev - ic][2] * fu + st2[lev - ic][3];
        acc[lev - ic][0] += wj * fu;
        acc[lev - ic][1] += wj * fd;
        if (lev > ic) acc[lev - ic][2] += wj * st2[lev - ic][4];
      }
    }
    if (act) acc[0][2] = W * fdir;   // the direct beam at level ic, common to all
  }

  // 3. levels L..0 in blocks of 4: below ic from acc, above it the joint
  // back-substitution of the lane's solves; sums over the segment's lanes
  double x = act ? acc[0][0] : 0.0;
  double vdir[8], o8[8];
#pragma unroll
  for (int k = 0; k < 8; k++) vdir[k] = 0.0;
  for (int blk = L / kSeg; blk >= 0; blk--) {
    const int l0 = blk * kSeg;
    double sa[kSeg], sb[kSeg], sr[kSeg], sd[kSeg], sf[kSeg];
#pragma unroll
    for (int k = 0; k < kSeg; k++) { sa[k] = 0.0; sb[k] = 0.0; sr[k] = 0.0; sd[k] = 0.0; sf[k] = 0.0; }
    if (l0 < ic1) {
      if (RECOMP) {
        // the segment's layers again from its checkpoint, each layer's
        // inputs loaded (unconditionally, index clamped) while the one above
        // is computed
        double fs = st[blk][0], rs = st[blk][1], ds = st[blk][2];
        LayerIn xa, xb;
        load_layer<MIXED, GEN, true>(src, l0, xa);
#pragma unroll
        for (int k = 0; k < kSeg; k++) {
          const int i = l0 + k;
          if (k + 1 < kSeg) load_layer<MIXED, GEN, true>(src, min(i + 1, ic1 - 1), xb);
          sr[k] = rs; sd[k] = ds; sf[k] = fs;
          if (i < ic1) layer_step<MIXED, GEN>(c, t, xa, fs, rs, ds, sa[k], sb[k]);
          xa = xb;
        }
      } else {
#pragma unroll
        for (int k = 0; k < kSeg; k++) {
          const int i = l0 + k;
          if (i < ic1) { sa[k] = st[i][0]; sb[k] = st[i][1]; sr[k] = st[i][2]; sd[k] = st[i][3]; sf[k] = st[i][4]; }
        }
      }
    }
    double v[8];
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const int k = kSeg - 1 - m;   // levels l0+3 .. l0
      const int lev = l0 + k;
      double up = 0.0, dw = 0.0, dr = 0.0;
      if (act && lev <= L) {
        if (lev >= ic) {
          up = acc[lev - ic][0]; dw = acc[lev - ic][1]; dr = acc[lev - ic][2];
        } else {
          x = sa[k] * x + sb[k] * W;
          up = x;
          dw = sr[k] * x + sd[k] * W;
          dr = W * sf[k];
        }
      }
      v[2 * m] = up;
      v[2 * m + 1] = dw;
      if (solar) vdir[(blk & 1) * 4 + k] = dr;
    }
    for (int sg = 0; sg < nseg; sg++) {
      seg_reduce8(v, seg, sg, nseg, o8, lane);
      if ((lane & 7) == 0) {
        const int jj = lane >> 3, lev = l0 + kSeg - 1 - (jj >> 1);
        if (lev <= L) out[(sg * 3 + (jj & 1)) * (L + 1) + lev] = o8[0];
      }
    }
    if (solar && (blk & 1) == 0) {
      // direct flux of levels l0 .. l0+7
      for (int sg = 0; sg < nseg; sg++) {
        seg_reduce8(vdir, seg, sg, nseg, o8, lane);
        if ((lane & 7) == 0 && l0 + (lane >> 3) <= L) out[(sg * 3 + 2) * (L + 1) + l0 + (lane >> 3)] = o8[0];
      }
#pragma unroll
      for (int k = 0; k < 8; k++) vdir[k] = 0.0;
    }
  }
  if (!solar) {
    for (int sg = 0; sg < nseg; sg++)
      for (int i = lane; i <= L; i += kWave) out[(sg * 3 + 2) * (L + 1) + i] = 0.0;
  }
}

// The same solve with one lane per item (a clear k-term or one McICA
// sample, each over the whole column): the layout the mixed form keeps for
// spectra with the general gas optics. The per-g-point layout above faulted
// on the device in that one combination (mixed form, general gas optics,
// stored elimination state) for a cause not yet found; every other
// combination runs the per-g-point layout.
template <bool MIXED, bool RECOMP, bool GEN>
__device__ __forceinline__ void solve_body_items(const CallDims &d, const SgSpec *__restrict__ spec,
                                           int l_rayleigh, const double *__restrict__ colrec,
                                           const double *__restrict__ grec,
                                           const double *__restrict__ colscal,
                                           const double *__restrict__ bandrec,
                                           const double *__restrict__ bandscal,
                                           const double *__restrict__ c_sub,
                                           const double *__restrict__ cldopt,
                                           const double *__restrict__ aerrec,
                                           double *__restrict__ partial, int wave, int lane) {
#pragma clang fp contract(on)
  const int L = d.L, NB = d.NB, IM = d.n_items;
  const long long g0 = (long long)wave * kWave, gmax = (long long)d.P * IM;
  if (g0 >= gmax) return;
  const int colA = (int)(g0 / IM);
  const long long glast = (g0 + kWave < gmax ? g0 + kWave : gmax) - 1;
  const int nseg = (int)(glast / IM) - colA + 1;  // wave-uniform, 1..kMaxSeg
  const long long g = g0 + lane;
  const bool inrange = g < gmax;
  const int col = inrange ? (int)(g / IM) : colA;
  const int item = inrange ? (int)(g - (long long)col * IM) : 0;
  const int seg = col - colA;
  const bool solar = d.isolir == SOCGPU_SOURCE_SOLAR;
  const double *cs = colscal + (size_t)col * SG_NS;
  const double fc = d.mcica ? cs[SG_S_FC] : 0.0;
  const int G = spec->n_g;
  int gi = 0, isub = 0;
  double wgt = 0.0;
  bool active = false;
  if (inrange) {
    if (item < G) {
      if (fc < 1.0) {
        active = true;
        gi = item;
        wgt = spec->g[gi].w * (1.0 - fc);
      }
    } else if (fc > 0.0 && item - G < (d.mcica ? spec->n_sub_items[d.mode] : 0)) {
      const int sI = item - G;
      active = true;
      gi = spec->sub_g[d.mode][sI];
      isub = spec->sub_isub[d.mode][sI];
      wgt = spec->g[gi].w * (fc * spec->sub_inv_s[d.mode][sI]);
    }
  }
  double *out = partial + (size_t)wave * kMaxSeg * 3 * (L + 1);
  const SgTerm t = spec->g[gi];
  const int b = t.band;
  const double *rec = colrec + (size_t)col * L * SG_NQ;
  const double *br = bandrec + ((size_t)col * NB + b) * L * SG_NBQ;
  const double *bs = bandscal + ((size_t)col * NB + b) * SG_NBQ;
  LaneCtx c;
  c.solar = solar;
  c.sec0 = cs[SG_S_SEC0];
  c.ksc_clear = (solar && l_rayleigh) ? spec->band[b].rayleigh : 0.0;
  c.ice = d.ice && !MIXED;
  c.nosc = term_nosc(d.scat, t.iscat);
  c.aer = GEN && aerrec != nullptr;
  LaneSrc src;
  src.aer = c.aer ? aerrec + ((size_t)col * NB + b) * L * 4 : nullptr;
  src.rec = rec;
  src.br = br;
  src.cop = d.mcica ? cldopt + ((size_t)col * NB + b) * L * d.co_stride : nullptr;
  src.co_stride = d.co_stride;
  src.ice = d.ice && !MIXED;
  src.ksc_clear = c.ksc_clear;
  src.csub = c_sub + (size_t)col * L * SG_MAXSUBCOL + (isub > 0 ? isub - 1 : 0);
  src.copf = d.mcica ? (const float *)cldopt + ((size_t)col * NB + b) * L * 4 : nullptr;
  src.csubf = (const float *)c_sub + (size_t)col * L * SG_MAXSUBCOL + (isub > 0 ? isub - 1 : 0);
  src.gas = t.gas;
  src.isub = isub;
  src.fnc = t.fnc;
  src.term = spec->g + gi;
  src.spec = spec;
  src.grec = grec ? grec + (size_t)col * L * SG_NGQ : nullptr;
  const double alb_diff = bs[2], alb_dir = bs[3];

  constexpr int kNst = RECOMP ? (SG_MAXL + kSeg - 1) / kSeg : SG_MAXL;
  double st[kNst][RECOMP ? 3 : 4];
  const double finc = solar ? bs[0] / c.sec0 : -bs[0];  // flux_inc_down, flux_inc_direct
  double fdir = solar ? finc : 0.0;
  double rho = 0.0, dn = finc;
  double vdir[8], o8[8];
  LayerIn cur, nxt;
  load_layer<MIXED, GEN>(src, 0, cur);
  for (int i = 0; i < L; i++) {
    if (solar) {
      // direct flux at the top of layer i, summed 8 levels at a time
      const int m = i & 7;
      double x = active ? wgt * fdir : 0.0;
#pragma unroll
      for (int k = 0; k < 8; k++) if (k == m) vdir[k] = x;
      if (m == 7) {
        for (int sg = 0; sg < nseg; sg++) {
          seg_reduce8(vdir, seg, sg, nseg, o8, lane);
          if ((lane & 7) == 0) out[(sg * 3 + 2) * (L + 1) + i - 7 + (lane >> 3)] = o8[0];
        }
      }
    }
    if (i + 1 < L) load_layer<MIXED, GEN>(src, i + 1, nxt);
    if (active) {
      if (RECOMP) {
        if (i % kSeg == 0) {
          st[i / kSeg][0] = fdir; st[i / kSeg][1] = rho; st[i / kSeg][2] = dn;
        }
      }
      double a, bb;
      const double rho0 = rho, dn0 = dn;
      layer_step<MIXED, GEN>(c, t, cur, fdir, rho, dn, a, bb);
      if (!RECOMP) {
        st[i][0] = a; st[i][1] = bb; st[i][2] = rho0; st[i][3] = dn0;
      }
    }
    cur = nxt;
  }
  if (solar) {
    // remaining direct-flux levels (the partial block and the surface)
    const int base = L & ~7;
    const int m = L & 7;
    double x = active ? wgt * fdir : 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (k == m) vdir[k] = x;
      else if (k > m) vdir[k] = 0.0;
    }
    for (int sg = 0; sg < nseg; sg++) {
      seg_reduce8(vdir, seg, sg, nseg, o8, lane);
      if ((lane & 7) == 0 && base + (lane >> 3) <= L) out[(sg * 3 + 2) * (L + 1) + base + (lane >> 3)] = o8[0];
    }
  } else {
    for (int sg = 0; sg < nseg; sg++)
      for (int i = lane; i <= L; i += kWave) out[(sg * 3 + 2) * (L + 1) + i] = 0.0;
  }
  double fu = 0.0, fd = 0.0;
  if (active) {
    double sgs = solar ? (alb_dir - alb_diff) * fdir
                       : (1.0 - alb_diff) * (bs[1] - br[(L - 1) * SG_NBQ + 3]);
    fu = (alb_diff * dn + sgs) / (1.0 - alb_diff * rho);
    fd = rho * fu + dn;
  }
  if (!RECOMP) {
    // back-substitution, levels L..0, 4 levels (x up/down) per reduction
    for (int top = L; top >= 0; top -= 4) {
      double v[8];
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const int lev = top - m;
        if (lev >= 0 && lev < L && active) {
          fu = st[lev][0] * fu + st[lev][1];
          fd = st[lev][2] * fu + st[lev][3];
        }
        v[2 * m] = lev >= 0 ? wgt * fu : 0.0;
        v[2 * m + 1] = lev >= 0 ? wgt * fd : 0.0;
      }
      for (int sg = 0; sg < nseg; sg++) {
        seg_reduce8(v, seg, sg, nseg, o8, lane);
        if ((lane & 7) == 0) {
          const int j = lane >> 3, lev = top - (j >> 1);
          if (lev >= 0) out[(sg * 3 + (j & 1)) * (L + 1) + lev] = o8[0];
        }
      }
    }
  } else {
    // surface level, then segments bottom-up: recompute the segment's
    // layers from its checkpoint into registers, back-substitute them
    {
      double v[8];
#pragma unroll
      for (int m = 0; m < 4; m++) { v[2 * m] = 0.0; v[2 * m + 1] = 0.0; }
      v[0] = wgt * fu;
      v[1] = wgt * fd;
      for (int sg = 0; sg < nseg; sg++) {
        seg_reduce8(v, seg, sg, nseg, o8, lane);
        if ((lane & 7) == 0 && (lane >> 3) < 2) out[(sg * 3 + (lane >> 3)) * (L + 1) + L] = o8[0];
      }
    }
    const int nsegL = (L + kSeg - 1) / kSeg;
    for (int sgi = nsegL - 1; sgi >= 0; sgi--) {
      const int i0 = sgi * kSeg;
      double sa[kSeg], sb[kSeg], sr[kSeg], sdn[kSeg];
      if (active) {
        double fdir_s = st[sgi][0], rho_s = st[sgi][1], dn_s = st[sgi][2];
#pragma unroll
        for (int k = 0; k < kSeg; k++) {
          const int i = i0 + k;
          sr[k] = rho_s; sdn[k] = dn_s; sa[k] = 0.0; sb[k] = 0.0;
          if (i < L) {
            LayerIn x;
            load_layer<MIXED, GEN>(src, i, x);
            layer_step<MIXED, GEN>(c, t, x, fdir_s, rho_s, dn_s, sa[k], sb[k]);
          }
        }
      }
      double v[8];
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const int k = kSeg - 1 - m;   // levels i0+3 .. i0
        const int lev = i0 + k;
        if (lev < L && active) {
          fu = sa[k] * fu + sb[k];
          fd = sr[k] * fu + sdn[k];
        }
        v[2 * m] = lev < L ? wgt * fu : 0.0;
        v[2 * m + 1] = lev < L ? wgt * fd : 0.0;
      }
      for (int sg = 0; sg < nseg; sg++) {
        seg_reduce8(v, seg, sg, nseg, o8, lane);
        if ((lane & 7) == 0) {
          const int j = lane >> 3, lev = i0 + kSeg - 1 - (j >> 1);
          if (lev < L) out[(sg * 3 + (j & 1)) * (L + 1) + lev] = o8[0];
        }
      }
    }
  }
}

// Sum of the per-wave partials, the Planck flux added back for the
// infra-red (adjust_ir_radiance.F90:186-200, weight_band = 1), output in the
// caller's [level][profile] layout. A column's partials come from the waves
// its items fall in (packed solve: partial[wave][segment], in wave order) or
// from its nwpc waves (triple-column solve: partial[col][chunk]).
__device__ __forceinline__ void finish_body(const CallDims &d, const double *__restrict__ partial,
                                            const double *__restrict__ bandrec,
                                            const double *__restrict__ bandscal, int col, int lev,
                                            double &su, double &sd, double &sx) {
  const int L = d.L;
  su = 0.0; sd = 0.0; sx = 0.0;
  if (d.triple) {
    for (int c = 0; c < d.nwpc; c++) {
      const double *o = partial + ((size_t)col * d.nwpc + c) * 3 * (L + 1);
      su += o[lev];
      sd += o[(L + 1) + lev];
      sx += o[2 * (L + 1) + lev];
    }
  } else {
    const long long IM = d.n_items;
    const long long wa = (long long)col * IM / kWave, wb = ((long long)col * IM + IM - 1) / kWave;
    for (long long w = wa; w <= wb; w++) {
      const int sg = col - (int)(w * kWave / IM);
      const double *o = partial + ((size_t)w * kMaxSeg + sg) * 3 * (L + 1);
      su += o[lev];
      sd += o[(L + 1) + lev];
      sx += o[2 * (L + 1) + lev];
    }
  }
  if (d.isolir == SOCGPU_SOURCE_THERMAL) {
    double bsum = 0.0;
    for (int b = 0; b < d.NB; b++) {
      double bl = lev == 0 ? bandscal[((size_t)col * d.NB + b) * SG_NBQ + 0]
                           : bandrec[(((size_t)col * d.NB + b) * L + lev - 1) * SG_NBQ + 3];
      bsum += bl;
    }
    su += bsum;
    sd += bsum;
  }
}

// Heating rate from flux divergence (socrates_set_diag.F90:140-210), K/s.
__device__ __forceinline__ void heating_body(int P, int L, const double *__restrict__ up,
                                             const double *__restrict__ down,
                                             const double *__restrict__ heat_cap,
                                             double *__restrict__ hr, int idx) {
  if (idx >= P * L) return;
  int col = idx % P, i = idx / P;  // layer i+1 lies between levels i and i+1
  double fdiv = down[(size_t)i * P + col] - down[(size_t)(i + 1) * P + col] +
                up[(size_t)(i + 1) * P + col] - up[(size_t)i * P + col];
  hr[idx] = heat_cap ? fdiv / heat_cap[idx] : fdiv;
}

// ------------------------------------------------------------ triple column
// Maximum-random overlap transfer coefficients of the triple-column
// representation (overlap_coupled.F90:150-453 with full correlation; the
// region fractions are those of the cloud types, aggregate_cloud.F90:
// 106-166) at interface j of one column (0 the top of the atmosphere, L the
// surface), regions 0 clear, 1 stratiform, 2 convective: ov[3a + b] is the
// fraction of the downward flux in region b above the interface that enters
// region a below, ov[9 + 3a + b] the fraction of the upward flux in region b
// below that enters region a above. Interfaces above the call's cloud top
// minus one are not used. ovl[col][L+1][18].
constexpr int kOv = 18;

__device__ __forceinline__ void overlap_body(const CallDims &d, const double *__restrict__ colrec,
                                             const int *__restrict__ krange, double dp_strat,
                                             double dp_conv, double *__restrict__ ovl, int idx) {
  const int P = d.P, L = d.L;
  if (idx >= P * (L + 1)) return;
  const int col = idx % P, j = idx / P;
  const int nct = krange[0] < L ? krange[0] : L;
  if (j < nct - 1) return;
  const double *rec = colrec + (size_t)col * L * SG_NQ;
  double up[3] = {1.0, 0.0, 0.0}, lo[3] = {1.0, 0.0, 0.0};
  if (j >= nct) {  // areas of layer j (1-based), the layer above
    const double *r = rec + (size_t)(j - 1) * SG_NQ;
    up[0] = 1.0 - r[SG_Q_WCLD];
    up[1] = r[SG_Q_WCLD] * r[SG_Q_FRAC1];
    up[2] = r[SG_Q_WCLD] * r[SG_Q_FRAC2];
  }
  if (j < L) {     // areas of layer j+1, the layer below
    const double *r = rec + (size_t)j * SG_NQ;
    lo[0] = 1.0 - r[SG_Q_WCLD];
    lo[1] = r[SG_Q_WCLD] * r[SG_Q_FRAC1];
    lo[2] = r[SG_Q_WCLD] * r[SG_Q_FRAC2];
  }
  const double tol = 100.0 * kEps;
  // correlation of like regions across the interface (overlap_coupled.F90:
  // 190-231): 1 for maximum-random overlap, 0 for random, and for
  // exponential-random exp((p_j - p_j+1) / dp) at interfaces 1 < j < L with
  // dp = 1 Pa for the clear region and the decorrelation pressures of the
  // stratiform and convective regions
  double corr[3] = {1.0, 1.0, 1.0};
  if (d.ioverlap == SOCGPU_OVERLAP_RANDOM) {
    corr[0] = 0.0; corr[1] = 0.0; corr[2] = 0.0;
  } else if (d.ioverlap == SOCGPU_OVERLAP_EXP_RAND && j > 1 && j < L) {
    const double dpl = rec[(size_t)(j - 1) * SG_NQ + SG_Q_P] - rec[(size_t)j * SG_NQ + SG_Q_P];
    corr[0] = exp(dpl / 1.0);
    // (a zero decorrelation pressure gives correlation 0, as exp(-inf) in
    // the reference for layers of distinct pressure)
    corr[1] = dp_strat > 0.0 ? exp(dpl / dp_strat) : 0.0;
    corr[2] = dp_conv > 0.0 ? exp(dpl / dp_conv) : 0.0;
  }
  double ao[3][3];  // ao[k][m]: area of region k above overlapping region m below
  for (int k = 0; k < 3; k++)
    for (int m = 0; m < 3; m++) ao[k][m] = k == m ? corr[k] * fmin(lo[k], up[k]) : 0.0;
  double art = 1.0 - ao[0][0];
  art = art - ao[1][1];
  art = art - ao[2][2];
  double aru[3], arl[3];
  for (int k = 0; k < 3; k++) {
    aru[k] = up[k] - ao[k][k];
    arl[k] = lo[k] - ao[k][k];
  }
  if (art > tol)
    for (int k = 0; k < 3; k++)
      for (int m = 0; m < 3; m++) ao[k][m] = ao[k][m] + aru[k] * arl[m] / art;
  double *o = ovl + ((size_t)col * (L + 1) + j) * kOv;
  for (int b = 0; b < 3; b++)    // region b above
    for (int a = 0; a < 3; a++) o[3 * a + b] = up[b] > tol ? ao[b][a] / up[b] : (a == b ? 1.0 : 0.0);
  for (int b = 0; b < 3; b++)    // region b below
    for (int a = 0; a < 3; a++) o[9 + 3 * a + b] = lo[b] > tol ? ao[a][b] / lo[b] : (a == b ? 1.0 : 0.0);
}

// Two-stream coefficients of one region of one layer for one lane: single
// scattering and delta-rescaling (single_scattering_all.F90,
// rescale_tau_omega.F90:73-80) of the region's grey optics kg (extinction,
// with the equivalent extinction), ks (scattering), g (rescaled asymmetry),
// fw (forward scattering) plus the lane's gas absorption kgas, then
// two_coeff (FP64) or two_coeff_f32 (mixed form, with the absorption formed
// in FP64 before rounding, as in layer_coef_mixed).
template <bool MIXED>
__device__ __forceinline__ void region_coef(const LaneCtx &c, double kgas, double mass, double kg, double ks,
                                            double g, double fw, Coef &o) {
#pragma clang fp contract(on)
  if (!MIXED) {
    double ktot = kg + kgas;
    double tau = ktot * mass;
    double om = fmin(ks / (ktot + kTiny), 1.0 - 32.0 * kEps);
    tau = tau * (1.0 - om * fw);
    om = om * (1.0 - fw) / (1.0 - om * fw);
    two_coeff(c.solar, tau, om, g, c.sec0, o);
  } else {
    const float ktot = (float)(kg + kgas);
    const float kab = (float)((kg - ks) + kgas);
    const float omc_min = (float)(32.0 * kEps);
    float tau = ktot * (float)mass;
    float omc = ktot > 0.0f ? fmaxf(SG_FDIV(kab, ktot), omc_min) : 1.0f;
    float om = 1.0f - omc;
    const float sc = 1.0f - om * (float)fw;
    tau = tau * sc;
    omc = SG_FDIV(omc, sc);
    om = 1.0f - omc;
    CoefF f;
    two_coeff_f32(c.solar, tau, om, omc, (float)g, (float)c.sec0, f);
    coef_from_f32(f, o);
  }
}

// Region coefficients weighted as two_coeff_region.F90:200-445 forms them:
// the type's coefficients times its fraction, summed from zero, times the
// reciprocal of the region fraction (the same fraction here); zero where
// the type is absent.
__device__ __forceinline__ void region_mix(double f, Coef &o) {
  if (f > kTiny) {
    const double inv = 1.0 / f;
    o.T = (0.0 + f * o.T) * inv;
    o.R = (0.0 + f * o.R) * inv;
    o.T0 = (0.0 + f * o.T0) * inv;
    o.s1 = (0.0 + f * o.s1) * inv;
    o.s2 = (0.0 + f * o.s2) * inv;
  } else {
    o.T = 0.0; o.R = 0.0; o.T0 = 0.0; o.s1 = 0.0; o.s2 = 0.0;
  }
}

template <bool MIXED, bool GEN>
__device__ __forceinline__ double lane_kgas(const SgTerm &t, const SgSpec *spec, const double *r,
                                            const double *g) {
  if (GEN && t.fnc != SG_FNC_POWER_LAW) {
    const double k = k_scaled(t, spec, r[SG_Q_LNP], r[SG_Q_P], r[SG_Q_T], g, r[SG_Q_MMR + t.gas]);
    return MIXED ? (double)(float)k : k;
  }
  if (MIXED) return (double)k_term_f32(t, (float)r[SG_Q_LNP], (float)r[SG_Q_LNT], (float)r[SG_Q_MMR + t.gas]);
  return k_term(t, r[SG_Q_LNP], r[SG_Q_LNT], r[SG_Q_MMR + t.gas]);
}

// Coefficients of the three regions of layer i (0-based) for one lane; the
// cloud regions only from the cloud top (1-based nct) down.
template <bool MIXED>
__device__ __forceinline__ void tri_layer_coef(const LaneCtx &c, double kgas, const double *r,
                                               const double *bq, const double *co, const double *ar,
                                               bool cloud_rows, Coef (&o)[3]) {
  const double mass = r[SG_Q_MASS];
  if (c.aer) {
    // clear region with aerosols: moments over the scattering, the phase
    // function delta-rescaled (grey_opt_prop.F90:1360-1440,
    // rescale_phase_fnc.F90:116-124)
    double ks = ar[0], pf = ar[1], fw = ar[2];
    {
#pragma clang fp contract(off)
      if (ks > kTiny) {
        const double inv = 1.0 / ks;
        fw = fw * inv;
        pf = pf * inv;
      }
      pf = (pf - fw) / (1.0 - fw);
    }
    region_coef<MIXED>(c, kgas, mass, bq[0], ks, pf, fw, o[0]);
  } else {
    region_coef<MIXED>(c, kgas, mass, bq[0], c.ksc_clear, 0.0, 0.0, o[0]);
  }
  if (cloud_rows) {
    const double f1 = r[SG_Q_FRAC1], f2 = r[SG_Q_FRAC2];
    if (f1 > kTiny) region_coef<MIXED>(c, kgas, mass, co[0], co[1], co[2], co[3], o[1]);
    region_mix(f1, o[1]);
    if (f2 > kTiny) region_coef<MIXED>(c, kgas, mass, co[4], co[5], co[6], co[7], o[2]);
    region_mix(f2, o[2]);
  }
}

// The triple-column solve of one column for one lane per major k-term
// (triple_column.F90:255-569): region coefficients (two_coeff_region), the
// direct beam carried through the overlap (triple_solar_source.F90:180-412)
// or the area-weighted infra-red sources (triple_column.F90:298-349), the
// surface sources, and solver_triple_hogan (solver_triple_hogan.F90:
// 207-411): an upward elimination of the region albedos and sources coupled
// through the overlap coefficients, then a downward substitution. Three
// sweeps over the layers: down (solar only; the direct flux entering each
// region of each layer, kept in tws), up (coefficients, sources and the
// elimination; per layer and region the substitution pair fu = c1 fd + c2,
// fd' = e1 fd + e2 kept in tws) and down again (fluxes). tws holds kTriQ
// doubles per layer and lane, lane-interleaved: [wave][layer][q][64].
// Output: partial[col][chunk][3][L+1] of this wave's k-term-weighted sums.
constexpr int kTriQ = 12;

template <bool MIXED, bool GEN>
__device__ __forceinline__ void tri_solve_body(const CallDims &d, const SgSpec *__restrict__ spec,
                                               int l_rayleigh, const double *__restrict__ colrec,
                                               const double *__restrict__ grec,
                                               const double *__restrict__ colscal,
                                               const double *__restrict__ bandrec,
                                               const double *__restrict__ bandscal,
                                               const double *__restrict__ cldopt,
                                               const double *__restrict__ ovl,
                                               const int *__restrict__ krange, double *__restrict__ tws,
                                               const double *__restrict__ aerrec,
                                               double *__restrict__ partial, int wave, int lane) {
  const int L = d.L, NB = d.NB;
  const int col = wave / d.nwpc, chunk = wave % d.nwpc;
  if (col >= d.P) return;
  const bool solar = d.isolir == SOCGPU_SOURCE_SOLAR;
  const int item = chunk * kWave + lane;
  const bool active = item < spec->n_g;
  const int gi = active ? item : 0;
  const double wgt = active ? spec->g[gi].w : 0.0;
  const SgTerm t = spec->g[gi];
  const int b = t.band;
  const int nct = krange[0] < L ? krange[0] : L;  // 1-based cloud top of the call
  const double *cs = colscal + (size_t)col * SG_NS;
  const double *rec = colrec + (size_t)col * L * SG_NQ;
  const double *br = bandrec + ((size_t)col * NB + b) * L * SG_NBQ;
  const double *bs = bandscal + ((size_t)col * NB + b) * SG_NBQ;
  const double *cop = cldopt + ((size_t)col * NB + b) * L * 8;
  const double *ov = ovl + (size_t)col * (L + 1) * kOv;
  double *ws = tws + (size_t)wave * L * kTriQ * kWave + lane;
  double *out = partial + ((size_t)col * d.nwpc + chunk) * 3 * (L + 1);
  LaneCtx c;
  c.solar = solar;
  c.sec0 = cs[SG_S_SEC0];
  c.ksc_clear = (solar && l_rayleigh) ? spec->band[b].rayleigh : 0.0;
  c.ice = false;
  c.nosc = 0;
  c.aer = GEN && aerrec != nullptr;
  const double *ab = c.aer ? aerrec + ((size_t)col * NB + b) * L * 4 : nullptr;
  const double alb_diff = bs[2], alb_dir = bs[3];
  const double finc = solar ? bs[0] / c.sec0 : -bs[0];  // flux_inc_down (= flux_inc_direct)
#define TWS(i, q) ws[((size_t)(i) * kTriQ + (q)) * kWave]
  // lanes past the last k-term neither store nor load the workspace (their
  // results carry weight 0; zeros keep them finite)
#define TWL(i, q) (active ? TWS(i, q) : 0.0)

  // --- sweep 1 (solar): direct flux entering each region of each layer
  double base0 = finc, base1 = 0.0, base2 = 0.0;  // direct flux leaving the layer above, per region
  if (solar) {
    double vdir[8];
    for (int i = 0; i < L; i++) {
      {  // direct flux at level i, summed 8 levels at a time
        const int m = i & 7;
        const double x = wgt * (base0 + base1 + base2);
#pragma unroll
        for (int k = 0; k < 8; k++) if (k == m) vdir[k] = x;
        if (m == 7) {
          reduce8(vdir, lane);
          if ((lane & 7) == 0) out[2 * (L + 1) + i - 7 + (lane >> 3)] = vdir[0];
        }
      }
      const double *r = rec + (size_t)i * SG_NQ;
      const double *bq = br + (size_t)i * SG_NBQ;
      const double adj = bq[1];
      const double kgas = lane_kgas<MIXED, GEN>(t, spec, r, grec ? grec + ((size_t)col * L + i) * SG_NGQ : nullptr);
      Coef o[3];
      const bool cloud_rows = i + 1 >= nct;
      tri_layer_coef<MIXED>(c, kgas, r, bq, cop + (size_t)i * 8, ab ? ab + (size_t)i * 4 : nullptr, cloud_rows, o);
      double top0, top1 = 0.0, top2 = 0.0;
      if (!cloud_rows) {
        top0 = base0;
        base0 = top0 * o[0].T0 * adj;
      } else {
        const double *v = ov + (size_t)i * kOv;  // interface at the top of layer i+1
        top0 = v[0] * base0 + v[1] * base1 + v[2] * base2;
        top1 = v[3] * base0 + v[4] * base1 + v[5] * base2;
        top2 = v[6] * base0 + v[7] * base1 + v[8] * base2;
        base0 = top0 * o[0].T0 * adj;
        base1 = top1 * o[1].T0 * adj;
        base2 = top2 * o[2].T0 * adj;
      }
      if (active) TWS(i, 0) = top0;
      if (active) TWS(i, 1) = top1;
      if (active) TWS(i, 2) = top2;
    }
    {  // remaining direct-flux levels (the partial block and the surface)
      const int bl = L & ~7, m = L & 7;
      const double x = wgt * (base0 + base1 + base2);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        if (k == m) vdir[k] = x;
        else if (k > m) vdir[k] = 0.0;
      }
      reduce8(vdir, lane);
      if ((lane & 7) == 0 && bl + (lane >> 3) <= L) out[2 * (L + 1) + bl + (lane >> 3)] = vdir[0];
    }
  } else {
    for (int i = lane; i <= L; i += kWave) out[2 * (L + 1) + i] = 0.0;
  }

  // --- sweep 2: upward elimination (solver_triple_hogan.F90:207-312)
  double a0 = alb_diff, a1 = alb_diff, a2 = alb_diff, g0, g1, g2;
  if (solar) {
    g0 = (alb_dir - alb_diff) * base0;
    g1 = (alb_dir - alb_diff) * base1;
    g2 = (alb_dir - alb_diff) * base2;
  } else {
    const double dps = bs[1] - br[(size_t)(L - 1) * SG_NBQ + 3];
    const double *u = ov + (size_t)L * kOv + 9;
    g0 = u[0] * (1.0 - alb_diff) * dps;
    g1 = u[3] * (1.0 - alb_diff) * dps;
    g2 = u[6] * (1.0 - alb_diff) * dps;
  }
  for (int i = L - 1; i >= 0; i--) {
    const double *r = rec + (size_t)i * SG_NQ;
    const double *bq = br + (size_t)i * SG_NBQ;
    const double kgas = lane_kgas<MIXED, GEN>(t, spec, r, grec ? grec + ((size_t)col * L + i) * SG_NGQ : nullptr);
    const bool cloud_rows = i + 1 >= nct;
    Coef o[3];
    tri_layer_coef<MIXED>(c, kgas, r, bq, cop + (size_t)i * 8, ab ? ab + (size_t)i * 4 : nullptr, cloud_rows, o);
    double su[3], sd[3];
    const int nreg = cloud_rows ? 3 : 1;
    if (solar) {
      const double adj = bq[1];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (k < nreg) {
          const double top = TWL(i, k);
          const double base = top * o[k].T0 * adj;
          su[k] = o[k].s1 * top;
          sd[k] = (o[k].s2 - o[k].T0) * top + base;
        }
      }
    } else {
      const double dpl2 = bq[1], dpl = bq[2];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (k < nreg) {
          su[k] = o[k].s1 * dpl + o[k].s2 * dpl2;
          sd[k] = -o[k].s1 * dpl + o[k].s2 * dpl2;
        }
      }
      if (cloud_rows) {
        const double w = r[SG_Q_WCLD], wf = 1.0 - w;
        const double w1 = w * r[SG_Q_FRAC1], w2 = w * r[SG_Q_FRAC2];
        sd[0] = wf * sd[0]; su[0] = wf * su[0];
        sd[1] = w1 * sd[1]; su[1] = w1 * su[1];
        sd[2] = w2 * sd[2]; su[2] = w2 * su[2];
      }
    }
    if (cloud_rows) {
      const double *vb = ov + (size_t)(i + 1) * kOv;  // interface at the bottom of layer i+1
      const double th[3] = {a0 * vb[0] + a1 * vb[3] + a2 * vb[6], a0 * vb[1] + a1 * vb[4] + a2 * vb[7],
                            a0 * vb[2] + a1 * vb[5] + a2 * vb[8]};
      const double gg[3] = {g0, g1, g2};
      double lsrc[3], aa[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double T = o[k].T, R = o[k].R;
        const double binv = 1.0 / (1.0 - th[k] * R);
        const double gam = th[k] * T;
        const double h = gg[k] + th[k] * sd[k];
        lsrc[k] = su[k] + h * T * binv;
        aa[k] = R + th[k] * T * T * binv;
        const double c1 = gam * binv, c2 = h * binv;
        if (active) TWS(i, 4 * k + 0) = c1;
        if (active) TWS(i, 4 * k + 1) = c2;
        if (active) TWS(i, 4 * k + 2) = T + R * c1;
        if (active) TWS(i, 4 * k + 3) = R * c2 + sd[k];
      }
      const double *ut = ov + (size_t)i * kOv + 9;    // interface at the top of layer i+1
      a0 = aa[0]; a1 = aa[1]; a2 = aa[2];
      g0 = ut[0] * lsrc[0] + ut[1] * lsrc[1] + ut[2] * lsrc[2];
      g1 = ut[3] * lsrc[0] + ut[4] * lsrc[1] + ut[5] * lsrc[2];
      g2 = ut[6] * lsrc[0] + ut[7] * lsrc[1] + ut[8] * lsrc[2];
    } else {
      double th = a0;
      if (i + 1 == nct - 1 && nct < L) {
        const double *vb = ov + (size_t)(i + 1) * kOv;
        th = a0 * vb[0] + a1 * vb[3] + a2 * vb[6];
      }
      const double T = o[0].T, R = o[0].R;
      const double binv = 1.0 / (1.0 - th * R);
      const double gam = th * T;
      const double h = g0 + th * sd[0];
      const double lam = T * binv;
      a0 = R + lam * gam;
      g0 = su[0] + lam * h;
      const double c1 = gam * binv, c2 = h * binv;
      if (active) TWS(i, 0) = c1;
      if (active) TWS(i, 1) = c2;
      if (active) TWS(i, 2) = T + R * c1;
      if (active) TWS(i, 3) = R * c2 + sd[0];
    }
  }

  // --- sweep 3: downward substitution, up and down fluxes at every level,
  // summed over the wave 4 levels at a time
  double v[8];
#pragma unroll
  for (int k = 0; k < 8; k++) v[k] = 0.0;
  auto put = [&](int lev, double fu, double fd) {
    const int m = lev & 3;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (k == m) { v[2 * k] = wgt * fu; v[2 * k + 1] = wgt * fd; }
      else if (k > m && lev == L) { v[2 * k] = 0.0; v[2 * k + 1] = 0.0; }
    }
    if (m == 3 || lev == L) {
      reduce8(v, lane);
      if ((lane & 7) == 0) {
        const int j = lane >> 3, lv = (lev & ~3) + (j >> 1);
        if (lv <= L) out[(j & 1) * (L + 1) + lv] = v[0];
      }
    }
  };
  double fdt = finc;
  {
    const double up0 = nct > 1 ? a0 * finc + g0 : g0 + finc * (ov[0] * a0 + ov[3] * a1 + ov[6] * a2);
    put(0, up0, finc);
  }
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;
  for (int i = 0; i < L; i++) {
    double fu, fd;
    if (i + 1 < nct) {
      fu = TWL(i, 0) * fdt + TWL(i, 1);
      fdt = TWL(i, 2) * fdt + TWL(i, 3);
      fd = fdt;
    } else {
      const double *vt = ov + (size_t)i * kOv;
      if (i + 1 == nct) {
        d0 = vt[0] * fdt; d1 = vt[3] * fdt; d2 = vt[6] * fdt;
      } else {
        const double x0 = d0, x1 = d1, x2 = d2;
        d0 = vt[0] * x0 + vt[1] * x1 + vt[2] * x2;
        d1 = vt[3] * x0 + vt[4] * x1 + vt[5] * x2;
        d2 = vt[6] * x0 + vt[7] * x1 + vt[8] * x2;
      }
      const double u0 = TWL(i, 0) * d0 + TWL(i, 1);
      const double u1 = TWL(i, 4) * d1 + TWL(i, 5);
      const double u2 = TWL(i, 8) * d2 + TWL(i, 9);
      d0 = TWL(i, 2) * d0 + TWL(i, 3);
      d1 = TWL(i, 6) * d1 + TWL(i, 7);
      d2 = TWL(i, 10) * d2 + TWL(i, 11);
      fu = u0 + u1 + u2;
      fd = d0 + d1 + d2;
      fdt = fd;
    }
    put(i + 1, fu, fd);
  }
#undef TWS
#undef TWL
}

// ------------------------------------------------------------ batches
// A batch is a list of radiation calls ("requests"), each with its own
// spectrum, source, shape and options, packed into one launch sequence:
// every kernel loops over the concatenated work of all requests and finds
// the request of a work item by binary search over per-request starts. A
// request's results are those of the same call made alone.
// W_MINOR: waves of the per-(column, band) kernel; W_LONG counts those of
// calls with more than 128 layers (it only decides whether their kernel is
// launched).
// W_GEN counts the calls whose spectrum needs the general gas optics (any
// scaling other than the power law, continua or further gases); a batch with
// any such call runs the kernels built for it
// W_PLAIN counts the calls that do not: the solve kernel of each build runs
// only the calls of its kind, so a call's solve is the one it gets alone.
enum { W_LEV, W_LAY, W_CG, W_MINOR, W_PREP, W_SOLVE, W_TOVL, W_TSOLVE, W_PTILE, W_CCOL, W_FTILE, W_LONG, W_GEN,
       W_PLAIN, W_GBAND, kNWork };

struct ReqDesc {
  CallDims d;
  const SgSpec *spec;
  const double *xcw;
  int n1, n2, l_rayleigh, pad;
  double dp_cf, dp_cw;
  SocgpuRunesArgs a;            // device input pointers
  double *up, *dn, *dir, *hr;   // device outputs (dir may be null)
  double *colrec, *tlev, *colscal, *bandrec, *bandscal, *csub, *grec, *cldopt, *partial;
  double *ovl, *tws;            // triple column: overlap coefficients, solve workspace
  double *aerrec;               // modal aerosols: clear-air scattering, moment products
  int *krange;
  long long s[kNWork];          // first work item of this request in each space
};

// err: 0, or 1 + the (descriptor-order) index of a request whose input the
// kernels found invalid (a cloud fraction above 1.001); the host writes 0
// with every descriptor upload and reads it back after the launch.
struct BatchHdr {
  int R, err;
  long long tot[kNWork];
};

// A call's arguments with every input pointer seen as global (G).
__device__ __forceinline__ SocgpuRunesArgs gargs(const SocgpuRunesArgs &a) {
  SocgpuRunesArgs g = a;
  g.p_layer = G(a.p_layer); g.t_layer = G(a.t_layer); g.mass = G(a.mass); g.density = G(a.density);
  g.layer_heat_capacity = G(a.layer_heat_capacity); g.h2o = G(a.h2o); g.o3 = G(a.o3);
  g.p_level = G(a.p_level); g.t_level = G(a.t_level); g.t_ground = G(a.t_ground);
  g.cos_zenith_angle = G(a.cos_zenith_angle); g.solar_irrad = G(a.solar_irrad);
  g.albedo_diff = G(a.albedo_diff); g.albedo_dir = G(a.albedo_dir);
  g.cloud_frac = G(a.cloud_frac); g.liq_mmr = G(a.liq_mmr); g.liq_dim = G(a.liq_dim);
  g.conv_frac = G(a.conv_frac); g.liq_conv_mmr = G(a.liq_conv_mmr); g.liq_conv_dim = G(a.liq_conv_dim);
  g.rand_seed = G(a.rand_seed); g.flux_ground = G(a.flux_ground); g.co2 = G(a.co2);
  g.ice_mmr = G(a.ice_mmr); g.ice_dim = G(a.ice_dim); g.ice_conv_mmr = G(a.ice_conv_mmr);
  g.ice_conv_dim = G(a.ice_conv_dim);
  g.aer_mix_ratio = G(a.aer_mix_ratio); g.aer_absorption = G(a.aer_absorption);
  g.aer_scattering = G(a.aer_scattering); g.aer_asymmetry = G(a.aer_asymmetry);
  return g;
}

__device__ __forceinline__ int find_req(const BatchHdr *h, const ReqDesc *rq, int w, long long g) {
  int lo = 0, hi = h->R - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (rq[mid].s[w] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void b_init_kernel(const BatchHdr *h, const ReqDesc *rq) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= h->R) return;
  rq[r].krange[0] = rq[r].d.L + 1;
  rq[r].krange[1] = 0;
}

// Caller arrays ([layer][profile]) to column records ([col][layer][SG_NQ]):
// one tile of kPackC columns x kPackL levels per block and iteration, read
// along the profiles, transposed through LDS and written as runs of
// kPackL records per column.
constexpr int kPackC = 64, kPackL = 4;

// Modal aerosols of one (column, layer), every band (grey_opt_prop.F90:
// 716-743, opt_prop_ukca_aerosol.F90:135-200, two-stream: one moment,
// forward scattering of order 2): clear-air scattering (Rayleigh, then the
// modes'), the products of the modes' scattering with the asymmetry and
// with its square, and their absorption, into aerrec[col][band][layer][4].
// The threads of a wave take consecutive profiles, so the caller's
// [band][mode][layer][profile] fields are read along the profiles.
__device__ __forceinline__ void aer_sums(const CallDims &d, const SgSpec *spec, int l_rayleigh, const AerIn &ae,
                                         int col, int i) {
#pragma clang fp contract(off)
  const int L = d.L, M = d.n_aer;
  const bool solar = d.isolir == SOCGPU_SOURCE_SOLAR;
  for (int b = 0; b < d.NB; b++) {
    double ksc = (solar && l_rayleigh) ? spec->band[b].rayleigh : 0.0, kaer = 0.0, pf = 0.0, fw = 0.0;
    for (int m = 0; m < M; m++) {
      const size_t im = ((size_t)m * L + i) * d.P + col;
      const size_t ib = (((size_t)b * M + m) * L + i) * d.P + col;
      const double mix = ae.mix[im];
      kaer = kaer + mix * ae.abs[ib];
      const double ks = mix * ae.sca[ib];
      ksc = ksc + ks;
      const double g = ae.asy[ib];
      pf = pf + ks * g;
      const double ks_phf = ks * g;
      fw = fw + ks_phf * g;
    }
    double *ar = ae.rec + (((size_t)col * d.NB + b) * L + i) * 4;
    ar[0] = ksc; ar[1] = pf; ar[2] = fw; ar[3] = kaer;
  }
}

__host__ __device__ inline int pack_tiles(const CallDims &d) {
  return ((d.P + kPackC - 1) / kPackC) * ((d.L + 1 + kPackL - 1) / kPackL);
}

__global__ __launch_bounds__(kPackC * kPackL) void b_pack_kernel(const BatchHdr *h, const ReqDesc *rq) {
  __shared__ double sh[kPackC * kPackL * SG_NQ];
  __shared__ double shl[kPackC * kPackL];
  const long long tot = h->tot[W_PTILE];
  const int tx = threadIdx.x % kPackC, ty = threadIdx.x / kPackC;
  for (long long tile = blockIdx.x; tile < tot; tile += gridDim.x) {
    const int r = __builtin_amdgcn_readfirstlane(find_req(h, rq, W_PTILE, tile));
    const ReqDesc &q = rq[r];
    const SocgpuRunesArgs a = gargs(q.a);
    const SgSpec *spec = G(q.spec);
    double *colrec = G(q.colrec), *grec = G(q.grec), *tlev = G(q.tlev);
    const int P = q.d.P, L = q.d.L;
    const int t = (int)(tile - q.s[W_PTILE]);
    const int ntc = (P + kPackC - 1) / kPackC;
    const int c0 = (t % ntc) * kPackC, i0 = (t / ntc) * kPackL;
    const int col = c0 + tx, i = i0 + ty;
    if (col < P && i < L) {
      if (pack_layer(q.d, spec, a, i * P + col, q.dp_cf, q.dp_cw, sh + (tx * kPackL + ty) * SG_NQ))
        atomicMax(&((BatchHdr *)h)->err, r + 1);
      if (q.d.n_aer > 0) {
        const AerIn ae{a.aer_mix_ratio, a.aer_absorption, a.aer_scattering, a.aer_asymmetry, G(q.aerrec)};
        aer_sums(q.d, spec, q.l_rayleigh, ae, col, i);
      }
      if (q.d.gx) {
        // gas record: look-up factors and the density (continua)
        double *g = grec + ((size_t)col * L + i) * SG_NGQ;
        const int idx = i * P + col;
        if (spec->has_lut) lookup_factors(spec, a.p_layer[idx], a.t_layer[idx], g);
        g[SG_G_RHO] = a.density ? a.density[idx] : 0.0;
      }
    }
    if (col < P && i <= L) shl[tx * kPackL + ty] = a.t_level ? a.t_level[(size_t)i * P + col] : 0.0;
    if (col < P && i == 0) pack_column(q.d, a, col, G(q.colscal), G(q.bandscal));
    __syncthreads();
    const int nl = L - i0 < kPackL ? L - i0 : kPackL;
    if (nl > 0)
      for (int e = threadIdx.x; e < kPackC * kPackL * SG_NQ; e += kPackC * kPackL) {
        const int cc = e / (kPackL * SG_NQ), off = e - cc * (kPackL * SG_NQ);
        if (c0 + cc < P && off < nl * SG_NQ)
          colrec[((size_t)(c0 + cc) * L + i0) * SG_NQ + off] = sh[cc * kPackL * SG_NQ + off];
      }
    {
      const int e = threadIdx.x, cc = e / kPackL, off = e % kPackL;
      if (c0 + cc < P && i0 + off <= L) tlev[(size_t)(c0 + cc) * (L + 1) + i0 + off] = shl[e];
    }
    __syncthreads();
  }
}

// Cloud top and base over all columns of each request, as cloud_gen finds
// them (socrates_cloud_gen.F90:224-245): first layer with any w > epsilon,
// last layer with any w >= epsilon; krange[0] = min, krange[1] = max
// (1-based). For the triple-column representation krange[0] is also the
// call's cloud top (set_cloud_geometry.F90:93-99: first layer with any
// w > 0, L when there is none, applied by the callers). w is 0 or above
// the 0.001 threshold, so both tests read "the layer has cloud", evaluated
// here on the caller's fractions exactly as the pack kernel forms w. One
// thread per (column, chunk of kRangeChunk layers), reading along the
// profiles; one atomic per wave where the wave lies within one request.
constexpr int kRangeChunk = 8;
__device__ __forceinline__ bool layer_cloudy(const CallDims &d, const SocgpuRunesArgs &a, size_t idx) {
  double w = fmax(a.cloud_frac[idx], 0.0);
  if (d.conv) w = w + (a.conv_frac ? fmax(a.conv_frac[idx], 0.0) : 0.0);
  return w > 0.001;
}

__global__ void b_cldrange_kernel(const BatchHdr *h, const ReqDesc *rq) {
  const long long tot = h->tot[W_CCOL];
  const int lane = threadIdx.x & (kWave - 1);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long base = (blockIdx.x * (long long)blockDim.x + threadIdx.x) - lane; base < tot;
       base += stride) {
    const long long g = base + lane;
    int r = -1, kmin = INT32_MAX, kmax = 0;
    if (g < tot) {
      r = find_req(h, rq, W_CCOL, g);
      const ReqDesc &q = rq[r];
      const SocgpuRunesArgs a = gargs(q.a);
      const int idx = (int)(g - q.s[W_CCOL]), P = q.d.P, L = q.d.L;
      const int col = idx % P, i0 = (idx / P) * kRangeChunk;
      const int i1 = i0 + kRangeChunk < L ? i0 + kRangeChunk : L;
      if (q.d.raw_range) {
        for (int i = i0; i < i1; i++) {
          const double w = a.cloud_frac[(size_t)i * P + col];
          if (w > kEps && kmin == INT32_MAX) kmin = i + 1;
          if (w >= kEps) kmax = i + 1;
        }
      } else {
        for (int i = i0; i < i1; i++) {
          if (layer_cloudy(q.d, a, (size_t)i * P + col)) {
            kmin = kmin < i + 1 ? kmin : i + 1;
            kmax = i + 1;
          }
        }
      }
    }
    const int r0 = __shfl(r, 0, kWave);
    if (__all(r == r0 || r < 0)) {
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        kmin = min(kmin, __shfl_xor(kmin, o, kWave));
        kmax = max(kmax, __shfl_xor(kmax, o, kWave));
      }
      if (lane == 0 && r0 >= 0) {
        if (kmin != INT32_MAX) atomicMin(&rq[r0].krange[0], kmin);
        if (kmax != 0) atomicMax(&rq[r0].krange[1], kmax);
      }
    } else if (r >= 0) {
      if (kmin != INT32_MAX) atomicMin(&rq[r].krange[0], kmin);
      if (kmax != 0) atomicMax(&rq[r].krange[1], kmax);
    }
  }
}

// kCgWaves columns per block and iteration, one wave each; the condensate
// table (xcw, n1 x n2 <= kXcwMax) of the block's first column is kept in
// LDS and reloaded only when it changes; a wave whose column uses another
// table reads it from global memory.
constexpr int kCgWaves = 4, kXcwMax = 4096;

__global__ __launch_bounds__(64 * kCgWaves) void b_cloudgen_kernel(const BatchHdr *h, const ReqDesc *rq) {
  __shared__ double sx[kXcwMax];
  const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const long long tot = h->tot[W_CG];
  const double *loaded = nullptr;
  for (long long base = (long long)blockIdx.x * kCgWaves; base < tot; base += (long long)gridDim.x * kCgWaves) {
    const ReqDesc &q0 = rq[__builtin_amdgcn_readfirstlane(find_req(h, rq, W_CG, base))];
    const int nx0 = q0.n1 * q0.n2;
    if (q0.xcw != loaded && nx0 <= kXcwMax) {
      const double *x0 = G(q0.xcw);
      for (int e = threadIdx.x; e < nx0; e += blockDim.x) sx[e] = x0[e];
      loaded = q0.xcw;
    }
    __syncthreads();
    const long long w = base + wv;
    if (w < tot) {
      const ReqDesc &q = rq[__builtin_amdgcn_readfirstlane(find_req(h, rq, W_CG, w))];
      const double *tab = (q.xcw == loaded && q.n1 * q.n2 <= kXcwMax) ? sx : q.xcw;
      cloudgen_body(q.d, G(q.colrec), G(q.a.rand_seed), q.dp_cf, q.dp_cw, G(q.krange), tab, q.n1, q.n2,
                    G(q.csub), G(q.colscal), (int)(w - q.s[W_CG]), lane);
    }
    __syncthreads();
  }
}

// waves per SIMD the band kernel is built for (calls of up to 128 layers)
#ifndef SG_BAND_OCC
#define SG_BAND_OCC 3
#endif

template <int NCH, bool GEN>
__global__ __launch_bounds__(64 * kMinorWaves, NCH == 2 ? SG_BAND_OCC : 2) void b_band_kernel(const BatchHdr *h, const ReqDesc *rq) {
  const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const long long tot = h->tot[W_MINOR];
  for (long long w = blockIdx.x * (long long)kMinorWaves + wv; w < tot; w += (long long)gridDim.x * kMinorWaves) {
    const int r = __builtin_amdgcn_readfirstlane(find_req(h, rq, W_MINOR, w));
    const ReqDesc &q = rq[r];
    const int nch = (q.d.L + kWave - 1) / kWave;
    if ((NCH == 2 && nch <= 2) || (NCH > 2 && nch > 2))
    {
      const AerIn ae{G(q.a.aer_mix_ratio), G(q.a.aer_absorption), G(q.a.aer_scattering), G(q.a.aer_asymmetry),
                     G(q.aerrec)};
      band_body<NCH, GEN>(q.d, G(q.spec), q.l_rayleigh, G(q.colrec), G(q.grec), q.a.gas_mix_ratio, q.a.gas_field,
                          G(q.tlev), G(q.colscal), G(q.bandrec), G(q.bandscal), G(q.cldopt), G(q.krange), ae,
                          (int)(w - q.s[W_MINOR]), lane);
    }
  }
}


template <bool MIXED, bool RECOMP, int OCC, bool GEN>
__global__ __launch_bounds__(256, OCC) void b_solve_kernel(const BatchHdr *h, const ReqDesc *rq) {
  const long long tot = h->tot[W_SOLVE];
  const int lane = threadIdx.x % kWave;
  const long long nw = (long long)gridDim.x * (blockDim.x / kWave);
  for (long long w = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / kWave; w < tot; w += nw) {
    const int r = __builtin_amdgcn_readfirstlane(find_req(h, rq, W_SOLVE, w));
    const ReqDesc &q = rq[r];
    if ((q.d.gen != 0) != GEN) continue;   // the other build's call
    if (MIXED && GEN)
      solve_body_items<MIXED, RECOMP, GEN>(q.d, G(q.spec), q.l_rayleigh, G(q.colrec), G(q.grec), G(q.colscal),
                                           G(q.bandrec), G(q.bandscal), G(q.csub), G(q.cldopt), G(q.aerrec),
                                           G(q.partial),
                                           (int)(w - q.s[W_SOLVE]), lane);
    else
      solve_body<MIXED, RECOMP, GEN>(q.d, G(q.spec), q.l_rayleigh, G(q.colrec), G(q.grec), G(q.colscal),
                                     G(q.bandrec), G(q.bandscal), G(q.csub), G(q.cldopt), G(q.aerrec),
                                     G(q.partial),
                                     (int)(w - q.s[W_SOLVE]), lane);
  }
}

// Outputs in tiles of kFinC columns x kFinL levels per block and iteration:
// partials read along the levels of a column, transposed through LDS,
// written along the profiles of a level.
constexpr int kFinC = 16, kFinL = 16;

__host__ __device__ inline int finish_tiles(const CallDims &d) {
  return ((d.P + kFinC - 1) / kFinC) * ((d.L + 1 + kFinL - 1) / kFinL);
}

__global__ __launch_bounds__(kFinC * kFinL) void b_finish_kernel(const BatchHdr *h, const ReqDesc *rq) {
  __shared__ double sh[3][kFinC][kFinL + 1];
  const long long tot = h->tot[W_FTILE];
  const int ta = threadIdx.x % kFinL, tb = threadIdx.x / kFinL;
  for (long long tile = blockIdx.x; tile < tot; tile += gridDim.x) {
    const int r = __builtin_amdgcn_readfirstlane(find_req(h, rq, W_FTILE, tile));
    const ReqDesc &q = rq[r];
    const int P = q.d.P, L = q.d.L;
    const int t = (int)(tile - q.s[W_FTILE]);
    const int ntc = (P + kFinC - 1) / kFinC;
    const int c0 = (t % ntc) * kFinC, l0 = (t / ntc) * kFinL;
    {
      const int col = c0 + tb, lev = l0 + ta;   // levels fastest
      if (col < P && lev <= L)
        finish_body(q.d, G(q.partial), G(q.bandrec), G(q.bandscal), col, lev, sh[0][tb][ta], sh[1][tb][ta],
                    sh[2][tb][ta]);
    }
    __syncthreads();
    {
      const int col = c0 + ta, lev = l0 + tb;   // profiles fastest
      if (col < P && lev <= L) {
        const size_t idx = (size_t)lev * P + col;
        G(q.up)[idx] = sh[0][ta][tb];
        G(q.dn)[idx] = sh[1][ta][tb];
        if (q.dir) G(q.dir)[idx] = sh[2][ta][tb];
      }
    }
    __syncthreads();
  }
}

__global__ void b_overlap_kernel(const BatchHdr *h, const ReqDesc *rq) {
  const long long tot = h->tot[W_TOVL];
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < tot;
       g += (long long)gridDim.x * blockDim.x) {
    const ReqDesc &q = rq[find_req(h, rq, W_TOVL, g)];
    overlap_body(q.d, G(q.colrec), G(q.krange), q.a.cloud_vertical_decorr, q.a.conv_vertical_decorr, G(q.ovl),
                 (int)(g - q.s[W_TOVL]));
  }
}

template <bool MIXED, bool GEN>
__global__ __launch_bounds__(256) void b_tsolve_kernel(const BatchHdr *h, const ReqDesc *rq) {
  const long long tot = h->tot[W_TSOLVE];
  const int lane = threadIdx.x % kWave;
  const long long nw = (long long)gridDim.x * (blockDim.x / kWave);
  for (long long w = (blockIdx.x * (long long)blockDim.x + threadIdx.x) / kWave; w < tot; w += nw) {
    const int r = __builtin_amdgcn_readfirstlane(find_req(h, rq, W_TSOLVE, w));
    const ReqDesc &q = rq[r];
    tri_solve_body<MIXED, GEN>(q.d, G(q.spec), q.l_rayleigh, G(q.colrec), G(q.grec), G(q.colscal), G(q.bandrec),
                               G(q.bandscal), G(q.cldopt), G(q.ovl), G(q.krange), G(q.tws), G(q.aerrec),
                               G(q.partial), (int)(w - q.s[W_TSOLVE]), lane);
  }
}

__global__ void b_heating_kernel(const BatchHdr *h, const ReqDesc *rq) {
  const long long tot = h->tot[W_LAY];
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < tot;
       g += (long long)gridDim.x * blockDim.x) {
    const ReqDesc &q = rq[find_req(h, rq, W_LAY, g)];
    if (q.hr)
      heating_body(q.d.P, q.d.L, G(q.up), G(q.dn), G(q.a.layer_heat_capacity), G(q.hr), (int)(g - q.s[W_LAY]));
  }
}

// Kernel-level two-stream solve on prescribed optical properties, one lane
// per (term, profile) (profile fastest, so every per-layer read and
// per-level write is coalesced), with the packed solve's per-layer
// functions (two_coeff or two_coeff_f32 with coef_from_f32, elim_step):
// two_coeff, the infra-red (ir_source.F90:
// 93-120) or solar (solar_source.F90:104-162) sources, the surface source
// (column_solver.F90:160-185) and the column solve (solver_homogen_direct.F90
// :93-127 as a top-down elimination and bottom-up back-substitution).
// Arrays [term][layer][profile] (optical properties, outputs [term][level]
// [profile]) and [layer][profile] / [profile] for the rest.
struct TwoStreamArgs {
  int P, L, G, isolir, l_scale_solar, l_quad;
  const double *tau, *omega, *asym, *sec0, *finc_down, *finc_dir, *alb_diff, *alb_dir;
  const double *dpl, *dpl2, *dsurf, *adjust;
  double *fdir, *fup, *fdown;
};

template <bool MIXED>
__global__ __launch_bounds__(256) void two_stream_kernel(TwoStreamArgs a) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.P * a.G) return;
  const int P = a.P, L = a.L;
  const int l = idx % P, g = idx / P;
  const bool solar = a.isolir == SOCGPU_SOURCE_SOLAR;
  const double sec0 = solar ? a.sec0[l] : 1.0;
  const size_t t0 = (size_t)g * L * P + l;
  const size_t o0 = (size_t)g * (L + 1) * P + l;
  double st[SG_MAXL][4];
  double fd = solar ? a.finc_dir[l] : 0.0;
  if (solar && a.fdir) a.fdir[o0] = fd;
  double rho = 0.0, dn = a.finc_down[l];
  for (int i = 0; i < L; i++) {
    const size_t k = t0 + (size_t)i * P;
    const size_t kl = (size_t)i * P + l;
    const double tau = a.tau[k], om = a.omega[k], asym = a.asym[k];
    Coef o;
    if (MIXED) {
      CoefF f;
      two_coeff_f32(solar, (float)tau, (float)om, (float)fmax(1.0 - om, 32.0 * kEps), (float)asym,
                    (float)sec0, f);
      coef_from_f32(f, o);
    } else {
      two_coeff(solar, tau, om, asym, sec0, o);
    }
    double su, sd;
    if (solar) {
      if (a.l_scale_solar) {
        double fnew = fd * o.T0 * a.adjust[kl];
        su = o.s1 * fd;
        sd = (o.s2 - o.T0) * fd + fnew;
        fd = fnew;
      } else {
        su = o.s1 * fd;
        sd = o.s2 * fd;
        fd = fd * o.T0;
      }
      if (a.fdir) a.fdir[o0 + (size_t)(i + 1) * P] = fd;
    } else if (a.l_quad) {
      su = o.s1 * a.dpl[kl] + o.s2 * a.dpl2[kl];
      sd = -o.s1 * a.dpl[kl] + o.s2 * a.dpl2[kl];
    } else {
      su = o.s1 * a.dpl[kl];
      sd = -su;
    }
    double aa, bb;
    const double rho0 = rho, dn0 = dn;
    elim_step<MIXED>(o, su, sd, rho, dn, aa, bb);
    st[i][0] = aa; st[i][1] = bb; st[i][2] = rho0; st[i][3] = dn0;
  }
  const double ad = a.alb_diff[l];
  const double sg = solar ? (a.alb_dir[l] - ad) * fd : (1.0 - ad) * a.dsurf[l];
  double fu = (ad * dn + sg) / (1.0 - ad * rho);
  double fdn = rho * fu + dn;
  a.fup[o0 + (size_t)L * P] = fu;
  a.fdown[o0 + (size_t)L * P] = fdn;
  for (int i = L - 1; i >= 0; i--) {
    fu = st[i][0] * fu + st[i][1];
    fdn = st[i][2] * fu + st[i][3];
    a.fup[o0 + (size_t)i * P] = fu;
    a.fdown[o0 + (size_t)i * P] = fdn;
  }
}

// ---------------------------------------------------------------- host side
// Device and pinned host buffers grow by at least half their size when they
// have to grow (freeing either synchronises the device, so growth should
// be rare in a long run) and are kept until released.
inline size_t grown(size_t bytes, size_t n) { return bytes > n + n / 2 ? bytes : n + n / 2; }

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  int ensure(size_t bytes) {
    if (bytes <= n) return 0;
    const size_t want = grown(bytes, n);
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, want) != hipSuccess) return -1;
    n = want;
    return 0;
  }
  template <class T> T *as() { return (T *)p; }
  ~DevBuf() {}
};

// Pinned host staging for small transfers: a call's (or batch's) host
// inputs of at most kPinnedMax bytes are gathered here and cross PCIe in one
// copy, and its outputs come back in one copy and are scattered to the
// caller's arrays. Larger transfers are bandwidth-bound and go array by
// array from the caller's memory (pinning hundreds of MB would cost more than
// it saves). use() gives the buffer or nullptr.
constexpr size_t kPinnedMax = 32u << 20;

struct HostBuf {
  char *p = nullptr;
  size_t n = 0;
  char *use(size_t bytes) {
    if (bytes == 0 || bytes > kPinnedMax) return nullptr;
    return ensure(bytes) == 0 ? p : nullptr;
  }
  int ensure(size_t bytes) {
    if (bytes <= n) return 0;
    const size_t want = grown(bytes, n);
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    if (hipHostMalloc((void **)&p, want, hipHostMallocDefault) != hipSuccess) return -1;
    n = want;
    return 0;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
};

struct SpecEntry {
  SgSpec host;
  SgSpec *dev = nullptr;
  double *tables = nullptr;  // look-up k-tables and Wenyi tables on the device
  double *xcw = nullptr;  // McICA condensate table on the device
  bool solar_hint = false;
};

// Fixed device workspace of the allocation-free entry point
// (socgpu_runes_noalloc): one call's descriptor followed by its workspace.
// Every GPU the library runs on holds its own copy until the process ends.
// In host code `scratch` only names the symbol: its device address comes
// from hipGetSymbolAddress at socgpu_init (Context::scratch_dev), and that
// address is what kernels, copies and fills take.
constexpr size_t kScratchBytes = SOCGPU_NOALLOC_MAX_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

// A second stream beside a launch stream, and the events that fork the cloud
// generator onto it after the cloud range is known and join it back before
// the solve: the generator (one wave per column) and the band kernel (one
// wave per column and band) depend on the same inputs but not on each
// other, and both spend most of their time waiting on memory, so they share
// the device. Streams are created with the default priority.
struct Fork {
  hipStream_t side = nullptr;
  hipEvent_t f = nullptr, j = nullptr;
  bool make() {
    return hipStreamCreateWithFlags(&side, hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&f, hipEventDisableTiming) == hipSuccess &&
           hipEventCreateWithFlags(&j, hipEventDisableTiming) == hipSuccess;
  }
  void drop() {
    if (f) (void)hipEventDestroy(f);
    if (j) (void)hipEventDestroy(j);
    if (side) (void)hipStreamDestroy(side);
    f = j = nullptr;
    side = nullptr;
  }
};

struct Context {
  int device = -1;
  hipStream_t stream = nullptr;
  Fork fork;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  std::map<std::string, std::unique_ptr<SpecEntry>> spectra;
  DevBuf ws, desc;
  DevBuf in_stage, out_stage;
  HostBuf in_host, out_host;
  double last_total_ms = 0.0, last_solve_ms = 0.0;
  unsigned char *scratch_dev = nullptr;  // device address of `scratch` on `device`
  int *err_host = nullptr;               // pinned word for the noalloc call's error flag
  std::mutex mu;
};
static Context g_ctx;
// solve-kernel tile per precision form ([0] FP64, [1] mixed); the defaults
// are the autotuner's picks on the 4096-column x 90-layer all-sky SW + LW
// batch
static SocgpuTile g_tile[2] = {{0, 1, 3}, {1, 1, 3}};

static int ensure_init() {
  if (g_ctx.device >= 0) return 0;
  return fail(SOCGPU_EINVAL, "socgpu_init has not been called");
}

}  // namespace sg

using namespace sg;

extern "C" {

const char *socgpu_last_error(void) { return g_err.c_str(); }

int socgpu_init(int device) {
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  if (g_ctx.device >= 0) return g_ctx.device == device ? 0 : fail(SOCGPU_EINVAL, "already initialised on device %d", g_ctx.device);
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(SOCGPU_EINVAL, "device %d not present (%d devices)", device, n);
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&g_ctx.ev0));
  HIPCHK(hipEventCreate(&g_ctx.ev1));
  HIPCHK(hipEventCreate(&g_ctx.ev2));
  HIPCHK(hipEventCreate(&g_ctx.ev3));
  if (!g_ctx.fork.make()) return fail(SOCGPU_EDEVICE, "stream or event creation failed");
  // the allocation-free entry point's workspace on this device, and a pinned
  // word its error flag comes back in (allocated here, once, so that the
  // entry point itself allocates nothing)
  void *sp = nullptr;
  HIPCHK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  g_ctx.scratch_dev = (unsigned char *)sp;
  HIPCHK(hipHostMalloc((void **)&g_ctx.err_host, 256, hipHostMallocDefault));
  g_ctx.device = device;
  return 0;
}

}  // extern "C"

static int register_spectrum(const char *name, const char *path, const std::vector<int> *retain) {
  if (int rc = ensure_init()) return rc;
  if (!name || !path) return fail(SOCGPU_EINVAL, "spectrum name and path are required");
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  auto e = std::make_unique<SpecEntry>();
  std::string err;
  SpecTables tab;
  int rc = read_spectrum(path, &e->host, &tab, &err, retain);
  if (rc) return fail(rc == -3 ? SOCGPU_EUNSUPPORTED : SOCGPU_EIO, "%s: %s", path, err.c_str());
  if (!tab.data.empty()) {
    HIPCHK(hipMalloc(&e->tables, tab.data.size() * sizeof(double)));
    HIPCHK(hipMemcpy(e->tables, tab.data.data(), tab.data.size() * sizeof(double), hipMemcpyHostToDevice));
    e->host.klut = e->host.has_lut ? e->tables : nullptr;
    e->host.wenyi = tab.wenyi_off >= 0 ? e->tables + tab.wenyi_off : nullptr;
  }
  HIPCHK(hipMalloc(&e->dev, sizeof(SgSpec)));
  HIPCHK(hipMemcpy(e->dev, &e->host, sizeof(SgSpec), hipMemcpyHostToDevice));
  auto it = g_ctx.spectra.find(name);
  if (it != g_ctx.spectra.end()) {
    // the replaced entry's tables; its condensate table may be shared with
    // another spectrum (socgpu_mcica_register) and is freed with its last user
    double *xcw = it->second->xcw;
    (void)hipFree(it->second->dev);
    if (it->second->tables) (void)hipFree(it->second->tables);
    it->second->xcw = nullptr;
    bool shared = false;
    for (auto &kv : g_ctx.spectra) shared |= kv.second->xcw == xcw;
    if (xcw && !shared) (void)hipFree(xcw);
  }
  g_ctx.spectra[name] = std::move(e);
  return 0;
}

extern "C" {

int socgpu_spectrum_register(const char *name, const char *path) {
  return register_spectrum(name, path, nullptr);
}

int socgpu_spectrum_register_gases(const char *name, const char *path, int n_type, const int *type_absorb) {
  if (n_type < 0) return register_spectrum(name, path, nullptr);
  if (n_type > 0 && !type_absorb) return fail(SOCGPU_EINVAL, "null absorber list");
  const std::vector<int> keep(type_absorb, type_absorb + n_type);
  return register_spectrum(name, path, &keep);
}

int socgpu_spectrum_query(const char *name, int *n_band, int *n_absorb, int *type_absorb,
                          int max_absorb) {
  if (!name) return fail(SOCGPU_EINVAL, "spectrum name is required");
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  auto it = g_ctx.spectra.find(name);
  if (it == g_ctx.spectra.end()) return fail(SOCGPU_ENOTFOUND, "Spectrum name not found: %s", name);
  const SgSpec &S = it->second->host;
  if (n_band) *n_band = S.n_band;
  if (n_absorb) *n_absorb = S.n_absorb;
  if (type_absorb)
    for (int i = 0; i < S.n_absorb && i < max_absorb; i++) type_absorb[i] = S.type_absorb[i];
  return 0;
}

int socgpu_mcica_register(const char *path, const char *sw_name, const char *lw_name) {
  if (int rc = ensure_init()) return rc;
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  McicaData m;
  std::string err;
  if (read_mcica(path, &m, &err)) return fail(SOCGPU_EIO, "%s: %s", path, err.c_str());
  const char *names[2] = {sw_name, lw_name};
  for (int k = 0; k < 2; k++)
    if (names[k] && g_ctx.spectra.find(names[k]) == g_ctx.spectra.end())
      return fail(SOCGPU_ENOTFOUND, "spectrum '%s' not registered", names[k]);
  // one device copy of the condensate table, shared by the spectra it is
  // associated with (the cloud generator keeps one table per block in LDS)
  double *xcw = nullptr;
  HIPCHK(hipMalloc(&xcw, m.xcw.size() * sizeof(double)));
  HIPCHK(hipMemcpy(xcw, m.xcw.data(), m.xcw.size() * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double *> old;
  for (int k = 0; k < 2; k++) {
    if (!names[k]) continue;
    SpecEntry &e = *g_ctx.spectra.find(names[k])->second;
    if (attach_mcica(&e.host, m, k == 0, &err)) return fail(SOCGPU_EUNSUPPORTED, "%s", err.c_str());
    HIPCHK(hipMemcpy(e.dev, &e.host, sizeof(SgSpec), hipMemcpyHostToDevice));
    if (e.xcw) old.push_back(e.xcw);
    e.xcw = xcw;
  }
  for (double *p : old) {
    bool used = false;
    for (auto &kv : g_ctx.spectra) used |= kv.second->xcw == p;
    if (!used) {
      (void)hipDeviceSynchronize();
      (void)hipFree(p);
    }
  }
  return 0;
}

// A call that asks for any clear-sky flux.
static bool wants_clear(const SocgpuDiag *g) {
  return g && (g->flux_up_clear || g->flux_down_clear || g->flux_direct_clear);
}

// Input arrays of one call in the order they are staged; n = element counts.
struct InField { const double **dst; const double *src; size_t n; };
constexpr int kMaxInFields = 48;

static size_t input_fields(const SocgpuRunesArgs *a, SocgpuRunesArgs *dv, int NB, InField *f) {
  const size_t P = a->n_profile, L = a->n_layer;
  InField g[] = {
      {&dv->p_layer, a->p_layer, L * P}, {&dv->t_layer, a->t_layer, L * P},
      {&dv->mass, a->mass, L * P}, {&dv->density, a->density, L * P},
      {&dv->layer_heat_capacity, a->layer_heat_capacity, L * P},
      {&dv->h2o, a->h2o, L * P}, {&dv->o3, a->o3, L * P},
      {&dv->p_level, a->p_level, (L + 1) * P}, {&dv->t_level, a->t_level, (L + 1) * P},
      {&dv->t_ground, a->t_ground, P}, {&dv->cos_zenith_angle, a->cos_zenith_angle, P},
      {&dv->solar_irrad, a->solar_irrad, P},
      {&dv->albedo_diff, a->albedo_diff, NB * P}, {&dv->albedo_dir, a->albedo_dir, NB * P},
      {&dv->cloud_frac, a->cloud_frac, L * P}, {&dv->liq_mmr, a->liq_mmr, L * P},
      {&dv->liq_dim, a->liq_dim, L * P}, {&dv->conv_frac, a->conv_frac, L * P},
      {&dv->liq_conv_mmr, a->liq_conv_mmr, L * P}, {&dv->liq_conv_dim, a->liq_conv_dim, L * P},
      {&dv->flux_ground, a->l_flux_ground ? a->flux_ground : nullptr, NB * P},
      {&dv->co2, a->l_co2_well_mixed ? nullptr : a->co2, L * P},
      {&dv->gas_field[0], a->gas_field[0], L * P}, {&dv->gas_field[1], a->gas_field[1], L * P},
      {&dv->gas_field[2], a->gas_field[2], L * P}, {&dv->gas_field[3], a->gas_field[3], L * P},
      {&dv->gas_field[4], a->gas_field[4], L * P}, {&dv->gas_field[5], a->gas_field[5], L * P},
      {&dv->gas_field[6], a->gas_field[6], L * P}, {&dv->gas_field[7], a->gas_field[7], L * P},
      {&dv->gas_field[8], a->gas_field[8], L * P}, {&dv->gas_field[9], a->gas_field[9], L * P},
      {&dv->ice_mmr, a->ice_mmr, L * P}, {&dv->ice_dim, a->ice_dim, L * P},
      {&dv->ice_conv_mmr, a->ice_conv_mmr, L * P}, {&dv->ice_conv_dim, a->ice_conv_dim, L * P},
      {&dv->aer_mix_ratio, a->n_aer_mode > 0 ? a->aer_mix_ratio : nullptr, (size_t)a->n_aer_mode * L * P},
      {&dv->aer_absorption, a->n_aer_mode > 0 ? a->aer_absorption : nullptr, NB * (size_t)a->n_aer_mode * L * P},
      {&dv->aer_scattering, a->n_aer_mode > 0 ? a->aer_scattering : nullptr, NB * (size_t)a->n_aer_mode * L * P},
      {&dv->aer_asymmetry, a->n_aer_mode > 0 ? a->aer_asymmetry : nullptr, NB * (size_t)a->n_aer_mode * L * P},
  };
  const size_t n = sizeof g / sizeof g[0];
  static_assert(sizeof g / sizeof g[0] <= kMaxInFields, "input field table too small");
  for (size_t k = 0; k < n; k++) f[k] = g[k];
  return n;
}

// Bytes of device staging one call's host inputs need (0 for device inputs).
static size_t stage_bytes(const SocgpuRunesArgs *a, int NB) {
  if (a->l_inputs_on_device) return 0;
  SocgpuRunesArgs dv = *a;
  InField f[kMaxInFields];
  size_t n = input_fields(a, &dv, NB, f), total = 0;
  for (size_t k = 0; k < n; k++) if (f[k].src) total += (f[k].n * 8 + 255) & ~(size_t)255;
  if (a->rand_seed) total += ((size_t)a->n_profile * 4 + 255) & ~(size_t)255;
  return total;
}

// Gathers one call's host inputs into the pinned block hbase + *off and
// points dv at the same offsets of the device block base (the caller copies
// the whole block across at once); device inputs are used in place.
// (hbase null: each array is copied across from the caller's memory on s.)
static int stage_inputs_at(const SocgpuRunesArgs *a, SocgpuRunesArgs *dv, int NB, char *base, char *hbase,
                           size_t *off, hipStream_t s) {
  *dv = *a;
  if (a->l_inputs_on_device) return 0;
  InField f[kMaxInFields];
  size_t n = input_fields(a, dv, NB, f);
  for (size_t k = 0; k < n; k++) {
    InField &x = f[k];
    if (!x.src) { *x.dst = nullptr; continue; }
    if (hbase) std::memcpy(hbase + *off, x.src, x.n * sizeof(double));
    else HIPCHK(hipMemcpyAsync(base + *off, x.src, x.n * sizeof(double), hipMemcpyHostToDevice, s));
    *x.dst = (const double *)(base + *off);
    *off += (x.n * 8 + 255) & ~(size_t)255;
  }
  if (a->rand_seed) {
    const size_t P = a->n_profile;
    if (hbase) std::memcpy(hbase + *off, a->rand_seed, P * sizeof(int));
    else HIPCHK(hipMemcpyAsync(base + *off, a->rand_seed, P * sizeof(int), hipMemcpyHostToDevice, s));
    dv->rand_seed = (const int *)(base + *off);
    *off += (P * 4 + 255) & ~(size_t)255;
  }
  return 0;
}

static int stage_inputs(const SocgpuRunesArgs *a, SocgpuRunesArgs *dv, int NB) {
  const size_t need = stage_bytes(a, NB) + 256;
  if (g_ctx.in_stage.ensure(need)) return fail(SOCGPU_ENOMEM, "input staging allocation failed");
  char *h = g_ctx.in_host.use(need);
  size_t off = 0;
  if (int rc = stage_inputs_at(a, dv, NB, g_ctx.in_stage.as<char>(), h, &off, g_ctx.stream)) return rc;
  if (h && off > 0) HIPCHK(hipMemcpyAsync(g_ctx.in_stage.p, h, off, hipMemcpyHostToDevice, g_ctx.stream));
  return 0;
}

// Lanes per column of the packed solve and its waves: one lane per g-point,
// or (per_item: the mixed form's calls that need the general gas optics) one
// lane per clear k-term and per McICA sample.
static void solve_layout(CallDims &d, const SgSpec &S, bool per_item) {
  const int items = S.n_g + ((per_item && d.mcica) ? S.n_sub_items[d.mode] : 0);
  d.n_items = items > kMinItems ? items : kMinItems;
  d.n_waves = (int)(((long long)d.P * d.n_items + kWave - 1) / kWave);
}

// Checks one call's arguments and resolves its options (socrates_set_control
// defaults on the two-stream path); fills the request's shapes.
static int resolve_request(const char *spectrum_name, const SocgpuRunesArgs *a, int precision,
                           SpecEntry **entry, CallDims *dout) {
  if (!spectrum_name || !a) return fail(SOCGPU_EINVAL, "null argument");
  auto it = g_ctx.spectra.find(spectrum_name);
  if (it == g_ctx.spectra.end()) return fail(SOCGPU_ENOTFOUND, "Spectrum name not found: %s", spectrum_name);
  SpecEntry &e = *it->second;
  const SgSpec &S = e.host;
  const int P = a->n_profile, L = a->n_layer, NB = S.n_band;
  if (P <= 0 || L <= 1) return fail(SOCGPU_EINVAL, "n_profile and n_layer must be positive (n_layer > 1)");
  if (L > SG_MAXL) return fail(SOCGPU_EUNSUPPORTED, "n_layer %d above %d", L, SG_MAXL);
  if (precision != SOCGPU_FP64 && precision != SOCGPU_MIXED) return fail(SOCGPU_EINVAL, "unknown precision %d", precision);
  if (a->i_source != SOCGPU_SOURCE_SOLAR && a->i_source != SOCGPU_SOURCE_THERMAL)
    return fail(SOCGPU_EINVAL, "Radiative source option unrecognised.");
  if (!a->p_layer || !a->t_layer || !a->mass) return fail(SOCGPU_EINVAL, "p_layer, t_layer and mass are required");
  const bool solar = a->i_source == SOCGPU_SOURCE_SOLAR;
  if (solar && (!a->cos_zenith_angle || !a->solar_irrad)) return fail(SOCGPU_EINVAL, "solar call needs cos_zenith_angle and solar_irrad");
  if (!solar && (!a->t_level || (!a->t_ground && !a->l_flux_ground)))
    return fail(SOCGPU_EINVAL, "thermal call needs t_level and t_ground (or l_flux_ground)");
  if (!solar && S.n_deg <= 0 && S.band[0].planck[0] == 0.0) return fail(SOCGPU_EUNSUPPORTED, "spectral file has no Planck data");
  if (S.has_cont && !a->density) return fail(SOCGPU_EINVAL, "density is required for continuum absorption");
  CallDims d;
  std::memset(&d, 0, sizeof d);
  d.P = P; d.L = L; d.NB = NB; d.isolir = a->i_source;
  d.gx = S.has_lut || S.has_cont;
  d.gen = d.gx;
  for (int t = 0; t < S.n_g; t++) d.gen |= S.g[t].fnc != SG_FNC_POWER_LAW;
  for (int t = 0; t < S.n_minor; t++) d.gen |= S.minor[t].fnc != SG_FNC_POWER_LAW || S.minor[t].gas >= SG_GAS_WM0;
  int cloud = a->i_cloud_representation;
  auto check_drop = [&](int t, const char *what) -> int {
    if (t <= 0) return fail(SOCGPU_EUNSUPPORTED, "Prescribed liquid cloud not yet implemented.");
    if (t > SG_NDROP || !S.drop[t - 1].present) return fail(SOCGPU_EINVAL, "%s cloud type not in spectral file.", what);
    if (S.drop[t - 1].scheme != 5 && S.drop[t - 1].scheme != 1) return fail(SOCGPU_EUNSUPPORTED, "droplet scheme %d", S.drop[t - 1].scheme);
    return 0;
  };
  if (cloud == SOCGPU_CLOUD_CONV_STRAT) {
    // stratiform and convective cloud: triple-column regions; McICA does not
    // apply to this representation (socrates_set_control.F90:338-356).
    // Random overlap selects ip_cloud_mix_random, which the reference's
    // mix_column rejects for the triple solver.
    if (a->i_overlap != SOCGPU_OVERLAP_MAX_RAND && a->i_overlap != SOCGPU_OVERLAP_EXP_RAND)
      return fail(SOCGPU_EUNSUPPORTED, "overlap %d with convective and stratiform cloud is not supported", a->i_overlap);
    if (!a->cloud_frac || !a->conv_frac) return fail(SOCGPU_EINVAL, "Cloud fraction not provided.");
    if (int rc = check_drop(a->i_st_water, "Liquid")) return rc;
    if (int rc = check_drop(a->i_cnv_water, "Convective liquid")) return rc;
    d.triple = 1;
    d.conv = 1;
    d.drop = a->i_st_water - 1;
    d.drop2 = a->i_cnv_water - 1;
    d.ioverlap = a->i_overlap;
  } else if (cloud == SOCGPU_CLOUD_HOMOGEN && a->i_inhom != SOCGPU_INHOM_MCICA) {
    // mixed-column representation (ip_cloud_mix_max / ip_cloud_mix_random /
    // ip_cloud_part_corr with solver_mix_direct_hogan, socrates_set_control.F90:
    // 319-334): its clear and cloudy regions are the clear and stratiform
    // regions of the triple-column solve with no convective cloud, which the
    // reference computes to the same bits
    if (a->i_overlap != SOCGPU_OVERLAP_MAX_RAND && a->i_overlap != SOCGPU_OVERLAP_EXP_RAND &&
        a->i_overlap != SOCGPU_OVERLAP_RANDOM)
      return fail(SOCGPU_EUNSUPPORTED, "overlap %d is not supported", a->i_overlap);
    if (!a->cloud_frac || !a->liq_mmr || !a->liq_dim) return fail(SOCGPU_EINVAL, "Cloud fraction not provided.");
    if (int rc = check_drop(a->i_st_water, "Liquid")) return rc;
    d.triple = 1;
    d.drop = a->i_st_water - 1;
    d.drop2 = a->i_st_water - 1;
    d.ioverlap = a->i_overlap;
  } else if (cloud != SOCGPU_CLOUD_OFF) {
    if (cloud != SOCGPU_CLOUD_HOMOGEN || a->i_inhom != SOCGPU_INHOM_MCICA)
      return fail(SOCGPU_EUNSUPPORTED, "cloud representation %d with i_inhom %d is not on the path", cloud, a->i_inhom);
    if (!S.has_mcica) return fail(SOCGPU_ENOTFOUND, "Spectrum name not associated with MCICA data.");
    if (a->i_mcica_sampling != SOCGPU_MCICA_SINGLE && a->i_mcica_sampling != SOCGPU_MCICA_OPTIMAL)
      return fail(SOCGPU_EUNSUPPORTED, "McICA sampling %d not supported", a->i_mcica_sampling);
    if (a->i_overlap != SOCGPU_OVERLAP_EXP_RAND && a->i_overlap != SOCGPU_OVERLAP_RANDOM &&
        a->i_overlap != SOCGPU_OVERLAP_MAX_RAND && a->i_overlap != SOCGPU_OVERLAP_EXP)
      return fail(SOCGPU_EUNSUPPORTED, "overlap %d not supported with McICA", a->i_overlap);
    if (!a->cloud_frac || !a->liq_mmr || !a->liq_dim) return fail(SOCGPU_EINVAL, "Cloud fraction not provided.");
    if (!a->rand_seed) return fail(SOCGPU_EINVAL, "rand_seed is required: default seeds depend on the position in the call");
    int drop = a->i_st_water;
    if (int rc = check_drop(drop, "Liquid")) return rc;
    d.mcica = 1;
    d.cg_pre = 1;
    d.f32 = precision == SOCGPU_MIXED;
    d.mode = a->i_mcica_sampling == SOCGPU_MCICA_OPTIMAL ? 1 : 0;
    if (S.n_req[d.mode] <= 0) return fail(SOCGPU_EUNSUPPORTED, "McICA data has no tables for this sampling");
    d.drop = drop - 1;
    d.n_gen = a->n_subcol_gen > 0 ? a->n_subcol_gen : S.n_gen;
    if (d.n_gen > SG_MAXSUBCOL) return fail(SOCGPU_EUNSUPPORTED, "n_subcol_gen above 64");
    d.n_fill = S.n_req[d.mode] < d.n_gen ? S.n_req[d.mode] : d.n_gen;
    d.ipph = S.ipph;
    d.ioverlap = a->i_overlap;
  }
  // the ice component of the cloud types (socrates_set_cld.F90:341-381):
  // the given ice types, i_cnv_ice defaulting to i_st_ice; without i_st_ice
  // there is no ice component (the reference's zero-ice call with an ice
  // fit that is zero at zero ice)
  d.co_stride = 4;
  if ((d.mcica || d.triple) && a->i_st_ice > 0) {
    auto check_ice = [&](int t) -> int {
      if (t > SG_NICE || !S.ice[t - 1].present) return fail(SOCGPU_EINVAL, "Ice cloud type not in spectral file.");
      const int sc = S.ice[t - 1].scheme;
      if (sc != 12 && sc != 19 && sc != 20)
        return fail(SOCGPU_EUNSUPPORTED, "ice parametrisation %d is not on this path", sc);
      return 0;
    };
    const int tc = a->i_cnv_ice > 0 ? a->i_cnv_ice : a->i_st_ice;
    if (int rc = check_ice(a->i_st_ice)) return rc;
    if (d.conv) {
      if (int rc = check_ice(tc)) return rc;
    }
    d.ice = 1;
    d.ice_t = a->i_st_ice - 1;
    d.ice_t2 = d.conv ? tc - 1 : d.ice_t;
    if (d.mcica && !d.f32) d.co_stride = 8;
  } else if (a->ice_mmr || a->ice_conv_mmr) {
    if (d.mcica || d.triple) return fail(SOCGPU_EINVAL, "ice condensate needs i_st_ice");
  }
  d.ir_lin = !solar && a->l_ir_source_linear;
  d.no_cont = a->l_no_continuum;
  // treatment of scattering (socrates_set_control.F90:247-277, 312-318,
  // 368-370; solve_band_k_eqv_scl.F90:1053-1058)
  d.scat = a->i_scatter_method > 0 ? a->i_scatter_method : SOCGPU_SCATTER_FULL;
  if (d.scat > SOCGPU_SCATTER_HYBRID) return fail(SOCGPU_EINVAL, "i_scatter_method %d unknown", d.scat);
  d.grey1 = d.scat != SOCGPU_SCATTER_FULL;
  for (int g = 0; g < S.n_g; g++) {
    const int m = d.scat == SOCGPU_SCATTER_HYBRID ? (S.g[g].iscat > 0 ? S.g[g].iscat : SOCGPU_SCATTER_FULL) : d.scat;
    if (term_nosc(d.scat, S.g[g].iscat) && solar)
      return fail(SOCGPU_EUNSUPPORTED, "scattering can be ignored in the infra-red only (k-term %d)", g + 1);
    if (term_nosc(d.scat, S.g[g].iscat) && d.triple)
      return fail(SOCGPU_EUNSUPPORTED, "the triple and mixed columns take full scattering only");
    if (m == SOCGPU_SCATTER_APPROX && d.triple)
      return fail(SOCGPU_EUNSUPPORTED, "approximate scattering in the triple and mixed columns is not supported");
  }
  // modal aerosols (socrates_set_control.F90:378-384: on when modes are
  // given; socrates_set_aer.F90:299-309): the general solve build carries
  // them; the band kernel's build follows the gas optics alone
  d.ggen = d.gen;
  if (a->n_aer_mode > 0) {
    if (!a->aer_mix_ratio || !a->aer_absorption || !a->aer_scattering || !a->aer_asymmetry)
      return fail(SOCGPU_EINVAL, "aerosol modes need aer_mix_ratio, aer_absorption, aer_scattering and aer_asymmetry");
    d.n_aer = a->n_aer_mode;
    d.gen = 1;
  }
  d.nwpc = (S.n_g + kWave - 1) / kWave;
  solve_layout(d, S, precision == SOCGPU_MIXED && d.gen);
  *entry = &e;
  *dout = d;
  return 0;
}

// Per-wave partial sums of a request's solve.
static size_t partial_doubles(const CallDims &d) {
  if (d.triple) return (size_t)d.P * d.nwpc * 3 * (d.L + 1);
  return (size_t)d.n_waves * kMaxSeg * 3 * (d.L + 1);
}

// Device workspace of one request, in doubles (8-byte units), and its work
// item counts per space.
static size_t request_ws_doubles(const CallDims &d, const SgSpec &S) {
  const size_t P = d.P, L = d.L, NB = d.NB;
  size_t n = P * L * SG_NQ + P * (L + 1) + P * SG_NS + P * NB * L * SG_NBQ + P * NB * SG_NBQ +
             partial_doubles(d) + 2;
  if (d.gx) n += P * L * SG_NGQ;
  if (d.mcica) n += P * L * SG_MAXSUBCOL + P * NB * L * d.co_stride;
  if (d.triple) n += P * NB * L * 8 + P * (L + 1) * kOv + P * d.nwpc * L * kTriQ * kWave;
  if (d.n_aer) n += P * NB * L * 4;
  return n;
}

static void request_work(const CallDims &d, const SgSpec &S, long long w[kNWork]) {
  const long long P = d.P, L = d.L;
  w[W_LEV] = P * (L + 1);
  w[W_LAY] = P * L;
  w[W_CG] = d.mcica ? P : 0;
  w[W_MINOR] = P * d.NB;
  w[W_LONG] = L > 2 * kWave ? P * d.NB : 0;
  w[W_PREP] = P * d.NB * L;
  w[W_SOLVE] = d.triple ? 0 : d.n_waves;
  w[W_TOVL] = d.triple ? P * (L + 1) : 0;
  w[W_TSOLVE] = d.triple ? P * d.nwpc : 0;
  w[W_PTILE] = pack_tiles(d);
  w[W_CCOL] = (d.mcica || d.triple) ? P * ((L + kRangeChunk - 1) / kRangeChunk) : 0;
  w[W_FTILE] = finish_tiles(d);
  w[W_GEN] = d.gen ? 1 : 0;
  w[W_PLAIN] = d.gen ? 0 : 1;
  w[W_GBAND] = d.ggen ? 1 : 0;
}

// Lays out the workspace of request r at *cursor (doubles, 32-byte aligned)
// and fills its descriptor's pointers.
static void place_request(ReqDesc &q, const SgSpec &S, double *base, size_t *cursor) {
  const CallDims &d = q.d;
  const size_t P = d.P, L = d.L, NB = d.NB;
  auto take = [&](size_t n) {
    double *p = base + *cursor;
    *cursor += (n + 3) & ~(size_t)3;
    return p;
  };
  q.colrec = take(P * L * SG_NQ);
  q.grec = d.gx ? take(P * L * SG_NGQ) : nullptr;
  q.tlev = take(P * (L + 1));
  q.colscal = take(P * SG_NS);
  q.bandrec = take(P * NB * L * SG_NBQ);
  q.bandscal = take(P * NB * SG_NBQ);
  q.partial = take(partial_doubles(d));
  q.krange = (int *)take(2);
  q.csub = d.mcica ? take(P * L * SG_MAXSUBCOL) : nullptr;
  q.cldopt = d.mcica ? take(P * NB * L * d.co_stride) : (d.triple ? take(P * NB * L * 8) : nullptr);
  q.ovl = d.triple ? take(P * (L + 1) * kOv) : nullptr;
  q.tws = d.triple ? take((size_t)P * d.nwpc * L * kTriQ * kWave) : nullptr;
  q.aerrec = d.n_aer ? take(P * NB * L * 4) : nullptr;
}

static size_t padded_ws_doubles(const CallDims &d, const SgSpec &S) {
  return request_ws_doubles(d, S) + 13 * 4;  // per-array alignment padding
}

static void fill_desc(ReqDesc &q, const CallDims &d, const SpecEntry &e, const SocgpuRunesArgs &dv,
                      double *up, double *dn, double *dir, double *hr) {
  std::memset(&q, 0, sizeof q);
  q.d = d;
  q.spec = e.dev;
  q.xcw = e.xcw;
  q.n1 = e.host.n1;
  q.n2 = e.host.n2;
  q.l_rayleigh = dv.l_rayleigh;
  q.dp_cf = dv.cloud_vertical_decorr;
  q.dp_cw = dv.cloud_vertical_decorr * 0.5;
  q.a = dv;
  q.up = up; q.dn = dn; q.dir = dir; q.hr = hr;
}

// Launches the kernel sequence of a packed batch on stream s. hdr/rq are
// device copies of the batch header and descriptors; tot is the host copy of
// the work totals (grid sizes); solve-kernel time between ev_a and ev_b.
// The solve kernel of a tile, for the power-law-only or the general gas
// optics.
using SolveKernel = void (*)(const BatchHdr *, const ReqDesc *);
static SolveKernel solve_kernel_for(bool mixed, bool recomp, int occ, bool gen) {
#define SG_SK(M, R, G) (occ == 3 ? b_solve_kernel<M, R, 3, G> : b_solve_kernel<M, R, 4, G>)
  if (gen) return mixed ? (recomp ? SG_SK(true, true, true) : SG_SK(true, false, true))
                        : (recomp ? SG_SK(false, true, true) : SG_SK(false, false, true));
  return mixed ? (recomp ? SG_SK(true, true, false) : SG_SK(true, false, false))
               : (recomp ? SG_SK(false, true, false) : SG_SK(false, false, false));
#undef SG_SK
}

static int launch_batch(const BatchHdr *hdr, const ReqDesc *rq, const long long tot[kNWork], int R,
                        int precision, hipStream_t s, const Fork *fk, hipEvent_t ev_a, hipEvent_t ev_b) {
  const int T = 256;
  const bool gen = tot[W_GEN] > 0;
  auto blocks = [](long long n, int per) {
    long long b = (n + per - 1) / per;
    return (int)(b < 1 ? 1 : (b > 1048576 ? 1048576 : b));
  };
  b_init_kernel<<<(R + 63) / 64, 64, 0, s>>>(hdr, rq);
  b_pack_kernel<<<blocks(tot[W_PTILE], 1), kPackC * kPackL, 0, s>>>(hdr, rq);
  if (tot[W_CCOL] > 0) b_cldrange_kernel<<<blocks(tot[W_CCOL], T), T, 0, s>>>(hdr, rq);
  // the cloud generator on the side stream when there is one, beside the
  // band kernel; joined before the solve
  const bool fork = fk && fk->side && tot[W_CG] > 0;
  if (fork) {
    HIPCHK(hipEventRecord(fk->f, s));
    HIPCHK(hipStreamWaitEvent(fk->side, fk->f, 0));
    b_cloudgen_kernel<<<blocks(tot[W_CG], kCgWaves), 64 * kCgWaves, 0, fk->side>>>(hdr, rq);
    HIPCHK(hipEventRecord(fk->j, fk->side));
  } else if (tot[W_CG] > 0) {
    b_cloudgen_kernel<<<blocks(tot[W_CG], kCgWaves), 64 * kCgWaves, 0, s>>>(hdr, rq);
  }
  {
    // calls of up to 128 layers, then the longer ones (each kernel skips the
    // other's waves)
    const bool gband = tot[W_GBAND] > 0;
    auto *bk = gband ? b_band_kernel<2, true> : b_band_kernel<2, false>;
    bk<<<blocks(tot[W_MINOR], kMinorWaves), 64 * kMinorWaves, 0, s>>>(hdr, rq);
    if (tot[W_LONG] > 0) {
      auto *bl = gband ? b_band_kernel<kKeqMaxChunks, true> : b_band_kernel<kKeqMaxChunks, false>;
      bl<<<blocks(tot[W_MINOR], kMinorWaves), 64 * kMinorWaves, 0, s>>>(hdr, rq);
    }
  }
  if (tot[W_TOVL] > 0) b_overlap_kernel<<<blocks(tot[W_TOVL], T), T, 0, s>>>(hdr, rq);
  if (fork) HIPCHK(hipStreamWaitEvent(s, fk->j, 0));
  if (ev_a) HIPCHK(hipEventRecord(ev_a, s));
  const SocgpuTile &tile = g_tile[precision == SOCGPU_MIXED ? 1 : 0];
  const int wpb = tile.waves_per_block;
  // general gas optics: the recomputing tile would evaluate the look-up
  // tables twice per layer; the storing tile is faster there (GA9 files,
  // 4096 x 90 all-sky: 5.27 ms per SW+LW step against 6.13 ms)
  if (tot[W_SOLVE] > 0) {
    // the power-law build for the calls that need no more, the general one
    // for the others (each skips the other's calls)
    if (tot[W_PLAIN] > 0) {
      auto *kern = solve_kernel_for(precision == SOCGPU_MIXED, tile.recompute != 0, tile.waves_per_simd, false);
      kern<<<blocks(tot[W_SOLVE], wpb), wpb * kWave, 0, s>>>(hdr, rq);
    }
    if (gen) {
      auto *kern = solve_kernel_for(precision == SOCGPU_MIXED, false, tile.waves_per_simd, true);
      kern<<<blocks(tot[W_SOLVE], wpb), wpb * kWave, 0, s>>>(hdr, rq);
    }
  }
  if (tot[W_TSOLVE] > 0) {
    auto *tk = precision == SOCGPU_MIXED ? (gen ? b_tsolve_kernel<true, true> : b_tsolve_kernel<true, false>)
                                         : (gen ? b_tsolve_kernel<false, true> : b_tsolve_kernel<false, false>);
    tk<<<blocks(tot[W_TSOLVE], 4), 4 * kWave, 0, s>>>(hdr, rq);
  }
  if (ev_b) HIPCHK(hipEventRecord(ev_b, s));
  b_finish_kernel<<<blocks(tot[W_FTILE], 1), kFinC * kFinL, 0, s>>>(hdr, rq);
  b_heating_kernel<<<blocks(tot[W_LAY], T), T, 0, s>>>(hdr, rq);
  HIPCHK(hipGetLastError());
  return 0;
}

int socgpu_runes(const char *spectrum_name, const SocgpuRunesArgs *a, SocgpuDiag *diag, int precision) {
  if (int rc = ensure_init()) return rc;
  if (!spectrum_name || !a || !diag) return fail(SOCGPU_EINVAL, "null argument");
  if (wants_clear(diag)) {
    // the call and its cloud-free companion, packed as one batch
    const SocgpuRequest r{spectrum_name, a, diag};
    return socgpu_runes_batch(1, &r, precision);
  }
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  SpecEntry *e = nullptr;
  CallDims d;
  if (int rc = resolve_request(spectrum_name, a, precision, &e, &d)) return rc;
  const SgSpec &S = e->host;
  const int P = d.P, L = d.L, NB = d.NB;

  SocgpuRunesArgs dv;
  if (int rc = stage_inputs(a, &dv, NB)) return rc;
  if (g_ctx.ws.ensure(padded_ws_doubles(d, S) * 8) ||
      g_ctx.desc.ensure(sizeof(BatchHdr) + sizeof(ReqDesc)) ||
      g_ctx.out_stage.ensure((size_t)P * (L + 1) * 3 * 8 + (size_t)P * L * 8))
    return fail(SOCGPU_ENOMEM, "device buffer allocation failed");
  double *oup = g_ctx.out_stage.as<double>();
  double *odn = oup + (size_t)P * (L + 1), *odir = odn + (size_t)P * (L + 1), *ohr = odir + (size_t)P * (L + 1);
  if (diag->l_outputs_on_device) {
    if (diag->flux_up) oup = diag->flux_up;
    if (diag->flux_down) odn = diag->flux_down;
    if (diag->flux_direct) odir = diag->flux_direct;
    if (diag->heating_rate) ohr = diag->heating_rate;
  }
  struct {
    BatchHdr h;
    ReqDesc q;
  } host;
  std::memset(&host, 0, sizeof host);
  fill_desc(host.q, d, *e, dv, oup, odn, odir, diag->heating_rate ? ohr : nullptr);
  size_t cursor = 0;
  place_request(host.q, S, g_ctx.ws.as<double>(), &cursor);
  long long w[kNWork];
  request_work(d, S, w);
  host.h.R = 1;
  for (int k = 0; k < kNWork; k++) { host.q.s[k] = 0; host.h.tot[k] = w[k]; }
  hipStream_t s = g_ctx.stream;
  HIPCHK(hipEventRecord(g_ctx.ev0, s));
  HIPCHK(hipMemcpyAsync(g_ctx.desc.p, &host, sizeof host, hipMemcpyHostToDevice, s));
  const BatchHdr *dh = (const BatchHdr *)g_ctx.desc.p;
  const ReqDesc *dq = (const ReqDesc *)((char *)g_ctx.desc.p + offsetof(decltype(host), q));
  if (int rc = launch_batch(dh, (ReqDesc *)dq, host.h.tot, 1, precision, s, &g_ctx.fork, g_ctx.ev1, g_ctx.ev2)) return rc;
  HIPCHK(hipEventRecord(g_ctx.ev3, s));
  int herr = 0;
  HIPCHK(hipMemcpyAsync(&herr, &dh->err, sizeof herr, hipMemcpyDeviceToHost, s));
  const size_t nout = (size_t)P * (L + 1) * 3 + (size_t)P * L;
  const size_t nl = (size_t)P * (L + 1);
  char *oh = diag->l_outputs_on_device ? nullptr : g_ctx.out_host.use(nout * 8);
  if (oh) {
    // one copy of the output block into pinned memory, scattered below
    HIPCHK(hipMemcpyAsync(oh, g_ctx.out_stage.p, nout * 8, hipMemcpyDeviceToHost, s));
  } else if (!diag->l_outputs_on_device) {
    if (diag->flux_up) HIPCHK(hipMemcpyAsync(diag->flux_up, oup, nl * 8, hipMemcpyDeviceToHost, s));
    if (diag->flux_down) HIPCHK(hipMemcpyAsync(diag->flux_down, odn, nl * 8, hipMemcpyDeviceToHost, s));
    if (diag->flux_direct) HIPCHK(hipMemcpyAsync(diag->flux_direct, odir, nl * 8, hipMemcpyDeviceToHost, s));
    if (diag->heating_rate) HIPCHK(hipMemcpyAsync(diag->heating_rate, ohr, (size_t)P * L * 8, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  if (herr) return fail(SOCGPU_EINVAL, "Cloud fraction greater than 1.");
  if (oh) {
    const double *h = (const double *)oh;
    if (diag->flux_up) std::memcpy(diag->flux_up, h, nl * 8);
    if (diag->flux_down) std::memcpy(diag->flux_down, h + nl, nl * 8);
    if (diag->flux_direct) std::memcpy(diag->flux_direct, h + 2 * nl, nl * 8);
    if (diag->heating_rate) std::memcpy(diag->heating_rate, h + 3 * nl, (size_t)P * L * 8);
  }
  float t_all = 0.f, t_solve = 0.f;
  HIPCHK(hipEventElapsedTime(&t_all, g_ctx.ev0, g_ctx.ev3));
  HIPCHK(hipEventElapsedTime(&t_solve, g_ctx.ev1, g_ctx.ev2));
  g_ctx.last_total_ms = t_all;
  g_ctx.last_solve_ms = t_solve;
  return 0;
}

}  // extern "C"

namespace sg {

// A one-call descriptor block written into device memory by a kernel that
// takes it by value (the allocation-free entry point copies nothing across
// with the runtime's copy engines: no staging buffer is involved).
struct DescBlob {
  BatchHdr h;
  ReqDesc q;
};
static_assert(sizeof(DescBlob) % 8 == 0, "descriptor block is copied in 8-byte words");

__global__ void b_desc_store_kernel(DescBlob v, DescBlob *dst) {
  const unsigned long long *src = (const unsigned long long *)&v;
  unsigned long long *d = (unsigned long long *)dst;
  for (int i = threadIdx.x; i < (int)(sizeof(DescBlob) / 8); i += blockDim.x) d[i] = src[i];
}

}  // namespace sg

extern "C" {

int socgpu_runes_noalloc(const char *spectrum_name, const SocgpuRunesArgs *a, SocgpuDiag *diag,
                         int precision) {
  if (int rc = ensure_init()) return rc;
  if (!spectrum_name || !a || !diag) return fail(SOCGPU_EINVAL, "null argument");
  if (!a->l_inputs_on_device || !diag->l_outputs_on_device)
    return fail(SOCGPU_EINVAL, "socgpu_runes_noalloc takes device inputs and outputs "
                               "(l_inputs_on_device and l_outputs_on_device set)");
  if (!diag->flux_up || !diag->flux_down)
    return fail(SOCGPU_EINVAL, "socgpu_runes_noalloc needs flux_up and flux_down");
  if (diag->flux_up_clear || diag->flux_down_clear || diag->flux_direct_clear)
    return fail(SOCGPU_EUNSUPPORTED, "socgpu_runes_noalloc gives no clear-sky fluxes (use socgpu_runes)");
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  if (!g_ctx.scratch_dev || !g_ctx.err_host) return fail(SOCGPU_EDEVICE, "fixed workspace not mapped");
  SpecEntry *e = nullptr;
  CallDims d;
  if (int rc = resolve_request(spectrum_name, a, precision, &e, &d)) return rc;
  const SgSpec &S = e->host;
  const size_t ws_off = (sizeof(DescBlob) + 255) & ~(size_t)255;
  const size_t need = ws_off + padded_ws_doubles(d, S) * 8;
  if (need > kScratchBytes)
    return fail(SOCGPU_ENOMEM, "call needs %zu bytes of workspace, above the %zu-byte fixed workspace of "
                               "socgpu_runes_noalloc (use socgpu_runes)", need, kScratchBytes);
  DescBlob host;
  std::memset(&host, 0, sizeof host);
  fill_desc(host.q, d, *e, *a, diag->flux_up, diag->flux_down, diag->flux_direct, diag->heating_rate);
  size_t cursor = 0;
  place_request(host.q, S, (double *)(g_ctx.scratch_dev + ws_off), &cursor);
  long long w[kNWork];
  request_work(d, S, w);
  host.h.R = 1;
  for (int k = 0; k < kNWork; k++) { host.q.s[k] = 0; host.h.tot[k] = w[k]; }
  hipStream_t s = g_ctx.stream;
  DescBlob *dd = (DescBlob *)g_ctx.scratch_dev;
  HIPCHK(hipEventRecord(g_ctx.ev0, s));
  b_desc_store_kernel<<<1, 64, 0, s>>>(host, dd);
  if (int rc = launch_batch(&dd->h, &dd->q, host.h.tot, 1, precision, s, &g_ctx.fork, g_ctx.ev1, g_ctx.ev2)) return rc;
  HIPCHK(hipEventRecord(g_ctx.ev3, s));
  HIPCHK(hipMemcpyAsync(g_ctx.err_host, &dd->h.err, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (*g_ctx.err_host) return fail(SOCGPU_EINVAL, "Cloud fraction greater than 1.");
  float t_all = 0.f, t_solve = 0.f;
  HIPCHK(hipEventElapsedTime(&t_all, g_ctx.ev0, g_ctx.ev3));
  HIPCHK(hipEventElapsedTime(&t_solve, g_ctx.ev1, g_ctx.ev2));
  g_ctx.last_total_ms = t_all;
  g_ctx.last_solve_ms = t_solve;
  return 0;
}

}  // extern "C"

namespace sg {

// One packed batch on the host: descriptors (ordered by decreasing layer
// count, so that the long columns start first), totals and the outputs to
// copy back to host memory after the launch.
// Clear-sky diagnostics (flux_up_clear, flux_down_clear, flux_direct_clear):
// the reference's clear-sky pass of a cloudy call gives, bit for bit, the
// fluxes of the same call with cloud off (its i_solver_clear is the
// homogeneous solver of the cloud-free call, socrates_set_control.F90:
// 310-375), so a call that asks for them gets a cloud-free companion request
// in the same batch that writes them.
struct Expanded {
  std::vector<SocgpuRequest> req;
  std::vector<int> caller;           // the caller's index of each request
  std::deque<SocgpuRunesArgs> args;  // companions' arguments (stable addresses)
  std::deque<SocgpuDiag> diag;
};

static int expand_clear(int n, const SocgpuRequest *rq, Expanded &x) {
  x.req.clear();
  x.caller.clear();
  x.args.clear();
  x.diag.clear();
  for (int i = 0; i < n; i++) {
    x.req.push_back(rq[i]);
    x.caller.push_back(i);
    const SocgpuDiag *g = rq[i].diag;
    if (!rq[i].args || !wants_clear(g)) continue;
    if (g->l_outputs_on_device && (!g->flux_up_clear || !g->flux_down_clear))
      return fail(SOCGPU_EINVAL, "request %d: device clear-sky outputs need flux_up_clear and flux_down_clear", i);
    SocgpuRunesArgs c = *rq[i].args;
    c.i_cloud_representation = SOCGPU_CLOUD_OFF;
    c.i_inhom = SOCGPU_INHOM_HOMOGENEOUS;
    c.cloud_frac = c.liq_mmr = c.liq_dim = nullptr;
    c.conv_frac = c.liq_conv_mmr = c.liq_conv_dim = nullptr;
    c.ice_mmr = c.ice_dim = c.ice_conv_mmr = c.ice_conv_dim = nullptr;
    c.i_st_ice = c.i_cnv_ice = 0;
    c.rand_seed = nullptr;
    x.args.push_back(c);
    SocgpuDiag cd;
    std::memset(&cd, 0, sizeof cd);
    cd.flux_up = g->flux_up_clear;
    cd.flux_down = g->flux_down_clear;
    cd.flux_direct = g->flux_direct_clear;
    cd.l_outputs_on_device = g->l_outputs_on_device;
    x.diag.push_back(cd);
    x.req.push_back(SocgpuRequest{rq[i].spectrum_name, &x.args.back(), &x.diag.back()});
    x.caller.push_back(i);
  }
  return 0;
}

struct Plan {
  Expanded x;              // the requests planned (the caller's and companions)
  int n() const { return (int)x.req.size(); }
  std::vector<ReqDesc> rq;
  BatchHdr h;
  size_t ws_doubles = 0, in_bytes = 0, out_doubles = 0, out_used = 0, in_used = 0;
  struct Back { double *dst; const double *src; size_t n; };
  std::vector<Back> back;
  std::vector<int> order;  // caller index of each descriptor
  long long need[kNWork];
};

// The failure a batch's error word (BatchHdr::err) reports, if any.
static int batch_err(const Plan &pl, int err) {
  if (err <= 0) return 0;
  const int j = err - 1;
  int i = j < (int)pl.order.size() ? pl.order[j] : j;
  if (i < (int)pl.x.caller.size()) i = pl.x.caller[i];
  return fail(SOCGPU_EINVAL, "request %d: Cloud fraction greater than 1.", i);
}

struct Resolved {
  SpecEntry *e;
  CallDims d;
};

// Sizes of a batch (workspace, staging, outputs) without placing it; the
// plan keeps the request list (with the clear-sky companions).
static int plan_sizes(int n_caller, const SocgpuRequest *caller_req, int precision, std::vector<Resolved> &res,
                      Plan &pl) {
  for (int i = 0; i < n_caller; i++)
    if (!caller_req[i].args || !caller_req[i].diag) return fail(SOCGPU_EINVAL, "request %d: null args or diag", i);
  if (int rc = expand_clear(n_caller, caller_req, pl.x)) return rc;
  const int n = pl.n();
  const SocgpuRequest *req = pl.x.req.data();
  res.resize(n);
  pl.ws_doubles = pl.in_bytes = pl.out_doubles = 0;
  for (int k = 0; k < kNWork; k++) pl.need[k] = 0;
  for (int i = 0; i < n; i++) {
    if (int rc = resolve_request(req[i].spectrum_name, req[i].args, precision, &res[i].e, &res[i].d))
      return rc;
    const CallDims &d = res[i].d;
    pl.ws_doubles += padded_ws_doubles(d, res[i].e->host);
    pl.in_bytes += stage_bytes(req[i].args, d.NB);
    if (!req[i].diag->l_outputs_on_device) pl.out_doubles += (size_t)d.P * (3 * (d.L + 1) + d.L) + 16;
    long long w[kNWork];
    request_work(d, res[i].e->host, w);
    for (int k = 0; k < kNWork; k++) pl.need[k] += w[k];
  }
  return 0;
}

// Places a sized batch into the given buffers; host inputs are gathered into
// the pinned block in_host and copied across in one copy on s, or (in_host
// null) copied array by array.
static int plan_place(const std::vector<Resolved> &res, Plan &pl, double *ws, char *in_stage, char *in_host,
                      double *out_stage, hipStream_t s) {
  const int n = pl.n();
  const SocgpuRequest *req = pl.x.req.data();
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return res[x].d.L > res[y].d.L; });
  pl.order = order;
  pl.rq.assign(n, ReqDesc());
  pl.back.clear();
  std::memset(&pl.h, 0, sizeof pl.h);
  pl.h.R = n;
  size_t ws_cur = 0, in_off = 0, out_cur = 0;
  long long start[kNWork] = {};
  for (int j = 0; j < n; j++) {
    const int i = order[j];
    const CallDims &d = res[i].d;
    const SpecEntry &e = *res[i].e;
    SocgpuRunesArgs dv;
    if (int rc = stage_inputs_at(req[i].args, &dv, d.NB, in_stage, in_host, &in_off, s)) return rc;
    const SocgpuDiag *g = req[i].diag;
    const size_t nl = (size_t)d.P * (d.L + 1), nh = (size_t)d.P * d.L;
    double *up, *dn, *dir, *hr;
    if (g->l_outputs_on_device) {
      up = g->flux_up; dn = g->flux_down; dir = g->flux_direct; hr = g->heating_rate;
      if (!up || !dn)
        return fail(SOCGPU_EINVAL, "request %d: device outputs need flux_up and flux_down", pl.x.caller[i]);
    } else {
      double *b = out_stage + out_cur;
      up = b; dn = b + nl; dir = b + 2 * nl; hr = b + 3 * nl;
      out_cur += 3 * nl + nh + 16;
      if (g->flux_up) pl.back.push_back({g->flux_up, up, nl});
      if (g->flux_down) pl.back.push_back({g->flux_down, dn, nl});
      if (g->flux_direct) pl.back.push_back({g->flux_direct, dir, nl});
      if (g->heating_rate) pl.back.push_back({g->heating_rate, hr, nh});
      if (!g->flux_direct) dir = nullptr;
      if (!g->heating_rate) hr = nullptr;
    }
    ReqDesc &q = pl.rq[j];
    fill_desc(q, d, e, dv, up, dn, dir, hr);
    place_request(q, e.host, ws, &ws_cur);
    long long w[kNWork];
    request_work(d, e.host, w);
    for (int k = 0; k < kNWork; k++) {
      q.s[k] = start[k];
      start[k] += w[k];
    }
  }
  for (int k = 0; k < kNWork; k++) pl.h.tot[k] = start[k];
  pl.out_used = out_cur;
  pl.in_used = in_off;
  if (in_host && in_off > 0) HIPCHK(hipMemcpyAsync(in_stage, in_host, in_off, hipMemcpyHostToDevice, s));
  return 0;
}

static size_t desc_bytes(int n) { return sizeof(BatchHdr) + (size_t)n * sizeof(ReqDesc); }

static void pack_desc(const Plan &pl, char *dst) {
  std::memcpy(dst, &pl.h, sizeof(BatchHdr));
  std::memcpy(dst + sizeof(BatchHdr), pl.rq.data(), pl.rq.size() * sizeof(ReqDesc));
}

// Host outputs: one copy of the used part of the output block into pinned
// memory (issue), then, once the stream has reached it, the scatter to the
// caller's arrays (finish); with out_host null, array by array directly.
static int copy_back_issue(const Plan &pl, const double *out_stage, char *out_host, hipStream_t s) {
  if (pl.back.empty()) return 0;
  if (out_host) {
    HIPCHK(hipMemcpyAsync(out_host, out_stage, pl.out_used * 8, hipMemcpyDeviceToHost, s));
  } else {
    for (const auto &b : pl.back) HIPCHK(hipMemcpyAsync(b.dst, b.src, b.n * 8, hipMemcpyDeviceToHost, s));
  }
  return 0;
}

static void copy_back_finish(const Plan &pl, const double *out_stage, const char *out_host) {
  if (!out_host) return;
  const double *h = (const double *)out_host;
  for (const auto &b : pl.back) std::memcpy(b.dst, h + (b.src - out_stage), b.n * 8);
}

}  // namespace sg

// Persistent batch: capacity-sized buffers and a captured HIP graph of the
// kernel sequence, replayed for every batch that fits (kernels read the
// batch's shape from the device descriptors and stride over it).
struct SocgpuBatch {
  int precision = SOCGPU_FP64;
  SocgpuTile tile{};
  size_t cap_ws = 0, cap_in = 0, cap_out = 0, cap_desc = 0;
  long long cap[sg::kNWork] = {};
  sg::DevBuf ws, in_stage, out_stage, desc;
  sg::HostBuf in_host, out_host;
  char *pinned = nullptr;
  hipStream_t stream = nullptr;
  Fork fork;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double last_ms = 0.0;
  int replays = 0, captures = 0;
};

extern "C" {

int socgpu_runes_batch(int n_request, const SocgpuRequest *requests, int precision) {
  if (int rc = ensure_init()) return rc;
  if (n_request <= 0 || !requests) return fail(SOCGPU_EINVAL, "empty batch");
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  std::vector<Resolved> res;
  Plan pl;
  if (int rc = plan_sizes(n_request, requests, precision, res, pl)) return rc;
  if (g_ctx.ws.ensure(pl.ws_doubles * 8) || g_ctx.in_stage.ensure(pl.in_bytes + 256) ||
      g_ctx.out_stage.ensure(pl.out_doubles * 8 + 64) || g_ctx.desc.ensure(desc_bytes(pl.n())))
    return fail(SOCGPU_ENOMEM, "device buffer allocation failed");
  char *ih = g_ctx.in_host.use(pl.in_bytes), *oh = g_ctx.out_host.use(pl.out_doubles * 8);
  hipStream_t s = g_ctx.stream;
  HIPCHK(hipEventRecord(g_ctx.ev0, s));
  if (int rc = plan_place(res, pl, g_ctx.ws.as<double>(), g_ctx.in_stage.as<char>(), ih,
                          g_ctx.out_stage.as<double>(), s))
    return rc;
  std::vector<char> blob(desc_bytes(pl.n()));
  pack_desc(pl, blob.data());
  HIPCHK(hipMemcpyAsync(g_ctx.desc.p, blob.data(), blob.size(), hipMemcpyHostToDevice, s));
  const BatchHdr *dh = (const BatchHdr *)g_ctx.desc.p;
  ReqDesc *dq = (ReqDesc *)((char *)g_ctx.desc.p + sizeof(BatchHdr));
  if (int rc = launch_batch(dh, dq, pl.h.tot, pl.n(), precision, s, &g_ctx.fork, g_ctx.ev1, g_ctx.ev2)) return rc;
  HIPCHK(hipEventRecord(g_ctx.ev3, s));
  int herr = 0;
  HIPCHK(hipMemcpyAsync(&herr, &dh->err, sizeof herr, hipMemcpyDeviceToHost, s));
  if (int rc = copy_back_issue(pl, g_ctx.out_stage.as<double>(), oh, s)) return rc;
  HIPCHK(hipStreamSynchronize(s));
  if (int rc = batch_err(pl, herr)) return rc;
  copy_back_finish(pl, g_ctx.out_stage.as<double>(), oh);
  float t_all = 0.f, t_solve = 0.f;
  HIPCHK(hipEventElapsedTime(&t_all, g_ctx.ev0, g_ctx.ev3));
  HIPCHK(hipEventElapsedTime(&t_solve, g_ctx.ev1, g_ctx.ev2));
  g_ctx.last_total_ms = t_all;
  g_ctx.last_solve_ms = t_solve;
  return 0;
}

int socgpu_batch_create(SocgpuBatch **batch, int precision) {
  if (int rc = ensure_init()) return rc;
  if (!batch) return fail(SOCGPU_EINVAL, "null batch");
  if (precision != SOCGPU_FP64 && precision != SOCGPU_MIXED) return fail(SOCGPU_EINVAL, "unknown precision %d", precision);
  auto *b = new SocgpuBatch();
  b->precision = precision;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess || !b->fork.make()) {
    delete b;
    return fail(SOCGPU_EDEVICE, "stream or event creation failed");
  }
  *batch = b;
  return 0;
}

static void batch_drop_graph(SocgpuBatch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int socgpu_batch_run(SocgpuBatch *b, int n_request, const SocgpuRequest *requests) {
  if (int rc = ensure_init()) return rc;
  if (!b || n_request <= 0 || !requests) return fail(SOCGPU_EINVAL, "null batch or empty request list");
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  std::vector<Resolved> res;
  Plan pl;
  if (int rc = plan_sizes(n_request, requests, b->precision, res, pl)) return rc;
  const SocgpuTile &tile = g_tile[b->precision == SOCGPU_MIXED ? 1 : 0];
  bool recapture = !b->exec || std::memcmp(&tile, &b->tile, sizeof tile) != 0;
  // grow capacity by a quarter over the need; a larger batch re-captures
  auto grow = [](size_t need, size_t &cap) {
    if (need <= cap) return false;
    cap = need + need / 4 + 64;
    return true;
  };
  bool g_ws = grow(pl.ws_doubles, b->cap_ws), g_in = grow(pl.in_bytes + 256, b->cap_in);
  bool g_out = grow(pl.out_doubles + 64, b->cap_out), g_desc = grow(desc_bytes(pl.n()), b->cap_desc);
  for (int k = 0; k < kNWork; k++)
    if (pl.need[k] > b->cap[k]) { b->cap[k] = pl.need[k] + pl.need[k] / 4 + 64; recapture = true; }
  if (g_ws || g_desc) recapture = true;
  if ((g_ws && b->ws.ensure(b->cap_ws * 8)) || (g_in && b->in_stage.ensure(b->cap_in)) ||
      (g_out && b->out_stage.ensure(b->cap_out * 8)) || (g_desc && b->desc.ensure(b->cap_desc)))
    return fail(SOCGPU_ENOMEM, "batch buffer allocation failed");
  char *ih = b->in_host.use(pl.in_bytes), *oh = b->out_host.use(pl.out_doubles * 8);
  if (g_desc || !b->pinned) {
    if (b->pinned) (void)hipHostFree(b->pinned);
    b->pinned = nullptr;
    HIPCHK(hipHostMalloc((void **)&b->pinned, b->cap_desc, hipHostMallocDefault));
  }
  hipStream_t s = b->stream;
  // the previous replay has finished (every run ends synchronised), so the
  // pinned descriptor block can be rewritten
  if (int rc = plan_place(res, pl, b->ws.as<double>(), b->in_stage.as<char>(), ih, b->out_stage.as<double>(), s))
    return rc;
  pack_desc(pl, b->pinned);
  HIPCHK(hipMemcpyAsync(b->desc.p, b->pinned, desc_bytes(pl.n()), hipMemcpyHostToDevice, s));
  const BatchHdr *dh = (const BatchHdr *)b->desc.p;
  ReqDesc *dq = (ReqDesc *)((char *)b->desc.p + sizeof(BatchHdr));
  if (recapture) {
    batch_drop_graph(b);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = launch_batch(dh, dq, b->cap, b->cap_desc / sizeof(ReqDesc) + 1, b->precision, s, &b->fork, nullptr, nullptr);
    hipGraph_t g = nullptr;
    hipError_t ec = hipStreamEndCapture(s, &g);
    if (rc) return rc;
    if (ec != hipSuccess) return fail(SOCGPU_EDEVICE, "graph capture failed: %s", hipGetErrorString(ec));
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->tile = tile;
    b->captures++;
  }
  HIPCHK(hipEventRecord(b->ev0, s));
  HIPCHK(hipGraphLaunch(b->exec, s));
  HIPCHK(hipEventRecord(b->ev1, s));
  // the error word back into the pinned descriptor block (outside the graph)
  HIPCHK(hipMemcpyAsync(&((BatchHdr *)b->pinned)->err, &dh->err, sizeof(int), hipMemcpyDeviceToHost, s));
  if (int rc = copy_back_issue(pl, b->out_stage.as<double>(), oh, s)) return rc;
  HIPCHK(hipStreamSynchronize(s));
  if (int rc = batch_err(pl, ((const BatchHdr *)b->pinned)->err)) return rc;
  copy_back_finish(pl, b->out_stage.as<double>(), oh);
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  b->replays++;
  return 0;
}

int socgpu_batch_stats(const SocgpuBatch *b, double *last_ms, int *captures, int *replays) {
  if (!b) return fail(SOCGPU_EINVAL, "null batch");
  if (last_ms) *last_ms = b->last_ms;
  if (captures) *captures = b->captures;
  if (replays) *replays = b->replays;
  return 0;
}

int socgpu_batch_destroy(SocgpuBatch *b) {
  if (!b) return 0;
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  for (DevBuf *x : {&b->ws, &b->in_stage, &b->out_stage, &b->desc})
    if (x->p) (void)hipFree(x->p);
  b->in_host.release();
  b->out_host.release();
  if (b->pinned) (void)hipHostFree(b->pinned);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  b->fork.drop();
  delete b;
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------- serving
// A request loop shaped like a host model's radiation calls: callers submit
// calls from any thread; a server thread admits them while their device
// footprint fits the HBM budget, packs everything admitted into one batch
// per free slot (a stream with its own buffers) and launches it, so calls
// that arrive while others are on the GPU join the next batch and run
// concurrently with the batch in flight (continuous batching).
namespace sg {

struct ServeItem {
  long long ticket;
  SocgpuRequest req;
  SocgpuRunesArgs args;  // copies: the caller's structs may go out of scope
  SocgpuDiag diag;
  std::string name;
  size_t footprint = 0;
  std::chrono::steady_clock::time_point t_submit;
  int status = 1;  // 1 pending, 0 done, < 0 error
};

struct ServeSlot {
  hipStream_t stream = nullptr;
  Fork fork;
  hipEvent_t done = nullptr;
  DevBuf ws, in_stage, out_stage, desc;
  char *out_host_used = nullptr;  // the pinned output block of the batch in flight, or null
  char *pinned = nullptr;
  size_t pinned_bytes = 0;
  std::vector<std::shared_ptr<ServeItem>> items;
  Plan plan;
  bool busy = false;
};

}  // namespace sg

struct SocgpuServer {
  SocgpuServeConfig cfg{};
  std::thread th;
  std::mutex mu;
  std::condition_variable cv, cv_done;
  bool stop = false;
  long long next_ticket = 1;
  std::deque<std::shared_ptr<sg::ServeItem>> pending, admitted;
  std::map<long long, std::shared_ptr<sg::ServeItem>> all;
  size_t used = 0;
  std::vector<sg::ServeSlot> slots;
  std::vector<double> lat_ms;
  long long completed = 0, batches = 0, batched_requests = 0, solves = 0;
  std::chrono::steady_clock::time_point t0;
  bool started = false;
};

namespace sg {

// Device bytes a call holds while admitted: workspace, staged inputs,
// staged outputs and its descriptor.
static size_t footprint_bytes(const SocgpuRequest &r, const CallDims &d, const SgSpec &S) {
  return padded_ws_doubles(d, S) * 8 + stage_bytes(r.args, d.NB) + ((size_t)d.P * (4 * (d.L + 1)) + 16) * 8 +
         sizeof(ReqDesc);
}

static void serve_finish_slot(SocgpuServer *sv, ServeSlot &sl, int status) {
  // an invalid input fails the call it belongs to; the batch's other calls
  // complete
  int bad = -1;
  if (status == 0 && sl.pinned) {
    const int err = ((const BatchHdr *)sl.pinned)->err;
    if (err > 0 && err - 1 < (int)sl.plan.order.size()) bad = sl.plan.x.caller[sl.plan.order[err - 1]];
  }
  if (status == 0) copy_back_finish(sl.plan, sl.out_stage.as<double>(), sl.out_host_used);
  const auto now = std::chrono::steady_clock::now();
  std::lock_guard<std::mutex> lk(sv->mu);
  for (size_t k = 0; k < sl.items.size(); k++) {
    auto &it = sl.items[k];
    it->status = (int)k == bad ? SOCGPU_EINVAL : status;
    sv->used -= it->footprint;
    if (status == 0) {
      sv->lat_ms.push_back(std::chrono::duration<double, std::milli>(now - it->t_submit).count());
      sv->completed++;
    }
  }
  sl.items.clear();
  sl.busy = false;
  sv->cv_done.notify_all();
  sv->cv.notify_all();
}

// Plans and launches one batch on a slot (the caller holds g_ctx.mu).
static int serve_launch(SocgpuServer *sv, ServeSlot &sl) {
  const int n = (int)sl.items.size();
  std::vector<SocgpuRequest> req(n);
  for (int i = 0; i < n; i++) req[i] = sl.items[i]->req;
  std::vector<Resolved> res;
  if (int rc = plan_sizes(n, req.data(), sv->cfg.precision, res, sl.plan)) return rc;
  const int nq = sl.plan.n();
  if (sl.ws.ensure(sl.plan.ws_doubles * 8) || sl.in_stage.ensure(sl.plan.in_bytes + 256) ||
      sl.out_stage.ensure(sl.plan.out_doubles * 8 + 64) || sl.desc.ensure(desc_bytes(nq)))
    return fail(SOCGPU_ENOMEM, "serving slot allocation failed");
  // host arrays cross PCIe array by array: gathering them into pinned
  // staging here would put the copies on this thread, which also launches
  // the next batch (the runtime's own pageable copies measured faster under
  // load)
  char *ih = nullptr;
  sl.out_host_used = nullptr;
  if (sl.pinned_bytes < desc_bytes(nq)) {
    if (sl.pinned) (void)hipHostFree(sl.pinned);
    sl.pinned = nullptr;
    sl.pinned_bytes = desc_bytes(nq) * 2;
    HIPCHK(hipHostMalloc((void **)&sl.pinned, sl.pinned_bytes, hipHostMallocDefault));
  }
  hipStream_t s = sl.stream;
  if (int rc = plan_place(res, sl.plan, sl.ws.as<double>(), sl.in_stage.as<char>(), ih, sl.out_stage.as<double>(), s))
    return rc;
  pack_desc(sl.plan, sl.pinned);
  HIPCHK(hipMemcpyAsync(sl.desc.p, sl.pinned, desc_bytes(nq), hipMemcpyHostToDevice, s));
  const BatchHdr *dh = (const BatchHdr *)sl.desc.p;
  ReqDesc *dq = (ReqDesc *)((char *)sl.desc.p + sizeof(BatchHdr));
  if (int rc = launch_batch(dh, dq, sl.plan.h.tot, nq, sv->cfg.precision, s, &sl.fork, nullptr, nullptr)) return rc;
  HIPCHK(hipMemcpyAsync(&((BatchHdr *)sl.pinned)->err, &dh->err, sizeof(int), hipMemcpyDeviceToHost, s));
  if (int rc = copy_back_issue(sl.plan, sl.out_stage.as<double>(), sl.out_host_used, s)) return rc;
  HIPCHK(hipEventRecord(sl.done, s));
  long long sol = 0;
  for (int i = 0; i < nq; i++) sol += (long long)res[i].d.P * res[i].e->host.n_g;
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->batches++;
  sv->batched_requests += n;
  sv->solves += sol;
  return 0;
}

static void serve_loop(SocgpuServer *sv) {
  (void)hipSetDevice(g_ctx.device);
  for (;;) {
    // completions
    for (auto &sl : sv->slots)
      if (sl.busy && hipEventQuery(sl.done) == hipSuccess) serve_finish_slot(sv, sl, 0);
    ServeSlot *free_slot = nullptr;
    std::vector<std::shared_ptr<ServeItem>> take;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      // admission by footprint, first come first served
      while (!sv->pending.empty() && sv->used + sv->pending.front()->footprint <= sv->cfg.hbm_budget_bytes) {
        auto it = sv->pending.front();
        sv->pending.pop_front();
        sv->used += it->footprint;
        sv->admitted.push_back(it);
      }
      for (auto &sl : sv->slots)
        if (!sl.busy) { free_slot = &sl; break; }
      if (free_slot && !sv->admitted.empty()) {
        while (!sv->admitted.empty() && (int)take.size() < sv->cfg.max_batch_requests) {
          take.push_back(sv->admitted.front());
          sv->admitted.pop_front();
        }
        free_slot->busy = true;
      }
      bool any_busy = false;
      for (auto &sl : sv->slots) any_busy |= sl.busy;
      if (sv->stop && sv->pending.empty() && sv->admitted.empty() && !any_busy && take.empty()) break;
      if (take.empty()) {
        if (any_busy) {
          lk.unlock();
          std::this_thread::sleep_for(std::chrono::microseconds(50));
        } else {
          sv->cv.wait_for(lk, std::chrono::milliseconds(5));
        }
        continue;
      }
    }
    free_slot->items = take;
    int rc;
    {
      std::lock_guard<std::mutex> g(g_ctx.mu);
      rc = serve_launch(sv, *free_slot);
    }
    if (rc) serve_finish_slot(sv, *free_slot, rc);
  }
}

}  // namespace sg

extern "C" {

int socgpu_serve_start(SocgpuServer **server, const SocgpuServeConfig *cfg) {
  if (int rc = ensure_init()) return rc;
  if (!server || !cfg) return fail(SOCGPU_EINVAL, "null argument");
  if (cfg->precision != SOCGPU_FP64 && cfg->precision != SOCGPU_MIXED) return fail(SOCGPU_EINVAL, "unknown precision");
  if (cfg->n_slots < 1 || cfg->n_slots > 8) return fail(SOCGPU_EINVAL, "n_slots must be 1..8");
  if (cfg->max_batch_requests < 1) return fail(SOCGPU_EINVAL, "max_batch_requests must be positive");
  auto *sv = new SocgpuServer();
  sv->cfg = *cfg;
  sv->slots.resize(cfg->n_slots);
  for (auto &sl : sv->slots) {
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess || !sl.fork.make()) {
      socgpu_serve_stop(sv);  // releases what the earlier slots hold
      return fail(SOCGPU_EDEVICE, "stream or event creation failed");
    }
  }
  sv->t0 = std::chrono::steady_clock::now();
  sv->th = std::thread(serve_loop, sv);
  *server = sv;
  return 0;
}

int socgpu_serve_submit(SocgpuServer *sv, const SocgpuRequest *request, long long *ticket) {
  if (!sv || !request || !request->args || !request->diag || !request->spectrum_name || !ticket)
    return fail(SOCGPU_EINVAL, "null argument");
  auto it = std::make_shared<ServeItem>();
  it->args = *request->args;
  it->diag = *request->diag;
  it->name = request->spectrum_name;
  {
    std::lock_guard<std::mutex> g(g_ctx.mu);
    SpecEntry *e = nullptr;
    CallDims d;
    if (int rc = resolve_request(it->name.c_str(), &it->args, sv->cfg.precision, &e, &d)) return rc;
    SocgpuRequest r{it->name.c_str(), &it->args, &it->diag};
    // with its clear-sky companion, if it asks for clear-sky fluxes
    Expanded x;
    if (int rc = expand_clear(1, &r, x)) return rc;
    for (const SocgpuRequest &q : x.req) {
      if (int rc = resolve_request(q.spectrum_name, q.args, sv->cfg.precision, &e, &d)) return rc;
      it->footprint += footprint_bytes(q, d, e->host);
    }
  }
  it->req = SocgpuRequest{it->name.c_str(), &it->args, &it->diag};
  if (it->footprint > sv->cfg.hbm_budget_bytes)
    return fail(SOCGPU_ENOMEM, "request needs %zu bytes, above the serving budget of %zu", it->footprint,
                sv->cfg.hbm_budget_bytes);
  it->t_submit = std::chrono::steady_clock::now();
  std::lock_guard<std::mutex> lk(sv->mu);
  if (sv->stop) return fail(SOCGPU_EINVAL, "server stopping");
  it->ticket = sv->next_ticket++;
  sv->all[it->ticket] = it;
  sv->pending.push_back(it);
  *ticket = it->ticket;
  sv->cv.notify_all();
  return 0;
}

int socgpu_serve_wait(SocgpuServer *sv, long long ticket, double timeout_s) {
  if (!sv) return fail(SOCGPU_EINVAL, "null server");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto f = sv->all.find(ticket);
  if (f == sv->all.end()) return fail(SOCGPU_EINVAL, "unknown ticket %lld", ticket);
  auto it = f->second;
  auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
  while (it->status == 1) {
    if (sv->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && it->status == 1)
      return fail(SOCGPU_EDEVICE, "ticket %lld not done after %.1f s", ticket, timeout_s);
  }
  sv->all.erase(ticket);
  if (it->status != 0) return fail(it->status, "request %lld failed", ticket);
  return 0;
}

int socgpu_serve_stats(SocgpuServer *sv, SocgpuServeStats *st) {
  if (!sv || !st) return fail(SOCGPU_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  std::memset(st, 0, sizeof *st);
  st->completed = sv->completed;
  st->batches = sv->batches;
  st->mean_batch = sv->batches ? (double)sv->batched_requests / sv->batches : 0.0;
  st->solves = sv->solves;
  st->elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - sv->t0).count();
  if (!sv->lat_ms.empty()) {
    std::vector<double> v = sv->lat_ms;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) { return v[std::min(v.size() - 1, (size_t)(p * (v.size() - 1) + 0.5))]; };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    st->max_ms = v.back();
  }
  return 0;
}

int socgpu_serve_reset_stats(SocgpuServer *sv) {
  if (!sv) return fail(SOCGPU_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->lat_ms.clear();
  sv->completed = sv->batches = sv->batched_requests = sv->solves = 0;
  sv->t0 = std::chrono::steady_clock::now();
  return 0;
}

int socgpu_serve_stop(SocgpuServer *sv) {
  if (!sv) return 0;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
    sv->cv.notify_all();
  }
  if (sv->th.joinable()) sv->th.join();
  for (auto &sl : sv->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    for (DevBuf *x : {&sl.ws, &sl.in_stage, &sl.out_stage, &sl.desc})
      if (x->p) (void)hipFree(x->p);
    if (sl.pinned) (void)hipHostFree(sl.pinned);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    sl.fork.drop();
  }
  delete sv;
  return 0;
}

int socgpu_serve_replay(SocgpuServer *sv, int n_pool, const SocgpuRequest *pool, int n_arrival,
                        const double *arrival_s, const int *pick, double timeout_s, double *wall_s) {
  if (!sv || n_pool <= 0 || !pool || n_arrival < 0 || (n_arrival > 0 && (!arrival_s || !pick)))
    return fail(SOCGPU_EINVAL, "null or empty argument");
  for (int i = 0; i < n_arrival; i++)
    if (pick[i] < 0 || pick[i] >= n_pool) return fail(SOCGPU_EINVAL, "pick %d out of range", i);
  std::vector<long long> tickets((size_t)n_arrival);
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < n_arrival; i++) {
    // sleep to within 200 us of the arrival, then spin
    const auto due = t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                              std::chrono::duration<double>(arrival_s[i]));
    for (;;) {
      const auto now = std::chrono::steady_clock::now();
      if (now >= due) break;
      if (due - now > std::chrono::microseconds(200)) std::this_thread::sleep_for(due - now - std::chrono::microseconds(150));
    }
    if (int rc = socgpu_serve_submit(sv, &pool[pick[i]], &tickets[i])) return rc;
  }
  int first_rc = 0;
  for (int i = 0; i < n_arrival; i++) {
    const int rc = socgpu_serve_wait(sv, tickets[i], timeout_s);
    if (rc && !first_rc) first_rc = rc;
  }
  if (wall_s) *wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return first_rc;
}

int socgpu_autotune(int n_request, const SocgpuRequest *requests, int precision, int repeats,
                    SocgpuTile *best) {
  if (int rc = ensure_init()) return rc;
  if (n_request <= 0 || !requests || !best) return fail(SOCGPU_EINVAL, "null or empty argument");
  if (precision != SOCGPU_FP64 && precision != SOCGPU_MIXED) return fail(SOCGPU_EINVAL, "unknown precision %d", precision);
  const int slot = precision == SOCGPU_MIXED ? 1 : 0;
  const SocgpuTile saved = g_tile[slot];
  SocgpuTile win = saved;
  double best_ms = 1e30;
  if (repeats < 1) repeats = 1;
  for (int rec = 0; rec <= 1; rec++)
    for (int wpb : {1, 2, 4})
      for (int occ : {3, 4}) {
        {
          std::lock_guard<std::mutex> lk(g_ctx.mu);
          g_tile[slot] = SocgpuTile{rec, wpb, occ};
        }
        double t = 1e30;
        for (int r = 0; r <= repeats; r++) {  // the first run warms up
          if (int rc = socgpu_runes_batch(n_request, requests, precision)) {
            std::lock_guard<std::mutex> lk(g_ctx.mu);
            g_tile[slot] = saved;
            return rc;
          }
          if (r > 0) t = std::min(t, g_ctx.last_solve_ms);
        }
        if (t < best_ms) { best_ms = t; win = SocgpuTile{rec, wpb, occ}; }
      }
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  g_tile[slot] = win;
  *best = win;
  return 0;
}

int socgpu_set_tile(int precision, const SocgpuTile *tile) {
  if (precision != SOCGPU_FP64 && precision != SOCGPU_MIXED) return fail(SOCGPU_EINVAL, "unknown precision %d", precision);
  if (!tile) return fail(SOCGPU_EINVAL, "null tile");
  const int w = tile->waves_per_block;
  if (w != 1 && w != 2 && w != 4) return fail(SOCGPU_EINVAL, "waves_per_block must be 1, 2 or 4");
  if (tile->waves_per_simd != 3 && tile->waves_per_simd != 4)
    return fail(SOCGPU_EINVAL, "waves_per_simd must be 3 or 4");
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  g_tile[precision] = *tile;
  g_tile[precision].recompute = tile->recompute != 0;
  return 0;
}

int socgpu_get_tile(int precision, SocgpuTile *tile) {
  if (precision != SOCGPU_FP64 && precision != SOCGPU_MIXED) return fail(SOCGPU_EINVAL, "unknown precision %d", precision);
  if (!tile) return fail(SOCGPU_EINVAL, "null tile");
  *tile = g_tile[precision];
  return 0;
}

int socgpu_last_timing(double *total_ms, double *solve_ms) {
  if (total_ms) *total_ms = g_ctx.last_total_ms;
  if (solve_ms) *solve_ms = g_ctx.last_solve_ms;
  return 0;
}

int socgpu_two_stream(int n_profile, int n_layer, int n_term, int isolir, int l_scale_solar,
                      int l_ir_source_quad, const double *tau, const double *omega,
                      const double *asymmetry, const double *sec_0, const double *flux_inc_down,
                      const double *flux_inc_direct, const double *diffuse_albedo,
                      const double *direct_albedo, const double *diff_planck,
                      const double *diff_planck_2, const double *d_planck_flux_surface,
                      const double *adjust_solar_ke, double *flux_direct, double *flux_up,
                      double *flux_down, int l_on_device, int precision) {
  if (int rc = ensure_init()) return rc;
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  const int P = n_profile, L = n_layer, G = n_term;
  if (P <= 0 || L <= 0 || G <= 0) return fail(SOCGPU_EINVAL, "n_profile, n_layer and n_term must be positive");
  if (L > SG_MAXL) return fail(SOCGPU_EUNSUPPORTED, "n_layer %d above %d", L, SG_MAXL);
  if ((size_t)P * G > (size_t)INT32_MAX) return fail(SOCGPU_EINVAL, "n_profile * n_term too large");
  if (precision != SOCGPU_FP64 && precision != SOCGPU_MIXED) return fail(SOCGPU_EINVAL, "unknown precision %d", precision);
  const bool solar = isolir == SOCGPU_SOURCE_SOLAR;
  if (!solar && isolir != SOCGPU_SOURCE_THERMAL) return fail(SOCGPU_EINVAL, "Radiative source option unrecognised.");
  if (!tau || !omega || !asymmetry || !flux_inc_down || !diffuse_albedo || !flux_up || !flux_down)
    return fail(SOCGPU_EINVAL, "tau, omega, asymmetry, flux_inc_down, diffuse_albedo and the outputs are required");
  if (solar && (!sec_0 || !flux_inc_direct || !direct_albedo || (l_scale_solar && !adjust_solar_ke)))
    return fail(SOCGPU_EINVAL, "solar solve needs sec_0, flux_inc_direct, direct_albedo (and adjust_solar_ke when scaled)");
  if (!solar && (!diff_planck || !d_planck_flux_surface || (l_ir_source_quad && !diff_planck_2)))
    return fail(SOCGPU_EINVAL, "infra-red solve needs diff_planck, d_planck_flux_surface (and diff_planck_2 when quadratic)");
  const size_t nopt = (size_t)G * L * P, nlay = (size_t)L * P, nout = (size_t)G * (L + 1) * P;
  TwoStreamArgs a;
  a.P = P; a.L = L; a.G = G; a.isolir = isolir; a.l_scale_solar = l_scale_solar != 0;
  a.l_quad = l_ir_source_quad != 0;
  hipStream_t s = g_ctx.stream;
  if (l_on_device) {
    a.tau = tau; a.omega = omega; a.asym = asymmetry; a.sec0 = sec_0; a.finc_down = flux_inc_down;
    a.finc_dir = flux_inc_direct; a.alb_diff = diffuse_albedo; a.alb_dir = direct_albedo;
    a.dpl = diff_planck; a.dpl2 = diff_planck_2; a.dsurf = d_planck_flux_surface; a.adjust = adjust_solar_ke;
    a.fdir = flux_direct; a.fup = flux_up; a.fdown = flux_down;
  } else {
    struct F { const double **dst; const double *src; size_t n; };
    F f[] = {{&a.tau, tau, nopt}, {&a.omega, omega, nopt}, {&a.asym, asymmetry, nopt},
             {&a.sec0, sec_0, (size_t)P}, {&a.finc_down, flux_inc_down, (size_t)P},
             {&a.finc_dir, flux_inc_direct, (size_t)P}, {&a.alb_diff, diffuse_albedo, (size_t)P},
             {&a.alb_dir, direct_albedo, (size_t)P}, {&a.dpl, diff_planck, nlay},
             {&a.dpl2, diff_planck_2, nlay}, {&a.dsurf, d_planck_flux_surface, (size_t)P},
             {&a.adjust, adjust_solar_ke, nlay}};
    size_t total = 0;
    for (auto &x : f) if (x.src) total += x.n;
    if (g_ctx.in_stage.ensure(total * 8 + 64) || g_ctx.out_stage.ensure(3 * nout * 8))
      return fail(SOCGPU_ENOMEM, "staging allocation failed");
    double *base = g_ctx.in_stage.as<double>();
    size_t off = 0;
    for (auto &x : f) {
      if (!x.src) { *x.dst = nullptr; continue; }
      HIPCHK(hipMemcpyAsync(base + off, x.src, x.n * 8, hipMemcpyHostToDevice, s));
      *x.dst = base + off;
      off += x.n;
    }
    double *ob = g_ctx.out_stage.as<double>();
    a.fup = ob; a.fdown = ob + nout; a.fdir = flux_direct ? ob + 2 * nout : nullptr;
  }
  if (!solar) a.fdir = nullptr;
  const int threads = 256, blocks = (int)(((size_t)P * G + threads - 1) / threads);
  HIPCHK(hipEventRecord(g_ctx.ev1, s));
  if (precision == SOCGPU_MIXED) two_stream_kernel<true><<<blocks, threads, 0, s>>>(a);
  else two_stream_kernel<false><<<blocks, threads, 0, s>>>(a);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(g_ctx.ev2, s));
  if (!l_on_device) {
    HIPCHK(hipMemcpyAsync(flux_up, a.fup, nout * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(flux_down, a.fdown, nout * 8, hipMemcpyDeviceToHost, s));
    if (a.fdir) HIPCHK(hipMemcpyAsync(flux_direct, a.fdir, nout * 8, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, g_ctx.ev1, g_ctx.ev2));
  g_ctx.last_total_ms = ms;
  g_ctx.last_solve_ms = ms;
  return 0;
}

int socgpu_cloud_gen(int n_profile, int n_layer, int n_subcol, int n_req, int ipph, int ioverlap,
                     const int *rand_seed, const double *rlc_cf, const double *rlc_cw,
                     const double *sigma_qcw, const double *avg_cf, const double *p_layer, int n1, int n2,
                     const double *xcw, double *c_sub, int *n_subcol_cld) {
  if (int rc = ensure_init()) return rc;
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  const int P = n_profile, L = n_layer;
  if (P <= 0 || L <= 0 || L > SG_MAXL || n_subcol <= 0 || n_subcol > SG_MAXSUBCOL || n1 < 2 || n2 < 2)
    return fail(SOCGPU_EINVAL, "bad cloud_gen sizes");
  std::vector<double> rec((size_t)P * L * SG_NQ, 0.0);
  for (int i = 0; i < L; i++)
    for (int l = 0; l < P; l++) {
      double *r = &rec[((size_t)l * L + i) * SG_NQ];
      r[SG_Q_WCLD] = avg_cf[(size_t)i * P + l];
      r[SG_Q_P] = p_layer[(size_t)i * P + l];
      r[SG_Q_RSD] = sigma_qcw[(size_t)i * P + l];
    }
  CallDims d;
  std::memset(&d, 0, sizeof d);
  d.P = P; d.L = L; d.n_gen = n_subcol; d.n_fill = n_req < n_subcol ? n_req : n_subcol;
  d.ipph = ipph; d.ioverlap = ioverlap; d.mcica = 1; d.raw_range = 1;
  DevBuf drec, dseed, dxcw, dcs, dscal, dk, ddesc, dcf;
  struct {
    BatchHdr h;
    ReqDesc q;
  } host;
  if (drec.ensure(rec.size() * 8) || dseed.ensure((size_t)P * 4) || dxcw.ensure((size_t)n1 * n2 * 8) ||
      dcs.ensure((size_t)P * L * SG_MAXSUBCOL * 8) || dscal.ensure((size_t)P * SG_NS * 8) || dk.ensure(8) ||
      ddesc.ensure(sizeof host) || dcf.ensure((size_t)P * L * 8))
    return fail(SOCGPU_ENOMEM, "allocation failed");
  std::memset(&host, 0, sizeof host);
  host.h.R = 1;
  host.h.tot[W_LAY] = (long long)P * L;
  host.h.tot[W_CG] = P;
  host.h.tot[W_CCOL] = (long long)P * ((L + kRangeChunk - 1) / kRangeChunk);
  host.q.a.cloud_frac = dcf.as<double>();
  host.q.d = d;
  host.q.colrec = drec.as<double>();
  host.q.a.rand_seed = dseed.as<int>();
  host.q.dp_cf = rlc_cf[0];
  host.q.dp_cw = rlc_cw[0];
  host.q.krange = dk.as<int>();
  host.q.xcw = dxcw.as<double>();
  host.q.n1 = n1;
  host.q.n2 = n2;
  host.q.csub = dcs.as<double>();
  host.q.colscal = dscal.as<double>();
  hipStream_t s = g_ctx.stream;
  HIPCHK(hipMemcpyAsync(drec.p, rec.data(), rec.size() * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dseed.p, rand_seed, (size_t)P * 4, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dcf.p, avg_cf, (size_t)P * L * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dxcw.p, xcw, (size_t)n1 * n2 * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(ddesc.p, &host, sizeof host, hipMemcpyHostToDevice, s));
  const BatchHdr *dh = (const BatchHdr *)ddesc.p;
  ReqDesc *dq = (ReqDesc *)((char *)ddesc.p + offsetof(decltype(host), q));
  b_init_kernel<<<1, 64, 0, s>>>(dh, dq);
  b_cldrange_kernel<<<(int)((host.h.tot[W_CCOL] + 255) / 256), 256, 0, s>>>(dh, dq);
  b_cloudgen_kernel<<<(P + kCgWaves - 1) / kCgWaves, 64 * kCgWaves, 0, s>>>(dh, dq);
  HIPCHK(hipGetLastError());
  std::vector<double> cs((size_t)P * L * SG_MAXSUBCOL), sc((size_t)P * SG_NS);
  HIPCHK(hipMemcpyAsync(cs.data(), dcs.p, cs.size() * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(sc.data(), dscal.p, sc.size() * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int l = 0; l < P; l++) {
    n_subcol_cld[l] = (int)sc[(size_t)l * SG_NS + SG_S_NCLD];
    for (int i = 0; i < L; i++)
      for (int j = 0; j < n_subcol; j++)
        c_sub[((size_t)j * L + i) * P + l] = cs[((size_t)l * L + i) * SG_MAXSUBCOL + j];
  }
  for (DevBuf *b : {&drec, &dseed, &dxcw, &dcs, &dscal, &dk, &ddesc, &dcf}) if (b->p) (void)hipFree(b->p);
  return 0;
}

int socgpu_finalize(void) {
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  if (g_ctx.device < 0) return 0;
  (void)hipStreamSynchronize(g_ctx.stream);
  std::vector<double *> xcws;
  for (auto &kv : g_ctx.spectra) {
    (void)hipFree(kv.second->dev);
    if (kv.second->tables) (void)hipFree(kv.second->tables);
    if (kv.second->xcw && std::find(xcws.begin(), xcws.end(), kv.second->xcw) == xcws.end())
      xcws.push_back(kv.second->xcw);
  }
  for (double *p : xcws) (void)hipFree(p);
  g_ctx.spectra.clear();
  g_ctx.in_host.release();
  g_ctx.out_host.release();
  for (DevBuf *b : {&g_ctx.ws, &g_ctx.desc, &g_ctx.in_stage, &g_ctx.out_stage}) {
    if (b->p) (void)hipFree(b->p);
    b->p = nullptr;
    b->n = 0;
  }
  if (g_ctx.err_host) (void)hipHostFree(g_ctx.err_host);
  g_ctx.err_host = nullptr;
  g_ctx.scratch_dev = nullptr;
  (void)hipEventDestroy(g_ctx.ev0); (void)hipEventDestroy(g_ctx.ev1);
  (void)hipEventDestroy(g_ctx.ev2); (void)hipEventDestroy(g_ctx.ev3);
  (void)hipStreamDestroy(g_ctx.stream);
  g_ctx.fork.drop();
  g_ctx.device = -1;
  return 0;
}

}  // extern "C"

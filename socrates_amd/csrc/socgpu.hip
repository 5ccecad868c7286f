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

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
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

// Butterfly reduce-scatter of 16 values per lane over the 64 lanes of a
// wave: afterwards lane 4j (j = 0..15) holds the wave total of value j in
// v[0] (lanes 4j+1..4j+3 hold the same total). 17 shuffles instead of the 96
// of sixteen separate all-reductions.
__device__ __forceinline__ void reduce16(double (&v)[16], int lane) {
  {
    const bool hi = lane & 32;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      double keep = hi ? v[k + 8] : v[k], give = hi ? v[k] : v[k + 8];
      v[k] = keep + __shfl_xor(give, 32, kWave);
    }
  }
  {
    const bool hi = lane & 16;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      double keep = hi ? v[k + 4] : v[k], give = hi ? v[k] : v[k + 4];
      v[k] = keep + __shfl_xor(give, 16, kWave);
    }
  }
  {
    const bool hi = lane & 8;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      double keep = hi ? v[k + 2] : v[k], give = hi ? v[k] : v[k + 2];
      v[k] = keep + __shfl_xor(give, 8, kWave);
    }
  }
  {
    const bool hi = lane & 4;
    double keep = hi ? v[1] : v[0], give = hi ? v[0] : v[1];
    v[0] = keep + __shfl_xor(give, 4, kWave);
  }
  v[0] += __shfl_xor(v[0], 2, kWave);
  v[0] += __shfl_xor(v[0], 1, kWave);
}

// Same for 8 values per lane: lane 8j (j = 0..7) holds the total of value j.
__device__ __forceinline__ void reduce8(double (&v)[8], int lane) {
  {
    const bool hi = lane & 32;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      double keep = hi ? v[k + 4] : v[k], give = hi ? v[k] : v[k + 4];
      v[k] = keep + __shfl_xor(give, 32, kWave);
    }
  }
  {
    const bool hi = lane & 16;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      double keep = hi ? v[k + 2] : v[k], give = hi ? v[k] : v[k + 2];
      v[k] = keep + __shfl_xor(give, 16, kWave);
    }
  }
  {
    const bool hi = lane & 8;
    double keep = hi ? v[1] : v[0], give = hi ? v[0] : v[1];
    v[0] = keep + __shfl_xor(give, 8, kWave);
  }
  v[0] += __shfl_xor(v[0], 4, kWave);
  v[0] += __shfl_xor(v[0], 2, kWave);
  v[0] += __shfl_xor(v[0], 1, kWave);
}

struct CallDims {
  int P, L, NB, isolir;
  int mcica, mode;   // McICA on, sampling table (0 single, 1 optimal)
  int drop;          // droplet type index (0-based)
  int nwpc;          // waves per column in the solve kernel
  int n_gen, n_fill, ipph, ioverlap;
};

// Caller arrays -> column records. set_atm (socrates_set_atm.F90:246-300),
// set_cld (socrates_set_cld.F90:817-852: in-cloud mixing ratio, w_cloud
// threshold 0.001, type-fraction normalisation), set_cld_dim
// (socrates_set_cld_dim.F90:214-220: radius clamp), cond_rsd of
// set_cld_mcica (socrates_set_cld_mcica.F90:168-180).
__global__ void pack_kernel(CallDims d, const SgSpec *__restrict__ spec, SocgpuRunesArgs a,
                            double *__restrict__ colrec, double *__restrict__ tlev,
                            double *__restrict__ colscal, double *__restrict__ bandscal) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int P = d.P, L = d.L;
  if (idx < P * L) {
    int col = idx % P, i = idx / P;
    double p = a.p_layer[idx], t = a.t_layer[idx];
    double *r = colrec + ((size_t)col * L + i) * SG_NQ;
    r[SG_Q_LNP] = log(p);
    r[SG_Q_LNT] = log(t);
    r[SG_Q_MASS] = a.mass[idx];
    r[SG_Q_MMR + SG_GAS_H2O] = a.h2o ? fmax(a.h2o[idx], 0.0) : 0.0;
    r[SG_Q_MMR + SG_GAS_CO2] = fmax(a.co2_mix_ratio, 0.0);
    r[SG_Q_MMR + SG_GAS_O3] = a.o3 ? fmax(a.o3[idx], 0.0) : 0.0;
    r[SG_Q_MMR + SG_GAS_ZERO] = 0.0;
    r[SG_Q_P] = p;
    r[SG_Q_T] = t;
    double w = 0.0, lwc = 0.0, re = 0.0, f1 = 0.0, rsd = 0.0;
    if (a.cloud_frac) {
      const double eps = kEps;
      double frac = fmax(a.cloud_frac[idx], 0.0);
      double mmr = a.liq_mmr ? fmax(a.liq_mmr[idx], 0.0) : 0.0;
      lwc = frac > eps ? mmr / frac : 0.0;
      w = frac;
      f1 = frac;
      if (w > 0.001) f1 = f1 / w;
      else { w = 0.0; f1 = 0.0; }
      if (w > 1.0) w = 1.0;
      const SgDrop &dr = spec->drop[d.drop];
      double r0 = a.liq_dim ? fmax(a.liq_dim[idx], 0.0) : 0.0;
      re = fmin(fmax(r0, dr.dim_min), dr.dim_max);
      // liquid component and the zero-mixing-ratio ice component, both with
      // the call's horizontal relative standard deviation
      double wl = fmax(eps, lwc * f1), wi = fmax(eps, 0.0 * f1);
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
  }
  if (idx < P * (L + 1)) {
    int col = idx % P, i = idx / P;
    tlev[(size_t)col * (L + 1) + i] = a.t_level ? a.t_level[idx] : 0.0;
  }
  if (idx < P) {
    int col = idx;
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
    }
  }
}

// Cloud top and base over all columns of the call, as cloud_gen finds them
// (socrates_cloud_gen.F90:224-245): first layer with any w > epsilon, last
// layer with any w >= epsilon. krange[0] = min, krange[1] = max (1-based).
__global__ void cldrange_kernel(CallDims d, const double *__restrict__ colrec, int *krange) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  int kmin = INT32_MAX, kmax = 0;
  if (idx < d.P * d.L) {
    int col = idx / d.L, i = idx % d.L;
    double w = colrec[((size_t)col * d.L + i) * SG_NQ + SG_Q_WCLD];
    if (w > kEps) kmin = i + 1;
    if (w >= kEps) kmax = i + 1;
  }
  // one atomic per wave: the whole call shares one cloud top and base
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    kmin = min(kmin, __shfl_xor(kmin, o));
    kmax = max(kmax, __shfl_xor(kmax, o));
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    if (kmin != INT32_MAX) atomicMin(&krange[0], kmin);
    if (kmax != 0) atomicMax(&krange[1], kmax);
  }
}

__global__ void fill_int_kernel(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
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

__global__ __launch_bounds__(64) void cloudgen_kernel(CallDims d, const double *__restrict__ colrec,
                                                      const int *__restrict__ seeds,
                                                      double dp_cf, double dp_cw,
                                                      const int *__restrict__ krange,
                                                      const double *__restrict__ xcw, int n1, int n2,
                                                      double *__restrict__ c_sub,
                                                      double *__restrict__ colscal) {
#pragma clang fp contract(off)
  const int col = blockIdx.x, lane = threadIdx.x, L = d.L, n_gen = d.n_gen;
  const double rm = 1.0 / (double)kLcgM;
  const int seed_max = (2147483647 - kLcgC) / kLcgA;
  double *cs = c_sub + (size_t)col * L * SG_MAXSUBCOL;
  const double *rec = colrec + (size_t)col * L * SG_NQ;
  const int k_top = krange[0], k_base = krange[1];
  const bool valid = lane < n_gen;
  // seeds of this sub-column (cloud_gen:180-205)
  int rs = abs(seeds[col]) % seed_max;
  for (int t = 0; t < n_gen; t++) rs = lcg_step(rs);
  int sx = 0, sy = 0;
  for (int t = 0; t < n_gen; t++) { rs = lcg_step(rs); if (t == lane) sx = rs; }
  for (int t = 0; t < n_gen; t++) { rs = lcg_step(rs); if (t == lane) sy = rs; }
  int ix = sx, iy = sy;
  double x = 0.0, y = 0.0;
  bool cloudy = false;
  const bool any = k_top <= L;
  for (int k = 1; k <= L; k++) {
    double val = 0.0;
    double cf = rec[(k - 1) * SG_NQ + SG_Q_WCLD];
    if (valid && any && k >= k_top && k <= k_base && cf > 0.0) {
      if (d.ioverlap == 0) {
        // maximum-random (cloud_gen:253-298)
        ix = lcg_step(ix); x = (double)ix * rm;
        iy = lcg_step(iy); y = (double)iy * rm;
        if (k == 1) {
          ix = lcg_step(ix); x = (double)ix * rm;
          iy = lcg_step(iy); y = (double)iy * rm;
        } else {
          double cfa = rec[(k - 2) * SG_NQ + SG_Q_WCLD];
          if (x < (1.0 - cfa)) {
            ix = lcg_step(ix); double x1 = (double)ix * rm;
            iy = lcg_step(iy); double y1 = (double)iy * rm;
            x = x1 * (1.0 - cfa);
            y = y1;
          }
        }
      } else {
        // random / exponential-random (cloud_gen:339-405)
        double alpha = 0.0, rcorr = 0.0;
        if (d.ioverlap != 1 && k - 1 >= 1) {
          double cfa = rec[(k - 2) * SG_NQ + SG_Q_WCLD];
          if (cfa > 0.0) {
            double dz = rec[(k - 2) * SG_NQ + SG_Q_P] - rec[(k - 1) * SG_NQ + SG_Q_P];
            alpha = dp_cf > 0.0 ? exp(dz / dp_cf) : 0.0;
            rcorr = dp_cw > 0.0 ? exp(dz / dp_cw) : 0.0;
          }
        }
        ix = lcg_step(ix); double x1 = (double)ix * rm;
        iy = lcg_step(iy); double y1 = (double)iy * rm;
        if (x1 >= alpha) { ix = lcg_step(ix); x = (double)ix * rm; }
        if (y1 >= rcorr) { iy = lcg_step(iy); y = (double)iy * rm; }
      }
      if (x < cf) {
        val = d.ipph == 0 ? 1.0
                          : zcw_lookup(y, rec[(k - 1) * SG_NQ + SG_Q_RSD], n1, n2, xcw) * 1.0;
        cloudy = true;
      }
    }
    cs[(size_t)(k - 1) * SG_MAXSUBCOL + lane] = val;
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
  __syncthreads();
  for (int k = 0; k < L; k++) {
    double v = src >= 0 ? cs[(size_t)k * SG_MAXSUBCOL + src] : 0.0;
    __syncthreads();
    cs[(size_t)k * SG_MAXSUBCOL + lane] = v;
    __syncthreads();
  }
  if (lane == 0) {
    colscal[(size_t)col * SG_NS + SG_S_NCLD] = (double)ncld;
    colscal[(size_t)col * SG_NS + SG_S_FC] = (double)ncld / (double)n_gen;
  }
}

// Wave-wide inclusive scans over the 64 lanes (lane order = scan order).
__device__ __forceinline__ double scan_mul(double x, int lane) {
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    double y = __shfl_up(x, off, kWave);
    if (lane >= off) x *= y;
  }
  return x;
}

// Composition of affine maps F -> A F + C in lane order: lane j ends with
// the map of lanes 0..j applied in turn.
__device__ __forceinline__ void scan_affine(double &A, double &C, int lane) {
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    double A0 = __shfl_up(A, off, kWave), C0 = __shfl_up(C, off, kWave);
    if (lane >= off) {
      C = A * C0 + C;
      A = A * A0;
    }
  }
}

// Minor-gas weights for the equivalent extinction, one wave per (column,
// minor k-term), lanes over layers (64 at a time with a carry). Solar: the
// direct-beam transmission of the term down to each layer
// (solve_band_k_eqv_scl.F90:654-739). Infra-red: the non-scattering fluxes
// of monochromatic_gas_flux (monochromatic_gas_flux.F90:115-196,
// diffusivity 1.66) as two affine recurrences, solved by wave scans.
// minor[col][term][layer][2]: solar (fdt, k fdt(surface)), infra-red
// (k |F| sum, |F| sum) with the term's weight. The scans associate the
// products differently from the reference's left-to-right sweep (relative
// differences of order 1e-16 in the weights).
constexpr int kMinorWaves = 4;

__global__ __launch_bounds__(64 * kMinorWaves) void minor_kernel(CallDims d, const SgSpec *__restrict__ spec,
                                                                const double *__restrict__ colrec,
                                                                const double *__restrict__ tlev,
                                                                const double *__restrict__ colscal,
                                                                const double *__restrict__ bandscal,
                                                                double *__restrict__ minor) {
  __shared__ double sh[kMinorWaves][4][SG_MAXL];
  const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int wave = blockIdx.x * kMinorWaves + wv;
  const int NT = spec->n_minor, L = d.L;
  if (wave >= d.P * NT) return;
  const int col = wave / NT, ti = wave % NT;
  const SgTerm &t = spec->minor[ti];
  const SgBand &B = spec->band[t.band];
  const double *rec = colrec + (size_t)col * L * SG_NQ;
  double *o = minor + ((size_t)col * NT + ti) * L * 2;
  double *kk = sh[wv][0];
  const int nch = (L + kWave - 1) / kWave;
  if (d.isolir == SOCGPU_SOURCE_SOLAR) {
    const double sec0 = colscal[(size_t)col * SG_NS + SG_S_SEC0];
    double fdt = t.w;
    for (int ch = 0; ch < nch; ch++) {
      const int i = ch * kWave + lane;
      double x = 1.0;
      if (i < L) {
        const double *r = rec + i * SG_NQ;
        double k = k_term(t, r[SG_Q_LNP], r[SG_Q_LNT], r[SG_Q_MMR + t.gas]);
        kk[i] = k;
        x = exp(-k * r[SG_Q_MASS] * sec0);
      }
      double f = fdt * scan_mul(x, lane);
      if (i < L) o[2 * i] = f;
      fdt = __shfl(f, kWave - 1, kWave);
    }
    __syncwarp();
    for (int i = lane; i < L; i += kWave) o[2 * i + 1] = kk[i] * fdt;
  } else {
    double *tr = sh[wv][1], *su = sh[wv][2], *fdl = sh[wv][3];
    const double *tl = tlev + (size_t)col * (L + 1);
    const double alb = bandscal[((size_t)col * d.NB + t.band) * SG_NBQ + 2];
    const double sqe = 1.4901161193847656e-08;
    const double b_top = planck_band(B, spec->n_deg, spec->t_ref_planck, tl[0]);
    double F = -b_top;  // differential downward flux at the top of the chunk
    for (int ch = 0; ch < nch; ch++) {
      const int i = ch * kWave + lane;
      double A = 1.0, C = 0.0;
      if (i < L) {
        const double *r = rec + i * SG_NQ;
        double k = k_term(t, r[SG_Q_LNP], r[SG_Q_LNT], r[SG_Q_MMR + t.gas]);
        double tau = k * r[SG_Q_MASS];
        double b0 = planck_band(B, spec->n_deg, spec->t_ref_planck, tl[i]);
        double b1 = planck_band(B, spec->n_deg, spec->t_ref_planck, tl[i + 1]);
        double trans = exp(-kDiffMinor * tau);
        double s = (1.0 - trans + sqe) * (b1 - b0) / (kDiffMinor * tau + sqe);
        kk[i] = k; tr[i] = trans; su[i] = s;
        A = trans;
        C = -s;
      }
      scan_affine(A, C, lane);
      double below = A * F + C;                       // F at the bottom of layer i
      double above = __shfl_up(below, 1, kWave);      // F at its top
      if (lane == 0) above = F;
      if (i < L) fdl[i] = above;
      F = __shfl(below, kWave - 1, kWave);
    }
    const double b_bot = planck_band(B, spec->n_deg, spec->t_ref_planck, tl[L]);
    const double bg = planck_band(B, spec->n_deg, spec->t_ref_planck,
                                  colscal[(size_t)col * SG_NS + SG_S_TGROUND]);
    double U = (bg - b_bot) + alb * F;                // upward flux at the surface
    __syncwarp();
    for (int ch = 0; ch < nch; ch++) {
      const int i = L - 1 - (ch * kWave + lane);      // bottom-up
      double A = 1.0, C = 0.0;
      if (i >= 0) { A = tr[i]; C = su[i]; }
      scan_affine(A, C, lane);
      double above = A * U + C;                       // F up at the top of layer i
      double below = __shfl_up(above, 1, kWave);      // at its bottom
      if (lane == 0) below = U;
      if (i >= 0) {
        double a = t.w * fabs(fdl[i]), bb = t.w * fabs(below);
        o[2 * i] = kk[i] * a + kk[i] * bb;
        o[2 * i + 1] = a + bb;
      }
      U = __shfl(above, kWave - 1, kWave);
    }
  }
}

// Per (column, band, layer): the grey extinction (Rayleigh, grey_opt_prop.F90:
// 476-541, 816-835) plus the equivalent extinction of the minor gases from
// the per-term sums (solve_band_k_eqv_scl.F90:826-846, 995-1019), the solar
// direct-beam adjustment (ibid. :848-891), and the Planck source terms
// (diff_planck_source_mod.F90:100-134). bandrec[col][band][layer][4].
__global__ __launch_bounds__(256) void prep_kernel(CallDims d, const SgSpec *__restrict__ spec,
                                                   int l_rayleigh, const double *__restrict__ colrec,
                                                   const double *__restrict__ tlev,
                                                   const double *__restrict__ colscal,
                                                   const double *__restrict__ minor,
                                                   double *__restrict__ bandrec,
                                                   double *__restrict__ bandscal,
                                                   double *__restrict__ cldopt) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int NB = d.NB, L = d.L, NT = spec->n_minor;
  if (idx >= d.P * NB * L) return;
  const int i = idx % L, b = (idx / L) % NB, col = idx / (L * NB);
  const SgBand &B = spec->band[b];
  const double *rec = colrec + (size_t)col * L * SG_NQ;
  const double *r = rec + i * SG_NQ;
  if (d.mcica) {
    // droplet optics of the layer's in-cloud water at unit condensate
    // scaling (they depend on the band, not on the k-term; the solve scales
    // them by each sub-column's c_sub)
    double *co = cldopt + (((size_t)col * NB + b) * L + i) * 4;
    double kext = 0.0, kscat = 0.0, phf = 0.0, fwd = 0.0;
    if (r[SG_Q_FRAC1] > 0.0) {
      const SgDrop &dr = spec->drop[d.drop];
      if (dr.scheme == 5) pade2(dr.parm[b], r[SG_Q_LWC], r[SG_Q_RE], kext, kscat, phf, fwd);
      else slingo(dr.parm[b], r[SG_Q_LWC], r[SG_Q_RE], kext, kscat, phf, fwd);
    }
    co[0] = kext; co[1] = kscat; co[2] = phf; co[3] = fwd;
  }
  double *br = bandrec + (((size_t)col * NB + b) * L + i) * SG_NBQ;
  double *bs = bandscal + ((size_t)col * NB + b) * SG_NBQ;
  const bool solar = d.isolir == SOCGPU_SOURCE_SOLAR;
  const double sec0 = colscal[(size_t)col * SG_NS + SG_S_SEC0];
  const double *mc = minor + (size_t)col * NT * L * 2;
  double keq = 0.0, adjust = 1.0;
  int t0 = B.minor0;
  for (int j = 0; j < B.n_minor_gas; j++) {
    const int nk = B.minor_nk[j];
    if (solar) {
      double num = 0.0, sflux = 0.0, fg = 0.0, fgp = 0.0;
      for (int q = 0; q < nk; q++) {
        const double *o = mc + (size_t)(t0 + q) * L * 2;
        num += o[2 * i + 1];
        sflux += o[2 * (L - 1)];
        fg += o[2 * i];
        fgp += i > 0 ? o[2 * (i - 1)] : 0.0;
      }
      if (i == 0) fgp = 1.0;
      double kq;
      if (sflux > 0.0) kq = num / sflux;
      else {
        kq = DBL_MAX;
        for (int q = 0; q < nk; q++)
          kq = fmin(kq, k_term(spec->minor[t0 + q], r[SG_Q_LNP], r[SG_Q_LNT],
                               r[SG_Q_MMR + spec->minor[t0 + q].gas]));
      }
      keq += kq;
      if (fgp > 0.0) adjust = adjust * fg / fgp;
    } else {
      double num = 0.0, den = 0.0;
      for (int q = 0; q < nk; q++) {
        const double *o = mc + ((size_t)(t0 + q) * L + i) * 2;
        num += o[0];
        den += o[1];
      }
      double kq;
      if (den > 0.0) kq = num / den;
      else {
        kq = DBL_MAX;
        for (int q = 0; q < nk; q++)
          kq = fmin(kq, k_term(spec->minor[t0 + q], r[SG_Q_LNP], r[SG_Q_LNT],
                               r[SG_Q_MMR + spec->minor[t0 + q].gas]));
      }
      keq += kq;
    }
    t0 += nk;
  }
  const double ray = (solar && l_rayleigh) ? B.rayleigh : 0.0;
  br[0] = (0.0 + ray) + keq;
  if (solar) {
    const double temp_max = 36.043653389117154;  // log(1/DBL_EPSILON)
    br[1] = adjust * exp(fmin(keq * r[SG_Q_MASS] * sec0, temp_max));
    br[2] = 0.0;
    br[3] = 0.0;
    if (i == 0) { bs[0] = colscal[(size_t)col * SG_NS + SG_S_SOL] * B.solar_flux; bs[1] = 0.0; }
  } else {
    const double *tl = tlev + (size_t)col * (L + 1);
    double b0 = planck_band(B, spec->n_deg, spec->t_ref_planck, tl[i]);
    double b1 = planck_band(B, spec->n_deg, spec->t_ref_planck, tl[i + 1]);
    double bt = planck_band(B, spec->n_deg, spec->t_ref_planck, r[SG_Q_T]);
    br[1] = 2.0 * (b1 + b0 - 2.0 * bt);  // diff_2
    br[2] = b1 - b0;                     // diff
    br[3] = b1;                          // planck at the bottom of the layer
    if (i == 0) {
      bs[0] = b0;
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
  double sec0, ksc_clear;
};

// co: the layer's droplet optics at unit condensate scaling (kext, kscat,
// first moment, forward scattering), from prep_kernel.
__device__ __forceinline__ void layer_coef(const LaneCtx &c, const SgTerm &t, const double *r,
                                           double kgt, double csub, const double *co, Coef &o) {
  double kgas = k_term(t, r[SG_Q_LNP], r[SG_Q_LNT], r[SG_Q_MMR + t.gas]);
  const double frac1 = r[SG_Q_FRAC1];
  if (csub * frac1 > 0.0) {
    const double kext = co[0], kscat = co[1], phf = co[2], fwd = co[3];
    double kg = kgt + kext * csub;
    double ks = c.ksc_clear + kscat * csub;
    double fw = fwd * csub;
    double pf = phf * csub;
    pf = (pf - fw) / fmax(ks - fw, kEps);
    fw = fw / fmax(ks, kEps);
    double ktot = kg + kgas;
    double tau = ktot * r[SG_Q_MASS];
    double om = fmin(ks / (ktot + kTiny), 1.0 - 32.0 * kEps);
    tau = tau * (1.0 - om * fw);
    om = om * (1.0 - fw) / (1.0 - om * fw);
    two_coeff(c.solar, tau, om, pf, c.sec0, o);
    o.T *= frac1; o.R *= frac1; o.T0 *= frac1; o.s1 *= frac1; o.s2 *= frac1;
  } else {
    double ktot = kgt + kgas;
    double tau = ktot * r[SG_Q_MASS];
    double om = fmin(c.ksc_clear / (ktot + kTiny), 1.0 - 32.0 * kEps);
    two_coeff(c.solar, tau, om, 0.0, c.sec0, o);
  }
}

// Mixed form of layer_coef: gas k-term, cloud optics, single scattering and
// two-stream coefficients in FP32 (two_coeff_f32), carrying the co-albedo
// 1 - omega from the absorption so the near-conservative limit keeps its
// accuracy; the reference's clamp omega <= 1 - 32 eps(FP64) is kept as
// 1 - omega >= 32 eps(FP64). No FP64 fallback: two_coeff_f32 has no
// resonance singularity.
__device__ __forceinline__ void layer_coef_mixed(const LaneCtx &c, const SgTerm &t, const double *r,
                                                double kgt, double csub, const double *co, Coef &o) {
  const float kgas = k_term_f32(t, (float)r[SG_Q_LNP], (float)r[SG_Q_LNT], (float)r[SG_Q_MMR + t.gas]);
  const float frac1 = (float)r[SG_Q_FRAC1];
  const float kag = (float)(kgt - c.ksc_clear);  // grey absorption (equivalent extinction)
  const float mass = (float)r[SG_Q_MASS];
  const float omc_min = (float)(32.0 * kEps);
  float tau, om, omc, g;
  bool cloudy = csub * frac1 > 0.0;
  if (cloudy) {
    // absorption kext - kscat formed in FP64 (relative accuracy ~1e-10 even
    // for weakly absorbing droplets), then rounded
    const float kext = (float)co[0], kabs = (float)(co[0] - co[1]);
    const float phf = (float)co[2], fwd = (float)co[3];
    const float cs = (float)csub;
    float ks = (float)c.ksc_clear + (kext - kabs) * cs;
    float fw = fwd * cs;
    float pf = phf * cs;
    pf = SG_FDIV(pf - fw, fmaxf(ks - fw, (float)kEps));
    fw = SG_FDIV(fw, fmaxf(ks, (float)kEps));
    float ktot = (float)c.ksc_clear + kag + kgas + kext * cs;
    float kab = kag + kgas + kabs * cs;
    tau = ktot * mass;
    omc = fmaxf(SG_FDIV(kab, ktot), omc_min);
    om = 1.0f - omc;
    float sc = 1.0f - om * fw;
    tau = tau * sc;
    omc = SG_FDIV(omc, sc);
    om = 1.0f - omc;
    g = pf;
  } else {
    float ktot = (float)c.ksc_clear + kag + kgas;
    tau = ktot * mass;
    omc = ktot > 0.0f ? fmaxf(SG_FDIV(kag + kgas, ktot), omc_min) : 1.0f;
    om = 1.0f - omc;
    g = 0.0f;
  }
  CoefF f;
  two_coeff_f32(c.solar, tau, om, omc, g, (float)c.sec0, f);
  coef_from_f32(f, o);
  if (cloudy) { o.T *= frac1; o.R *= frac1; o.T0 *= frac1; o.s1 *= frac1; o.s2 *= frac1; }
}

// One layer of the top-down elimination for one lane: the layer's
// coefficients (FP64 layer_coef or the mixed layer_coef_mixed), its sources
// (solar_source.F90:104-162 with the k_eqv direct-beam adjustment, or
// ir_source.F90:93-120), and the update of the reflectance of the atmosphere
// above (rho) and the diffuse flux it sends down (dn). Returns the
// back-substitution pair fu(i) = a fu(i+1) + bb.
template <bool MIXED>
__device__ __forceinline__ void layer_step(const LaneCtx &c, const SgTerm &t, const double *r,
                                           const double *q, double cv, const double *co,
                                           double &fdir, double &rho, double &dn, double &a,
                                           double &bb) {
  Coef o;
  if (MIXED) layer_coef_mixed(c, t, r, q[0], cv, co, o);
  else layer_coef(c, t, r, q[0], cv, co, o);
  double su, sd;
  if (c.solar) {
    double fnew = fdir * o.T0 * q[1];
    su = o.s1 * fdir;
    sd = (o.s2 - o.T0) * fdir + fnew;
    fdir = fnew;
  } else {
    su = o.s1 * q[2] + o.s2 * q[1];
    sd = -o.s1 * q[2] + o.s2 * q[1];
  }
  double bp;
  if (MIXED) {
    // reciprocal from the hardware estimate and two Newton steps (a few ulp)
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

// The column solve, one wave per column (or per chunk of 64 items of a
// column), one lane per item: a clear-sky k-term (weight w_k (1 - frac_cloudy))
// or a McICA sub-column sample (weight w_k frac_cloudy / subcol_k)
// (mcica_sample.F90:453-930; monochromatic_radiance.F90:429-691;
// mcica_column.F90:255-622; solver_homogen_direct.F90:93-127). The two-stream
// system is eliminated top-down (reflectance of the atmosphere above and the
// downward diffuse flux it emits) and back-substituted bottom-up. Output:
// partial[col][chunk][3][L+1] (up, down, direct) of this wave's weighted sum.
//
// RECOMP = false keeps every layer's elimination state (a, bb, rho, dn) in
// lane-private memory between the sweeps: 64 B per lane and layer of
// scratch traffic. RECOMP = true keeps only (fdir, rho, dn) at the top of
// every kSeg-th layer and, going up, recomputes each segment's layers
// (coefficients and elimination, bit-identical to the first sweep) into
// registers before back-substituting it: a second pass of arithmetic in
// place of the memory traffic, which pays where the coefficients are cheap
// (the mixed form's FP32 arithmetic).
constexpr int kSeg = 4;

template <bool MIXED, bool RECOMP>
__global__ __launch_bounds__(256, 4) void solve_kernel(CallDims d, const SgSpec *__restrict__ spec,
                                                    int l_rayleigh,
                                                    const double *__restrict__ colrec,
                                                    const double *__restrict__ colscal,
                                                    const double *__restrict__ bandrec,
                                                    const double *__restrict__ bandscal,
                                                    const double *__restrict__ c_sub,
                                                    const double *__restrict__ cldopt,
                                                    double *__restrict__ partial) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x % kWave;
  const int L = d.L, NB = d.NB;
  const int col = wave / d.nwpc, chunk = wave % d.nwpc;
  if (col >= d.P) return;
  const bool solar = d.isolir == SOCGPU_SOURCE_SOLAR;
  const double *cs = colscal + (size_t)col * SG_NS;
  const double fc = d.mcica ? cs[SG_S_FC] : 0.0;
  const int G = spec->n_g;
  const int n_clear = fc < 1.0 ? G : 0;
  const int n_sub = (d.mcica && fc > 0.0) ? spec->n_sub_items[d.mode] : 0;
  const int item = chunk * kWave + lane;
  const bool active = item < n_clear + n_sub;
  double *out = partial + ((size_t)col * d.nwpc + chunk) * 3 * (L + 1);
  int gi = 0, isub = 0;
  double wgt = 0.0;
  if (active) {
    if (item < n_clear) {
      gi = item;
      wgt = spec->g[gi].w * (1.0 - fc);
    } else {
      int s = item - n_clear;
      gi = spec->sub_g[d.mode][s];
      isub = spec->sub_isub[d.mode][s];
      wgt = spec->g[gi].w * (fc * spec->sub_inv_s[d.mode][s]);
    }
  }
  const SgTerm t = spec->g[gi];
  const int b = t.band;
  const double *rec = colrec + (size_t)col * L * SG_NQ;
  const double *br = bandrec + ((size_t)col * NB + b) * L * SG_NBQ;
  const double *bs = bandscal + ((size_t)col * NB + b) * SG_NBQ;
  LaneCtx c;
  c.solar = solar;
  c.sec0 = cs[SG_S_SEC0];
  c.ksc_clear = (solar && l_rayleigh) ? spec->band[b].rayleigh : 0.0;
  const double *cop = d.mcica ? cldopt + ((size_t)col * NB + b) * L * 4 : nullptr;
  const double *csub = c_sub + (size_t)col * L * SG_MAXSUBCOL + (isub > 0 ? isub - 1 : 0);
  const double alb_diff = bs[2], alb_dir = bs[3];

  constexpr int kNst = RECOMP ? (SG_MAXL + kSeg - 1) / kSeg : SG_MAXL;
  double st[kNst][RECOMP ? 3 : 4];
  const double finc = solar ? bs[0] / c.sec0 : -bs[0];  // flux_inc_down, flux_inc_direct
  double fdir = solar ? finc : 0.0;
  double rho = 0.0, dn = finc;
  double vdir[8];
  for (int i = 0; i < L; i++) {
    if (solar) {
      // direct flux at the top of layer i, summed 8 levels at a time
      const int m = i & 7;
      double x = active ? wgt * fdir : 0.0;
#pragma unroll
      for (int k = 0; k < 8; k++) if (k == m) vdir[k] = x;
      if (m == 7) {
        reduce8(vdir, lane);
        if ((lane & 7) == 0) out[2 * (L + 1) + i - 7 + (lane >> 3)] = vdir[0];
      }
    }
    if (active) {
      if (RECOMP) {
        if (i % kSeg == 0) {
          st[i / kSeg][0] = fdir; st[i / kSeg][1] = rho; st[i / kSeg][2] = dn;
        }
      }
      double cv = isub > 0 ? csub[(size_t)i * SG_MAXSUBCOL] : 0.0;
      double a, bb;
      const double rho0 = rho, dn0 = dn;
      layer_step<MIXED>(c, t, rec + i * SG_NQ, br + i * SG_NBQ, cv, cop + i * 4, fdir, rho, dn, a, bb);
      if (!RECOMP) {
        st[i][0] = a; st[i][1] = bb; st[i][2] = rho0; st[i][3] = dn0;
      }
    }
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
    reduce8(vdir, lane);
    if ((lane & 7) == 0 && base + (lane >> 3) <= L) out[2 * (L + 1) + base + (lane >> 3)] = vdir[0];
  } else if (lane < 64) {
    for (int i = lane; i <= L; i += kWave) out[2 * (L + 1) + i] = 0.0;
  }
  double fu = 0.0, fd = 0.0;
  if (active) {
    double sg = solar ? (alb_dir - alb_diff) * fdir
                      : (1.0 - alb_diff) * (bs[1] - br[(L - 1) * SG_NBQ + 3]);
    fu = (alb_diff * dn + sg) / (1.0 - alb_diff * rho);
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
      reduce8(v, lane);
      if ((lane & 7) == 0) {
        const int j = lane >> 3, lev = top - (j >> 1);
        if (lev >= 0) out[(j & 1) * (L + 1) + lev] = v[0];
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
      reduce8(v, lane);
      if ((lane & 7) == 0 && (lane >> 3) < 2) out[(lane >> 3) * (L + 1) + L] = v[0];
    }
    const int nseg = (L + kSeg - 1) / kSeg;
    for (int sgi = nseg - 1; sgi >= 0; sgi--) {
      const int i0 = sgi * kSeg;
      double sa[kSeg], sb[kSeg], sr[kSeg], sdn[kSeg];
      if (active) {
        double fdir_s = st[sgi][0], rho_s = st[sgi][1], dn_s = st[sgi][2];
#pragma unroll
        for (int k = 0; k < kSeg; k++) {
          const int i = i0 + k;
          sr[k] = rho_s; sdn[k] = dn_s; sa[k] = 0.0; sb[k] = 0.0;
          if (i < L) {
            double cv = isub > 0 ? csub[(size_t)i * SG_MAXSUBCOL] : 0.0;
            layer_step<MIXED>(c, t, rec + i * SG_NQ, br + i * SG_NBQ, cv, cop + i * 4, fdir_s, rho_s,
                              dn_s, sa[k], sb[k]);
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
      reduce8(v, lane);
      if ((lane & 7) == 0) {
        const int j = lane >> 3, lev = i0 + kSeg - 1 - (j >> 1);
        if (lev < L) out[(j & 1) * (L + 1) + lev] = v[0];
      }
    }
  }
}

// Sum of the per-wave partials, the Planck flux added back for the
// infra-red (adjust_ir_radiance.F90:186-200, weight_band = 1), output in the
// caller's [level][profile] layout.
__global__ void finish_kernel(CallDims d, const double *__restrict__ partial,
                              const double *__restrict__ bandrec,
                              const double *__restrict__ bandscal, double *__restrict__ up,
                              double *__restrict__ down, double *__restrict__ direct) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int P = d.P, L = d.L;
  if (idx >= P * (L + 1)) return;
  int col = idx % P, lev = idx / P;
  double su = 0.0, sd = 0.0, sx = 0.0;
  for (int c = 0; c < d.nwpc; c++) {
    const double *o = partial + ((size_t)col * d.nwpc + c) * 3 * (L + 1);
    su += o[lev];
    sd += o[(L + 1) + lev];
    sx += o[2 * (L + 1) + lev];
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
  up[idx] = su;
  down[idx] = sd;
  if (direct) direct[idx] = sx;
}

// Heating rate from flux divergence (socrates_set_diag.F90:140-210), K/s.
__global__ void heating_kernel(int P, int L, const double *__restrict__ up,
                               const double *__restrict__ down,
                               const double *__restrict__ heat_cap, double *__restrict__ hr) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P * L) return;
  int col = idx % P, i = idx / P;  // layer i+1 lies between levels i and i+1
  double fdiv = down[(size_t)i * P + col] - down[(size_t)(i + 1) * P + col] +
                up[(size_t)(i + 1) * P + col] - up[(size_t)i * P + col];
  hr[idx] = heat_cap ? fdiv / heat_cap[idx] : fdiv;
}

// Kernel-level two-stream solve on prescribed optical properties, one lane
// per (term, profile) (profile fastest, so every per-layer read and
// per-level write is coalesced): two_coeff, the infra-red (ir_source.F90:
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
    double bp = 1.0 / (1.0 - o.R * rho);
    double aa = o.T * bp;
    double bb = (o.R * dn + su) * bp;
    st[i][0] = aa; st[i][1] = bb; st[i][2] = rho; st[i][3] = dn;
    double rho_n = o.R + o.T * rho * aa;
    dn = o.T * (rho * bb + dn) + sd;
    rho = rho_n;
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
struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  int ensure(size_t bytes) {
    if (bytes <= n) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, bytes) != hipSuccess) return -1;
    n = bytes;
    return 0;
  }
  template <class T> T *as() { return (T *)p; }
  ~DevBuf() {}
};

struct SpecEntry {
  SgSpec host;
  SgSpec *dev = nullptr;
  double *xcw = nullptr;  // McICA condensate table on the device
  bool solar_hint = false;
};

struct Context {
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  std::map<std::string, std::unique_ptr<SpecEntry>> spectra;
  DevBuf colrec, tlev, colscal, bandrec, bandscal, csub, partial, krange, minor, cldopt;
  DevBuf in_stage, out_stage;
  double last_total_ms = 0.0, last_solve_ms = 0.0;
  std::mutex mu;
};
static Context g_ctx;
// solve-kernel tile per precision form ([0] FP64, [1] mixed)
static SocgpuTile g_tile[2] = {{0, 2}, {0, 2}};

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
  g_ctx.device = device;
  return 0;
}

int socgpu_spectrum_register(const char *name, const char *path) {
  if (int rc = ensure_init()) return rc;
  if (!name || !path) return fail(SOCGPU_EINVAL, "spectrum name and path are required");
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  auto e = std::make_unique<SpecEntry>();
  std::string err;
  int rc = read_spectrum(path, &e->host, &err);
  if (rc) return fail(rc == -3 ? SOCGPU_EUNSUPPORTED : SOCGPU_EIO, "%s: %s", path, err.c_str());
  HIPCHK(hipMalloc(&e->dev, sizeof(SgSpec)));
  HIPCHK(hipMemcpy(e->dev, &e->host, sizeof(SgSpec), hipMemcpyHostToDevice));
  auto it = g_ctx.spectra.find(name);
  if (it != g_ctx.spectra.end()) {
    (void)hipFree(it->second->dev);
    if (it->second->xcw) (void)hipFree(it->second->xcw);
  }
  g_ctx.spectra[name] = std::move(e);
  return 0;
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
  for (int k = 0; k < 2; k++) {
    if (!names[k]) continue;
    auto it = g_ctx.spectra.find(names[k]);
    if (it == g_ctx.spectra.end()) return fail(SOCGPU_ENOTFOUND, "spectrum '%s' not registered", names[k]);
    SpecEntry &e = *it->second;
    if (attach_mcica(&e.host, m, k == 0, &err)) return fail(SOCGPU_EUNSUPPORTED, "%s", err.c_str());
    HIPCHK(hipMemcpy(e.dev, &e.host, sizeof(SgSpec), hipMemcpyHostToDevice));
    if (e.xcw) (void)hipFree(e.xcw);
    HIPCHK(hipMalloc(&e.xcw, m.xcw.size() * sizeof(double)));
    HIPCHK(hipMemcpy(e.xcw, m.xcw.data(), m.xcw.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return 0;
}

static int stage_inputs(const SocgpuRunesArgs *a, SocgpuRunesArgs *dv, int NB) {
  *dv = *a;
  if (a->l_inputs_on_device) return 0;
  const size_t P = a->n_profile, L = a->n_layer;
  struct F { const double **dst; const double *src; size_t n; };
  F f[] = {
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
  };
  size_t total = 0;
  for (auto &x : f) if (x.src) total += x.n;
  size_t ints = a->rand_seed ? P : 0;
  if (g_ctx.in_stage.ensure(total * sizeof(double) + ints * sizeof(int) + 256))
    return fail(SOCGPU_ENOMEM, "input staging allocation failed");
  char *base = g_ctx.in_stage.as<char>();
  size_t off = 0;
  for (auto &x : f) {
    if (!x.src) { *x.dst = nullptr; continue; }
    HIPCHK(hipMemcpyAsync(base + off, x.src, x.n * sizeof(double), hipMemcpyHostToDevice, g_ctx.stream));
    *x.dst = (const double *)(base + off);
    off += x.n * sizeof(double);
  }
  if (ints) {
    HIPCHK(hipMemcpyAsync(base + off, a->rand_seed, P * sizeof(int), hipMemcpyHostToDevice, g_ctx.stream));
    dv->rand_seed = (const int *)(base + off);
  }
  return 0;
}

int socgpu_runes(const char *spectrum_name, const SocgpuRunesArgs *a, SocgpuDiag *diag, int precision) {
  if (int rc = ensure_init()) return rc;
  if (!spectrum_name || !a || !diag) return fail(SOCGPU_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(g_ctx.mu);
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
  if (!solar && (!a->t_level || !a->t_ground)) return fail(SOCGPU_EINVAL, "thermal call needs t_level and t_ground");
  if (!solar && S.n_deg <= 0 && S.band[0].planck[0] == 0.0) return fail(SOCGPU_EUNSUPPORTED, "spectral file has no Planck data");
  CallDims d;
  std::memset(&d, 0, sizeof d);
  d.P = P; d.L = L; d.NB = NB; d.isolir = a->i_source;
  int cloud = a->i_cloud_representation;
  if (cloud != SOCGPU_CLOUD_OFF) {
    if (cloud != SOCGPU_CLOUD_HOMOGEN || a->i_inhom != SOCGPU_INHOM_MCICA)
      return fail(SOCGPU_EUNSUPPORTED, "cloud representation %d with i_inhom %d is not on the McICA path", cloud, a->i_inhom);
    if (!S.has_mcica) return fail(SOCGPU_ENOTFOUND, "Spectrum name not associated with MCICA data.");
    if (a->i_mcica_sampling != SOCGPU_MCICA_SINGLE && a->i_mcica_sampling != SOCGPU_MCICA_OPTIMAL)
      return fail(SOCGPU_EUNSUPPORTED, "McICA sampling %d not supported", a->i_mcica_sampling);
    if (a->i_overlap != SOCGPU_OVERLAP_EXP_RAND && a->i_overlap != SOCGPU_OVERLAP_RANDOM && a->i_overlap != SOCGPU_OVERLAP_MAX_RAND)
      return fail(SOCGPU_EUNSUPPORTED, "overlap %d not supported with McICA", a->i_overlap);
    if (!a->cloud_frac || !a->liq_mmr || !a->liq_dim) return fail(SOCGPU_EINVAL, "Cloud fraction not provided.");
    if (!a->rand_seed) return fail(SOCGPU_EINVAL, "rand_seed is required: default seeds depend on the position in the call");
    int drop = a->i_st_water;
    if (drop < 1 || drop > SG_NDROP || !S.drop[drop - 1].present) return fail(SOCGPU_EINVAL, "Liquid cloud type not in spectral file.");
    if (S.drop[drop - 1].scheme != 5 && S.drop[drop - 1].scheme != 1) return fail(SOCGPU_EUNSUPPORTED, "droplet scheme %d", S.drop[drop - 1].scheme);
    d.mcica = 1;
    d.mode = a->i_mcica_sampling == SOCGPU_MCICA_OPTIMAL ? 1 : 0;
    if (S.n_req[d.mode] <= 0) return fail(SOCGPU_EUNSUPPORTED, "McICA data has no tables for this sampling");
    d.drop = drop - 1;
    d.n_gen = a->n_subcol_gen > 0 ? a->n_subcol_gen : S.n_gen;
    if (d.n_gen > SG_MAXSUBCOL) return fail(SOCGPU_EUNSUPPORTED, "n_subcol_gen above 64");
    d.n_fill = S.n_req[d.mode] < d.n_gen ? S.n_req[d.mode] : d.n_gen;
    d.ipph = S.ipph;
    d.ioverlap = a->i_overlap;
  }
  const int items = S.n_g + (d.mcica ? S.n_sub_items[d.mode] : 0);
  d.nwpc = (items + kWave - 1) / kWave;

  SocgpuRunesArgs dv;
  if (int rc = stage_inputs(a, &dv, NB)) return rc;
  if (g_ctx.colrec.ensure((size_t)P * L * SG_NQ * 8) || g_ctx.tlev.ensure((size_t)P * (L + 1) * 8) ||
      g_ctx.colscal.ensure((size_t)P * SG_NS * 8) || g_ctx.bandrec.ensure((size_t)P * NB * L * SG_NBQ * 8) ||
      g_ctx.bandscal.ensure((size_t)P * NB * SG_NBQ * 8) ||
      g_ctx.partial.ensure((size_t)P * d.nwpc * 3 * (L + 1) * 8) || g_ctx.krange.ensure(64) ||
      g_ctx.minor.ensure((size_t)P * (S.n_minor > 0 ? S.n_minor : 1) * L * 2 * 8) ||
      (d.mcica && g_ctx.csub.ensure((size_t)P * L * SG_MAXSUBCOL * 8)) ||
      (d.mcica && g_ctx.cldopt.ensure((size_t)P * NB * L * 4 * 8)) ||
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
  hipStream_t s = g_ctx.stream;
  HIPCHK(hipEventRecord(g_ctx.ev0, s));
  const int T = 256;
  int nlev = P * (L + 1);
  pack_kernel<<<(nlev + T - 1) / T, T, 0, s>>>(d, e.dev, dv, g_ctx.colrec.as<double>(), g_ctx.tlev.as<double>(),
                                              g_ctx.colscal.as<double>(), g_ctx.bandscal.as<double>());
  if (d.mcica) {
    fill_int_kernel<<<1, 64, 0, s>>>(g_ctx.krange.as<int>(), 1, L + 1);
    fill_int_kernel<<<1, 64, 0, s>>>(g_ctx.krange.as<int>() + 1, 1, 0);
    cldrange_kernel<<<(P * L + T - 1) / T, T, 0, s>>>(d, g_ctx.colrec.as<double>(), g_ctx.krange.as<int>());
    cloudgen_kernel<<<P, 64, 0, s>>>(d, g_ctx.colrec.as<double>(), dv.rand_seed, a->cloud_vertical_decorr,
                                     a->cloud_vertical_decorr * 0.5, g_ctx.krange.as<int>(), e.xcw, S.n1, S.n2,
                                     g_ctx.csub.as<double>(), g_ctx.colscal.as<double>());
  }
  if (S.n_minor > 0)
    minor_kernel<<<(P * S.n_minor + kMinorWaves - 1) / kMinorWaves, 64 * kMinorWaves, 0, s>>>(d, e.dev, g_ctx.colrec.as<double>(), g_ctx.tlev.as<double>(),
                                                          g_ctx.colscal.as<double>(), g_ctx.bandscal.as<double>(),
                                                          g_ctx.minor.as<double>());
  prep_kernel<<<(P * NB * L + T - 1) / T, T, 0, s>>>(d, e.dev, a->l_rayleigh, g_ctx.colrec.as<double>(),
                                                     g_ctx.tlev.as<double>(), g_ctx.colscal.as<double>(),
                                                     g_ctx.minor.as<double>(), g_ctx.bandrec.as<double>(),
                                                     g_ctx.bandscal.as<double>(), g_ctx.cldopt.as<double>());
  HIPCHK(hipEventRecord(g_ctx.ev1, s));
  int waves = P * d.nwpc;
  const SocgpuTile &tile = g_tile[precision == SOCGPU_MIXED ? 1 : 0];
  const int wpb = tile.waves_per_block;
  const int grid = (waves + wpb - 1) / wpb;
  auto *kern = precision == SOCGPU_MIXED
                   ? (tile.recompute ? solve_kernel<true, true> : solve_kernel<true, false>)
                   : (tile.recompute ? solve_kernel<false, true> : solve_kernel<false, false>);
  kern<<<grid, wpb * kWave, 0, s>>>(d, e.dev, a->l_rayleigh, g_ctx.colrec.as<double>(), g_ctx.colscal.as<double>(),
                                    g_ctx.bandrec.as<double>(), g_ctx.bandscal.as<double>(),
                                    d.mcica ? g_ctx.csub.as<double>() : nullptr,
                                    d.mcica ? g_ctx.cldopt.as<double>() : nullptr, g_ctx.partial.as<double>());
  HIPCHK(hipEventRecord(g_ctx.ev2, s));
  finish_kernel<<<(nlev + T - 1) / T, T, 0, s>>>(d, g_ctx.partial.as<double>(), g_ctx.bandrec.as<double>(),
                                                g_ctx.bandscal.as<double>(), oup, odn, odir);
  if (diag->heating_rate)
    heating_kernel<<<(P * L + T - 1) / T, T, 0, s>>>(P, L, oup, odn, dv.layer_heat_capacity, ohr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(g_ctx.ev3, s));
  if (!diag->l_outputs_on_device) {
    size_t nl = (size_t)P * (L + 1) * 8;
    if (diag->flux_up) HIPCHK(hipMemcpyAsync(diag->flux_up, oup, nl, hipMemcpyDeviceToHost, s));
    if (diag->flux_down) HIPCHK(hipMemcpyAsync(diag->flux_down, odn, nl, hipMemcpyDeviceToHost, s));
    if (diag->flux_direct) HIPCHK(hipMemcpyAsync(diag->flux_direct, odir, nl, hipMemcpyDeviceToHost, s));
    if (diag->heating_rate) HIPCHK(hipMemcpyAsync(diag->heating_rate, ohr, (size_t)P * L * 8, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  float t_all = 0.f, t_solve = 0.f;
  HIPCHK(hipEventElapsedTime(&t_all, g_ctx.ev0, g_ctx.ev3));
  HIPCHK(hipEventElapsedTime(&t_solve, g_ctx.ev1, g_ctx.ev2));
  g_ctx.last_total_ms = t_all;
  g_ctx.last_solve_ms = t_solve;
  return 0;
}

int socgpu_set_tile(int precision, const SocgpuTile *tile) {
  if (precision != SOCGPU_FP64 && precision != SOCGPU_MIXED) return fail(SOCGPU_EINVAL, "unknown precision %d", precision);
  if (!tile) return fail(SOCGPU_EINVAL, "null tile");
  const int w = tile->waves_per_block;
  if (w != 1 && w != 2 && w != 4) return fail(SOCGPU_EINVAL, "waves_per_block must be 1, 2 or 4");
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
  if (ioverlap == 2) return fail(SOCGPU_EUNSUPPORTED, "ioverlap 2 carries draws across sub-columns");
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
  d.ipph = ipph; d.ioverlap = ioverlap;
  DevBuf drec, dseed, dxcw, dcs, dscal, dk;
  if (drec.ensure(rec.size() * 8) || dseed.ensure((size_t)P * 4) || dxcw.ensure((size_t)n1 * n2 * 8) ||
      dcs.ensure((size_t)P * L * SG_MAXSUBCOL * 8) || dscal.ensure((size_t)P * SG_NS * 8) || dk.ensure(8))
    return fail(SOCGPU_ENOMEM, "allocation failed");
  hipStream_t s = g_ctx.stream;
  HIPCHK(hipMemcpyAsync(drec.p, rec.data(), rec.size() * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dseed.p, rand_seed, (size_t)P * 4, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dxcw.p, xcw, (size_t)n1 * n2 * 8, hipMemcpyHostToDevice, s));
  fill_int_kernel<<<1, 64, 0, s>>>(dk.as<int>(), 1, L + 1);
  fill_int_kernel<<<1, 64, 0, s>>>(dk.as<int>() + 1, 1, 0);
  cldrange_kernel<<<(P * L + 255) / 256, 256, 0, s>>>(d, drec.as<double>(), dk.as<int>());
  cloudgen_kernel<<<P, 64, 0, s>>>(d, drec.as<double>(), dseed.as<int>(), rlc_cf[0], rlc_cw[0], dk.as<int>(),
                                   dxcw.as<double>(), n1, n2, dcs.as<double>(), dscal.as<double>());
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
  for (DevBuf *b : {&drec, &dseed, &dxcw, &dcs, &dscal, &dk}) if (b->p) (void)hipFree(b->p);
  return 0;
}

int socgpu_finalize(void) {
  std::lock_guard<std::mutex> lk(g_ctx.mu);
  if (g_ctx.device < 0) return 0;
  (void)hipStreamSynchronize(g_ctx.stream);
  for (auto &kv : g_ctx.spectra) {
    (void)hipFree(kv.second->dev);
    if (kv.second->xcw) (void)hipFree(kv.second->xcw);
  }
  g_ctx.spectra.clear();
  for (DevBuf *b : {&g_ctx.colrec, &g_ctx.tlev, &g_ctx.colscal, &g_ctx.bandrec, &g_ctx.bandscal, &g_ctx.csub,
                    &g_ctx.partial, &g_ctx.krange, &g_ctx.minor, &g_ctx.cldopt, &g_ctx.in_stage,
                    &g_ctx.out_stage}) {
    if (b->p) (void)hipFree(b->p);
    b->p = nullptr;
    b->n = 0;
  }
  (void)hipEventDestroy(g_ctx.ev0); (void)hipEventDestroy(g_ctx.ev1);
  (void)hipEventDestroy(g_ctx.ev2); (void)hipEventDestroy(g_ctx.ev3);
  (void)hipStreamDestroy(g_ctx.stream);
  g_ctx.device = -1;
  return 0;
}

}  // extern "C"

// Host build of the device physics (socrates_amd/csrc/physics.h) for the CPU
// tests: the FP64 and the mixed-precision two-stream coefficients and a
// one-column two-stream solve in either form, as the device kernels compute
// them (the device's exp/expf may differ from the host's in the last bits).
// Test infrastructure only; compiled by tests/test_mixed_formulas.py.
#define __host__
#define __device__
#define __forceinline__ inline
#include <cmath>
#include <vector>

#include "../../socrates_amd/csrc/physics.h"

using namespace sg;

extern "C" {

void host_two_coeff(int n, int solar, const double *tau, const double *om, const double *g,
                    const double *sec0, double *out /* [n][5] */) {
  for (int i = 0; i < n; i++) {
    Coef o;
    two_coeff(solar != 0, tau[i], om[i], g[i], sec0[i], o);
    double *r = out + 5 * i;
    r[0] = o.T; r[1] = o.R; r[2] = o.T0; r[3] = o.s1; r[4] = o.s2;
  }
}

void host_two_coeff_f32(int n, int solar, const double *tau, const double *om, const double *g,
                        const double *sec0, double *out /* [n][7] */) {
  for (int i = 0; i < n; i++) {
    CoefF f;
    two_coeff_f32(solar != 0, (float)tau[i], (float)om[i], (float)fmax(1.0 - om[i], 32.0 * kEps),
                  (float)g[i], (float)sec0[i], f);
    double *r = out + 7 * i;
    r[0] = f.T; r[1] = f.R; r[2] = f.T0; r[3] = f.s1; r[4] = f.s2; r[5] = f.u; r[6] = f.em0;
  }
}

// One column, one term, arrays [layer]: the kernel-level solve of
// socrates_amd/csrc/socgpu.hip (two_stream_kernel) in FP64 (mixed = 0) or
// the mixed form (mixed = 1).
void host_two_stream(int L, int solar, int lss, int lq, int mixed, const double *tau,
                     const double *om, const double *g, double sec0, double finc_down,
                     double finc_dir, double ad, double adir, const double *dpl,
                     const double *dpl2, double dsurf, const double *adj, double *fup,
                     double *fdn) {
  std::vector<double> st(4 * (size_t)L);
  double fd = solar ? finc_dir : 0.0, rho = 0.0, dn = finc_down;
  for (int i = 0; i < L; i++) {
    Coef o;
    if (mixed) {
      CoefF f;
      two_coeff_f32(solar != 0, (float)tau[i], (float)om[i], (float)fmax(1.0 - om[i], 32.0 * kEps),
                    (float)g[i], (float)sec0, f);
      coef_from_f32(f, o);
    } else {
      two_coeff(solar != 0, tau[i], om[i], g[i], sec0, o);
    }
    double su, sd;
    if (solar) {
      if (lss) {
        double fn = fd * o.T0 * adj[i];
        su = o.s1 * fd;
        sd = (o.s2 - o.T0) * fd + fn;
        fd = fn;
      } else {
        su = o.s1 * fd;
        sd = o.s2 * fd;
        fd = fd * o.T0;
      }
    } else if (lq) {
      su = o.s1 * dpl[i] + o.s2 * dpl2[i];
      sd = -o.s1 * dpl[i] + o.s2 * dpl2[i];
    } else {
      su = o.s1 * dpl[i];
      sd = -su;
    }
    double bp = 1.0 / (1.0 - o.R * rho), a = o.T * bp, bb = (o.R * dn + su) * bp;
    st[4 * i] = a; st[4 * i + 1] = bb; st[4 * i + 2] = rho; st[4 * i + 3] = dn;
    double rn = o.R + o.T * rho * a;
    dn = o.T * (rho * bb + dn) + sd;
    rho = rn;
  }
  double sg = solar ? (adir - ad) * fd : (1.0 - ad) * dsurf;
  double fu = (ad * dn + sg) / (1.0 - ad * rho), fx = rho * fu + dn;
  fup[L] = fu;
  fdn[L] = fx;
  for (int i = L - 1; i >= 0; i--) {
    fu = st[4 * i] * fu + st[4 * i + 1];
    fx = st[4 * i + 2] * fu + st[4 * i + 3];
    fup[i] = fu;
    fdn[i] = fx;
  }
}

}

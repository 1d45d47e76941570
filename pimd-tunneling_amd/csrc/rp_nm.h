// rp_nm.h -- normal-mode set-up (host) and the shared FP64 MFMA transform.
#pragma once
#include <cmath>
#include <vector>

#include "rp_internal.h"

namespace rp {

typedef double d4 __attribute__((ext_vector_type(4)));

// Host arrays of init_nm (verletmodule.f90:306-338) and the Langevin c1/c2 table
// (:380-386).  T[i*n+l] (symmetric), lam[i], bm[i*natom+j] = beadmass(j,i),
// bv[dof*n+i] = beadvec(i,dof), c1/c2[i*natom+j].
struct NmHost {
  std::vector<double> T, lam, bm, bv, c1, c2;
};

inline void nm_host_beadvec(NmHost& nm, int n, int ndim, int natom, double betan,
                            const double* a, const double* b) {
  const int D = ndim * natom;
  nm.bv.assign((size_t)n * D, 0.0);
  for (int i = 1; i <= n; ++i)
    for (int j = 0; j < ndim; ++j)
      for (int k = 0; k < natom; ++k) {
        int dofi = k * ndim + j;
        double v = a[dofi] * std::sin((double)i * kPiRef / (double)(n + 1)) +
                   b[dofi] * std::sin((double)(n * i) * kPiRef / (double)(n + 1));
        v = v * std::sqrt(2.0 / (double)(n + 1));
        double lb = nm.lam[i - 1] * betan;
        nm.bv[(size_t)dofi * n + (i - 1)] = v / (lb * lb);
      }
}

inline void nm_host_init(NmHost& nm, int n, int ndim, int natom, const double* mass,
                         double betan, double tau, double dt, double gamma, const double* a,
                         const double* b) {
  nm.T.assign((size_t)n * n, 0.0);
  nm.lam.assign(n, 0.0);
  nm.bm.assign((size_t)n * natom, 0.0);
  nm.c1.assign((size_t)n * natom, 0.0);
  nm.c2.assign((size_t)n * natom, 0.0);
  for (int i = 1; i <= n; ++i) {
    nm.lam[i - 1] = 2.0 * std::sin((double)i * kPiRef / (double)(2 * n + 2)) / betan;
    double lt = nm.lam[i - 1] * tau;
    for (int j = 0; j < natom; ++j) nm.bm[(size_t)(i - 1) * natom + j] = mass[j] * (lt * lt);
    for (int l = i; l <= n; ++l) {
      double t = std::sin((double)(i * l) * kPiRef / (double)(n + 1)) * std::sqrt(2.0 / (double)(n + 1));
      nm.T[(size_t)(i - 1) * n + (l - 1)] = t;
      nm.T[(size_t)(l - 1) * n + (i - 1)] = t;
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < natom; ++j) {
      double c1 = std::exp(-gamma * dt * nm.lam[i] * std::sqrt(mass[j] / nm.bm[(size_t)i * natom + j]));
      nm.c1[(size_t)i * natom + j] = c1;
      nm.c2[(size_t)i * natom + j] = std::sqrt(1.0 - c1 * c1);
    }
  nm_host_beadvec(nm, n, ndim, natom, betan, a, b);
}

// Natural cubic spline second derivatives of every dof of a path
// (spline/tridag, instantonmod.f90:397-480); path[dof*npath + k].
inline void path_spline_host(int npath, int ndof, const double* x, const double* path,
                             std::vector<double>& y2all) {
  y2all.assign((size_t)npath * ndof, 0.0);
  const int n = npath;
  std::vector<double> a(n), bb(n), c(n), r(n), gam(n);
  for (int d = 0; d < ndof; ++d) {
    const double* y = path + (size_t)d * npath;
    double* y2 = y2all.data() + (size_t)d * npath;
    for (int i = 0; i < n - 1; ++i) c[i] = x[i + 1] - x[i];
    for (int i = 0; i < n - 1; ++i) r[i] = 6.0 * ((y[i + 1] - y[i]) / c[i]);
    for (int i = n - 2; i >= 1; --i) r[i] = r[i] - r[i - 1];
    for (int i = 1; i < n - 1; ++i) a[i] = c[i - 1];
    for (int i = 1; i < n - 1; ++i) bb[i] = 2.0 * (c[i] + a[i]);
    bb[0] = 1.0;
    bb[n - 1] = 1.0;
    r[0] = 0.0;  // natural ends (yp1 = ypn = 1e31 in every reference call)
    c[0] = 0.0;
    r[n - 1] = 0.0;
    a[n - 1] = 0.0;
    double bet = bb[0];
    y2[0] = r[0] / bet;
    for (int j = 1; j < n; ++j) {
      gam[j] = c[j - 1] / bet;
      bet = bb[j] - a[j] * gam[j];
      y2[j] = (r[j] - a[j] * y2[j - 1]) / bet;
    }
    for (int j = n - 2; j >= 0; --j) y2[j] = y2[j] - gam[j + 1] * y2[j + 1];
  }
}

// splint with locate's bisection (instantonmod.f90:499-528, 565-603)
__host__ __device__ inline double splint_dev(const double* xa, const double* ya,
                                             const double* y2a, int n, double x) {
  bool ascnd = xa[n - 1] >= xa[0];
  int jl = 0, ju = n + 1;
  while (ju - jl > 1) {
    int jm = (ju + jl) / 2;
    if (ascnd == (x >= xa[jm - 1]))
      jl = jm;
    else
      ju = jm;
  }
  int loc = (x == xa[0]) ? 1 : (x == xa[n - 1]) ? n - 1 : jl;
  int klo = loc > n - 1 ? n - 1 : loc;
  klo = (klo < 1 ? 1 : klo) - 1;
  int khi = klo + 1;
  double h = xa[khi] - xa[klo];
  double a = (xa[khi] - x) / h, b = (x - xa[klo]) / h;
  return a * ya[klo] + b * ya[khi] +
         ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2a[khi]) * (h * h) / 6.0;
}

// out[col][i] = sum_l T[i][l] in'[col][l] for ncols n-vectors ([col][bead] layout).
// backward (1): in' = in + beadvec(:, shift[col]) when shift[col] >= 0;
// forward (0): out -= beadvec(:, shift[col]) when shift[col] >= 0.
int launch_nm_gemm(const double* T, int n, int ncols, const double* in, double* out,
                   const int* shift, const double* bv, int backward, hipStream_t s);

}  // namespace rp

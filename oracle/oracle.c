/*
 * oracle.c -- CPU restatement of the pimd-tunneling ring-polymer hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Reference citations are to files of
 * christophevaillant/pimd-tunneling.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ===================================================================== PES */
/*
 * Parameter layouts (shared with the HIP library's potentials):
 *   QUARTIC   dp = {h, x0}
 *   COUPLED2D dp = {V0, a, m, w, c}             ndim = 2, natom = 1
 *   EVB       dp = {kappa, Q[D*D] (column-major Q(i,j) = Q[i + D*j]), lam[D], cub[D], xa[D]}
 *             ip = perm[natom]; (Pi x)(atom a) = x(atom perm[a])
 *   WATER     dp = {De, alpha, re, kbend, theta0, eps, sigma, qH}; atoms O,H,H per molecule
 *   HARMONIC  dp = {k[D], c[D]}
 */

static double evb_va(int D, const double *Q, const double *lam, const double *cub,
                     const double *xa, const double *x, double *g) {
  /* VA = 1/2 y.Lam.y + sum c y^3, y = Q^T (x - xa); g = Q (Lam y + 3 c y^2) */
  double va = 0.0;
  double *y = (double *)alloca(sizeof(double) * D);
  for (int i = 0; i < D; ++i) {
    double s = 0.0;
    for (int j = 0; j < D; ++j) s += Q[j + D * i] * (x[j] - xa[j]);
    y[i] = s;
    va += 0.5 * lam[i] * s * s + cub[i] * s * s * s;
  }
  if (g) {
    for (int j = 0; j < D; ++j) g[j] = 0.0;
    for (int i = 0; i < D; ++i) {
      double h = lam[i] * y[i] + 3.0 * cub[i] * y[i] * y[i];
      for (int j = 0; j < D; ++j) g[j] += Q[j + D * i] * h;
    }
  }
  return va;
}

static void evb_parts(const or_pes *pes, const double *x, double *va, double *vb, double *ga,
                      double *gb, double *px) {
  int D = pes->ndof, nd = pes->ndim;
  const double *Q = pes->dp + 1, *lam = Q + D * D, *cub = lam + D, *xa = cub + D;
  const int *perm = pes->ip;
  for (int a = 0; a < pes->natom; ++a)
    for (int d = 0; d < nd; ++d) px[a * nd + d] = x[perm[a] * nd + d];
  *va = evb_va(D, Q, lam, cub, xa, x, ga);
  double *gpx = (double *)alloca(sizeof(double) * D);
  *vb = evb_va(D, Q, lam, cub, xa, px, gb ? gpx : NULL);
  if (gb)
    for (int a = 0; a < pes->natom; ++a)
      for (int d = 0; d < nd; ++d) gb[perm[a] * nd + d] = gpx[a * nd + d];
}

static double water_eval(const or_pes *pes, const double *x, double *g, double *pg) {
  const double De = pes->dp[0], al = pes->dp[1], re = pes->dp[2], kb = pes->dp[3],
               th0 = pes->dp[4], eps = pes->dp[5], sig = pes->dp[6], qH = pes->dp[7];
  int na = pes->natom, nmol = na / 3;
  double E = 0.0, Ecs = 0.0, dDe = 0.0;
  if (g) memset(g, 0, sizeof(double) * pes->ndof);
  for (int m = 0; m < nmol; ++m) {
    const double *O = x + 9 * m;
    double u[2][3], r[2];
    for (int h = 0; h < 2; ++h) {
      const double *H = x + 9 * m + 3 * (h + 1);
      double rr = 0.0;
      for (int d = 0; d < 3; ++d) {
        u[h][d] = H[d] - O[d];
        rr += u[h][d] * u[h][d];
      }
      r[h] = sqrt(rr);
      double e = exp(-al * (r[h] - re));
      E += De * (1.0 - e) * (1.0 - e);
      dDe += (1.0 - e) * (1.0 - e);
      if (g) {
        double dEdr = 2.0 * De * (1.0 - e) * al * e;
        for (int d = 0; d < 3; ++d) {
          g[9 * m + 3 * (h + 1) + d] += dEdr * u[h][d] / r[h];
          g[9 * m + d] -= dEdr * u[h][d] / r[h];
        }
      }
    }
    double dot = u[0][0] * u[1][0] + u[0][1] * u[1][1] + u[0][2] * u[1][2];
    double cs = dot / (r[0] * r[1]);
    double th = acos(cs);
    E += 0.5 * kb * (th - th0) * (th - th0);
    if (g) {
      double dEdc = kb * (th - th0) * (-1.0 / sqrt(1.0 - cs * cs));
      for (int d = 0; d < 3; ++d) {
        double d0 = u[1][d] / (r[0] * r[1]) - cs * u[0][d] / (r[0] * r[0]);
        double d1 = u[0][d] / (r[0] * r[1]) - cs * u[1][d] / (r[1] * r[1]);
        g[9 * m + 3 + d] += dEdc * d0;
        g[9 * m + 6 + d] += dEdc * d1;
        g[9 * m + d] -= dEdc * (d0 + d1);
      }
    }
  }
  for (int i = 0; i < na; ++i) {
    for (int j = i + 1; j < na; ++j) {
      if (i / 3 == j / 3) continue;
      double dx[3], rr = 0.0;
      for (int d = 0; d < 3; ++d) {
        dx[d] = x[3 * i + d] - x[3 * j + d];
        rr += dx[d] * dx[d];
      }
      double r = sqrt(rr);
      double si = (i % 3 == 0) ? -2.0 : 1.0, sj = (j % 3 == 0) ? -2.0 : 1.0;
      double cs = si * sj / r;
      Ecs += cs;
      double dEdr = -qH * qH * si * sj / rr;
      if (i % 3 == 0 && j % 3 == 0) {
        double s2 = sig * sig / rr, s6 = s2 * s2 * s2;
        E += 4.0 * eps * (s6 * s6 - s6);
        dEdr += 4.0 * eps * (-12.0 * s6 * s6 + 6.0 * s6) / r;
      }
      if (g)
        for (int d = 0; d < 3; ++d) {
          g[3 * i + d] += dEdr * dx[d] / r;
          g[3 * j + d] -= dEdr * dx[d] / r;
        }
    }
  }
  E += qH * qH * Ecs;
  if (pg) {
    pg[0] = 2.0 * qH * Ecs;
    pg[1] = dDe;
  }
  return E;
}

static double pes_eval(const or_pes *pes, const double *x, double *g) {
  int D = pes->ndof;
  const double *p = pes->dp;
  double v = 0.0;
  switch (pes->kind) {
  case OR_PES_QUARTIC: {
    double h = p[0], x0 = p[1];
    for (int d = 0; d < D; ++d) {
      double t = (x[d] / x0) * (x[d] / x0) - 1.0;
      v += h * t * t;
      if (g) g[d] = t * 4.0 * h * x[d] / (x0 * x0);
    }
    break;
  }
  case OR_PES_COUPLED2D: {
    double V0 = p[0], a = p[1], m = p[2], w = p[3], c = p[4];
    double X = x[0], Y = x[1];
    double t = X * X / (a * a) - 1.0, s = Y + c * X * X;
    v = V0 * t * t + 0.5 * m * w * w * s * s;
    if (g) {
      g[0] = 4.0 * V0 * t * X / (a * a) + m * w * w * s * 2.0 * c * X;
      g[1] = m * w * w * s;
    }
    break;
  }
  case OR_PES_EVB: {
    double kappa = p[0], va, vb;
    double *ga = (double *)alloca(sizeof(double) * D), *gb = (double *)alloca(sizeof(double) * D);
    double *px = (double *)alloca(sizeof(double) * D);
    evb_parts(pes, x, &va, &vb, g ? ga : NULL, g ? gb : NULL, px);
    double dl = va - vb, R = sqrt(0.25 * dl * dl + kappa * kappa);
    v = 0.5 * (va + vb) - R;
    if (g)
      for (int d = 0; d < D; ++d) g[d] = 0.5 * (ga[d] + gb[d]) - dl * (ga[d] - gb[d]) / (4.0 * R);
    break;
  }
  case OR_PES_WATER:
    v = water_eval(pes, x, g, NULL);
    break;
  case OR_PES_HARMONIC: {
    const double *k = p, *c = p + D;
    for (int d = 0; d < D; ++d) {
      v += 0.5 * k[d] * (x[d] - c[d]) * (x[d] - c[d]);
      if (g) g[d] = k[d] * (x[d] - c[d]);
    }
    break;
  }
  default:
    v = NAN;
  }
  return v - pes->v0;
}

double or_pes_V(const or_pes *pes, const double *x) { return pes_eval(pes, x, NULL); }
void or_pes_Vprime(const or_pes *pes, const double *x, double *grad) { pes_eval(pes, x, grad); }
double or_pes_potforce(const or_pes *pes, const double *x, double *grad) {
  return pes_eval(pes, x, grad);
}

int or_pes_param_grad(const or_pes *pes, const double *x, double *out) {
  if (pes->kind == OR_PES_WATER) {
    water_eval(pes, x, NULL, out);
    return 2;
  }
  return 0;
}

/* hess[(i) + D*(k)] = d grad_k / d x_i  (hess(i,j,k,l) in the reference's 4-index form) */
void or_pes_Vdoubleprime(const or_pes *pes, const double *x, double *hess) {
  int D = pes->ndof;
  const double *p = pes->dp;
  memset(hess, 0, sizeof(double) * D * D);
  switch (pes->kind) {
  case OR_PES_QUARTIC: {
    double h = p[0], x0 = p[1];
    for (int d = 0; d < D; ++d)
      hess[d + D * d] = 4.0 * h * (3.0 * x[d] * x[d] / (x0 * x0) - 1.0) / (x0 * x0);
    return;
  }
  case OR_PES_COUPLED2D: {
    double V0 = p[0], a = p[1], m = p[2], w = p[3], c = p[4];
    double X = x[0], Y = x[1], s = Y + c * X * X;
    hess[0] = 4.0 * V0 * (3.0 * X * X / (a * a) - 1.0) / (a * a) +
              m * w * w * (4.0 * c * c * X * X + 2.0 * c * s);
    hess[1] = hess[2] = m * w * w * 2.0 * c * X;
    hess[3] = m * w * w;
    return;
  }
  case OR_PES_HARMONIC:
    for (int d = 0; d < D; ++d) hess[d + D * d] = p[d];
    return;
  case OR_PES_EVB: {
    /* V = (VA+VB)/2 - R, R = sqrt(dl^2/4 + kappa^2), dl = VA - VB.
     * Hess = (HA+HB)/2 - [g g^T + dl (HA-HB)]/(4R) + dl^2 g g^T/(16 R^3), g = gA - gB */
    int nd = pes->ndim;
    const double *Q = p + 1, *lam = Q + D * D, *cub = lam + D, *xa = cub + D;
    const int *perm = pes->ip;
    double kappa = p[0], va, vb;
    double *ga = (double *)alloca(sizeof(double) * D), *gb = (double *)alloca(sizeof(double) * D);
    double *px = (double *)alloca(sizeof(double) * D);
    evb_parts(pes, x, &va, &vb, ga, gb, px);
    double dl = va - vb, R = sqrt(0.25 * dl * dl + kappa * kappa);
    double *HA = (double *)malloc(sizeof(double) * D * D), *HB = (double *)malloc(sizeof(double) * D * D);
    double *ya = (double *)alloca(sizeof(double) * D), *yb = (double *)alloca(sizeof(double) * D);
    for (int i = 0; i < D; ++i) {
      double s = 0.0, t = 0.0;
      for (int j = 0; j < D; ++j) {
        s += Q[j + D * i] * (x[j] - xa[j]);
        t += Q[j + D * i] * (px[j] - xa[j]);
      }
      ya[i] = s;
      yb[i] = t;
    }
    /* HA(j,k) = sum_i Q(j,i) (lam_i + 6 c_i y_i) Q(k,i) */
    for (int j = 0; j < D; ++j)
      for (int k = 0; k < D; ++k) {
        double sa = 0.0, sb = 0.0;
        for (int i = 0; i < D; ++i) {
          sa += Q[j + D * i] * (lam[i] + 6.0 * cub[i] * ya[i]) * Q[k + D * i];
          sb += Q[j + D * i] * (lam[i] + 6.0 * cub[i] * yb[i]) * Q[k + D * i];
        }
        HA[j + D * k] = sa;
        HB[j + D * k] = sb; /* still in permuted coordinates */
      }
    /* HB in x coordinates: HB_x(perm[a]d, perm[b]e) = HB_px(ad, be) */
    double *HBx = (double *)malloc(sizeof(double) * D * D);
    for (int a = 0; a < pes->natom; ++a)
      for (int d = 0; d < nd; ++d)
        for (int b = 0; b < pes->natom; ++b)
          for (int e = 0; e < nd; ++e)
            HBx[(perm[a] * nd + d) + D * (perm[b] * nd + e)] = HB[(a * nd + d) + D * (b * nd + e)];
    for (int j = 0; j < D; ++j)
      for (int k = 0; k < D; ++k) {
        double gj = ga[j] - gb[j], gk = ga[k] - gb[k];
        hess[j + D * k] = 0.5 * (HA[j + D * k] + HBx[j + D * k]) -
                          (gj * gk + dl * (HA[j + D * k] - HBx[j + D * k])) / (4.0 * R) +
                          dl * dl * gj * gk / (16.0 * R * R * R);
      }
    free(HA);
    free(HB);
    free(HBx);
    return;
  }
  default: {
    /* central finite differences of the gradient, eps = 1e-4, as the reference's
     * plugins do (mcmod_waterdimer.f90:83-103) */
    const double e = 1e-4;
    double *xp = (double *)malloc(sizeof(double) * D);
    double *gp = (double *)malloc(sizeof(double) * D), *gm = (double *)malloc(sizeof(double) * D);
    memcpy(xp, x, sizeof(double) * D);
    for (int i = 0; i < D; ++i) {
      xp[i] = x[i] + e;
      pes_eval(pes, xp, gp);
      xp[i] = x[i] - e;
      pes_eval(pes, xp, gm);
      xp[i] = x[i];
      for (int k = 0; k < D; ++k) hess[i + D * k] = (gp[k] - gm[k]) / (2.0 * e);
    }
    free(xp);
    free(gp);
    free(gm);
    return;
  }
  }
}

/* ===================================================================== noise */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
 * 1, 2, 3", SC11), written from the published description. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void philox_block(uint64_t seed, uint64_t stream, uint64_t block, double *u0, double *u1) {
  uint32_t ctr[4] = {(uint32_t)block, (uint32_t)(block >> 32), (uint32_t)stream,
                     (uint32_t)(stream >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  or_philox4x32_10(ctr, key, o);
  uint64_t a = ((uint64_t)o[1] << 32) | o[0], b = ((uint64_t)o[3] << 32) | o[2];
  *u0 = ((double)(a >> 11) + 0.5) * 0x1.0p-53; /* in (0,1) */
  *u1 = ((double)(b >> 11) + 0.5) * 0x1.0p-53;
}

/* Box-Muller: draw 2k uses cos, draw 2k+1 uses sin of the same block k. */
double or_philox_normal(uint64_t seed, uint64_t stream, int64_t index) {
  double u0, u1;
  philox_block(seed, stream, (uint64_t)index >> 1, &u0, &u1);
  double r = sqrt(-2.0 * log(u0));
  double ang = 6.283185307179586 * u1;
  return (index & 1) ? r * sin(ang) : r * cos(ang);
}

/* Poisson(mean) from Philox: inversion below mean 16, rounded normal approximation
 * above (the reference's VSL POISNORM is also a normal approximation for large
 * means).  Uses the stream's high half to stay disjoint from the normals. */
int32_t or_philox_poisson(uint64_t seed, uint64_t stream, int64_t index, double mean) {
  uint64_t s = stream ^ 0x8000000000000000ull;
  double u0, u1;
  philox_block(seed, s, (uint64_t)index, &u0, &u1);
  if (mean < 16.0) {
    double pk = exp(-mean), cdf = pk;
    int32_t k = 0;
    while (u0 > cdf && k < 1000) {
      ++k;
      pk *= mean / k;
      cdf += pk;
    }
    return k;
  }
  double z = sqrt(-2.0 * log(u0)) * cos(6.283185307179586 * u1);
  double k = floor(mean + sqrt(mean) * z + 0.5);
  return k < 0.0 ? 0 : (int32_t)k;
}

void or_rng_normals(or_rng *r, int64_t count, double stdev, double *out) {
  for (int64_t i = 0; i < count; ++i) {
    double z;
    if (r->normals) {
      if (r->cursor >= r->nnormals) {
        r->error = 1;
        z = 0.0;
      } else {
        z = r->normals[r->cursor];
      }
    } else {
      z = or_philox_normal(r->seed, r->stream, r->cursor);
    }
    r->cursor++;
    out[i] = stdev * z;
  }
}

int32_t or_rng_poisson(or_rng *r, double mean) {
  int32_t k;
  if (r->kicks) {
    if (r->kcursor >= r->nkicks) {
      r->error = 1;
      k = 1;
    } else {
      k = r->kicks[r->kcursor];
    }
  } else {
    k = or_philox_poisson(r->seed, r->stream, r->kcursor, mean);
  }
  r->kcursor++;
  return k;
}

/* ===================================================================== normal modes */
int or_nm_init(or_nm *nm, int n, int ndim, int natom, const double *mass, double betan,
               double tau, double dt, double gamma, int cayley, const double *a,
               const double *b) {
  int ndof = ndim * natom;
  nm->n = n; nm->ndim = ndim; nm->natom = natom; nm->ndof = ndof;
  nm->betan = betan; nm->tau = tau; nm->dt = dt; nm->gamma = gamma; nm->cayley = cayley;
  nm->mass = mass;
  nm->lam = (double *)malloc(sizeof(double) * n);
  nm->beadmass = (double *)malloc(sizeof(double) * n * natom);
  nm->T = (double *)malloc(sizeof(double) * n * n);
  nm->beadvec = (double *)malloc(sizeof(double) * n * ndof);
  nm->c1 = (double *)malloc(sizeof(double) * n * natom);
  nm->c2 = (double *)malloc(sizeof(double) * n * natom);
  nm->work = (double *)malloc(sizeof(double) * 4 * n * ndof);
  /* verletmodule.f90:310-336 */
  for (int i = 1; i <= n; ++i) {
    for (int j = 0; j < natom; ++j) {
      nm->lam[i - 1] = 2.0 * sin((double)i * OR_PI / (double)(2 * n + 2)) / betan;
      double lt = nm->lam[i - 1] * tau;
      nm->beadmass[(i - 1) * natom + j] = mass[j] * (lt * lt);
    }
    for (int l = i; l <= n; ++l) {
      double t = sin((double)(i * l) * OR_PI / (double)(n + 1)) * sqrt(2.0 / (double)(n + 1));
      nm->T[(i - 1) * n + (l - 1)] = t;
      nm->T[(l - 1) * n + (i - 1)] = t;
      if (t != t) return 3;
    }
    for (int dim = 0; dim < ndim; ++dim)
      for (int k = 0; k < natom; ++k) {
        int dofi = k * ndim + dim;
        double bv = a[dofi] * sin((double)i * OR_PI / (double)(n + 1)) +
                    b[dofi] * sin((double)(n * i) * OR_PI / (double)(n + 1));
        bv = bv * sqrt(2.0 / (double)(n + 1));
        double lb = nm->lam[i - 1] * betan;
        bv = bv / (lb * lb);
        nm->beadvec[dofi * n + (i - 1)] = bv;
      }
  }
  /* verletmodule.f90:380-386 */
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < natom; ++j) {
      double c1 = exp(-gamma * dt * nm->lam[i] * sqrt(mass[j] / nm->beadmass[i * natom + j]));
      nm->c1[i * natom + j] = c1;
      nm->c2[i * natom + j] = sqrt(1.0 - c1 * c1);
    }
  return 0;
}

void or_nm_free(or_nm *nm) {
  free(nm->lam); free(nm->beadmass); free(nm->T); free(nm->beadvec);
  free(nm->c1); free(nm->c2); free(nm->work);
  memset(nm, 0, sizeof(*nm));
}

/* dsymv('U', n, 1, T, n, x, 1, 0, y, 1) -- y = T x, T symmetric */
/* The transforms' symmetric mat-vec.  The reference calls BLAS dsymv('U', ...)
 * (verletmodule.f90:259, 279); the restatement loops by default, and
 * or_use_blas_dsymv swaps in a BLAS dsymv_ (e.g. MKL's) so the CPU baseline can be
 * calibrated against the library the reference links (makefile:6).  NULL restores the loop. */
typedef void (*or_dsymv_fn)(const char *uplo, const int *n, const double *alpha, const double *a,
                            const int *lda, const double *x, const int *incx, const double *beta,
                            double *y, const int *incy);
static or_dsymv_fn g_blas_dsymv = NULL;
void or_use_blas_dsymv(void *f) { g_blas_dsymv = (or_dsymv_fn)f; }

static void symv(int n, const double *T, const double *x, double *y) {
  if (g_blas_dsymv) {
    const double one = 1.0, zero = 0.0;
    const int inc = 1;
    g_blas_dsymv("U", &n, &one, T, &n, x, &inc, &zero, y, &inc);
    return;
  }
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int l = 0; l < n; ++l) s += T[i * n + l] * x[l];
    y[i] = s;
  }
}

void or_nm_forward(const or_nm *nm, const double *x, double *q, int dof) {
  int n = nm->n;
  symv(n, nm->T, x, q);
  if (dof >= 0)
    for (int i = 0; i < n; ++i) q[i] -= nm->beadvec[dof * n + i];
}

void or_nm_backward(const or_nm *nm, const double *q, double *x, int dof) {
  int n = nm->n;
  if (dof >= 0) {
    double *t = (double *)alloca(sizeof(double) * n);
    for (int i = 0; i < n; ++i) t[i] = q[i] + nm->beadvec[dof * n + i];
    symv(n, nm->T, t, x);
  } else {
    symv(n, nm->T, q, x);
  }
}

int or_step_v(const or_nm *nm, const or_pes *pes, double h, double *x, double *p) {
  int n = nm->n, D = nm->ndof;
  double *xb = (double *)alloca(sizeof(double) * D), *g = (double *)alloca(sizeof(double) * D);
  for (int i = 0; i < n; ++i) {
    for (int d = 0; d < D; ++d) xb[d] = x[d * n + i];
    or_pes_Vprime(pes, xb, g);
    /* loops j=1..ndim, k=1..natom (verletmodule.f90:574-582) */
    for (int j = 0; j < nm->ndim; ++j)
      for (int k = 0; k < nm->natom; ++k) {
        int d = k * nm->ndim + j;
        p[d * n + i] = p[d * n + i] - g[d] * h;
        if (p[d * n + i] != p[d * n + i]) return 3;
      }
  }
  return 0;
}

int or_step_nm(const or_nm *nm, double h, double *x, double *p) {
  int n = nm->n, D = nm->ndof;
  double *q = nm->work, *pip = nm->work + n * D;
  for (int d = 0; d < D; ++d) {
    or_nm_forward(nm, p + d * n, pip + d * n, -1);
    or_nm_forward(nm, x + d * n, q + d * n, d);
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < nm->ndim; ++j)
      for (int k = 0; k < nm->natom; ++k) {
        int d = k * nm->ndim + j;
        double bm = nm->beadmass[i * nm->natom + k];
        double om = sqrt(nm->mass[k] / bm) * nm->lam[i];
        double pi0 = pip[d * n + i], q0 = q[d * n + i], np, nq;
        if (nm->cayley) {
          np = pi0 * (4.0 - om * om * h * h) - 4.0 * q0 * bm * om * om * h;
          np = np / (4.0 + om * om * h * h);
          nq = q0 * (4.0 - om * om * h * h) + 4.0 * pi0 * h / bm;
          nq = nq / (4.0 + om * om * h * h);
        } else {
          np = pi0 * cos(h * om) - q0 * om * bm * sin(om * h);
          nq = q0 * cos(h * om) + pi0 * sin(om * h) / (om * bm);
        }
        if (np != np) return 3;
        pip[d * n + i] = np;
        q[d * n + i] = nq;
      }
  for (int d = 0; d < D; ++d) {
    or_nm_backward(nm, pip + d * n, p + d * n, -1);
    or_nm_backward(nm, q + d * n, x + d * n, d);
  }
  return 0;
}

int or_step_langevin(const or_nm *nm, or_rng *rng, double *p) {
  int n = nm->n, D = nm->ndof;
  double *pk = nm->work, *pt = nm->work + n * D;
  or_rng_normals(rng, (int64_t)n * D, 1.0, pk);
  for (int i = 0; i < nm->ndim; ++i)
    for (int j = 0; j < nm->natom; ++j) {
      int dofi = j * nm->ndim + i;
      or_nm_forward(nm, p + dofi * n, pt, -1);
      for (int k = 0; k < n; ++k) {
        double c1 = nm->c1[k * nm->natom + j], c2 = nm->c2[k * nm->natom + j];
        double bm = nm->beadmass[k * nm->natom + j];
        pt[k] = (c1 * c1) * pt[k] + sqrt(bm / nm->betan) * c2 * sqrt(1.0 + c1 * c1) * pk[dofi * n + k];
      }
      or_nm_backward(nm, pt, p + dofi * n, -1);
    }
  return rng->error ? 4 : 0;
}

int or_time_step_pile(const or_nm *nm, const or_pes *pes, or_rng *rng, double *x, double *p) {
  int rc;
  if ((rc = or_step_v(nm, pes, nm->dt, x, p))) return rc;
  if ((rc = or_step_nm(nm, 0.5 * nm->dt, x, p))) return rc;
  if ((rc = or_step_langevin(nm, rng, p))) return rc;
  return or_step_nm(nm, 0.5 * nm->dt, x, p);
}

int or_time_step_nm(const or_nm *nm, const or_pes *pes, double *x, double *p) {
  int rc;
  if ((rc = or_step_nm(nm, 0.5 * nm->dt, x, p))) return rc;
  if ((rc = or_step_v(nm, pes, nm->dt, x, p))) return rc;
  return or_step_nm(nm, 0.5 * nm->dt, x, p);
}

void or_andersen_resample(const or_nm *nm, or_rng *rng, double *p) {
  int n = nm->n;
  double *pp = nm->work, *tp = nm->work + n;
  for (int j = 0; j < nm->ndim; ++j)
    for (int k = 0; k < nm->natom; ++k) {
      int dofi = k * nm->ndim + j;
      or_rng_normals(rng, n, sqrt(1.0 / nm->betan), pp);
      for (int l = 0; l < n; ++l) tp[l] = pp[l] * sqrt(nm->beadmass[l * nm->natom + k]);
      or_nm_backward(nm, tp, p + dofi * n, -1);
    }
}

/* ===================================================================== path */
/* spline + tridag (instantonmod.f90:397-480), Numerical Recipes' natural cubic spline */
void or_spline(const double *x, const double *y, int n, double yp1, double ypn, double *y2) {
  double *a = (double *)malloc(sizeof(double) * n), *b = (double *)malloc(sizeof(double) * n);
  double *c = (double *)malloc(sizeof(double) * n), *r = (double *)malloc(sizeof(double) * n);
  double *gam = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n - 1; ++i) c[i] = x[i + 1] - x[i];
  for (int i = 0; i < n - 1; ++i) r[i] = 6.0 * ((y[i + 1] - y[i]) / c[i]);
  for (int i = n - 2; i >= 1; --i) r[i] = r[i] - r[i - 1]; /* r(2:n-1)=r(2:n-1)-r(1:n-2), vector op */
  for (int i = 1; i < n - 1; ++i) a[i] = c[i - 1];
  for (int i = 1; i < n - 1; ++i) b[i] = 2.0 * (c[i] + a[i]);
  b[0] = 1.0;
  b[n - 1] = 1.0;
  if (yp1 > 0.99e30) {
    r[0] = 0.0;
    c[0] = 0.0;
  } else {
    r[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
    c[0] = 0.5;
  }
  if (ypn > 0.99e30) {
    r[n - 1] = 0.0;
    a[n - 1] = 0.0;
  } else {
    r[n - 1] = (-3.0 / (x[n - 1] - x[n - 2])) * ((y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]) - ypn);
    a[n - 1] = 0.5;
  }
  /* tridag(a(2:n), b(1:n), c(1:n-1), r, y2): subdiagonal aa(j) = a(j+1) */
  double bet = b[0];
  y2[0] = r[0] / bet;
  for (int j = 1; j < n; ++j) {
    gam[j] = c[j - 1] / bet;
    bet = b[j] - a[j] * gam[j];
    y2[j] = (r[j] - a[j] * y2[j - 1]) / bet;
  }
  for (int j = n - 2; j >= 0; --j) y2[j] = y2[j] - gam[j + 1] * y2[j + 1];
  free(a); free(b); free(c); free(r); free(gam);
}

int or_locate(const double *xx, int n, double x) {
  int ascnd = (xx[n - 1] >= xx[0]);
  int jl = 0, ju = n + 1;
  while (ju - jl > 1) {
    int jm = (ju + jl) / 2;
    if (ascnd == (x >= xx[jm - 1]))
      jl = jm;
    else
      ju = jm;
  }
  if (x == xx[0]) return 1;
  if (x == xx[n - 1]) return n - 1;
  return jl;
}

static int klo_of(const double *xa, int n, double x) {
  int k = or_locate(xa, n, x);
  if (k > n - 1) k = n - 1;
  if (k < 1) k = 1;
  return k - 1; /* 0-based */
}

double or_splint(const double *xa, const double *ya, const double *y2a, int n, double x) {
  int klo = klo_of(xa, n, x), khi = klo + 1;
  double h = xa[khi] - xa[klo];
  double a = (xa[khi] - x) / h, b = (x - xa[klo]) / h;
  return a * ya[klo] + b * ya[khi] +
         ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2a[khi]) * (h * h) / 6.0;
}

double or_splin_grad(const double *xa, const double *ya, const double *y2a, int n, double x) {
  int klo = klo_of(xa, n, x), khi = klo + 1;
  double h = xa[khi] - xa[klo];
  double a = (xa[khi] - x) / h, b = (x - xa[klo]) / h;
  return ((ya[khi] - ya[klo]) / h) +
         ((1.0 - 3.0 * a * a) * y2a[klo] + (3.0 * b * b - 1.0) * y2a[khi]) * h / 6.0;
}

int or_path_init(or_path *pth, int npath, int ndof, const double *lampath, const double *path) {
  pth->npath = npath;
  pth->ndof = ndof;
  pth->lampath = lampath;
  pth->path = path;
  pth->y2 = (double *)malloc(sizeof(double) * npath * ndof);
  for (int d = 0; d < ndof; ++d)
    or_spline(lampath, path + d * npath, npath, 1.0e31, 1.0e31, pth->y2 + d * npath);
  return 0;
}

void or_path_free(or_path *pth) {
  free(pth->y2);
  pth->y2 = NULL;
}

void or_init_path(const or_nm *nm, const or_path *pth, or_rng *rng, double xi, double *x, double *p) {
  int n = nm->n, np = pth->npath;
  for (int i = 0; i < nm->ndim; ++i)
    for (int j = 0; j < nm->natom; ++j) {
      int d = j * nm->ndim + i;
      for (int k = 1; k <= n; ++k) {
        double xe = (double)(k - 1) * xi / (double)(n - 1);
        x[d * n + (k - 1)] = or_splint(pth->lampath, pth->path + d * np, pth->y2 + d * np, np, xe);
      }
    }
  or_andersen_resample(nm, rng, p); /* same draw pattern as verletmodule.f90:102-115 */
}

/* ===================================================================== TI loops */
static double ti_contr(const or_nm *nm, const double *dbdl, const double *x) {
  int n = nm->n;
  double contr = 0.0;
  for (int j = 0; j < nm->ndim; ++j)
    for (int k = 0; k < nm->natom; ++k) {
      int d = k * nm->ndim + j;
      contr = contr + nm->mass[k] * (-x[d * n + (n - 1)]) * dbdl[d];
    }
  return contr;
}

int or_propagate_pile(const or_nm *nm, const or_pes *pes, or_rng *rng, or_ti *ti, double *x,
                      double *p, double *dhdr) {
  double sum = ti->restartnmc > 0 ? ti->dhdr_sum : 0.0;
  ti->nreinit = 0;
  for (int64_t ii = 1; ii <= ti->nmc; ++ii) {
    int rc = or_time_step_pile(nm, pes, rng, x, p);
    if (rc) return rc;
    if (ii > ti->imin) {
      double contr = ti_contr(nm, ti->dbdl, x);
      if (fabs(contr) < ti->dhdrlimit || ti->dhdrlimit < 0.0) {
        sum = sum + contr;
      } else {
        or_init_path(nm, ti->path, rng, ti->xi, x, p);
        ti->nreinit++;
      }
    }
  }
  ti->dhdr_sum = sum;
  *dhdr = sum / (double)(ti->nmc + ti->restartnmc - ti->imin);
  return rng->error ? 4 : 0;
}

int or_propagate_andersen(const or_nm *nm, const or_pes *pes, or_rng *rng, or_ti *ti,
                          int64_t noutput, double *x, double *p, double *dhdr) {
  double sum = ti->restartnmc > 0 ? ti->dhdr_sum : 0.0;
  int64_t count = 0;
  int32_t rkick = or_rng_poisson(rng, (double)noutput);
  for (int64_t ii = 1; ii <= ti->nmc; ++ii) {
    count = count + 1;
    if (count >= rkick) {
      or_andersen_resample(nm, rng, p);
      rkick = or_rng_poisson(rng, (double)noutput);
      count = 0;
    }
    int rc = or_time_step_nm(nm, pes, x, p);
    if (rc) return rc;
    if (ii > ti->imin) sum = sum + ti_contr(nm, ti->dbdl, x);
  }
  ti->dhdr_sum = sum;
  *dhdr = sum / (double)(ti->nmc + ti->restartnmc - ti->imin);
  return rng->error ? 4 : 0;
}

void or_gauleg(double x1, double x2, double *x, double *w, int n) {
  const double EPS = 3.0e-14;
  int m = (n + 1) / 2;
  double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= m; ++i) {
    double z = cos(OR_PI * (i - 0.25) / (n + 0.5)), z1, pp, p1, p2, p3;
    do {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (fabs(z - z1) > EPS);
    x[i - 1] = xm - xl * z;
    x[n - i] = xm + xl * z;
    w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

/* ===================================================================== instanton */
static void bead_of(const or_chain *c, const double *x, int i, double *xb) {
  for (int d = 0; d < c->ndof; ++d) xb[d] = x[d * c->n + i];
}

double or_UM(const or_chain *c, const or_pes *pes, const double *x, const double *a, const double *b) {
  int n = c->n, nd = c->ndim;
  double *xb = (double *)alloca(sizeof(double) * c->ndof);
  double um = 0.0, bn2 = c->betan * c->betan;
  for (int i = 0; i < n - 1; ++i) {
    bead_of(c, x, i, xb);
    um = um + or_pes_V(pes, xb);
    for (int j = 0; j < nd; ++j)
      for (int k = 0; k < c->natom; ++k) {
        int d = k * nd + j;
        double dx = x[d * n + i + 1] - x[d * n + i];
        um = um + (0.5 * c->mass[k] / bn2) * dx * dx;
      }
  }
  bead_of(c, x, n - 1, xb);
  um = um + or_pes_V(pes, xb);
  if (c->fixedends)
    for (int j = 0; j < nd; ++j)
      for (int k = 0; k < c->natom; ++k) {
        int d = k * nd + j;
        double d1 = x[d * n] - a[d], d2 = b[d] - x[d * n + n - 1];
        um = um + (0.5 * c->mass[k] / bn2) * d1 * d1;
        um = um + (0.5 * c->mass[k] / bn2) * d2 * d2;
      }
  return um;
}

static void spring_grad(const or_chain *c, const double *x, const double *a, const double *b, int i,
                        double *g) {
  int n = c->n, nd = c->ndim;
  double bn2 = c->betan * c->betan;
  for (int j = 0; j < nd; ++j)
    for (int k = 0; k < c->natom; ++k) {
      int d = k * nd + j;
      const double *xd = x + d * n;
      double m = c->mass[k], v;
      if (i == 0) {
        v = c->fixedends ? m * (2.0 * xd[0] - a[d] - xd[1]) / bn2 : m * (xd[0] - xd[1]) / bn2;
      } else if (i == n - 1) {
        v = c->fixedends ? m * (2.0 * xd[n - 1] - xd[n - 2] - b[d]) / bn2 : m * (xd[n - 1] - xd[n - 2]) / bn2;
      } else {
        v = m * (2.0 * xd[i] - xd[i - 1] - xd[i + 1]) / bn2;
      }
      g[d * n + i] = v;
    }
}

void or_UMprime(const or_chain *c, const or_pes *pes, const double *x, const double *a, const double *b, double *g) {
  int n = c->n, D = c->ndof;
  double *xb = (double *)alloca(sizeof(double) * D), *gr = (double *)alloca(sizeof(double) * D);
  for (int i = 0; i < n; ++i) {
    spring_grad(c, x, a, b, i, g);
    bead_of(c, x, i, xb);
    or_pes_Vprime(pes, xb, gr);
    for (int d = 0; d < D; ++d) g[d * n + i] = g[d * n + i] + gr[d];
  }
}

double or_UMforceenergy(const or_chain *c, const or_pes *pes, const double *x, const double *a, const double *b, double *g) {
  int n = c->n, D = c->ndof, nd = c->ndim;
  double *xb = (double *)alloca(sizeof(double) * D), *gr = (double *)alloca(sizeof(double) * D);
  double um = 0.0, bn2 = c->betan * c->betan;
  for (int i = 0; i < n; ++i) {
    bead_of(c, x, i, xb);
    double e = or_pes_potforce(pes, xb, gr);
    um = um + e;
    for (int j = 0; j < nd; ++j)
      for (int k = 0; k < c->natom; ++k) {
        int d = k * nd + j;
        if (i < n - 1) {
          double dx = x[d * n + i + 1] - x[d * n + i];
          um = um + (0.5 * c->mass[k] / bn2) * dx * dx;
        }
      }
    spring_grad(c, x, a, b, i, g);
    for (int d = 0; d < D; ++d) g[d * n + i] = g[d * n + i] + gr[d];
  }
  if (c->fixedends)
    for (int j = 0; j < nd; ++j)
      for (int k = 0; k < c->natom; ++k) {
        int d = k * nd + j;
        double d1 = x[d * n] - a[d], d2 = b[d] - x[d * n + n - 1];
        um = um + (0.5 * c->mass[k] / bn2) * d1 * d1;
        um = um + (0.5 * c->mass[k] / bn2) * d2 * d2;
      }
  return um;
}

void or_UMhessian(const or_chain *c, const or_pes *pes, const double *x, int singlewell, double *band) {
  int n = c->n, D = c->ndof, nd = c->ndim, tot = n * D;
  double bn2 = c->betan * c->betan;
  double *xb = (double *)alloca(sizeof(double) * D);
  double *hess = (double *)calloc((size_t)D * D, sizeof(double));
  memset(band, 0, sizeof(double) * (size_t)(D + 1) * tot);
  for (int i = 1; i <= n; ++i) {
    if ((i == 1 && singlewell) || !singlewell) {
      bead_of(c, x, i - 1, xb);
      or_pes_Vdoubleprime(pes, xb, hess);
    }
    for (int j1 = 0; j1 < nd; ++j1)
      for (int k1 = 0; k1 < c->natom; ++k1)
        for (int j2 = 0; j2 < nd; ++j2)
          for (int k2 = 0; k2 < c->natom; ++k2) {
            int idof1 = k1 * nd + j1, idof2 = k2 * nd + j2;
            int f1 = D * (i - 1) + idof1, f2 = D * (i - 1) + idof2;
            if (f2 < f1) continue;
            /* hess(j2,k2,j1,k1) = d grad(j1,k1) / d x(j2,k2) */
            double h = hess[idof2 + D * idof1] / sqrt(c->mass[k1] * c->mass[k2]);
            if (idof1 == idof2) {
              band[0 * tot + f1] = 2.0 / bn2 + h;
              /* coupling to the next bead is set only for i > 1 (instantonmod.f90:203);
               * for i = n it falls outside the matrix and is never read */
              if (i > 1) band[D * tot + f1] = -1.0 / bn2;
            } else {
              int index = f2 - f1;
              band[index * tot + f1] = h;
            }
          }
  }
  free(hess);
}

/* ===================================================================== CPU baseline */
typedef struct {
  const or_pes *pes;
  int n, ndim, natom;
  const double *mass, *a, *b, *dbdl, *x0;
  double betan, tau, dt, gamma;
  int64_t nsteps;
  uint64_t seed;
  int first, last;
  double *out;
} bench_job;

static void *bench_worker(void *arg) {
  bench_job *j = (bench_job *)arg;
  int D = j->ndim * j->natom, n = j->n;
  or_nm nm;
  or_nm_init(&nm, n, j->ndim, j->natom, j->mass, j->betan, j->tau, j->dt, j->gamma, 0, j->a, j->b);
  double *x = (double *)malloc(sizeof(double) * n * D), *p = (double *)malloc(sizeof(double) * n * D);
  for (int t = j->first; t < j->last; ++t) {
    or_rng rng;
    memset(&rng, 0, sizeof(rng));
    rng.seed = j->seed;
    rng.stream = (uint64_t)t;
    memcpy(x, j->x0, sizeof(double) * n * D);
    or_andersen_resample(&nm, &rng, p);
    or_ti ti;
    memset(&ti, 0, sizeof(ti));
    ti.nmc = j->nsteps;
    ti.imin = 0;
    ti.dhdrlimit = -1.0;
    ti.dbdl = j->dbdl;
    double dh = 0.0;
    or_propagate_pile(&nm, j->pes, &rng, &ti, x, p, &dh);
    j->out[t] = dh;
  }
  free(x);
  free(p);
  or_nm_free(&nm);
  return NULL;
}

double or_bench_pile(const or_pes *pes, int n, int ndim, int natom, const double *mass,
                     double betan, double tau, double dt, double gamma, const double *a,
                     const double *b, const double *dbdl, const double *x0, int ntraj,
                     int64_t nsteps, int nthreads, uint64_t seed, double *dhdr_out) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > ntraj) nthreads = ntraj;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  bench_job *jobs = (bench_job *)malloc(sizeof(bench_job) * nthreads);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int t = 0; t < nthreads; ++t) {
    bench_job *j = &jobs[t];
    j->pes = pes; j->n = n; j->ndim = ndim; j->natom = natom; j->mass = mass;
    j->a = a; j->b = b; j->dbdl = dbdl; j->x0 = x0;
    j->betan = betan; j->tau = tau; j->dt = dt; j->gamma = gamma;
    j->nsteps = nsteps; j->seed = seed; j->out = dhdr_out;
    j->first = (int)((int64_t)ntraj * t / nthreads);
    j->last = (int)((int64_t)ntraj * (t + 1) / nthreads);
    pthread_create(&th[t], NULL, bench_worker, j);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(th);
  free(jobs);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

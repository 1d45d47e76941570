// rp_pes.h -- device potentials behind the mcmod_mass plugin interface.
//
// Interface restated (mcmod_waterdimer.f90:12-103): for one bead x(ndim, natom)
// (dof order dof = atom*ndim + dim, bohr) return V(x) - V0 (Hartree) and the
// gradient +dV/dx.  Every potential here is energy + hand-written reverse-mode
// gradient, templated on the scalar so Dual seeds give exact Hessian-vector
// products.  Parameter layouts match the oracle's (oracle/oracle.c):
//   QUARTIC   dp = {h, x0}                       sum_d h((x_d/x0)^2-1)^2  (mcmod_1d.f90:20)
//   COUPLED2D dp = {V0, a, m, w, c}              V0(x^2/a^2-1)^2 + m w^2 (y+c x^2)^2/2
//   EVB       dp = {kappa, Q[D*D], lam[D], cub[D], xa[D]}, ip = perm[natom]
//             VA = y.Lam.y/2 + sum c y^3, y = Q^T (x - xa); VB(x) = VA(Pi x);
//             V = (VA+VB)/2 - sqrt((VA-VB)^2/4 + kappa^2)
//             (expanded on the device to the padded layout of evb_eval_at)
//   WATER     dp = {De, alpha, re, kbend, theta0, eps, sigma, qH}; atoms O,H,H per
//             molecule; Morse OH + harmonic HOH bend + LJ O-O + charges (qO = -2 qH)
//   HARMONIC  dp = {k[D], c[D]}
#pragma once
#include "rp_dual.h"

namespace rp {

enum { PES_QUARTIC = 1, PES_COUPLED2D = 2, PES_EVB = 3, PES_WATER = 4, PES_HARMONIC = 5,
       PES_HOST = 6 /* callbacks into a host plugin; no device form */ };

struct PesDev {
  int kind, ndim, natom, ndof;
  double v0;
  const double* __restrict__ dp;  // device copy of the real parameters
  const int* __restrict__ ip;     // device copy of the integer parameters
};

template <int MAXD, class S>
__device__ __forceinline__ S quartic_eval(const PesDev& P, const S* x, S* g) {
  const double h = P.dp[0], x0 = P.dp[1], ix2 = 1.0 / (x0 * x0);
  S v = 0.0;
#pragma unroll
  for (int d = 0; d < MAXD; ++d) {
    if (d < P.ndof) {
      S t = x[d] * x[d] * ix2 - 1.0;
      v += h * t * t;
      g[d] = t * (4.0 * h) * x[d] * ix2;
    }
  }
  return v;
}

template <class S>
__device__ __forceinline__ S coupled_eval(const PesDev& P, const S* x, S* g) {
  const double V0 = P.dp[0], a = P.dp[1], m = P.dp[2], w = P.dp[3], c = P.dp[4];
  const double ia2 = 1.0 / (a * a), mw2 = m * w * w;
  S t = x[0] * x[0] * ia2 - 1.0, s = x[1] + c * x[0] * x[0];
  g[0] = (4.0 * V0) * t * x[0] * ia2 + mw2 * s * (2.0 * c) * x[0];
  g[1] = mw2 * s;
  return V0 * t * t + (0.5 * mw2) * s * s;
}

template <int MAXD, class S>
__device__ __forceinline__ S harmonic_eval(const PesDev& P, const S* x, S* g) {
  const int D = P.ndof;
  S v = 0.0;
#pragma unroll
  for (int d = 0; d < MAXD; ++d) {
    if (d < D) {
      S u = x[d] - P.dp[D + d];
      v += 0.5 * P.dp[d] * u * u;
      g[d] = P.dp[d] * u;
    }
  }
  return v;
}

// One harmonic+cubic well: VW = y.Lam.y/2 + sum c y^3 with y = W^T (x - xw), and its
// gradient W (Lam y + 3 c y^2).  Device layout (expanded on the host at creation, rows
// zero-padded to the compile-time stride M so no loop needs a bound check):
//   Wt[i*M + j] = W(j, i)   (row i = column i of W),   cw[i] = sum_j W(j, i) xw_j
// so y_i = Wt_i . x - cw_i.  Well A uses (Q, xa); well B = VA(Pi x) is the same form with
// the rows of Q permuted (W = Pi^T Q) and xw = Pi^T xa.  Padded rows have lam = c = 0.
template <int M, class S>
__device__ __forceinline__ S evb_well(const double* __restrict__ Wt, const double* __restrict__ cw,
                                      const double* __restrict__ lam,
                                      const double* __restrict__ cub, const S* x, S* gx) {
#pragma unroll
  for (int j = 0; j < M; ++j) gx[j] = 0.0;
  S vw = 0.0;
#pragma unroll 1
  for (int i = 0; i < M; ++i) {
    const double* Wi = Wt + i * M;
    S y = -cw[i];
#pragma unroll
    for (int j = 0; j < M; ++j) y += Wi[j] * x[j];
    const double li = lam[i], ci = cub[i];
    S y2 = y * y;
    vw += (0.5 * li) * y2 + ci * y2 * y;
    S h = li * y + (3.0 * ci) * y2;
#pragma unroll
    for (int j = 0; j < M; ++j) gx[j] += Wi[j] * h;
  }
  return vw;
}

// Device layout of the EVB parameters with row stride M (see evb_well):
//   {kappa, WtA[M*M], cwA[M], WtB[M*M], cwB[M], lam[M], cub[M]}
template <int M>
__host__ __device__ constexpr int evb_dev_size() { return 1 + 2 * (M * M + M) + 2 * M; }

template <int M, class S>
__device__ __forceinline__ S evb_eval_at(const double* __restrict__ dp, const S* x, S* g) {
  const double kappa = dp[0];
  const double* WA = dp + 1;
  const double* cA = WA + M * M;
  const double* WB = cA + M;
  const double* cB = WB + M * M;
  const double* lam = cB + M;
  const double* cub = lam + M;
  S gb[M];
  S va = evb_well<M>(WA, cA, lam, cub, x, g);
  S vb = evb_well<M>(WB, cB, lam, cub, x, gb);
  S dl = va - vb;
  S R = sqrt(0.25 * dl * dl + kappa * kappa);
  S f = dl / (4.0 * R);
  S wa = 0.5 - f, wb = 0.5 + f;
#pragma unroll
  for (int j = 0; j < M; ++j) g[j] = wa * g[j] + wb * gb[j];
  return 0.5 * (va + vb) - R;
}

template <int M, class S>
__device__ __forceinline__ S evb_eval(const PesDev& P, const S* x, S* g) {
  return evb_eval_at<M>(P.dp, x, g);
}

template <int NMOL, class S>
__device__ __forceinline__ S water_eval(const PesDev& P, const S* x, S* g, double* pgrad = nullptr) {
  const double De = P.dp[0], al = P.dp[1], re = P.dp[2], kb = P.dp[3], th0 = P.dp[4],
               eps = P.dp[5], sig = P.dp[6], qH = P.dp[7];
  constexpr int NA = 3 * NMOL;
  S E = 0.0, Ecs = 0.0, dDe = 0.0;
#pragma unroll
  for (int j = 0; j < 3 * NA; ++j) g[j] = 0.0;
#pragma unroll
  for (int m = 0; m < NMOL; ++m) {
    S u[2][3], r[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      S rr = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        u[h][d] = x[9 * m + 3 * (h + 1) + d] - x[9 * m + d];
        rr += u[h][d] * u[h][d];
      }
      r[h] = sqrt(rr);
      S e = exp(-al * (r[h] - re));
      S om = 1.0 - e;
      E += De * om * om;
      dDe += om * om;
      S dEdr = (2.0 * De * al) * om * e / r[h];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        S t = dEdr * u[h][d];
        g[9 * m + 3 * (h + 1) + d] += t;
        g[9 * m + d] -= t;
      }
    }
    S dot = u[0][0] * u[1][0] + u[0][1] * u[1][1] + u[0][2] * u[1][2];
    S ir01 = 1.0 / (r[0] * r[1]);
    S cs = dot * ir01;
    S th = acos(cs);
    S dth = th - th0;
    E += (0.5 * kb) * dth * dth;
    S dEdc = -kb * dth / sqrt(1.0 - cs * cs);
    S i00 = cs / (r[0] * r[0]), i11 = cs / (r[1] * r[1]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      S d0 = dEdc * (u[1][d] * ir01 - i00 * u[0][d]);
      S d1 = dEdc * (u[0][d] * ir01 - i11 * u[1][d]);
      g[9 * m + 3 + d] += d0;
      g[9 * m + 6 + d] += d1;
      g[9 * m + d] -= d0 + d1;
    }
  }
#pragma unroll
  for (int i = 0; i < NA; ++i) {
#pragma unroll
    for (int j = i + 1; j < NA; ++j) {
      if (i / 3 == j / 3) continue;
      S dx[3], rr = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        dx[d] = x[3 * i + d] - x[3 * j + d];
        rr += dx[d] * dx[d];
      }
      S ir2 = 1.0 / rr;
      S ir = sqrt(ir2);
      const double sij = ((i % 3 == 0) ? -2.0 : 1.0) * ((j % 3 == 0) ? -2.0 : 1.0);
      Ecs += sij * ir;
      S dEr = (-qH * qH * sij) * ir2 * ir;  // (dE/dr) / r
      if (i % 3 == 0 && j % 3 == 0) {
        S s2 = (sig * sig) * ir2;
        S s6 = s2 * s2 * s2;
        E += (4.0 * eps) * (s6 * s6 - s6);
        dEr += (4.0 * eps) * (-12.0 * s6 * s6 + 6.0 * s6) * ir2;
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        S t = dEr * dx[d];
        g[3 * i + d] += t;
        g[3 * j + d] -= t;
      }
    }
  }
  if (pgrad) {
    pgrad[0] = 2.0 * qH * value(Ecs);
    pgrad[1] = value(dDe);
  }
  return E + (qH * qH) * Ecs;
}

// Dispatcher for kernels that take any small potential (MAXD >= ndof).
template <int MAXD, class S>
__device__ __forceinline__ S pes_eval(const PesDev& P, const S* x, S* g) {
  S v;
  switch (P.kind) {
    case PES_QUARTIC: v = quartic_eval<MAXD>(P, x, g); break;
    case PES_COUPLED2D: v = coupled_eval(P, x, g); break;
    case PES_HARMONIC: v = harmonic_eval<MAXD>(P, x, g); break;
    case PES_EVB: v = evb_eval<MAXD>(P, x, g); break;
    default: v = 0.0; break;
  }
  return v - P.v0;
}

}  // namespace rp

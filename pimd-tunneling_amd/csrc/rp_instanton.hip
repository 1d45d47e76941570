// rp_instanton.hip -- the instanton search (instantonmod.f90:679-777) for a batch of
// chains, resident on the device.
//
// The reference drives L-BFGS-B (Byrd, Lu, Nocedal, Zhu 1995; v3.0 with Morales-Nocedal
// 2011) with nbd = 0, i.e. no bounds, m = 8 corrections, factr = 1e6 and pgtol = eps2.
// Without bounds L-BFGS-B's generalized Cauchy point is the unconstrained step and the
// method is limited-memory BFGS; this file implements that method from the published
// description (Nocedal & Wright, Numerical Optimization, 2nd ed.: two-loop recursion,
// Alg. 7.4; strong-Wolfe line search with zoom, Alg. 3.5/3.6) with the reference's two
// stopping tests:
//   (f_k - f_{k+1}) / max(|f_k|, |f_{k+1}|, 1) <= factr * epsmch,   |g|_inf <= pgtol.
//
// Every vector of every chain (iterate, gradient, direction, the gradients kept by the line
// search, the m correction pairs) lives in HBM; one workgroup per chain runs the chain's
// line-search state machine on scalars that every lane holds identically, and the vector
// work (dot products, the two-loop recursion, trial points) over the chain's coordinates.
// One search round is three launches for all chains at once: the ring's U_M + adjoint
// gradient at every trial point, lb_feed_kernel (line-search step or acceptance and the
// stopping tests), lb_direction_kernel (two-loop recursion and the first trial point of the
// next iteration).  The host only launches rounds and reads a status word per chain every
// few rounds; no coordinates cross PCIe during the search.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "rp_ring.h"

using namespace rp;

namespace {

enum : int { LS_BRACKET = 0, LS_ZOOM = 1 };
enum : int { ST_RUN = -1, ST_CONV = 0, ST_ABNORMAL = 1, ST_MAXITER = 2 };
enum : int { OP_NONE = 0, OP_DIRECTION = 1 };
enum : int { ACT_NONE = 0, ACT_TRIAL = 1, ACT_ACCEPT_CUR = 2, ACT_ACCEPT_LO = 3 };
// gradient copies that go with a line-search step
enum : int { CP_GT_TO_PREV = 1, CP_GT_TO_LO = 2, CP_PREV_TO_LO = 4, CP_G_TO_LO = 8 };

constexpr int kMaxM = 32;         // largest number of corrections
constexpr int kLbThreads = 1024;  // one workgroup of 16 waves per chain
constexpr int kLbWaves = kLbThreads / 64;
constexpr double kC1 = 1e-3;      // sufficient decrease (L-BFGS-B's ftol)
constexpr double kC2 = 0.9;       // curvature (L-BFGS-B's gtol)
constexpr int kMaxLS = 40;        // line-search evaluations (the reference's maxiter argument)
constexpr int kCheckEvery = 4;    // rounds between status reads
constexpr long kMaxRounds = 200000;  // evaluations when the caller sets no iteration cap

// Per chain scalar state.  The correction pairs sit in a ring of m + 1 slots: the history is
// `count` pairs from slot `head`, and slot (head + count) % (m + 1) is always free for the
// candidate pair of an acceptance, so a rejected pair never overwrites a kept one.  The
// gradient at a_prev is the iterate's own gradient G while prev_is_g is set (a_prev = 0),
// else the saved copy Gprev.
struct LbChain {
  double f, f0, dg0, a_cur, a_prev, f_prev, dg_prev, a_lo, f_lo, dg_lo, a_hi, f_hi, dg_hi;
  double ginf, gam;
  int iter, status, nls, phase, count, head, op, prev_is_g;
};

struct LbArgs {
  int nchain, m, maxiter, D;
  double factr, pgtol;
  int64_t tot;
  const int64_t* off;
  const int* n;
  double *X, *G, *Dv, *Glo, *Gprev, *S, *Y;  // iterate, gradient, direction, saved gradients, pairs
  double *XT, *GT;                            // the ring's evaluation point and gradient
  const double* E;                            // the ring's energies
  LbChain* st;
  double* rho;                                // 1/(s.y) per chain and slot, (kMaxM + 1) each
  int* status;                                // status word per chain for the host
  const double* pctab;                        // preconditioner (ring_pc), or nullptr
  const int64_t* pcoff;
  int ndim;
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
// workgroup reductions in a fixed order (deterministic); red holds kLbWaves doubles
__device__ double lb_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kLbWaves; ++w) t += red[w];
  return t;
}
__device__ double lb_max(double v, double* red) {
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kLbWaves; ++w) t = fmax(t, red[w]);
  return t;
}

__device__ double cubic_min(double a, double fa, double da, double b, double fb, double db) {
  // minimiser of the cubic interpolating (a, fa, da), (b, fb, db), safeguarded
  const double d1 = da + db - 3.0 * (fa - fb) / (a - b);
  const double disc = d1 * d1 - da * db;
  double t;
  if (disc >= 0.0) {
    const double d2 = copysign(sqrt(disc), b - a);
    t = b - (b - a) * (db + d2 - d1) / (db - da + 2.0 * d2);
  } else {
    t = 0.5 * (a + b);
  }
  const double lo = fmin(a, b), hi = fmax(a, b), w = hi - lo;
  if (!(t > lo + 0.1 * w && t < hi - 0.1 * w)) t = 0.5 * (a + b);
  return t;
}

// first function and gradient: X = x0, G = g(x0), convergence of the start point
__global__ void __launch_bounds__(kLbThreads) lb_init_kernel(LbArgs A) {
  __shared__ double red[kLbWaves];
  const int c = blockIdx.x;
  const int64_t base = A.off[c];
  const int L = A.n[c] * A.D;
  double gm = 0.0;
  for (int j = threadIdx.x; j < L; j += kLbThreads) {
    const double g = A.GT[base + j];
    A.X[base + j] = A.XT[base + j];
    A.G[base + j] = g;
    gm = fmax(gm, fabs(g));
  }
  const double ginf = lb_max(gm, red);
  if (threadIdx.x == 0) {
    LbChain s = {};
    s.f = A.E[c];
    s.ginf = ginf;
    s.gam = 1.0;
    s.status = ginf <= A.pgtol ? ST_CONV : ST_RUN;
    s.op = s.status == ST_RUN ? OP_DIRECTION : OP_NONE;
    A.st[c] = s;
    A.status[c] = s.status;
  }
}

// f(q) for q = Q0 .. P-1 (compile-time), in chunks of 8 (4 above P = 16) with a scheduling
// barrier after each:
// the loads of one chunk are in flight together and the registers they take stay bounded
template <int Q0, int P, typename F>
__device__ __forceinline__ void each(F&& f) {
  constexpr int C = P > 16 ? 4 : 8;
  if constexpr (Q0 < P) {
#pragma unroll
    for (int q = Q0; q < Q0 + C && q < P; ++q) f(q);
    __builtin_amdgcn_sched_barrier(0);
    each<Q0 + C, P>(f);
  }
}

// d = -H g by the two-loop recursion, the descent check, and the first trial point.  Each
// pass over the chain fuses one correction's update with the next correction's dot product,
// so the recursion reads the direction 2k + 1 times for k pairs.
// P > 0: the direction stays in registers (at most P elements per thread, element
// j = threadIdx.x + q kLbThreads in slot q) through the 2k passes of the recursion, which then
// read only the pairs, with every load of a pass in flight at once; the arithmetic and its
// order are those of the streaming form (P = 0), which keeps D in HBM for longer chains.
template <int P>
__global__ void __launch_bounds__(kLbThreads) lb_direction_kernel(LbArgs A) {
  __shared__ double red[kLbWaves];
  __shared__ double alpha[kMaxM + 1];
  const int c = blockIdx.x;
  if (A.st[c].op != OP_DIRECTION) return;  // uniform over the workgroup
  const int k = A.st[c].count, head = A.st[c].head;
  const double gam = A.st[c].gam, ginf = A.st[c].ginf;
  const double* rho = A.rho + (int64_t)c * (kMaxM + 1);
  const int64_t base = A.off[c], tot = A.tot;
  const int L = A.n[c] * A.D;
  const int m1 = A.m + 1;
  double* D = A.Dv + base;
  const double* G = A.G + base;
  const double* X = A.X + base;
  double* XT = A.XT + base;
  const double* S0 = A.S + base;
  const double* Y0 = A.Y + base;
  auto slot = [=](int i) { return (head + i) % m1; };
  auto Sv = [=](int i) { return S0 + (int64_t)slot(i) * tot; };
  auto Yv = [=](int i) { return Y0 + (int64_t)slot(i) * tot; };
  double part = 0.0, a0 = 1.0;
  if (k == 0) {
    a0 = fmin(1.0, 1.0 / fmax(ginf, 1e-300));
    for (int j = threadIdx.x; j < L; j += kLbThreads) {
      const double v = -G[j];
      D[j] = v;
      part += G[j] * v;
      XT[j] = X[j] + a0 * v;
    }
  } else if constexpr (P > 0) {
    // loads at a clamped index and selects, no branch per element: the loads of a chunk
    // issue back to back.  Slots past the chain end keep d = 0 and add nothing.  The thread
    // index is re-read from an opaque copy in every pass, so the element offsets are
    // recomputed where they are used instead of being held (and spilled) across passes.
    int t = threadIdx.x;
    auto fresh = [&]() __attribute__((always_inline)) {
      t = threadIdx.x;
      asm volatile("" : "+v"(t));
    };
    auto at = [&](int q) __attribute__((always_inline)) { return min(t + q * kLbThreads, L - 1); };
    auto in = [&](int q) __attribute__((always_inline)) { return t + q * kLbThreads < L; };
    double d[P];
    {
      const double* Sk = Sv(k - 1);
      fresh();
      each<0, P>([&](int q) __attribute__((always_inline)) {
        const int j = at(q);
        const double v = Sk[j] * G[j] + part;
        d[q] = 0.0;
        part = in(q) ? v : part;
      });
    }
    for (int i = k - 1; i >= 0; --i) {
      const double al = rho[slot(i)] * lb_sum(part, red);
      if (threadIdx.x == 0) alpha[i] = al;
      const double* Yp = Yv(i);
      const bool first = i == k - 1;  // the first backward update reads the gradient
      part = 0.0;
      if (i > 0) {
        const double* Sn = Sv(i - 1);
        if (first) {
          fresh();
          each<0, P>([&](int q) __attribute__((always_inline)) {
            const int j = at(q);
            const double v = G[j] - al * Yp[j];
            const double w = Sn[j] * v + part;
            d[q] = in(q) ? v : d[q];
            part = in(q) ? w : part;
          });
        } else {
          fresh();
          each<0, P>([&](int q) __attribute__((always_inline)) {
            const int j = at(q);
            const double v = d[q] - al * Yp[j];
            const double w = Sn[j] * v + part;
            d[q] = in(q) ? v : d[q];
            part = in(q) ? w : part;
          });
        }
      } else {
        // last backward update, the initial scaling, and the first forward dot product
        fresh();
        each<0, P>([&](int q) __attribute__((always_inline)) {
          const int j = at(q);
          const double y = Yp[j];
          const double src = first ? G[j] : d[q];
          const double v = (src - al * y) * gam;
          const double w = y * v + part;
          d[q] = in(q) ? v : d[q];
          part = in(q) ? w : part;
        });
      }
    }
    __syncthreads();  // alpha[] complete before the forward loop reads it
    for (int i = 0; i < k; ++i) {
      const double coef = alpha[i] - rho[slot(i)] * lb_sum(part, red);
      const double* Sp = Sv(i);
      part = 0.0;
      if (i < k - 1) {
        const double* Yn = Yv(i + 1);
        fresh();
        each<0, P>([&](int q) __attribute__((always_inline)) {
          const int j = at(q);
          const double v = d[q] + coef * Sp[j];
          const double w = Yn[j] * v + part;
          d[q] = in(q) ? v : d[q];
          part = in(q) ? w : part;
        });
      } else {
        // last forward update, the sign, the slope along d and the unit trial step
        fresh();
        each<0, P>([&](int q) __attribute__((always_inline)) {
          const int j = at(q);
          const double v = -(d[q] + coef * Sp[j]);
          const double w = G[j] * v + part;
          const double xt = X[j] + v;
          part = in(q) ? w : part;
          if (in(q)) {
            D[j] = v;
            XT[j] = xt;
          }
        });
      }
    }
  } else {
    {
      const double* Sk = Sv(k - 1);
      for (int j = threadIdx.x; j < L; j += kLbThreads) part += Sk[j] * G[j];
    }
    for (int i = k - 1; i >= 0; --i) {
      const double al = rho[slot(i)] * lb_sum(part, red);
      if (threadIdx.x == 0) alpha[i] = al;
      const double* Yp = Yv(i);
      const double* src = i == k - 1 ? G : D;
      part = 0.0;
      if (i > 0) {
        const double* Sn = Sv(i - 1);
        for (int j = threadIdx.x; j < L; j += kLbThreads) {
          const double d = src[j] - al * Yp[j];
          D[j] = d;
          part += Sn[j] * d;
        }
      } else {
        // last backward update, the initial scaling, and the first forward dot product
        for (int j = threadIdx.x; j < L; j += kLbThreads) {
          const double d = (src[j] - al * Yp[j]) * gam;
          D[j] = d;
          part += Yp[j] * d;
        }
      }
    }
    __syncthreads();  // alpha[] complete before the forward loop reads it
    for (int i = 0; i < k; ++i) {
      const double coef = alpha[i] - rho[slot(i)] * lb_sum(part, red);
      const double* Sp = Sv(i);
      part = 0.0;
      if (i < k - 1) {
        const double* Yn = Yv(i + 1);
        for (int j = threadIdx.x; j < L; j += kLbThreads) {
          const double d = D[j] + coef * Sp[j];
          D[j] = d;
          part += Yn[j] * d;
        }
      } else {
        // last forward update, the sign, the slope along d and the unit trial step
        for (int j = threadIdx.x; j < L; j += kLbThreads) {
          const double d = -(D[j] + coef * Sp[j]);
          D[j] = d;
          part += G[j] * d;
          XT[j] = X[j] + d;
        }
      }
    }
  }
  double dg0 = lb_sum(part, red);
  const bool reset = !(dg0 < 0.0);  // not a descent direction: steepest descent, no history
  if (reset && k > 0) {
    a0 = fmin(1.0, 1.0 / fmax(ginf, 1e-300));
    part = 0.0;
    for (int j = threadIdx.x; j < L; j += kLbThreads) {
      const double v = -G[j];
      D[j] = v;
      part += G[j] * v;
      XT[j] = X[j] + a0 * v;
    }
    dg0 = lb_sum(part, red);
  }
  if (threadIdx.x == 0) {
    LbChain t = A.st[c];
    if (reset) {
      t.count = 0;
      t.head = 0;
    }
    t.dg0 = dg0;
    t.f0 = t.f;
    t.phase = LS_BRACKET;
    t.a_prev = 0.0;
    t.f_prev = t.f;
    t.dg_prev = dg0;
    t.prev_is_g = 1;
    t.nls = 0;
    t.a_cur = a0;
    t.op = OP_NONE;
    A.st[c] = t;
  }
}

// v <- P^-1 v for one chain's vector in HBM (dof columns of n beads, the ring layout), P the
// spring tridiagonal of rp_ring_set_preconditioner.  A wave solves one dof column: each lane
// takes a run of at most kPcRun consecutive beads, and both Thomas sweeps -- the forward
// d'_i = alpha_i d'_{i-1} + inv_i r_i and the backward x_i = alpha_i x_{i+1} + d'_i -- are
// affine recurrences, so a lane composes its run into one affine map, a wave scan of those
// maps (xor-free shuffles up / down) gives each lane its incoming value, and the lane
// replays its run.  The callers bracket it with barriers.
constexpr int kPcRun = 16;  // beads per lane: n <= 64 kPcRun = 1024
__device__ void pc_solve(double* v, int n, int D, int ndim, const double* __restrict__ tab) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int run = (n + 63) >> 6;
  const int i0 = lane * run;
  for (int col = wave; col < D; col += kLbWaves) {
    const double* t = tab + (size_t)(col / ndim) * n * 2;
    double* vc = v + (size_t)col * n;
    double w[kPcRun];
    // forward sweep
    double A = 1.0, B = 0.0;
#pragma unroll
    for (int q = 0; q < kPcRun; ++q) {
      const int i = i0 + q;
      w[q] = 0.0;
      if (q < run && i < n) {
        const double al = t[2 * i + 1];
        w[q] = vc[i];
        B = al * B + t[2 * i] * w[q];
        A = al * A;
      }
    }
    for (int o = 1; o < 64; o <<= 1) {  // inclusive scan: lanes below come first
      const double A2 = __shfl_up(A, o), B2 = __shfl_up(B, o);
      if (lane >= o) {
        B = A * B2 + B;
        A = A * A2;
      }
    }
    double x = __shfl_up(B, 1);
    if (lane == 0) x = 0.0;
#pragma unroll
    for (int q = 0; q < kPcRun; ++q) {
      const int i = i0 + q;
      if (q < run && i < n) {
        x = t[2 * i + 1] * x + t[2 * i] * w[q];
        w[q] = x;  // d'_i
      }
    }
    // backward sweep
    A = 1.0;
    B = 0.0;
#pragma unroll
    for (int q = kPcRun - 1; q >= 0; --q) {
      const int i = i0 + q;
      if (q < run && i < n) {
        const double al = t[2 * i + 1];
        B = al * B + w[q];
        A = al * A;
      }
    }
    for (int o = 1; o < 64; o <<= 1) {  // inclusive scan: lanes above come first
      const double A2 = __shfl_down(A, o), B2 = __shfl_down(B, o);
      if (lane + o < 64) {
        B = A * B2 + B;
        A = A * A2;
      }
    }
    x = __shfl_down(B, 1);
    if (lane == 63) x = 0.0;
#pragma unroll
    for (int q = kPcRun - 1; q >= 0; --q) {
      const int i = i0 + q;
      if (q < run && i < n) {
        x = t[2 * i + 1] * x + w[q];
        vc[i] = x;
      }
    }
  }
}

// The direction of the preconditioned search: d = -H g by the two-loop recursion with the
// initial matrix gamma P^-1 (gamma = s.y / y.P^-1.y of the newest pair; the first and the
// reset directions are -P^-1 g, tried at unit step).  Plain passes over HBM: with the
// preconditioner a search takes tens of iterations where the unpreconditioned one takes
// thousands, so this kernel is not the one tuned for bandwidth.  XT serves as the scratch
// vector of the y.P^-1.y solve before it receives the trial point.
__global__ void __launch_bounds__(kLbThreads) lb_direction_pc_kernel(LbArgs A) {
  __shared__ double red[kLbWaves];
  __shared__ double alpha[kMaxM + 1];
  const int c = blockIdx.x;
  if (A.st[c].op != OP_DIRECTION) return;  // uniform over the workgroup
  const int k = A.st[c].count, head = A.st[c].head;
  const double* rho = A.rho + (int64_t)c * (kMaxM + 1);
  const int64_t base = A.off[c], tot = A.tot;
  const int n = A.n[c], L = n * A.D;
  const int m1 = A.m + 1;
  const double* tab = A.pctab + A.pcoff[c];
  double* D = A.Dv + base;
  const double* G = A.G + base;
  const double* X = A.X + base;
  double* XT = A.XT + base;
  const double* S0 = A.S + base;
  const double* Y0 = A.Y + base;
  auto slot = [=](int i) { return (head + i) % m1; };
  auto Sv = [=](int i) { return S0 + (int64_t)slot(i) * tot; };
  auto Yv = [=](int i) { return Y0 + (int64_t)slot(i) * tot; };
  // d = -P^-1 g: the first direction and the fallback of a non-descent one
  auto steepest = [&]() -> double {
    for (int j = threadIdx.x; j < L; j += kLbThreads) D[j] = -G[j];
    __syncthreads();
    pc_solve(D, n, A.D, A.ndim, tab);
    __syncthreads();
    double part = 0.0;
    for (int j = threadIdx.x; j < L; j += kLbThreads) {
      part += G[j] * D[j];
      XT[j] = X[j] + D[j];
    }
    return lb_sum(part, red);
  };
  double dg0;
  bool reset = false;
  if (k == 0) {
    dg0 = steepest();
  } else {
    // backward loop: q = g - sum alpha_i y_i, q in D
    for (int j = threadIdx.x; j < L; j += kLbThreads) D[j] = G[j];
    for (int i = k - 1; i >= 0; --i) {
      const double* Si = Sv(i);
      const double* Yi = Yv(i);
      double part = 0.0;
      for (int j = threadIdx.x; j < L; j += kLbThreads) part += Si[j] * D[j];
      const double al = rho[slot(i)] * lb_sum(part, red);
      if (threadIdx.x == 0) alpha[i] = al;
      for (int j = threadIdx.x; j < L; j += kLbThreads) D[j] -= al * Yi[j];
    }
    // gamma = s.y / y.P^-1.y of the newest pair, with XT as the solve's scratch
    const double* Yk = Yv(k - 1);
    for (int j = threadIdx.x; j < L; j += kLbThreads) XT[j] = Yk[j];
    __syncthreads();
    pc_solve(XT, n, A.D, A.ndim, tab);
    pc_solve(D, n, A.D, A.ndim, tab);
    __syncthreads();
    double part = 0.0;
    for (int j = threadIdx.x; j < L; j += kLbThreads) part += Yk[j] * XT[j];
    const double ypy = lb_sum(part, red);
    const double gam = (1.0 / rho[slot(k - 1)]) / ypy;
    for (int j = threadIdx.x; j < L; j += kLbThreads) D[j] *= gam;
    // forward loop: r += s_i (alpha_i - rho_i y_i.r)
    for (int i = 0; i < k; ++i) {
      const double* Si = Sv(i);
      const double* Yi = Yv(i);
      part = 0.0;
      for (int j = threadIdx.x; j < L; j += kLbThreads) part += Yi[j] * D[j];
      const double coef = alpha[i] - rho[slot(i)] * lb_sum(part, red);
      for (int j = threadIdx.x; j < L; j += kLbThreads) D[j] += coef * Si[j];
    }
    part = 0.0;
    for (int j = threadIdx.x; j < L; j += kLbThreads) {
      const double d = -D[j];
      D[j] = d;
      part += G[j] * d;
      XT[j] = X[j] + d;
    }
    dg0 = lb_sum(part, red);
    reset = !(dg0 < 0.0);  // not a descent direction: the preconditioned steepest descent
    if (reset) {
      __syncthreads();
      dg0 = steepest();
    }
  }
  if (threadIdx.x == 0) {
    LbChain t = A.st[c];
    if (reset) {
      t.count = 0;
      t.head = 0;
    }
    t.dg0 = dg0;
    t.f0 = t.f;
    t.phase = LS_BRACKET;
    t.a_prev = 0.0;
    t.f_prev = t.f;
    t.dg_prev = dg0;
    t.prev_is_g = 1;
    t.nls = 0;
    t.a_cur = 1.0;
    t.op = OP_NONE;
    A.st[c] = t;
  }
}

// one line-search evaluation: the value and slope at the trial point decide the next trial
// point or the acceptance, which stores the correction pair and tests convergence
__global__ void __launch_bounds__(kLbThreads) lb_feed_kernel(LbArgs A) {
  __shared__ double red[kLbWaves];
  const int c = blockIdx.x;
  LbChain s = A.st[c];
  if (s.status != ST_RUN) return;  // uniform over the workgroup
  const int64_t base = A.off[c];
  const int L = A.n[c] * A.D;
  double part = 0.0;
  {
    const double* __restrict__ gt = A.GT + base;
    const double* __restrict__ dv = A.Dv + base;
#pragma unroll 8
    for (int j = threadIdx.x; j < L; j += kLbThreads) part += gt[j] * dv[j];
  }
  const double dgt = lb_sum(part, red);
  const double ft = A.E[c];
  // every lane runs the same scalar decision
  int act = ACT_NONE, cp = 0;
  double a_new = 0.0;
  const double a = s.a_cur;
  ++s.nls;
  if (s.nls >= kMaxLS) {
    // out of evaluations: keep the best point if it decreased f
    if (ft < s.f0 - 1e-300) act = ACT_ACCEPT_CUR;
    else if (s.phase == LS_ZOOM && s.f_lo < s.f0) act = ACT_ACCEPT_LO;
    else s.status = ST_ABNORMAL;
  } else if (s.phase == LS_BRACKET) {
    if (ft > s.f0 + kC1 * a * s.dg0 || (s.nls > 1 && ft >= s.f_prev)) {
      s.phase = LS_ZOOM;
      s.a_lo = s.a_prev; s.f_lo = s.f_prev; s.dg_lo = s.dg_prev;
      cp |= s.prev_is_g ? CP_G_TO_LO : CP_PREV_TO_LO;
      s.a_hi = a; s.f_hi = ft; s.dg_hi = dgt;
      act = ACT_TRIAL;
      a_new = cubic_min(s.a_lo, s.f_lo, s.dg_lo, s.a_hi, s.f_hi, s.dg_hi);
    } else if (fabs(dgt) <= -kC2 * s.dg0) {
      act = ACT_ACCEPT_CUR;
    } else if (dgt >= 0.0) {
      s.phase = LS_ZOOM;
      s.a_lo = a; s.f_lo = ft; s.dg_lo = dgt; cp |= CP_GT_TO_LO;
      s.a_hi = s.a_prev; s.f_hi = s.f_prev; s.dg_hi = s.dg_prev;
      act = ACT_TRIAL;
      a_new = cubic_min(s.a_lo, s.f_lo, s.dg_lo, s.a_hi, s.f_hi, s.dg_hi);
    } else {
      s.a_prev = a; s.f_prev = ft; s.dg_prev = dgt; s.prev_is_g = 0; cp |= CP_GT_TO_PREV;
      act = ACT_TRIAL;
      a_new = 2.0 * a;  // extrapolate
    }
  } else {
    bool done = false;
    if (ft > s.f0 + kC1 * a * s.dg0 || ft >= s.f_lo) {
      s.a_hi = a; s.f_hi = ft; s.dg_hi = dgt;
    } else if (fabs(dgt) <= -kC2 * s.dg0) {
      act = ACT_ACCEPT_CUR;
      done = true;
    } else {
      if (dgt * (s.a_hi - s.a_lo) >= 0.0) {
        s.a_hi = s.a_lo; s.f_hi = s.f_lo; s.dg_hi = s.dg_lo;
      }
      s.a_lo = a; s.f_lo = ft; s.dg_lo = dgt; cp |= CP_GT_TO_LO;
    }
    if (!done) {
      if (fabs(s.a_hi - s.a_lo) <= 1e-14 * fmax(1.0, s.a_lo)) {
        // interval collapsed: take the best point found
        if (s.f_lo < s.f0) act = ACT_ACCEPT_LO;
        else s.status = ST_ABNORMAL;
      } else {
        act = ACT_TRIAL;
        a_new = cubic_min(s.a_lo, s.f_lo, s.dg_lo, s.a_hi, s.f_hi, s.dg_hi);
      }
    }
  }
  // distinct vectors: the loads of several iterations may run ahead of earlier stores
  double* __restrict__ X = A.X + base;
  double* __restrict__ G = A.G + base;
  double* __restrict__ Glo = A.Glo + base;
  double* __restrict__ Gp = A.Gprev + base;
  double* __restrict__ XT = A.XT + base;
  const double* __restrict__ GT = A.GT + base;
  const double* __restrict__ Dv = A.Dv + base;
#pragma unroll 4
  for (int j = threadIdx.x; j < L; j += kLbThreads) {
    if (cp & CP_PREV_TO_LO) Glo[j] = Gp[j];
    if (cp & CP_G_TO_LO) Glo[j] = G[j];
    if (cp & CP_GT_TO_LO) Glo[j] = GT[j];
    if (cp & CP_GT_TO_PREV) Gp[j] = GT[j];
  }
  if (act == ACT_TRIAL) {
#pragma unroll 8
    for (int j = threadIdx.x; j < L; j += kLbThreads) XT[j] = X[j] + a_new * Dv[j];
    s.a_cur = a_new;
  } else if (act == ACT_ACCEPT_CUR || act == ACT_ACCEPT_LO) {
    const int m1 = A.m + 1;
    const int slot = (s.head + s.count) % m1;
    double* __restrict__ Sp = A.S + (int64_t)slot * A.tot + base;
    double* __restrict__ Yp = A.Y + (int64_t)slot * A.tot + base;
    const bool cur = act == ACT_ACCEPT_CUR;
    const double alo = s.a_lo;
    double psy = 0.0, pyy = 0.0, pg = 0.0;
#pragma unroll 4
    for (int j = threadIdx.x; j < L; j += kLbThreads) {
      const double xa = cur ? XT[j] : X[j] + alo * Dv[j];
      const double ga = cur ? GT[j] : Glo[j];
      const double sj = xa - X[j], yj = ga - G[j];
      Sp[j] = sj;
      Yp[j] = yj;
      psy += sj * yj;
      pyy += yj * yj;
      pg = fmax(pg, fabs(ga));
      X[j] = xa;
      G[j] = ga;
      XT[j] = xa;
    }
    const double sy = lb_sum(psy, red);
    const double yy = lb_sum(pyy, red);
    const double ginf = lb_max(pg, red);
    const double fa = cur ? ft : s.f_lo;
    if (sy > DBL_EPSILON * yy) {
      if (threadIdx.x == 0) A.rho[(int64_t)c * (kMaxM + 1) + slot] = 1.0 / sy;
      s.gam = sy / yy;
      if (s.count < A.m) ++s.count;
      else s.head = (s.head + 1) % m1;
    }
    const double fold = s.f;
    s.f = fa;
    s.ginf = ginf;
    ++s.iter;
    if ((fold - fa) / fmax(fmax(fabs(fold), fabs(fa)), 1.0) <= A.factr * DBL_EPSILON ||
        ginf <= A.pgtol)
      s.status = ST_CONV;
    else if (A.maxiter > 0 && s.iter >= A.maxiter)
      s.status = ST_MAXITER;
    else
      s.op = OP_DIRECTION;
  }
  if (threadIdx.x == 0) {
    A.st[c] = s;
    A.status[c] = s.status;
  }
}

__global__ void lb_cap_kernel(LbArgs A) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < A.nchain && A.st[c].status == ST_RUN) {
    A.st[c].status = ST_MAXITER;
    A.status[c] = ST_MAXITER;
  }
}

}  // namespace

extern "C" int rp_instanton(rp_ring* r, double pgtol, double factr, int m, int maxiter,
                            double* x_inout, double* action, int* iters, int* status) {
  RP_REQUIRE(r && x_inout && action && m > 0 && m <= kMaxM,
             "rp_instanton: bad arguments (m must be 1..%d)", kMaxM);
  const int nc = ring_nchain(r);
  const int64_t tot = ring_tot(r);
  const RingBufs rb = ring_bufs(r);
  DevBuf<double> X, G, Dv, Glo, Gp, S, Y;
  DevBuf<LbChain> st;
  DevBuf<double> rho;
  DevBuf<int> stat;
  int rc;
  if ((rc = X.alloc(tot)) || (rc = G.alloc(tot)) || (rc = Dv.alloc(tot)) || (rc = Glo.alloc(tot)) ||
      (rc = Gp.alloc(tot)) || (rc = S.alloc((size_t)(m + 1) * tot)) ||
      (rc = Y.alloc((size_t)(m + 1) * tot)) || (rc = st.alloc(nc)) || (rc = stat.alloc(nc)) ||
      (rc = rho.alloc((size_t)nc * (kMaxM + 1))))
    return rc;
  LbArgs A;
  A.nchain = nc;
  A.m = m;
  A.maxiter = maxiter;
  A.D = rb.D;
  A.factr = factr;
  A.pgtol = pgtol;
  A.tot = tot;
  A.off = rb.off;
  A.n = rb.n;
  A.X = X.p; A.G = G.p; A.Dv = Dv.p; A.Glo = Glo.p; A.Gprev = Gp.p; A.S = S.p; A.Y = Y.p;
  A.XT = rb.x;
  A.GT = rb.g;
  A.E = rb.e;
  A.st = st.p;
  A.rho = rho.p;
  A.status = stat.p;
  const RingPc pc = ring_pc(r);
  A.pctab = pc.tab;
  A.pcoff = pc.off;
  A.ndim = pc.ndim;
  // initial function and gradient (instantonmod.f90:725-739)
  if ((rc = ring_upload_x(r, x_inout)) || (rc = ring_eval(r, true))) return rc;
  lb_init_kernel<<<nc, kLbThreads, 0, rb.s>>>(A);
  RP_HIP(hipGetLastError());
  // direction kernel by the longest chain's elements per thread
  const int nmax = *std::max_element(ring_n(r).begin(), ring_n(r).end());
  const int per = (nmax * rb.D + kLbThreads - 1) / kLbThreads;
  auto direction = A.pctab ? lb_direction_pc_kernel
                 : per <= 8 ? lb_direction_kernel<8> : per <= 16 ? lb_direction_kernel<16>
                 : per <= 28 ? lb_direction_kernel<28> : lb_direction_kernel<0>;
  RP_REQUIRE(!A.pctab || nmax <= 64 * kPcRun,
             "rp_instanton: the preconditioner takes chains of at most %d beads", 64 * kPcRun);
  direction<<<nc, kLbThreads, 0, rb.s>>>(A);
  RP_HIP(hipGetLastError());
  std::vector<int> hs(nc);
  const long cap = maxiter > 0 ? (long)maxiter * kMaxLS + 1 : kMaxRounds;
  bool capped = true;
  for (long round = 0; round < cap; ++round) {
    if (round % kCheckEvery == 0) {
      if ((rc = stat.download(hs.data(), nc, rb.s))) return rc;
      RP_HIP(hipStreamSynchronize(rb.s));
      if (std::none_of(hs.begin(), hs.end(), [](int v) { return v == ST_RUN; })) {
        capped = false;
        break;
      }
    }
    if ((rc = ring_eval(r, true))) return rc;
    lb_feed_kernel<<<nc, kLbThreads, 0, rb.s>>>(A);
    RP_HIP(hipGetLastError());
    direction<<<nc, kLbThreads, 0, rb.s>>>(A);
    RP_HIP(hipGetLastError());
  }
  if (capped) {
    lb_cap_kernel<<<(nc + 255) / 256, 256, 0, rb.s>>>(A);
    RP_HIP(hipGetLastError());
  }
  // results; the optimum stays on the device for the Hessian / log-determinant calls
  std::vector<LbChain> hst(nc);
  RP_HIP(hipMemcpyAsync(rb.x, X.p, tot * sizeof(double), hipMemcpyDeviceToDevice, rb.s));
  if ((rc = X.download(x_inout, tot, rb.s)) || (rc = st.download(hst.data(), nc, rb.s))) return rc;
  RP_HIP(hipStreamSynchronize(rb.s));
  for (int c = 0; c < nc; ++c) {
    action[c] = hst[c].f;
    if (iters) iters[c] = hst[c].iter;
    if (status) status[c] = hst[c].status;
  }
  return RP_OK;
}

// rp_instanton.hip -- the instanton search (instantonmod.f90:679-777) for a batch of
// chains: one quasi-Newton state machine per chain on the host, every function and
// gradient request of all chains served by one device evaluation of U_M and its adjoint
// gradient (ring_um_kernel).
//
// The reference drives L-BFGS-B (Byrd, Lu, Nocedal, Zhu 1995; v3.0 with Morales-Nocedal
// 2011) with nbd = 0, i.e. no bounds, m = 8 corrections, factr = 1e6 and pgtol = eps2.
// Without bounds L-BFGS-B's generalized Cauchy point is the unconstrained step and the
// method is limited-memory BFGS; this file implements that method from the published
// description (Nocedal & Wright, Numerical Optimization, 2nd ed.: two-loop recursion,
// Alg. 7.4; strong-Wolfe line search with zoom, Alg. 3.5/3.6) with the reference's two
// stopping tests:
//   (f_k - f_{k+1}) / max(|f_k|, |f_{k+1}|, 1) <= factr * epsmch,   |g|_inf <= pgtol.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "rp_ring.h"

using namespace rp;

namespace {

enum Phase { LS_BRACKET, LS_ZOOM };
enum Status { ST_RUN = -1, ST_CONV = 0, ST_ABNORMAL = 1, ST_MAXITER = 2 };

struct Chain {
  int n = 0;                       // number of coordinates of this chain
  std::vector<double> x, g, d, xt;  // iterate, gradient, direction, trial point
  std::vector<std::vector<double>> S, Y;
  std::vector<double> rho;
  double f = 0.0;
  int iter = 0, status = ST_RUN, nls = 0;
  // line search
  Phase ph = LS_BRACKET;
  double a_cur = 0, a_prev = 0, f_prev = 0, dg_prev = 0;
  double a_lo = 0, f_lo = 0, dg_lo = 0, a_hi = 0, f_hi = 0, dg_hi = 0;
  double f0 = 0, dg0 = 0;
  std::vector<double> g_lo, g_prev;  // gradients at a_lo and a_prev
};

constexpr double kC1 = 1e-3;   // sufficient decrease (L-BFGS-B's ftol)
constexpr double kC2 = 0.9;    // curvature (L-BFGS-B's gtol)
constexpr int kMaxLS = 40;     // line-search evaluations (the reference's maxiter argument)

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

double infnorm(const std::vector<double>& a) {
  double s = 0.0;
  for (double v : a) s = std::max(s, std::fabs(v));
  return s;
}

// d = -H g by the two-loop recursion
void direction(Chain& c) {
  const int k = (int)c.S.size();
  c.d = c.g;
  std::vector<double> alpha(k);
  for (int i = k - 1; i >= 0; --i) {
    alpha[i] = c.rho[i] * dot(c.S[i], c.d);
    for (int j = 0; j < c.n; ++j) c.d[j] -= alpha[i] * c.Y[i][j];
  }
  double gamma = 1.0;
  if (k > 0) gamma = dot(c.S[k - 1], c.Y[k - 1]) / dot(c.Y[k - 1], c.Y[k - 1]);
  for (double& v : c.d) v *= gamma;
  for (int i = 0; i < k; ++i) {
    const double beta = c.rho[i] * dot(c.Y[i], c.d);
    for (int j = 0; j < c.n; ++j) c.d[j] += c.S[i][j] * (alpha[i] - beta);
  }
  for (double& v : c.d) v = -v;
}

double cubic_min(double a, double fa, double da, double b, double fb, double db) {
  // minimiser of the cubic interpolating (a, fa, da), (b, fb, db), safeguarded
  const double d1 = da + db - 3.0 * (fa - fb) / (a - b);
  const double disc = d1 * d1 - da * db;
  double t;
  if (disc >= 0.0) {
    const double d2 = std::copysign(std::sqrt(disc), b - a);
    t = b - (b - a) * (db + d2 - d1) / (db - da + 2.0 * d2);
  } else {
    t = 0.5 * (a + b);
  }
  const double lo = std::min(a, b), hi = std::max(a, b), w = hi - lo;
  if (!(t > lo + 0.1 * w && t < hi - 0.1 * w)) t = 0.5 * (a + b);
  return t;
}

void set_trial(Chain& c, double a) {
  c.a_cur = a;
  for (int j = 0; j < c.n; ++j) c.xt[j] = c.x[j] + a * c.d[j];
}

// begin an iteration: direction + first trial step
void start_iteration(Chain& c) {
  direction(c);
  c.dg0 = dot(c.g, c.d);
  if (!(c.dg0 < 0.0)) {  // not a descent direction: restart from steepest descent
    c.S.clear();
    c.Y.clear();
    c.rho.clear();
    c.d = c.g;
    for (double& v : c.d) v = -v;
    c.dg0 = dot(c.g, c.d);
  }
  c.f0 = c.f;
  c.ph = LS_BRACKET;
  c.a_prev = 0.0;
  c.f_prev = c.f;
  c.dg_prev = c.dg0;
  c.g_prev = c.g;
  c.nls = 0;
  const double a0 = c.S.empty() ? std::min(1.0, 1.0 / std::max(infnorm(c.g), 1e-300)) : 1.0;
  set_trial(c, a0);
}

// accept the trial point (f, g) and test convergence; returns false when finished
bool accept(Chain& c, double ft, const std::vector<double>& gt, double factr, double pgtol, int m,
            int maxiter) {
  std::vector<double> s(c.n), y(c.n);
  for (int j = 0; j < c.n; ++j) {
    s[j] = c.xt[j] - c.x[j];
    y[j] = gt[j] - c.g[j];
  }
  const double sy = dot(s, y);
  if (sy > DBL_EPSILON * dot(y, y)) {
    if ((int)c.S.size() == m) {
      c.S.erase(c.S.begin());
      c.Y.erase(c.Y.begin());
      c.rho.erase(c.rho.begin());
    }
    c.S.push_back(s);
    c.Y.push_back(y);
    c.rho.push_back(1.0 / sy);
  }
  const double fold = c.f;
  c.x = c.xt;
  c.g = gt;
  c.f = ft;
  ++c.iter;
  if ((fold - ft) / std::max(std::max(std::fabs(fold), std::fabs(ft)), 1.0) <= factr * DBL_EPSILON ||
      infnorm(c.g) <= pgtol) {
    c.status = ST_CONV;
    return false;
  }
  if (maxiter > 0 && c.iter >= maxiter) {
    c.status = ST_MAXITER;
    return false;
  }
  start_iteration(c);
  return true;
}

// feed the value at the current trial point; either request another point or accept
void feed(Chain& c, double ft, const std::vector<double>& gt, double factr, double pgtol, int m,
          int maxiter) {
  const double dgt = dot(gt, c.d);
  ++c.nls;
  const double a = c.a_cur;
  auto zoom_step = [&]() {
    set_trial(c, cubic_min(c.a_lo, c.f_lo, c.dg_lo, c.a_hi, c.f_hi, c.dg_hi));
  };
  if (c.nls >= kMaxLS) {
    // the line search is out of evaluations: keep the best point if it decreased f
    if (ft < c.f0 - 1e-300) {
      accept(c, ft, gt, factr, pgtol, m, maxiter);
    } else if (c.ph == LS_ZOOM && c.f_lo < c.f0) {
      set_trial(c, c.a_lo);
      accept(c, c.f_lo, c.g_lo, factr, pgtol, m, maxiter);
    } else {
      c.status = ST_ABNORMAL;
    }
    return;
  }
  if (c.ph == LS_BRACKET) {
    if (ft > c.f0 + kC1 * a * c.dg0 || (c.nls > 1 && ft >= c.f_prev)) {
      c.ph = LS_ZOOM;
      c.a_lo = c.a_prev; c.f_lo = c.f_prev; c.dg_lo = c.dg_prev; c.g_lo = c.g_prev;
      c.a_hi = a; c.f_hi = ft; c.dg_hi = dgt;
      zoom_step();
      return;
    }
    if (std::fabs(dgt) <= -kC2 * c.dg0) {
      accept(c, ft, gt, factr, pgtol, m, maxiter);
      return;
    }
    if (dgt >= 0.0) {
      c.ph = LS_ZOOM;
      c.a_lo = a; c.f_lo = ft; c.dg_lo = dgt; c.g_lo = gt;
      c.a_hi = c.a_prev; c.f_hi = c.f_prev; c.dg_hi = c.dg_prev;
      zoom_step();
      return;
    }
    c.a_prev = a; c.f_prev = ft; c.dg_prev = dgt; c.g_prev = gt;
    set_trial(c, 2.0 * a);  // extrapolate
    return;
  }
  // zoom
  if (ft > c.f0 + kC1 * a * c.dg0 || ft >= c.f_lo) {
    c.a_hi = a; c.f_hi = ft; c.dg_hi = dgt;
  } else {
    if (std::fabs(dgt) <= -kC2 * c.dg0) {
      accept(c, ft, gt, factr, pgtol, m, maxiter);
      return;
    }
    if (dgt * (c.a_hi - c.a_lo) >= 0.0) {
      c.a_hi = c.a_lo; c.f_hi = c.f_lo; c.dg_hi = c.dg_lo;
    }
    c.a_lo = a; c.f_lo = ft; c.dg_lo = dgt; c.g_lo = gt;
  }
  if (std::fabs(c.a_hi - c.a_lo) <= 1e-14 * std::max(1.0, c.a_lo)) {
    // interval collapsed: take the best point found
    if (c.f_lo < c.f0) {
      set_trial(c, c.a_lo);
      accept(c, c.f_lo, c.g_lo, factr, pgtol, m, maxiter);
    } else {
      c.status = ST_ABNORMAL;
    }
    return;
  }
  zoom_step();
}

void parallel_for(int n, const std::function<void(int)>& fn) {
  const int nt = std::max(1, std::min(n, (int)std::min(16u, std::thread::hardware_concurrency())));
  if (nt <= 1) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      for (int i = t; i < n; i += nt) fn(i);
    });
  for (auto& t : th) t.join();
}

}  // namespace

extern "C" int rp_instanton(rp_ring* r, double pgtol, double factr, int m, int maxiter,
                            double* x_inout, double* action, int* iters, int* status) {
  RP_REQUIRE(r && x_inout && action && m > 0, "rp_instanton: bad arguments");
  const int nc = ring_nchain(r), tot = ring_tot(r);
  const auto& off = ring_off(r);
  const auto& nn = ring_n(r);
  const int D = ring_D(r);
  std::vector<Chain> ch(nc);
  for (int c = 0; c < nc; ++c) {
    ch[c].n = nn[c] * D;
    ch[c].x.assign(x_inout + off[c], x_inout + off[c] + ch[c].n);
    ch[c].xt = ch[c].x;
  }
  std::vector<double> xbuf(x_inout, x_inout + tot), gbuf(tot), ebuf(nc);
  int rc;
  // initial function and gradient (instantonmod.f90:725-739)
  if ((rc = ring_upload_x(r, xbuf.data())) || (rc = ring_eval(r, true)) ||
      (rc = ring_download(r, ebuf.data(), gbuf.data())))
    return rc;
  int active = 0;
  for (int c = 0; c < nc; ++c) {
    ch[c].f = ebuf[c];
    ch[c].g.assign(gbuf.begin() + off[c], gbuf.begin() + off[c] + ch[c].n);
    if (infnorm(ch[c].g) <= pgtol) {
      ch[c].status = ST_CONV;
    } else {
      start_iteration(ch[c]);
      ++active;
    }
  }
  while (active > 0) {
    for (int c = 0; c < nc; ++c)
      std::copy(ch[c].status == ST_RUN ? ch[c].xt.begin() : ch[c].x.begin(),
                ch[c].status == ST_RUN ? ch[c].xt.end() : ch[c].x.end(), xbuf.begin() + off[c]);
    if ((rc = ring_upload_x(r, xbuf.data())) || (rc = ring_eval(r, true)) ||
        (rc = ring_download(r, ebuf.data(), gbuf.data())))
      return rc;
    parallel_for(nc, [&](int c) {
      if (ch[c].status != ST_RUN) return;
      std::vector<double> gt(gbuf.begin() + off[c], gbuf.begin() + off[c] + ch[c].n);
      feed(ch[c], ebuf[c], gt, factr, pgtol, m, maxiter);
    });
    active = 0;
    for (int c = 0; c < nc; ++c) active += ch[c].status == ST_RUN;
  }
  for (int c = 0; c < nc; ++c) {
    std::copy(ch[c].x.begin(), ch[c].x.end(), x_inout + off[c]);
    std::copy(ch[c].x.begin(), ch[c].x.end(), xbuf.begin() + off[c]);
    action[c] = ch[c].f;
    if (iters) iters[c] = ch[c].iter;
    if (status) status[c] = ch[c].status;
  }
  // leave the optimum on the device for the Hessian / log-determinant calls
  if ((rc = ring_upload_x(r, xbuf.data()))) return rc;
  return ring_download(r, nullptr, nullptr);
}

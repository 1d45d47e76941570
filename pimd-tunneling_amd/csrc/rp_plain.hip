// rp_plain.hip -- the plain per-trajectory GPU path (the drop-in).
//
// One trajectory's verletint state lives on the device; every reference routine
// (step_v, step_nm, step_langevin, ... in verletmodule.f90) is one short launch
// sequence on the chain's stream, called once per step by the host, exactly as the
// reference's driver calls them.  The normal-mode transforms are FP64 MFMA tiles.
#include <cmath>
#include <cstring>

#include "rp_dispatch.h"
#include "rp_nm.h"
#include "rp_philox.h"

using namespace rp;

struct rp_chain {
  rp_pes* pes = nullptr;
  int n = 0, ndim = 0, natom = 0, ndof = 0;
  double betan = 0, tau = 0, dt = 0, gamma = 0;
  int cayley = 0;
  NmHost nm;                      // host copy of the init_nm arrays
  std::vector<double> mass;
  hipStream_t s = nullptr;
  DevBuf<double> d_T, d_lam, d_bm, d_mass, d_c1, d_c2, d_bv;
  DevBuf<double> d_state;         // [x (ndof*n) | p (ndof*n)]
  DevBuf<double> d_nm;            // [q | p~] work
  DevBuf<double> d_noise;         // injected draws
  DevBuf<double> d_ti;            // {sum}
  DevBuf<double> d_dbdl;
  DevBuf<int> d_flag;             // {nan, reinit}
  DevBuf<int> d_shift;            // per column: dof for positions, -1 for momenta
  // path spline (init_path / re-initialisation)
  int npath = 0;
  DevBuf<double> d_lampath, d_path, d_y2;
  // host copies for the fused propagation (a one-polymer batch): ends a, b and the path
  std::vector<double> ha, hb, hlampath, hpath;
  uint64_t seed = 0, stream = 0;
  int64_t cursor = 0;   // next normal of the stream
  int64_t kcursor = 0;  // next Poisson interval of the stream (Andersen)
};

// ---------------------------------------------------------------- kernels
// rotation of every (mode, dof) pair by h (verletmodule.f90:515-539)
__global__ void rotate_kernel(int n, int ndim, int natom, const double* __restrict__ lam,
                              const double* __restrict__ bm, const double* __restrict__ mass,
                              double h, int cayley, double* __restrict__ q,
                              double* __restrict__ pt, int* __restrict__ flag) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int ndof = ndim * natom;
  if (t >= n * ndof) return;
  int dof = t / n, i = t - dof * n;
  int k = dof / ndim;
  double b = bm[i * natom + k];
  double om = sqrt(mass[k] / b) * lam[i];
  double pi0 = pt[t], q0 = q[t], np, nq;
  if (cayley) {
    double w = om * om * h * h;
    np = (pi0 * (4.0 - w) - 4.0 * q0 * b * om * om * h) / (4.0 + w);
    nq = (q0 * (4.0 - w) + 4.0 * pi0 * h / b) / (4.0 + w);
  } else {
    double c = cos(h * om), sn = sin(om * h);
    np = pi0 * c - q0 * om * b * sn;
    nq = q0 * c + pi0 * sn / (om * b);
  }
  if (np != np) atomicOr(flag, 1);
  pt[t] = np;
  q[t] = nq;
}

// Langevin O-step in normal-mode space (verletmodule.f90:650-653)
__global__ void ostep_kernel(int n, int ndim, int natom, const double* __restrict__ c1,
                             const double* __restrict__ c2, const double* __restrict__ bm,
                             double betan, double* __restrict__ pt,
                             const double* __restrict__ noise, uint64_t seed, uint64_t stream,
                             int64_t cursor) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int ndof = ndim * natom;
  if (t >= n * ndof) return;
  int dofi = t / n, k = t - dofi * n;
  int j = dofi / ndim;
  double z = noise ? noise[t] : philox_normal(seed, stream, cursor + t);
  double a = c1[k * natom + j], b = c2[k * natom + j];
  pt[t] = (a * a) * pt[t] + sqrt(bm[k * natom + j] / betan) * b * sqrt(1.0 + a * a) * z;
}

// NM-space momenta of a resample (verletmodule.f90:211-219 and 102-109): draws come in
// (dim outer, atom inner) blocks of n
__global__ void andersen_kernel(int n, int ndim, int natom, const double* __restrict__ bm,
                                double stdev, double* __restrict__ out,
                                const double* __restrict__ noise, uint64_t seed,
                                uint64_t stream, int64_t cursor) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int ndof = ndim * natom;
  if (t >= n * ndof) return;
  int dof = t / n, l = t - dof * n;
  int atom = dof / ndim, dim = dof - atom * ndim;
  int64_t draw = (int64_t)(dim * natom + atom) * n + l;
  double z = noise ? noise[draw] : philox_normal(seed, stream, cursor + draw);
  out[t] = stdev * z * sqrt(bm[l * natom + atom]);
}

// TI estimator Sum m (-x_N) db/dxi with the dHdrlimit test (verletmodule.f90:397-410)
__global__ void ti_kernel(int n, int ndim, int natom, const double* __restrict__ mass,
                          const double* __restrict__ x, const double* __restrict__ dbdl,
                          double limit, double* __restrict__ ti, int* __restrict__ flag) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double contr = 0.0;
  for (int j = 0; j < ndim; ++j)
    for (int k = 0; k < natom; ++k) {
      int d = k * ndim + j;
      contr = contr + mass[k] * (-x[(size_t)d * n + n - 1]) * dbdl[d];
    }
  if (fabs(contr) < limit || limit < 0.0)
    ti[0] = ti[0] + contr;
  else
    flag[1] = 1;
}

// init_path positions: x_k = splint(path, (k-1) xi/(n-1)) (verletmodule.f90:40-48)
__global__ void init_x_kernel(int n, int ndof, int npath, const double* __restrict__ lampath,
                              const double* __restrict__ path, const double* __restrict__ y2,
                              double xi, double* __restrict__ x) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * ndof) return;
  int d = t / n, k = t - d * n;
  double xe = (double)k * xi / (double)(n - 1);
  x[t] = splint_dev(lampath, path + (size_t)d * npath, y2 + (size_t)d * npath, npath, xe);
}

// ---------------------------------------------------------------- helpers
static int chain_flag_check(rp_chain* c, int* flags) {
  RP_HIP(hipMemcpyAsync(flags, c->d_flag.p, 2 * sizeof(int), hipMemcpyDeviceToHost, c->s));
  RP_HIP(hipStreamSynchronize(c->s));
  return RP_OK;
}

static int nm_fwd(rp_chain* c, int ncols, const double* in, double* out, const int* shift) {
  return launch_nm_gemm(c->d_T.p, c->n, ncols, in, out, shift, c->d_bv.p, 0, c->s);
}
static int nm_bwd(rp_chain* c, int ncols, const double* in, double* out, const int* shift) {
  return launch_nm_gemm(c->d_T.p, c->n, ncols, in, out, shift, c->d_bv.p, 1, c->s);
}

static inline int grid1(int tot) { return (tot + 255) / 256; }

static int do_step_nm(rp_chain* c, double h) {
  const int n = c->n, D = c->ndof, nd = n * D;
  double* x = c->d_state.p;
  double* q = c->d_nm.p;
  int rc;
  // [q | p~] = T [x | p] - [beadvec | 0]
  if ((rc = nm_fwd(c, 2 * D, x, q, c->d_shift.p))) return rc;
  rotate_kernel<<<grid1(nd), 256, 0, c->s>>>(n, c->ndim, c->natom, c->d_lam.p, c->d_bm.p,
                                              c->d_mass.p, h, c->cayley, q, q + nd, c->d_flag.p);
  RP_HIP(hipGetLastError());
  return nm_bwd(c, 2 * D, q, x, c->d_shift.p);
}

static int do_step_v(rp_chain* c, double h) {
  const int nd = c->n * c->ndof;
  return launch_pes_kick(c->pes, c->n, 1, c->d_state.p, c->d_state.p + nd, h, c->d_flag.p, c->s);
}

// noise: device pointer to n*ndof draws, or nullptr for the Philox stream
static int do_step_langevin(rp_chain* c, const double* dnoise) {
  const int n = c->n, D = c->ndof, nd = n * D;
  double* p = c->d_state.p + nd;
  double* pt = c->d_nm.p;
  int rc;
  if ((rc = nm_fwd(c, D, p, pt, c->d_shift.p + D))) return rc;
  ostep_kernel<<<grid1(nd), 256, 0, c->s>>>(n, c->ndim, c->natom, c->d_c1.p, c->d_c2.p,
                                             c->d_bm.p, c->betan, pt, dnoise, c->seed,
                                             c->stream, c->cursor);
  RP_HIP(hipGetLastError());
  if (!dnoise) c->cursor += nd;
  return nm_bwd(c, D, pt, p, c->d_shift.p + D);
}

static int do_resample(rp_chain* c, const double* dnoise) {
  const int n = c->n, D = c->ndof, nd = n * D;
  double* pt = c->d_nm.p;
  andersen_kernel<<<grid1(nd), 256, 0, c->s>>>(n, c->ndim, c->natom, c->d_bm.p,
                                                sqrt(1.0 / c->betan), pt, dnoise, c->seed,
                                                c->stream, c->cursor);
  RP_HIP(hipGetLastError());
  if (!dnoise) c->cursor += nd;
  return nm_bwd(c, D, pt, c->d_state.p + nd, c->d_shift.p + D);
}

static int do_init_path(rp_chain* c, double xi, const double* dnoise) {
  RP_REQUIRE(c->npath >= 2, "init_path: no path set (rp_chain_set_path)");
  const int nd = c->n * c->ndof;
  init_x_kernel<<<grid1(nd), 256, 0, c->s>>>(c->n, c->ndof, c->npath, c->d_lampath.p,
                                              c->d_path.p, c->d_y2.p, xi, c->d_state.p);
  RP_HIP(hipGetLastError());
  return do_resample(c, dnoise);
}

static int nan_status(rp_chain* c) {
  int flags[2];
  int rc = chain_flag_check(c, flags);
  if (rc) return rc;
  if (flags[0]) {
    int zero[2] = {0, 0};
    RP_HIP(hipMemcpy(c->d_flag.p, zero, sizeof(zero), hipMemcpyHostToDevice));
    return set_error(RP_ERR_NAN, "NaN in the momenta (the reference stops here)");
  }
  return RP_OK;
}

static int upload_noise(rp_chain* c, const double* noise, size_t count) {
  int rc;
  if (c->d_noise.n < count && (rc = c->d_noise.alloc(count))) return rc;
  return c->d_noise.upload(noise, count, c->s);
}

// ---------------------------------------------------------------- C ABI
extern "C" int rp_nm_setup(rp_pes* pes, int n, int ndim, int natom, const double* mass,
                           double betan, double tau, double dt, double gamma, int cayley,
                           const double* a, const double* b, rp_chain** out) {
  RP_REQUIRE(out, "rp_nm_setup: out is NULL");
  *out = nullptr;
  RP_REQUIRE(pes && mass && a && b, "rp_nm_setup: NULL argument");
  RP_REQUIRE(n >= 2 && n <= 4096, "rp_nm_setup: n=%d outside 2..4096", n);
  RP_REQUIRE(ndim * natom == pes->ndof, "rp_nm_setup: ndim*natom=%d but the potential has %d",
             ndim * natom, pes->ndof);
  RP_REQUIRE(betan > 0 && tau > 0, "rp_nm_setup: betan and tau must be positive");
  int rc = ensure_init();
  if (rc) return rc;
  rp_chain* c = new rp_chain();
  c->pes = pes;
  c->n = n;
  c->ndim = ndim;
  c->natom = natom;
  c->ndof = ndim * natom;
  c->betan = betan;
  c->tau = tau;
  c->dt = dt;
  c->gamma = gamma;
  c->cayley = cayley;
  c->mass.assign(mass, mass + natom);
  c->ha.assign(a, a + ndim * natom);
  c->hb.assign(b, b + ndim * natom);
  nm_host_init(c->nm, n, ndim, natom, mass, betan, tau, dt, gamma, a, b);
  const int D = c->ndof, nd = n * D;
  std::vector<int> shift(2 * D);
  for (int d = 0; d < D; ++d) {
    shift[d] = d;
    shift[D + d] = -1;
  }
  int zero[2] = {0, 0};
  double dz = 0.0;
  if ((rc = [&]() -> int {
        RP_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
        int r;
        if ((r = c->d_T.alloc((size_t)n * n)) || (r = c->d_T.upload(c->nm.T.data(), (size_t)n * n, c->s)))
          return r;
        if ((r = c->d_lam.alloc(n)) || (r = c->d_lam.upload(c->nm.lam.data(), n, c->s))) return r;
        if ((r = c->d_bm.alloc(nd / ndim)) || (r = c->d_bm.upload(c->nm.bm.data(), (size_t)n * natom, c->s)))
          return r;
        if ((r = c->d_mass.alloc(natom)) || (r = c->d_mass.upload(mass, natom, c->s))) return r;
        if ((r = c->d_c1.alloc((size_t)n * natom)) || (r = c->d_c1.upload(c->nm.c1.data(), (size_t)n * natom, c->s)))
          return r;
        if ((r = c->d_c2.alloc((size_t)n * natom)) || (r = c->d_c2.upload(c->nm.c2.data(), (size_t)n * natom, c->s)))
          return r;
        if ((r = c->d_bv.alloc(nd)) || (r = c->d_bv.upload(c->nm.bv.data(), nd, c->s))) return r;
        if ((r = c->d_state.alloc(2 * (size_t)nd)) || (r = c->d_nm.alloc(2 * (size_t)nd))) return r;
        RP_HIP(hipMemsetAsync(c->d_state.p, 0, 2 * (size_t)nd * sizeof(double), c->s));
        if ((r = c->d_ti.alloc(1)) || (r = c->d_ti.upload(&dz, 1, c->s))) return r;
        if ((r = c->d_dbdl.alloc(D))) return r;
        if ((r = c->d_flag.alloc(2)) || (r = c->d_flag.upload(zero, 2, c->s))) return r;
        if ((r = c->d_shift.alloc(2 * D)) || (r = c->d_shift.upload(shift.data(), 2 * D, c->s))) return r;
        RP_HIP(hipStreamSynchronize(c->s));
        return RP_OK;
      }())) {
    rp_chain_destroy(c);
    return rc;
  }
  *out = c;
  return RP_OK;
}

extern "C" int rp_chain_destroy(rp_chain* c) {
  if (!c) return RP_OK;
  if (c->s) (void)hipStreamSynchronize(c->s);
  for (DevBuf<double>* b : {&c->d_T, &c->d_lam, &c->d_bm, &c->d_mass, &c->d_c1, &c->d_c2, &c->d_bv,
                            &c->d_state, &c->d_nm, &c->d_noise, &c->d_ti, &c->d_dbdl, &c->d_lampath,
                            &c->d_path, &c->d_y2})
    b->free();
  c->d_flag.free();
  c->d_shift.free();
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
  return RP_OK;
}

extern "C" int rp_chain_set_ends(rp_chain* c, const double* a, const double* b) {
  RP_REQUIRE(c && a && b, "rp_chain_set_ends: NULL argument");
  nm_host_beadvec(c->nm, c->n, c->ndim, c->natom, c->betan, a, b);
  c->ha.assign(a, a + c->ndof);
  c->hb.assign(b, b + c->ndof);
  int rc = c->d_bv.upload(c->nm.bv.data(), (size_t)c->n * c->ndof, c->s);
  if (rc) return rc;
  RP_HIP(hipStreamSynchronize(c->s));
  return RP_OK;
}

extern "C" int rp_chain_set_state(rp_chain* c, const double* x, const double* p) {
  RP_REQUIRE(c, "rp_chain_set_state: NULL chain");
  const size_t nd = (size_t)c->n * c->ndof;
  RP_HIP(hipMemcpyAsync(c->d_state.p, x, nd * sizeof(double), hipMemcpyHostToDevice, c->s));
  RP_HIP(hipMemcpyAsync(c->d_state.p + nd, p, nd * sizeof(double), hipMemcpyHostToDevice, c->s));
  RP_HIP(hipStreamSynchronize(c->s));
  return RP_OK;
}

extern "C" int rp_chain_get_state(rp_chain* c, double* x, double* p) {
  RP_REQUIRE(c, "rp_chain_get_state: NULL chain");
  const size_t nd = (size_t)c->n * c->ndof;
  if (x) RP_HIP(hipMemcpyAsync(x, c->d_state.p, nd * sizeof(double), hipMemcpyDeviceToHost, c->s));
  if (p) RP_HIP(hipMemcpyAsync(p, c->d_state.p + nd, nd * sizeof(double), hipMemcpyDeviceToHost, c->s));
  RP_HIP(hipStreamSynchronize(c->s));
  return RP_OK;
}

extern "C" int rp_chain_get_nm(rp_chain* c, double* T, double* lam, double* beadmass,
                               double* beadvec) {
  RP_REQUIRE(c, "rp_chain_get_nm: NULL chain");
  const size_t n = c->n;
  if (T) memcpy(T, c->nm.T.data(), n * n * sizeof(double));
  if (lam) memcpy(lam, c->nm.lam.data(), n * sizeof(double));
  if (beadmass) memcpy(beadmass, c->nm.bm.data(), n * c->natom * sizeof(double));
  if (beadvec) memcpy(beadvec, c->nm.bv.data(), n * c->ndof * sizeof(double));
  return RP_OK;
}

extern "C" int rp_chain_set_rng(rp_chain* c, uint64_t seed, uint64_t stream, int64_t cursor) {
  RP_REQUIRE(c && cursor >= 0, "rp_chain_set_rng: bad arguments");
  c->seed = seed;
  c->stream = stream;
  c->cursor = cursor;
  c->kcursor = 0;
  return RP_OK;
}

extern "C" int rp_chain_get_cursor(rp_chain* c, int64_t* cursor) {
  RP_REQUIRE(c && cursor, "rp_chain_get_cursor: NULL argument");
  *cursor = c->cursor;
  return RP_OK;
}

extern "C" int rp_chain_set_kick_cursor(rp_chain* c, int64_t kick_cursor) {
  RP_REQUIRE(c && kick_cursor >= 0, "rp_chain_set_kick_cursor: bad arguments");
  c->kcursor = kick_cursor;
  return RP_OK;
}

extern "C" int rp_chain_get_kick_cursor(rp_chain* c, int64_t* kick_cursor) {
  RP_REQUIRE(c && kick_cursor, "rp_chain_get_kick_cursor: NULL argument");
  *kick_cursor = c->kcursor;
  return RP_OK;
}

extern "C" int rp_nm_transform(rp_chain* c, int direction, int ncols, const int* shift_dof,
                               const double* in, double* out) {
  RP_REQUIRE(c && in && out && ncols > 0, "rp_nm_transform: bad arguments");
  RP_REQUIRE(direction == 0 || direction == 1, "rp_nm_transform: direction must be 0 or 1");
  const size_t n = c->n, tot = n * ncols;
  std::vector<int> sh(ncols, -1);
  if (shift_dof)
    for (int k = 0; k < ncols; ++k) {
      RP_REQUIRE(shift_dof[k] < c->ndof, "rp_nm_transform: shift dof %d out of range", shift_dof[k]);
      sh[k] = shift_dof[k];
    }
  DevBuf<double> din, dout;
  DevBuf<int> dsh;
  int rc;
  if ((rc = din.alloc(tot)) || (rc = dout.alloc(tot)) || (rc = dsh.alloc(ncols))) return rc;
  if ((rc = din.upload(in, tot, c->s)) || (rc = dsh.upload(sh.data(), ncols, c->s))) return rc;
  rc = launch_nm_gemm(c->d_T.p, c->n, ncols, din.p, dout.p, dsh.p, c->d_bv.p, direction, c->s);
  if (!rc) rc = dout.download(out, tot, c->s);
  if (!rc) RP_HIP(hipStreamSynchronize(c->s));
  din.free();
  dout.free();
  dsh.free();
  return rc;
}

extern "C" int rp_step_v(rp_chain* c, double h) {
  RP_REQUIRE(c, "rp_step_v: NULL chain");
  int rc = do_step_v(c, h);
  return rc ? rc : nan_status(c);
}

extern "C" int rp_step_nm(rp_chain* c, double h) {
  RP_REQUIRE(c, "rp_step_nm: NULL chain");
  int rc = do_step_nm(c, h);
  return rc ? rc : nan_status(c);
}

extern "C" int rp_step_langevin(rp_chain* c, const double* noise) {
  RP_REQUIRE(c, "rp_step_langevin: NULL chain");
  int rc;
  const size_t nd = (size_t)c->n * c->ndof;
  if (noise && (rc = upload_noise(c, noise, nd))) return rc;
  if ((rc = do_step_langevin(c, noise ? c->d_noise.p : nullptr))) return rc;
  RP_HIP(hipStreamSynchronize(c->s));
  return RP_OK;
}

static int pile_step(rp_chain* c, const double* dnoise) {
  int rc;
  if ((rc = do_step_v(c, c->dt))) return rc;
  if ((rc = do_step_nm(c, 0.5 * c->dt))) return rc;
  if ((rc = do_step_langevin(c, dnoise))) return rc;
  return do_step_nm(c, 0.5 * c->dt);
}

extern "C" int rp_time_step_pile(rp_chain* c, const double* noise) {
  RP_REQUIRE(c, "rp_time_step_pile: NULL chain");
  int rc;
  const size_t nd = (size_t)c->n * c->ndof;
  if (noise && (rc = upload_noise(c, noise, nd))) return rc;
  if ((rc = pile_step(c, noise ? c->d_noise.p : nullptr))) return rc;
  return nan_status(c);
}

extern "C" int rp_time_step_nm(rp_chain* c) {
  RP_REQUIRE(c, "rp_time_step_nm: NULL chain");
  int rc;
  if ((rc = do_step_nm(c, 0.5 * c->dt)) || (rc = do_step_v(c, c->dt)) ||
      (rc = do_step_nm(c, 0.5 * c->dt)))
    return rc;
  return nan_status(c);
}

extern "C" int rp_andersen_resample(rp_chain* c, const double* noise) {
  RP_REQUIRE(c, "rp_andersen_resample: NULL chain");
  int rc;
  const size_t nd = (size_t)c->n * c->ndof;
  if (noise && (rc = upload_noise(c, noise, nd))) return rc;
  if ((rc = do_resample(c, noise ? c->d_noise.p : nullptr))) return rc;
  RP_HIP(hipStreamSynchronize(c->s));
  return RP_OK;
}

extern "C" int rp_chain_set_path(rp_chain* c, int npath, const double* lampath,
                                 const double* path) {
  RP_REQUIRE(c && lampath && path && npath >= 2, "rp_chain_set_path: bad arguments");
  std::vector<double> y2;
  path_spline_host(npath, c->ndof, lampath, path, y2);
  int rc;
  const size_t np = (size_t)npath * c->ndof;
  if ((rc = c->d_lampath.alloc(npath)) || (rc = c->d_path.alloc(np)) || (rc = c->d_y2.alloc(np)))
    return rc;
  if ((rc = c->d_lampath.upload(lampath, npath, c->s)) || (rc = c->d_path.upload(path, np, c->s)) ||
      (rc = c->d_y2.upload(y2.data(), np, c->s)))
    return rc;
  c->npath = npath;
  c->hlampath.assign(lampath, lampath + npath);
  c->hpath.assign(path, path + np);
  RP_HIP(hipStreamSynchronize(c->s));
  return RP_OK;
}

extern "C" int rp_init_path(rp_chain* c, double xi, const double* noise) {
  RP_REQUIRE(c, "rp_init_path: NULL chain");
  int rc;
  const size_t nd = (size_t)c->n * c->ndof;
  if (noise && (rc = upload_noise(c, noise, nd))) return rc;
  if ((rc = do_init_path(c, xi, noise ? c->d_noise.p : nullptr))) return rc;
  RP_HIP(hipStreamSynchronize(c->s));
  return RP_OK;
}

// propagate_pimd_pile / propagate_pimd_nm as a one-polymer batch: the fused step kernel
// keeps the trajectory in one workgroup's LDS for all nmc steps (a launch per 8-step
// block, replayed from a captured graph), instead of a launch sequence and, with
// dHdrlimit, a host synchronisation per step.  Same draws, step order and TI rules as the
// per-step loop below (the batched engine follows the same reference routines); the
// stream positions come back to the chain.  Used for every device potential the batched
// engine supports (batch_supports).
static int propagate_fused(rp_chain* c, int thermostat, int64_t nmc, int64_t imin,
                           double dhdrlimit, int64_t noutput, const double* dbdl, double xi,
                           const double* noise, int64_t nnoise, const int32_t* kicks,
                           int64_t nkicks, double* dhdr, int64_t* nreinit) {
  const int n = c->n, D = c->ndof;
  const size_t nd = (size_t)n * D;
  rp_polymer_desc desc;
  desc.n = n;
  desc.betan = c->betan;
  desc.xi = xi;
  desc.a = c->ha.data();
  desc.b = c->hb.data();
  desc.dbdl = dbdl;
  desc.stream = c->stream;
  rp_batch* b = nullptr;
  int rc = rp_batch_create(c->pes, c->mass.data(), 1, &desc, c->tau, c->dt, c->gamma, thermostat,
                           noutput, dhdrlimit, c->cayley, c->seed, &b);
  if (rc) return rc;
  std::vector<double> x(nd), p(nd);
  double sum = 0.0;
  int64_t cnt = 0, nre = 0, cur = c->cursor, kcur = c->kcursor;
  rc = [&]() -> int {
    int r;
    if (c->npath >= 2 &&
        (r = rp_batch_set_path(b, c->npath, c->hlampath.data(), c->hpath.data())))
      return r;
    RP_HIP(hipMemcpyAsync(x.data(), c->d_state.p, nd * sizeof(double), hipMemcpyDeviceToHost, c->s));
    RP_HIP(hipMemcpyAsync(p.data(), c->d_state.p + nd, nd * sizeof(double), hipMemcpyDeviceToHost, c->s));
    RP_HIP(hipStreamSynchronize(c->s));
    if ((r = rp_batch_set_state(b, 0, x.data(), p.data()))) return r;
    if ((r = rp_batch_set_cursor(b, &cur, &kcur))) return r;
    if (kicks && (r = rp_batch_set_kicks(b, kicks, nkicks))) return r;
    if ((r = rp_batch_run(b, nmc, imin, noise, noise ? nnoise : 0))) return r;
    if ((r = rp_batch_ti_readout(b, &sum, &cnt, &nre))) return r;
    if ((r = rp_batch_get_state(b, 0, x.data(), p.data()))) return r;
    if ((r = rp_batch_get_cursor(b, &cur)) || (r = rp_batch_get_kick_cursor(b, &kcur))) return r;
    RP_HIP(hipMemcpyAsync(c->d_state.p, x.data(), nd * sizeof(double), hipMemcpyHostToDevice, c->s));
    RP_HIP(hipMemcpyAsync(c->d_state.p + nd, p.data(), nd * sizeof(double), hipMemcpyHostToDevice, c->s));
    RP_HIP(hipStreamSynchronize(c->s));
    return RP_OK;
  }();
  rp_batch_destroy(b);
  if (rc) return rc;
  // injected draws leave the chain's Philox positions where they were, as the loop does
  if (!noise) c->cursor = cur;
  if (!kicks && thermostat == RP_THERMO_ANDERSEN) c->kcursor = kcur;
  *dhdr = nmc > imin ? sum / (double)(nmc - imin) : 0.0;
  if (nreinit) *nreinit = nre;
  return RP_OK;
}

extern "C" int rp_propagate_pile(rp_chain* c, int64_t nmc, int64_t imin, double dhdrlimit,
                                 const double* dbdl, double xi, const double* noise,
                                 int64_t nnoise, double* dhdr, int64_t* nreinit) {
  RP_REQUIRE(c && dbdl && dhdr && nmc >= imin && imin >= 0, "rp_propagate_pile: bad arguments");
  RP_REQUIRE(!noise || nnoise >= 0, "rp_propagate_pile: negative nnoise");
  if (batch_supports(c->pes, c->n))
    return propagate_fused(c, RP_THERMO_LANGEVIN, nmc, imin, dhdrlimit, 1, dbdl, xi, noise, nnoise,
                           nullptr, 0, dhdr, nreinit);
  const int n = c->n, D = c->ndof;
  const size_t nd = (size_t)n * D;
  int rc;
  // injected draws: the trajectory consumes them sequentially, re-initialisations included
  RP_REQUIRE(!noise || nnoise >= 0, "rp_propagate_pile: negative nnoise");
  const size_t cap = noise ? (size_t)nnoise : 0;
  if (noise && (rc = upload_noise(c, noise, cap))) return rc;
  size_t ncur = 0;
  double zero = 0.0;
  if ((rc = c->d_ti.upload(&zero, 1, c->s)) || (rc = c->d_dbdl.upload(dbdl, D, c->s))) return rc;
  int64_t reinit = 0;
  for (int64_t ii = 1; ii <= nmc; ++ii) {
    const double* dn = nullptr;
    if (noise) {
      if (ncur + nd > cap) return set_error(RP_ERR_NOISE, "injected noise exhausted at step %lld", (long long)ii);
      dn = c->d_noise.p + ncur;
      ncur += nd;
    }
    if ((rc = pile_step(c, dn))) return rc;
    if (ii > imin) {
      ti_kernel<<<1, 64, 0, c->s>>>(n, c->ndim, c->natom, c->d_mass.p, c->d_state.p, c->d_dbdl.p,
                                    dhdrlimit, c->d_ti.p, c->d_flag.p);
      RP_HIP(hipGetLastError());
      if (dhdrlimit >= 0.0) {
        int flags[2];
        if ((rc = chain_flag_check(c, flags))) return rc;
        if (flags[0]) return nan_status(c);
        if (flags[1]) {
          int zf = 0;
          RP_HIP(hipMemcpyAsync(c->d_flag.p + 1, &zf, sizeof(int), hipMemcpyHostToDevice, c->s));
          const double* rn = nullptr;
          if (noise) {
            if (ncur + nd > cap) return set_error(RP_ERR_NOISE, "injected noise exhausted at re-init");
            rn = c->d_noise.p + ncur;
            ncur += nd;
          }
          if ((rc = do_init_path(c, xi, rn))) return rc;
          ++reinit;
        }
      }
    }
  }
  double sum = 0.0;
  if ((rc = c->d_ti.download(&sum, 1, c->s))) return rc;
  if ((rc = nan_status(c))) return rc;
  *dhdr = nmc > imin ? sum / (double)(nmc - imin) : 0.0;
  if (nreinit) *nreinit = reinit;
  return RP_OK;
}

extern "C" int rp_propagate_nm(rp_chain* c, int64_t nmc, int64_t imin, int64_t noutput,
                               const double* dbdl, const double* noise, int64_t nnoise,
                               const int32_t* kicks, int64_t nkicks, double* dhdr) {
  RP_REQUIRE(c && dbdl && dhdr && nmc >= imin && imin >= 0 && noutput > 0,
             "rp_propagate_nm: bad arguments");
  if (batch_supports(c->pes, c->n))
    return propagate_fused(c, RP_THERMO_ANDERSEN, nmc, imin, -1.0, noutput, dbdl, 0.0, noise,
                           nnoise, kicks, nkicks, dhdr, nullptr);
  const int n = c->n, D = c->ndof;
  const size_t nd = (size_t)n * D;
  int rc;
  if (noise && (rc = upload_noise(c, noise, (size_t)nnoise))) return rc;
  size_t ncur = 0;
  int64_t kcur = 0;  // injected intervals restart per call; the Philox ones continue
  auto next_kick = [&](int32_t* k) -> int {
    if (kicks) {
      if (kcur >= nkicks) return set_error(RP_ERR_NOISE, "injected Poisson draws exhausted");
      *k = kicks[kcur++];
    } else {
      *k = philox_poisson(c->seed, c->stream, c->kcursor++, (double)noutput);
    }
    return RP_OK;
  };
  double zero = 0.0;
  if ((rc = c->d_ti.upload(&zero, 1, c->s)) || (rc = c->d_dbdl.upload(dbdl, D, c->s))) return rc;
  int32_t rkick;
  if ((rc = next_kick(&rkick))) return rc;
  int64_t count = 0;
  for (int64_t ii = 1; ii <= nmc; ++ii) {
    count = count + 1;
    if (count >= rkick) {
      const double* dn = nullptr;
      if (noise) {
        if (ncur + nd > (size_t)nnoise) return set_error(RP_ERR_NOISE, "injected noise exhausted");
        dn = c->d_noise.p + ncur;
        ncur += nd;
      }
      if ((rc = do_resample(c, dn))) return rc;
      if ((rc = next_kick(&rkick))) return rc;
      count = 0;
    }
    if ((rc = do_step_nm(c, 0.5 * c->dt)) || (rc = do_step_v(c, c->dt)) ||
        (rc = do_step_nm(c, 0.5 * c->dt)))
      return rc;
    if (ii > imin) {
      ti_kernel<<<1, 64, 0, c->s>>>(n, c->ndim, c->natom, c->d_mass.p, c->d_state.p, c->d_dbdl.p,
                                    -1.0, c->d_ti.p, c->d_flag.p);
      RP_HIP(hipGetLastError());
    }
  }
  double sum = 0.0;
  if ((rc = c->d_ti.download(&sum, 1, c->s))) return rc;
  if ((rc = nan_status(c))) return rc;
  *dhdr = nmc > imin ? sum / (double)(nmc - imin) : 0.0;
  return RP_OK;
}

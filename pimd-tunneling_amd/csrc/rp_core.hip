// rp_core.hip -- runtime, errors, and the potential (mcmod_mass) objects.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "rp_dispatch.h"

namespace rp {

static thread_local std::string g_err;
static int g_device = -1;
static std::mutex g_init_mu;

// Staging memory of rp_pes_potforce_noalloc, the entry point that allocates nothing: one
// array per device (each device holds its own copy until the process ends), 2^20 doubles =
// 8 MiB.  Host code never passes the array's name to a kernel or a copy (on the host it is
// only a stand-in): rp_init takes each device's address of it once, with that device
// current, and every use goes through g_scratch_ptr.
constexpr size_t kScratchDoubles = size_t(1) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kScratchDoubles];
static double* g_scratch_ptr[kMaxDevices] = {};

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int ensure_init() {
  if (g_device >= 0) {
    RP_HIP(hipSetDevice(g_device));
    return RP_OK;
  }
  return rp_init(0);
}

// ---------------------------------------------------------------- kernels
// One thread per bead; x/grad in [dof][bead] layout with leading dimension ld.
template <int KIND, int MAXD>
__global__ void __launch_bounds__(128) pes_eval_kernel(PesDev P, int nbead, int ld,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ grad,
                                                       double* __restrict__ energy) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbead) return;
  double xs[MAXD], gs[MAXD];
#pragma unroll
  for (int d = 0; d < MAXD; ++d) xs[d] = (d < P.ndof) ? x[(size_t)d * ld + b] : 0.0;
  double v = eval_kind<KIND, MAXD>(P, xs, gs);
  if (energy) energy[b] = v;
  if (grad) {
#pragma unroll
    for (int d = 0; d < MAXD; ++d)
      if (d < P.ndof) grad[(size_t)d * ld + b] = gs[d];
  }
}

template <int KIND, int MAXD>
__global__ void __launch_bounds__(128) pes_hvp_kernel(PesDev P, int nbead, int ld,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ v,
                                                      double* __restrict__ hv) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbead) return;
  Dual xs[MAXD], gs[MAXD];
#pragma unroll
  for (int d = 0; d < MAXD; ++d)
    xs[d] = (d < P.ndof) ? Dual(x[(size_t)d * ld + b], v[(size_t)d * ld + b]) : Dual(0.0);
  (void)eval_kind<KIND, MAXD>(P, xs, gs);
#pragma unroll
  for (int d = 0; d < MAXD; ++d)
    if (d < P.ndof) hv[(size_t)d * ld + b] = gs[d].d;
}

// Hessian rows of one bead: hess[b][i][k] = d grad_k / d x_i from ndof HVPs with e_i.
template <int KIND, int MAXD>
__global__ void __launch_bounds__(64) pes_hess_kernel(PesDev P, int nbead,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ hess) {
  int b = blockIdx.x;
  int i = threadIdx.x;
  if (b >= nbead || i >= P.ndof) return;
  Dual xs[MAXD], gs[MAXD];
#pragma unroll
  for (int d = 0; d < MAXD; ++d)
    xs[d] = (d < P.ndof) ? Dual(x[(size_t)d * nbead + b], d == i ? 1.0 : 0.0) : Dual(0.0);
  (void)eval_kind<KIND, MAXD>(P, xs, gs);
  const int D = P.ndof;
#pragma unroll
  for (int k = 0; k < MAXD; ++k)
    if (k < D) hess[((size_t)b * D + i) * D + k] = gs[k].d;
}

// p -= h * grad V(x), chains [chain][dof][bead] with n beads each.
template <int KIND, int MAXD>
__global__ void __launch_bounds__(128) pes_kick_kernel(PesDev P, int n, int nchain,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ p, double h,
                                                       int* __restrict__ nan_flag) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nchain) return;
  int c = t / n, b = t - c * n;
  const size_t base = (size_t)c * P.ndof * n + b;
  double xs[MAXD], gs[MAXD];
#pragma unroll
  for (int d = 0; d < MAXD; ++d) xs[d] = (d < P.ndof) ? x[base + (size_t)d * n] : 0.0;
  (void)eval_kind<KIND, MAXD>(P, xs, gs);
  bool bad = false;
#pragma unroll
  for (int d = 0; d < MAXD; ++d) {
    if (d < P.ndof) {
      double np = p[base + (size_t)d * n] - gs[d] * h;
      p[base + (size_t)d * n] = np;
      bad |= (np != np);
    }
  }
  if (bad) atomicOr(nan_flag, 1);
}

template <int NMOL>
__global__ void __launch_bounds__(128) water_param_grad_kernel(PesDev P, int nbead,
                                                               const double* __restrict__ x,
                                                               double* __restrict__ out) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbead) return;
  double xs[9 * NMOL], gs[9 * NMOL], pg[2];
#pragma unroll
  for (int d = 0; d < 9 * NMOL; ++d) xs[d] = x[(size_t)d * nbead + b];
  (void)water_eval<NMOL>(P, xs, gs, pg);
  out[2 * b] = pg[0];
  out[2 * b + 1] = pg[1];
}

// p -= h * g elementwise (the kick of a potential whose gradient g came from the host)
__global__ void __launch_bounds__(256) kick_grad_kernel(int64_t count, const double* __restrict__ g,
                                                        double* __restrict__ p, double h,
                                                        int* __restrict__ nan_flag) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  const double np = p[e] - g[e] * h;
  p[e] = np;
  if (np != np) atomicOr(nan_flag, 1);
}

int host_pes_eval(const rp_pes* pes, int nbead, int ld, const double* x, double* grad,
                  double* energy, hipStream_t s) {
  if (nbead <= 0) return RP_OK;
  const int D = pes->ndof;
  const size_t nx = (size_t)nbead * D;
  pes->hx.resize(nx);
  pes->hg.resize(nx);
  pes->he.resize(nbead);
  // [dof][bead] of stride ld on the device <-> (nbead, ndim, natom) column-major on the host
  RP_HIP(hipMemcpy2DAsync(pes->hx.data(), nbead * sizeof(double), x, ld * sizeof(double),
                          nbead * sizeof(double), D, hipMemcpyDeviceToHost, s));
  RP_HIP(hipStreamSynchronize(s));
  if (pes->host_pf(pes->host_user, nbead, pes->hx.data(), pes->hg.data(), pes->he.data()) != 0)
    return set_error(RP_ERR_ARG, "the host potential's potforce callback reported an error");
  if (grad)
    RP_HIP(hipMemcpy2DAsync(grad, ld * sizeof(double), pes->hg.data(), nbead * sizeof(double),
                            nbead * sizeof(double), D, hipMemcpyHostToDevice, s));
  if (energy) {
    for (double& e : pes->he) e -= pes->v0;
    RP_HIP(hipMemcpyAsync(energy, pes->he.data(), nbead * sizeof(double), hipMemcpyHostToDevice, s));
  }
  RP_HIP(hipStreamSynchronize(s));  // the staging vectors are reused by the next call
  return RP_OK;
}

int host_pes_hessian(const rp_pes* pes, int nbead, const double* x, double* hess) {
  const int D = pes->ndof;
  if (nbead <= 0) return RP_OK;
  if (pes->host_hs) {
    if (pes->host_hs(pes->host_user, nbead, x, hess) != 0)
      return set_error(RP_ERR_ARG, "the host potential's hessian callback reported an error");
    return RP_OK;
  }
  // central differences of the gradient, step 1e-4 (mcmod_waterdimer.f90:83-103); a bead's
  // gradient depends on that bead alone, so dof i of every bead moves at once
  const double eps = 1e-4;
  const size_t nx = (size_t)nbead * D;
  std::vector<double> xp(x, x + nx), gp(nx), gm(nx), e(nbead);
  for (int i = 0; i < D; ++i) {
    for (int b = 0; b < nbead; ++b) xp[(size_t)i * nbead + b] = x[(size_t)i * nbead + b] + eps;
    if (pes->host_pf(pes->host_user, nbead, xp.data(), gp.data(), e.data()) != 0)
      return set_error(RP_ERR_ARG, "the host potential's potforce callback reported an error");
    for (int b = 0; b < nbead; ++b) xp[(size_t)i * nbead + b] = x[(size_t)i * nbead + b] - eps;
    if (pes->host_pf(pes->host_user, nbead, xp.data(), gm.data(), e.data()) != 0)
      return set_error(RP_ERR_ARG, "the host potential's potforce callback reported an error");
    for (int b = 0; b < nbead; ++b) {
      xp[(size_t)i * nbead + b] = x[(size_t)i * nbead + b];
      for (int k = 0; k < D; ++k)
        hess[((size_t)b * D + i) * D + k] =
            (gp[(size_t)k * nbead + b] - gm[(size_t)k * nbead + b]) / (2.0 * eps);
    }
  }
  return RP_OK;
}

int launch_pes_eval(const rp_pes* pes, int nbead, int ld, const double* x, double* grad,
                    double* energy, hipStream_t s) {
  if (nbead <= 0) return RP_OK;
  if (pes->kind == PES_HOST) return host_pes_eval(pes, nbead, ld, x, grad, energy, s);
  PesDev P = pes->dev();
  return dispatch_pes(pes, [&]<int K, int M>() -> int {
    pes_eval_kernel<K, M><<<(nbead + 127) / 128, 128, 0, s>>>(P, nbead, ld, x, grad, energy);
    RP_HIP(hipGetLastError());
    return RP_OK;
  });
}

int launch_pes_hvp(const rp_pes* pes, int nbead, int ld, const double* x, const double* v,
                   double* hv, hipStream_t s) {
  if (nbead <= 0) return RP_OK;
  PesDev P = pes->dev();
  return dispatch_pes(pes, [&]<int K, int M>() -> int {
    pes_hvp_kernel<K, M><<<(nbead + 127) / 128, 128, 0, s>>>(P, nbead, ld, x, v, hv);
    RP_HIP(hipGetLastError());
    return RP_OK;
  });
}

int launch_pes_kick(const rp_pes* pes, int n, int nchain, const double* x, double* p, double h,
                    int* nan_flag, hipStream_t s) {
  int tot = n * nchain;
  if (tot <= 0) return RP_OK;
  if (pes->kind == PES_HOST) {
    // the plugin's gradient for every bead of each chain, then the kick on the device
    const int64_t count = (int64_t)tot * pes->ndof;
    int rc;
    if (pes->d_hg.n < (size_t)count && (rc = pes->d_hg.alloc(count))) return rc;
    for (int c = 0; c < nchain; ++c) {
      const size_t o = (size_t)c * pes->ndof * n;
      if ((rc = host_pes_eval(pes, n, n, x + o, pes->d_hg.p + o, nullptr, s))) return rc;
    }
    kick_grad_kernel<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(count, pes->d_hg.p, p, h, nan_flag);
    RP_HIP(hipGetLastError());
    return RP_OK;
  }
  PesDev P = pes->dev();
  return dispatch_pes(pes, [&]<int K, int M>() -> int {
    pes_kick_kernel<K, M><<<(tot + 127) / 128, 128, 0, s>>>(P, n, nchain, x, p, h, nan_flag);
    RP_HIP(hipGetLastError());
    return RP_OK;
  });
}

int pes_maxd(const rp_pes* pes) {
  int D = pes->ndof;
  if (pes->kind == PES_WATER) return D;
  if (D <= 16) return 16;
  if (D <= 32) return 32;
  if (D <= 64) return 64;
  return 0;
}

}  // namespace rp

using namespace rp;

// ---------------------------------------------------------------- C ABI
extern "C" int rp_init(int device) {
  std::lock_guard<std::mutex> lk(g_init_mu);
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return set_error(RP_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  if (device < 0 || device >= count)
    return set_error(RP_ERR_ARG, "device %d out of range (count %d)", device, count);
  RP_HIP(hipSetDevice(device));
  if (device < kMaxDevices && !g_scratch_ptr[device]) {
    void* ptr = nullptr;
    RP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<double*>(ptr);
  }
  g_device = device;
  return RP_OK;
}

extern "C" int rp_finalize(void) {
  if (g_device >= 0) RP_HIP(hipDeviceSynchronize());
  return RP_OK;
}

extern "C" int rp_last_error(char* buf, int len) {
  if (!buf || len <= 0) return RP_ERR_ARG;
  snprintf(buf, (size_t)len, "%s", g_err.c_str());
  return RP_OK;
}

extern "C" int rp_version(void) { return 1; }

extern "C" int rp_pes_create(int kind, int ndim, int natom, const double* dp, int ndp,
                             const int* ip, int nip, rp_pes** out) {
  RP_REQUIRE(out != nullptr, "rp_pes_create: out is NULL");
  *out = nullptr;
  RP_REQUIRE(ndim > 0 && natom > 0, "rp_pes_create: ndim=%d natom=%d", ndim, natom);
  int rc = ensure_init();
  if (rc) return rc;
  const int D = ndim * natom;
  int need = 0;
  switch (kind) {
    case PES_QUARTIC: need = 2; break;
    case PES_COUPLED2D:
      RP_REQUIRE(D == 2, "COUPLED2D needs ndof = 2 (got %d)", D);
      need = 5;
      break;
    case PES_EVB:
      need = 1 + D * D + 3 * D;
      RP_REQUIRE(ip && nip == natom, "EVB needs the atom permutation (nip = natom)");
      for (int a = 0; a < natom; ++a)
        RP_REQUIRE(ip[a] >= 0 && ip[a] < natom, "EVB permutation entry %d out of range", a);
      break;
    case PES_WATER:
      RP_REQUIRE(ndim == 3 && (natom == 3 || natom == 6 || natom == 18),
                 "WATER needs ndim=3 and 1, 2 or 6 molecules (natom=%d)", natom);
      need = 8;
      break;
    case PES_HARMONIC: need = 2 * D; break;
    default: return set_error(RP_ERR_ARG, "unknown potential kind %d", kind);
  }
  RP_REQUIRE(dp && ndp >= need, "potential kind %d needs %d parameters (got %d)", kind, need, ndp);
  rp_pes* P = new rp_pes();
  P->kind = kind;
  P->ndim = ndim;
  P->natom = natom;
  P->ndof = D;
  P->v0 = 0.0;
  P->dp.assign(dp, dp + need);
  if (ip && nip > 0) P->ip.assign(ip, ip + nip);
  std::vector<double> dev = P->dp;
  if (kind == PES_EVB) {
    // API {kappa, Q, lam, cub, xa} -> device {kappa, WtA, cwA, WtB, cwB, lam, cub} with row
    // stride M = evb_stride(D) (rp_pes.h).  (Pi x)_j = x_{pi(j)}, pi(a*nd+d) = perm[a]*nd+d;
    // WB(k, i) = Q(pi^-1(k), i), xb_k = xa_{pi^-1(k)}.
    const double* Q = dp + 1;
    const double* lam = Q + D * D;
    const double* cub = lam + D;
    const double* xa = cub + D;
    const int M = evb_stride(D);
    std::vector<int> pinv(D);
    for (int a = 0; a < natom; ++a)
      for (int d = 0; d < ndim; ++d) pinv[ip[a] * ndim + d] = a * ndim + d;
    dev.assign(1 + 2 * (M * M + M) + 2 * M, 0.0);
    dev[0] = dp[0];
    double* WA = dev.data() + 1;
    double* cA = WA + M * M;
    double* WB = cA + M;
    double* cB = WB + M * M;
    double* L = cB + M;
    double* Cc = L + M;
    for (int i = 0; i < D; ++i) {
      double sa = 0.0, sb = 0.0;
      for (int j = 0; j < D; ++j) {
        WA[i * M + j] = Q[j + D * i];
        WB[i * M + j] = Q[pinv[j] + D * i];
        sa += Q[j + D * i] * xa[j];
        sb += Q[pinv[j] + D * i] * xa[pinv[j]];
      }
      cA[i] = sa;
      cB[i] = sb;
      L[i] = lam[i];
      Cc[i] = cub[i];
    }
  }
  RP_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
  if ((rc = P->d_dp.alloc(dev.size())) || (rc = P->d_dp.upload(dev.data(), dev.size(), P->stream))) {
    delete P;
    return rc;
  }
  std::vector<int> ipv = P->ip.empty() ? std::vector<int>{0} : P->ip;
  if ((rc = P->d_ip.alloc(ipv.size())) || (rc = P->d_ip.upload(ipv.data(), ipv.size(), P->stream))) {
    delete P;
    return rc;
  }
  RP_HIP(hipStreamSynchronize(P->stream));
  *out = P;
  return RP_OK;
}

extern "C" int rp_pes_create_host(int ndim, int natom, rp_host_potforce_fn potforce,
                                  rp_host_hessian_fn hessian, void* user, rp_pes** out) {
  RP_REQUIRE(out != nullptr, "rp_pes_create_host: out is NULL");
  *out = nullptr;
  RP_REQUIRE(ndim > 0 && natom > 0, "rp_pes_create_host: ndim=%d natom=%d", ndim, natom);
  RP_REQUIRE(potforce != nullptr, "rp_pes_create_host: potforce callback is NULL");
  int rc = ensure_init();
  if (rc) return rc;
  rp_pes* P = new rp_pes();
  P->kind = PES_HOST;
  P->ndim = ndim;
  P->natom = natom;
  P->ndof = ndim * natom;
  P->v0 = 0.0;  // the plugin subtracts its own V0
  P->host_pf = potforce;
  P->host_hs = hessian;
  P->host_user = user;
  if (hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking) != hipSuccess) {
    delete P;
    return set_error(RP_ERR_HIP, "rp_pes_create_host: stream creation failed");
  }
  *out = P;
  return RP_OK;
}

extern "C" int rp_pes_destroy(rp_pes* pes) {
  if (!pes) return RP_OK;
  (void)hipStreamSynchronize(pes->stream);
  pes->d_dp.free();
  pes->d_ip.free();
  pes->d_x.free();
  pes->d_g.free();
  pes->d_e.free();
  pes->d_hg.free();
  (void)hipStreamDestroy(pes->stream);
  delete pes;
  return RP_OK;
}

extern "C" int rp_pes_set_v0(rp_pes* pes, double v0) {
  RP_REQUIRE(pes, "rp_pes_set_v0: NULL potential");
  pes->v0 = v0;
  return RP_OK;
}

static int pes_scratch(rp_pes* pes, size_t nx, size_t ng, size_t ne) {
  int rc;
  if (pes->d_x.n < nx && (rc = pes->d_x.alloc(nx))) return rc;
  if (pes->d_g.n < ng && (rc = pes->d_g.alloc(ng))) return rc;
  if (pes->d_e.n < ne && (rc = pes->d_e.alloc(ne))) return rc;
  return RP_OK;
}

extern "C" int rp_pes_potforce(rp_pes* pes, int nbead, const double* x, double* grad,
                               double* energy) {
  RP_REQUIRE(pes && x && nbead >= 0, "rp_pes_potforce: bad arguments");
  if (nbead == 0) return RP_OK;
  int rc;
  const size_t nx = (size_t)nbead * pes->ndof;
  if (pes->kind == PES_HOST) {  // the caller's own plugin, called directly
    std::vector<double> g(nx), e(nbead);
    if (pes->host_pf(pes->host_user, nbead, x, g.data(), e.data()) != 0)
      return set_error(RP_ERR_ARG, "the host potential's potforce callback reported an error");
    if (grad) memcpy(grad, g.data(), nx * sizeof(double));
    if (energy)
      for (int b = 0; b < nbead; ++b) energy[b] = e[b] - pes->v0;
    return RP_OK;
  }
  if ((rc = ensure_init()) || (rc = pes_scratch(pes, nx, nx, nbead))) return rc;
  if ((rc = pes->d_x.upload(x, nx, pes->stream))) return rc;
  if ((rc = launch_pes_eval(pes, nbead, nbead, pes->d_x.p, grad ? pes->d_g.p : nullptr,
                            pes->d_e.p, pes->stream)))
    return rc;
  if (grad && (rc = pes->d_g.download(grad, nx, pes->stream))) return rc;
  if (energy && (rc = pes->d_e.download(energy, nbead, pes->stream))) return rc;
  RP_HIP(hipStreamSynchronize(pes->stream));
  return RP_OK;
}

// rp_pes_potforce with its staging in the static scratch array instead of the potential's
// own device buffers: no allocation, so callers that cannot allocate (graph capture, tight
// serving loops) can call it.  x, grad and energy are host arrays as for rp_pes_potforce.
// Every call on a device shares that device's array: calls that overlap on one device (two
// threads, or two potentials' streams) race, and nothing of a call survives the next.
extern "C" int rp_pes_potforce_noalloc(rp_pes* pes, int nbead, const double* x, double* grad,
                                       double* energy) {
  RP_REQUIRE(pes && x && nbead >= 0, "rp_pes_potforce_noalloc: bad arguments");
  RP_REQUIRE(pes->kind != PES_HOST,
             "rp_pes_potforce_noalloc: a potential that lives on the host has nothing to stage "
             "on the device (call rp_pes_potforce)");
  if (nbead == 0) return RP_OK;
  const size_t nx = (size_t)nbead * pes->ndof;
  const size_t need = 2 * nx + (size_t)nbead;
  RP_REQUIRE(need <= kScratchDoubles,
             "rp_pes_potforce_noalloc: %d beads of %d dof need %zu doubles of staging; the "
             "limit is %zu (at most %zu beads)", nbead, pes->ndof, need, kScratchDoubles,
             kScratchDoubles / (2 * (size_t)pes->ndof + 1));
  int rc;
  if ((rc = ensure_init())) return rc;
  RP_REQUIRE(g_device >= 0 && g_device < kMaxDevices && g_scratch_ptr[g_device],
             "rp_pes_potforce_noalloc: no staging array on device %d", g_device);
  double* dx = g_scratch_ptr[g_device];
  double* dg = dx + nx;
  double* de = dg + nx;
  hipStream_t s = pes->stream;
  RP_HIP(hipMemcpyAsync(dx, x, nx * sizeof(double), hipMemcpyHostToDevice, s));
  if ((rc = launch_pes_eval(pes, nbead, nbead, dx, grad ? dg : nullptr, de, s))) return rc;
  if (grad) RP_HIP(hipMemcpyAsync(grad, dg, nx * sizeof(double), hipMemcpyDeviceToHost, s));
  if (energy) RP_HIP(hipMemcpyAsync(energy, de, nbead * sizeof(double), hipMemcpyDeviceToHost, s));
  RP_HIP(hipStreamSynchronize(s));
  return RP_OK;
}

extern "C" int rp_pes_hessian(rp_pes* pes, int nbead, const double* x, double* hess) {
  RP_REQUIRE(pes && x && hess && nbead >= 0, "rp_pes_hessian: bad arguments");
  if (nbead == 0) return RP_OK;
  if (pes->kind == PES_HOST) return host_pes_hessian(pes, nbead, x, hess);
  RP_REQUIRE(pes->ndof <= 64, "rp_pes_hessian: ndof %d > 64", pes->ndof);
  int rc;
  const size_t nx = (size_t)nbead * pes->ndof, nh = nx * pes->ndof;
  if ((rc = ensure_init()) || (rc = pes_scratch(pes, nx, nh, 1))) return rc;
  if ((rc = pes->d_x.upload(x, nx, pes->stream))) return rc;
  PesDev P = pes->dev();
  hipStream_t s = pes->stream;
  const double* dx = pes->d_x.p;
  double* dh = pes->d_g.p;
  rc = dispatch_pes(pes, [&]<int K, int M>() -> int {
    pes_hess_kernel<K, M><<<nbead, 64, 0, s>>>(P, nbead, dx, dh);
    RP_HIP(hipGetLastError());
    return RP_OK;
  });
  if (rc) return rc;
  if ((rc = pes->d_g.download(hess, nh, s))) return rc;
  RP_HIP(hipStreamSynchronize(s));
  return RP_OK;
}

extern "C" int rp_pes_param_grad(rp_pes* pes, int nbead, const double* x, double* out,
                                 int* nparam) {
  RP_REQUIRE(pes && x && out && nbead >= 0, "rp_pes_param_grad: bad arguments");
  if (pes->kind != PES_WATER)
    return set_error(RP_ERR_ARG, "potential kind %d has no parameter sensitivities", pes->kind);
  if (nparam) *nparam = 2;
  if (nbead == 0) return RP_OK;
  int rc;
  const size_t nx = (size_t)nbead * pes->ndof;
  if ((rc = ensure_init()) || (rc = pes_scratch(pes, nx, 2 * (size_t)nbead, 1))) return rc;
  if ((rc = pes->d_x.upload(x, nx, pes->stream))) return rc;
  PesDev P = pes->dev();
  int nb = (nbead + 127) / 128;
  switch (pes->natom) {
    case 3: water_param_grad_kernel<1><<<nb, 128, 0, pes->stream>>>(P, nbead, pes->d_x.p, pes->d_g.p); break;
    case 6: water_param_grad_kernel<2><<<nb, 128, 0, pes->stream>>>(P, nbead, pes->d_x.p, pes->d_g.p); break;
    default: water_param_grad_kernel<6><<<nb, 128, 0, pes->stream>>>(P, nbead, pes->d_x.p, pes->d_g.p); break;
  }
  RP_HIP(hipGetLastError());
  if ((rc = pes->d_g.download(out, 2 * (size_t)nbead, pes->stream))) return rc;
  RP_HIP(hipStreamSynchronize(pes->stream));
  return RP_OK;
}

// rp_internal.h -- shared host-side plumbing of libpimdtunnel.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rp_hip.h"
#include "rp_pes.h"

namespace rp {

constexpr double kPiRef = 3.14159265358979;  // instantonmod.f90:4 (truncated literal)

int set_error(int code, const char* fmt, ...);

#define RP_HIP(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return ::rp::set_error(RP_ERR_HIP, "%s failed: %s (%s:%d)", #expr,              \
                             hipGetErrorString(e_), __FILE__, __LINE__);               \
  } while (0)

#define RP_REQUIRE(cond, ...)                                                          \
  do {                                                                                 \
    if (!(cond)) return ::rp::set_error(RP_ERR_ARG, __VA_ARGS__);                      \
  } while (0)

int ensure_init();  // rp_init(0) on first use

// Device buffer of T with explicit ownership; freed by the owning handle.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    free();
    if (count == 0) return RP_OK;
    RP_HIP(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
    return RP_OK;
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int upload(const T* h, size_t count, hipStream_t s) {
    RP_REQUIRE(count <= n, "upload of %zu elements into a buffer of %zu", count, n);
    RP_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    return RP_OK;
  }
  int download(T* h, size_t count, hipStream_t s) const {
    RP_REQUIRE(count <= n, "download of %zu elements from a buffer of %zu", count, n);
    RP_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    return RP_OK;
  }
};

}  // namespace rp

struct rp_pes {
  int kind, ndim, natom, ndof;
  double v0;
  std::vector<double> dp;   // API-layout parameters (as passed to rp_pes_create)
  std::vector<int> ip;
  rp::DevBuf<double> d_dp;  // device layout (EVB expanded)
  rp::DevBuf<int> d_ip;
  hipStream_t stream;
  rp::DevBuf<double> d_x, d_g, d_e;  // scratch for the per-call entry points
  // a potential that lives on the host (kind PES_HOST): the plugin's callbacks, the host
  // staging of their arguments and the device copy of their gradients
  rp_host_potforce_fn host_pf = nullptr;
  rp_host_hessian_fn host_hs = nullptr;
  void* host_user = nullptr;
  mutable std::vector<double> hx, hg, he;
  mutable rp::DevBuf<double> d_hg;
  rp::PesDev dev() const {
    rp::PesDev P;
    P.kind = kind;
    P.ndim = ndim;
    P.natom = natom;
    P.ndof = ndof;
    P.v0 = v0;
    P.dp = d_dp.p;
    P.ip = d_ip.p;
    return P;
  }
};

namespace rp {
// Launch helpers shared by the engines (defined in rp_core.hip).
// Gradient kick p -= h * grad V(x) for every bead of nchain chains stored
// [chain][dof][bead] with per-chain bead count n (uniform); sets *nan_flag on NaN.
int launch_pes_kick(const rp_pes* pes, int nbead_per_chain, int nchain, const double* x,
                    double* p, double h, int* nan_flag, hipStream_t s);
// energy / gradient of nbead beads in layout [dof][bead] (stride ld between dofs)
int launch_pes_eval(const rp_pes* pes, int nbead, int ld, const double* x, double* grad,
                    double* energy, hipStream_t s);
// Hessian-vector products of nbead beads: hv = H(x) v, layout as launch_pes_eval
int launch_pes_hvp(const rp_pes* pes, int nbead, int ld, const double* x, const double* v,
                   double* hv, hipStream_t s);
// A potential that lives on the host: energies and gradients of nbead beads from its
// potforce callback.  x and grad are device buffers in [dof][bead] layout of stride ld (grad
// may be null), energy a device array of nbead (may be null).  Synchronises s.
int host_pes_eval(const rp_pes* pes, int nbead, int ld, const double* x, double* grad,
                  double* energy, hipStream_t s);
// Host potential: hess[b][i][k] = d grad_k / d x_i of nbead beads at the host chain x
// ((nbead, ndim, natom) column-major), from the hessian callback or central differences.
int host_pes_hessian(const rp_pes* pes, int nbead, const double* x, double* hess);
// MAXD bucket a potential needs (16, 32, 64) or 0 when unsupported
int pes_maxd(const rp_pes* pes);
// Whether the batched engine runs a polymer of n beads on this potential: a device
// potential of a supported kind and size, n in 2..256, its workgroup within a CU's LDS.
bool batch_supports(const rp_pes* pes, int n);
// row stride of the device EVB layout (the MAXD bucket dispatch_pes picks)
inline int evb_stride(int D) { return D <= 16 ? 16 : D == 27 ? 27 : 32; }
}  // namespace rp

// rp_ring.hip -- ring-polymer action, adjoint gradient, Hessian and log-determinant
// for batches of open fixed-end chains of uneven length (instantonmod.f90:17-217,
// 782-827), the device side of the instanton search.
//
// Layout: chains concatenated, chain c at offset off[c] = ndof * sum_{c'<c} n_c', each
// chain (n, ndim, natom) column-major: x[off + dof*n + bead].  One workgroup per chain.
#include <cmath>
#include <cstring>

#include "rp_dispatch.h"
#include "rp_ring.h"
#include <algorithm>

using namespace rp;

namespace {
constexpr int kRingThreads = 256;

struct RingArgs {
  PesDev pes;
  int nchain, D, ndim, natom, fixedends;
  const int* n;
  const int64_t* off;      // per chain, in doubles
  const int64_t* boff;     // per chain, first global bead index
  const double* betan;
  const double* ab;        // per chain [a | b], D each
  const double* massd;     // per dof
};

// A barrier that orders LDS only, so global loads and stores stay in flight across it (a
// one-wave workgroup needs no hardware barrier at all: its LDS operations complete in issue
// order).
template <int T>
__device__ __forceinline__ void lds_sync() {
  if constexpr (T == 64) {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  }
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = kRingThreads / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

constexpr int kRingTile = 128;  // beads per workgroup of the per-bead ring kernels

// The potential's parameters are read from global memory at wave-uniform addresses (the
// EVB row i of a well, every dof j), which the compiler turns into scalar loads: the FMAs
// take them from SGPRs.  A copy staged in LDS made every lane read every element, and the
// LDS return bandwidth, not the FP64 pipes, then bounded the EVB evaluation.

// U_M and its adjoint gradient (UMforceenergy, instantonmod.f90:93-151):
// U = sum_i V(x_i) + sum_{i<n} (m/2bn^2)|x_{i+1}-x_i|^2 [+ (m/2bn^2)(|x_1-a|^2 + |b-x_n|^2)]
// Grid (bead tiles, chains), one bead per thread; each workgroup adds its partial sum to
// energy[c], which the host clears before the launch.
template <int KIND, int MAXD>
__global__ void __launch_bounds__(kRingTile) ring_um_kernel(RingArgs A, const double* __restrict__ x,
                                                            double* __restrict__ grad,
                                                            double* __restrict__ energy) {
  __shared__ double red[kRingTile];
  const int c = blockIdx.y;
  const int n = A.n[c], D = A.D;
  const int i = blockIdx.x * kRingTile + threadIdx.x;
  if (blockIdx.x * kRingTile >= n) return;  // the same for the whole workgroup
  const PesDev P = A.pes;
  const double bn2 = A.betan[c] * A.betan[c];
  const double* xc = x + A.off[c];
  const double* a = A.ab + (size_t)c * 2 * D;
  const double* b = a + D;
  double e = 0.0;
  if (i < n) {
    double xs[MAXD], gs[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) xs[d] = (d < D) ? xc[(size_t)d * n + i] : 0.0;
    e += eval_kind<KIND, MAXD>(P, xs, gs);
    // compile-time dof indices keep xs / gs in registers
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
      if (d >= D) continue;
      const double m = A.massd[d], xi = xs[d];
      const double* xd = xc + (size_t)d * n;
      if (i < n - 1) {
        const double dx = xd[i + 1] - xi;
        e += (0.5 * m / bn2) * dx * dx;
      }
      double sg;
      if (n == 1)
        sg = A.fixedends ? m * (2.0 * xi - a[d] - b[d]) / bn2 : 0.0;
      else if (i == 0)
        sg = A.fixedends ? m * (2.0 * xi - a[d] - xd[1]) / bn2 : m * (xi - xd[1]) / bn2;
      else if (i == n - 1)
        sg = A.fixedends ? m * (2.0 * xi - xd[n - 2] - b[d]) / bn2 : m * (xi - xd[n - 2]) / bn2;
      else
        sg = m * (2.0 * xi - xd[i - 1] - xd[i + 1]) / bn2;
      if (A.fixedends && i == 0) e += (0.5 * m / bn2) * (xi - a[d]) * (xi - a[d]);
      if (A.fixedends && i == n - 1) e += (0.5 * m / bn2) * (b[d] - xi) * (b[d] - xi);
      if (grad) grad[A.off[c] + (size_t)d * n + i] = sg + gs[d];
    }
  }
  // wave sums, then the waves' through LDS behind an LDS-only barrier (no wait for the
  // gradient stores)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  lds_sync<kRingTile>();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kRingTile / 64; ++w) t += red[w];
    atomicAdd(&energy[c], t);
  }
}

// U_M and its adjoint gradient for the EVB on the matrix cores.  A wave takes a tile of 16
// beads: both wells' coordinates y = Wt x - cw are one GEMM (rows: 2 x 16 RT well rows; K:
// the dofs) and the mixed gradient wa WtA^T hA + wb WtB^T hB a second GEMM whose B operand
// is the first one's accumulators, scaled in place (item j of row tile rt holds row
// 16 rt + g + 4 j of bead m, the operand of K-chunk 4 rt + j).  The spring terms are then
// added per (bead, dof) as in ring_um_kernel.  The sums run in another order than that
// kernel's per-bead loop, so the two agree to rounding, not bit for bit.
typedef double rd4 __attribute__((ext_vector_type(4)));
constexpr int kUmWaves = 4;

// Persistent: the grid is a few workgroups per CU, each staging the EVB block once and then
// taking bead tiles from the task list (tasks[2t] = chain, tasks[2t + 1] = first bead).
template <int M>
__global__ void __launch_bounds__(kUmWaves * 64) ring_um_evb_kernel(RingArgs A,
                                                                    const double* __restrict__ x,
                                                                    double* __restrict__ grad,
                                                                    double* __restrict__ energy,
                                                                    const int* __restrict__ tasks,
                                                                    int ntask) {
  constexpr int RT = (M + 15) / 16;  // row tiles per well, and dof tiles
  constexpr int DC = 16 * RT;        // padded dofs
  constexpr int LX = DC + 1;         // LDS stride of one bead
  constexpr int KY = 4 * RT;         // K-chunks of the y GEMM
  constexpr int KG = 8 * RT;         // K-chunks of the gradient GEMM
  constexpr int NT = kUmWaves * 64;
  constexpr int TB = kUmWaves * 16;  // beads per workgroup
  __shared__ double pl[evb_dev_size<M>()];
  __shared__ double xs[(TB + 2) * LX];  // x of beads b0-1 .. b0+TB (row 0 is bead b0-1)
  __shared__ double gp[TB * LX];        // the potential gradient of beads b0 .. b0+TB-1
  __shared__ double vpot[TB];
  __shared__ double red[NT];
  __shared__ double sm[3 * M];  // per dof: mass, end bead a, end bead b
  const int D = A.D;
  const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  {
    const double* __restrict__ dp = A.pes.dp;
    constexpr int NPL = (evb_dev_size<M>() + NT - 1) / NT;
    double vpl[NPL];
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int e = tid + q * NT;
      vpl[q] = e < evb_dev_size<M>() ? dp[e] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NPL; ++q)
      if (tid + q * NT < evb_dev_size<M>()) pl[tid + q * NT] = vpl[q];
  }
  for (int task = blockIdx.x; task < ntask; task += gridDim.x) {
    const int c = tasks[2 * task], b0 = tasks[2 * task + 1];
    const int n = A.n[c];
    const double* __restrict__ xc = x + A.off[c];
    {
      // every staging load of a thread is issued before the first one is used (compile-time
      // trip counts, registers in between): one memory latency for the whole staging, and it
      // overlaps the previous tile's last barrier
      const double* __restrict__ abc = A.ab + (size_t)c * 2 * D;
      constexpr int NXS = ((TB + 2) * DC + NT - 1) / NT;
      double vxs[NXS], vsm = 0.0;
#pragma unroll
      for (int q = 0; q < NXS; ++q) {
        const int e = tid + q * NT;
        // consecutive threads, consecutive beads
        const int d = e / (TB + 2), bl = e - d * (TB + 2);
        const int i = b0 - 1 + bl;
        vxs[q] = (d < D && i >= 0 && i < n) ? xc[(size_t)d * n + i] : 0.0;
      }
      if (tid < 3 * D) vsm = tid < D ? A.massd[tid] : abc[tid - D];
      lds_sync<NT>();  // the previous tile is done with xs, gp, vpot and sm
#pragma unroll
      for (int q = 0; q < NXS; ++q) {
        const int e = tid + q * NT;
        const int d = e / (TB + 2), bl = e - d * (TB + 2);
        if (e < (TB + 2) * DC) xs[bl * LX + d] = vxs[q];
      }
      if (tid < 3 * D) sm[tid] = vsm;
    }
    lds_sync<NT>();
    {
      const double kappa = pl[0];
      const double* WA = pl + 1;
      const double* cA = WA + M * M;
      const double* WB = cA + M;
      const double* cB = WB + M * M;
      const double* lam = cB + M;
      const double* cub = lam + M;
      const int tb = wave * 16;
      rd4 ay[2 * RT];
#pragma unroll
      for (int r = 0; r < 2 * RT; ++r) ay[r] = rd4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kc = 0; kc < KY; ++kc) {
        const int j = 4 * kc + g;
        const double bx = xs[(tb + m + 1) * LX + j];
#pragma unroll
        for (int r = 0; r < 2 * RT; ++r) {
          const int i = (r % RT) * 16 + m;
          const double* Wt = r < RT ? WA : WB;
          const double w = Wt[min(i, M - 1) * M + min(j, M - 1)];
          ay[r] = __builtin_amdgcn_mfma_f64_16x16x4f64((i < M && j < M) ? w : 0.0, bx, ay[r], 0,
                                                       0, 0);
        }
      }
      double va = 0.0, vb = 0.0;
#pragma unroll
      for (int r = 0; r < 2 * RT; ++r)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int i = (r % RT) * 16 + g + 4 * jj;
          const bool ok = i < M;
          const int ic = min(i, M - 1);
          const double cw = ok ? (r < RT ? cA[ic] : cB[ic]) : 0.0;
          const double li = ok ? lam[ic] : 0.0, ci = ok ? cub[ic] : 0.0;
          const double y = ay[r][jj] - cw, y2 = y * y;
          const double v = (0.5 * li) * y2 + ci * y2 * y;
          if (r < RT)
            va += v;
          else
            vb += v;
          ay[r][jj] = li * y + (3.0 * ci) * y2;
        }
      va += __shfl_xor(va, 16);
      va += __shfl_xor(va, 32);
      vb += __shfl_xor(vb, 16);
      vb += __shfl_xor(vb, 32);
      const double dl = va - vb, R = sqrt(0.25 * dl * dl + kappa * kappa), f = dl / (4.0 * R);
      const double wa = 0.5 - f, wb = 0.5 + f;
#pragma unroll
      for (int r = 0; r < 2 * RT; ++r) ay[r] *= (r < RT ? wa : wb);
      rd4 ag[RT];
#pragma unroll
      for (int d = 0; d < RT; ++d) ag[d] = rd4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kc = 0; kc < KG; ++kc) {
        const int r = kc / 4, jj = kc % 4;
        const int i = (r % RT) * 16 + 4 * jj + g;
        const double* Wt = r < RT ? WA : WB;
        const double bh = ay[r][jj];
#pragma unroll
        for (int d = 0; d < RT; ++d) {
          const int dof = d * 16 + m;
          const double w = Wt[min(i, M - 1) * M + min(dof, M - 1)];
          ag[d] = __builtin_amdgcn_mfma_f64_16x16x4f64((i < M && dof < M) ? w : 0.0, bh, ag[d],
                                                       0, 0, 0);
        }
      }
      if (g == 0) vpot[tb + m] = (0.5 * (va + vb) - R) - A.pes.v0;  // as eval_kind
#pragma unroll
      for (int d = 0; d < RT; ++d)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) gp[(tb + m) * LX + d * 16 + g + 4 * jj] = ag[d][jj];
    }
    lds_sync<NT>();
    // springs per (bead, dof), as in ring_um_kernel; the neighbours, masses and end beads come
    // from LDS
    const double bn2 = A.betan[c] * A.betan[c];
    const double* a = sm + D;
    const double* b = sm + 2 * D;
    double e = 0.0;
    double* __restrict__ gout = grad ? grad + A.off[c] : nullptr;  // loaded once, not per store
    const bool fixedends = A.fixedends;
    constexpr int NSP = (TB * M + NT - 1) / NT;
#pragma unroll
    for (int q = 0; q < NSP; ++q) {
      const int t = tid + q * NT;
      const int d = t / TB, bl = t - d * TB;
      const int i = b0 + bl;
      if (t >= TB * D || i >= n) continue;
      const double m_ = sm[d];
      const double xi = xs[(bl + 1) * LX + d];
      const double xl = xs[bl * LX + d];        // bead i - 1 (0 before the first bead)
      const double xr = xs[(bl + 2) * LX + d];  // bead i + 1 (0 after the last bead)
      if (d == 0) e += vpot[bl];
      if (i < n - 1) {
        const double dx = xr - xi;
        e += (0.5 * m_ / bn2) * dx * dx;
      }
      double sg;
      if (n == 1)
        sg = fixedends ? m_ * (2.0 * xi - a[d] - b[d]) / bn2 : 0.0;
      else if (i == 0)
        sg = fixedends ? m_ * (2.0 * xi - a[d] - xr) / bn2 : m_ * (xi - xr) / bn2;
      else if (i == n - 1)
        sg = fixedends ? m_ * (2.0 * xi - xl - b[d]) / bn2 : m_ * (xi - xl) / bn2;
      else
        sg = m_ * (2.0 * xi - xl - xr) / bn2;
      if (fixedends && i == 0) e += (0.5 * m_ / bn2) * (xi - a[d]) * (xi - a[d]);
      if (fixedends && i == n - 1) e += (0.5 * m_ / bn2) * (b[d] - xi) * (b[d] - xi);
      if (gout) gout[(size_t)d * n + i] = sg + gp[bl * LX + d];
    }
    // the energy: a wave sum, then the waves' sums through LDS behind an LDS-only barrier (a
    // full barrier would first wait for every gradient store above to be acknowledged)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
    if (lane == 0) red[wave] = e;
    lds_sync<NT>();
    if (tid == 0) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < kUmWaves; ++w) t += red[w];
      atomicAdd(&energy[c], t);
    }
  }  // tasks
}

// Hessian-vector product of U_M: springs exactly, potential by forward-over-reverse.
template <int KIND, int MAXD>
__global__ void __launch_bounds__(kRingTile) ring_hvp_kernel(RingArgs A, const double* __restrict__ x,
                                                             const double* __restrict__ v,
                                                             double* __restrict__ hv) {
  const int c = blockIdx.y;
  const int n = A.n[c], D = A.D;
  const int i = blockIdx.x * kRingTile + threadIdx.x;
  if (blockIdx.x * kRingTile >= n) return;  // the same for the whole workgroup
  const PesDev P = A.pes;
  const double bn2 = A.betan[c] * A.betan[c];
  const int64_t o = A.off[c];
  if (i >= n) return;
  Dual xs[MAXD], gs[MAXD];
#pragma unroll
  for (int d = 0; d < MAXD; ++d)
    xs[d] = (d < D) ? Dual(x[o + (size_t)d * n + i], v[o + (size_t)d * n + i]) : Dual(0.0);
  (void)eval_kind<KIND, MAXD>(P, xs, gs);
#pragma unroll
  for (int d = 0; d < MAXD; ++d) {
    if (d >= D) continue;
    const double m = A.massd[d];
    const double* vd = v + o + (size_t)d * n;
    double s;
    if (i == 0)
      s = A.fixedends ? 2.0 * vd[0] - (n > 1 ? vd[1] : 0.0) : vd[0] - (n > 1 ? vd[1] : 0.0);
    else if (i == n - 1)
      s = A.fixedends ? 2.0 * vd[i] - vd[i - 1] : vd[i] - vd[i - 1];
    else
      s = 2.0 * vd[i] - vd[i - 1] - vd[i + 1];
    hv[o + (size_t)d * n + i] = m * s / bn2 + gs[d].d;
  }
}

// Potential Hessian blocks of every bead from D forward-over-reverse sweeps:
// blocks[gbead][i][k] = d grad_k / d x_i.  Thread (bead, i).
template <int KIND, int MAXD>
__global__ void __launch_bounds__(128) ring_hblock_kernel(RingArgs A, const double* __restrict__ x,
                                                          int single_first,
                                                          double* __restrict__ blocks) {
  const int c = blockIdx.y;
  const int n = A.n[c], D = A.D;
  if ((int)(blockIdx.x * blockDim.x) >= n * D) return;  // the same for the whole workgroup
  const PesDev P = A.pes;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int bead = t / D, i = t - bead * D;
  if (bead >= n) return;
  const int src = single_first ? 0 : bead;  // singlewell: bead 1's Hessian everywhere
  const int64_t o = A.off[c];
  Dual xs[MAXD], gs[MAXD];
#pragma unroll
  for (int d = 0; d < MAXD; ++d)
    xs[d] = (d < D) ? Dual(x[o + (size_t)d * n + src], d == i ? 1.0 : 0.0) : Dual(0.0);
  (void)eval_kind<KIND, MAXD>(P, xs, gs);
  double* blk = blocks + ((size_t)(A.boff[c] + bead) * D + i) * D;
#pragma unroll
  for (int k = 0; k < MAXD; ++k)
    if (k < D) blk[k] = gs[k].d;
}

// U_M and its gradient for a potential that lives on the host: the bead energies pe[gbead]
// and the potential's gradients (already in grad) come from the plugin; this adds the
// springs exactly as ring_um_kernel does and sums each chain's energy.
__global__ void __launch_bounds__(kRingTile) ring_spring_kernel(RingArgs A, const double* __restrict__ x,
                                                                const double* __restrict__ pe,
                                                                double* __restrict__ grad,
                                                                double* __restrict__ energy) {
  __shared__ double red[kRingTile];
  const int c = blockIdx.y;
  const int n = A.n[c], D = A.D;
  const int i = blockIdx.x * kRingTile + threadIdx.x;
  if (blockIdx.x * kRingTile >= n) return;  // the same for the whole workgroup
  const double bn2 = A.betan[c] * A.betan[c];
  const double* xc = x + A.off[c];
  const double* a = A.ab + (size_t)c * 2 * D;
  const double* b = a + D;
  double e = 0.0;
  if (i < n) {
    e = pe[A.boff[c] + i];
    for (int d = 0; d < D; ++d) {
      const double m = A.massd[d];
      const double* xd = xc + (size_t)d * n;
      const double xi = xd[i];
      if (i < n - 1) e += (0.5 * m / bn2) * (xd[i + 1] - xi) * (xd[i + 1] - xi);
      double sg;
      if (i == 0)
        sg = A.fixedends ? m * (2.0 * xi - a[d] - xd[1]) / bn2 : m * (xi - xd[1]) / bn2;
      else if (i == n - 1)
        sg = A.fixedends ? m * (2.0 * xi - xd[n - 2] - b[d]) / bn2 : m * (xi - xd[n - 2]) / bn2;
      else
        sg = m * (2.0 * xi - xd[i - 1] - xd[i + 1]) / bn2;
      if (A.fixedends && i == 0) e += (0.5 * m / bn2) * (xi - a[d]) * (xi - a[d]);
      if (A.fixedends && i == n - 1) e += (0.5 * m / bn2) * (b[d] - xi) * (b[d] - xi);
      if (grad) grad[A.off[c] + (size_t)d * n + i] += sg;
    }
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int st = kRingTile / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&energy[c], red[0]);
}

// Hessian-vector product of U_M from stored potential Hessian blocks (a potential that lives
// on the host): the springs of ring_hvp_kernel plus sum_i blocks[gbead][i][d] v_i.
__global__ void __launch_bounds__(kRingTile) ring_hvp_blocks_kernel(RingArgs A,
                                                                    const double* __restrict__ blocks,
                                                                    const double* __restrict__ v,
                                                                    double* __restrict__ hv) {
  const int c = blockIdx.y;
  const int n = A.n[c], D = A.D;
  const int i = blockIdx.x * kRingTile + threadIdx.x;
  if (i >= n) return;
  const double bn2 = A.betan[c] * A.betan[c];
  const int64_t o = A.off[c];
  const double* blk = blocks + (size_t)(A.boff[c] + i) * D * D;
  for (int d = 0; d < D; ++d) {
    const double* vd = v + o + (size_t)d * n;
    double s;
    if (i == 0)
      s = A.fixedends ? 2.0 * vd[0] - (n > 1 ? vd[1] : 0.0) : vd[0] - (n > 1 ? vd[1] : 0.0);
    else if (i == n - 1)
      s = A.fixedends ? 2.0 * vd[i] - vd[i - 1] : vd[i] - vd[i - 1];
    else
      s = 2.0 * vd[i] - vd[i - 1] - vd[i + 1];
    double pot = 0.0;
    for (int k = 0; k < D; ++k) pot += blk[(size_t)k * D + d] * v[o + (size_t)k * n + i];
    hv[o + (size_t)d * n + i] = A.massd[d] * s / bn2 + pot;
  }
}

// UMhessian's band (instantonmod.f90:155-217) from the blocks of one chain, in the
// reference's array answer(ndof+1, n*ndof) (LAPACK lower band storage, column-major):
// band[col*(D+1) + r] = A(col + r, col), r = 0..D.  The inter-bead coupling
// -1/bn^2 is written only for beads i > 1 (1-based), as the reference does (:203).
__global__ void ring_band_kernel(int n, int D, double bn2, const double* __restrict__ massd,
                                 const double* __restrict__ blocks, double* __restrict__ band) {
  const int64_t tot = (int64_t)n * D;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (D + 1) * tot) return;
  const int64_t col = t / (D + 1);
  const int r = (int)(t - col * (D + 1));
  const int bead = (int)(col / D), i1 = (int)(col - (int64_t)bead * D);
  double v = 0.0;
  if (r == 0) {
    v = 2.0 / bn2 + blocks[((size_t)bead * D + i1) * D + i1] / sqrt(massd[i1] * massd[i1]);
  } else if (r == D) {
    if (bead > 0) v = -1.0 / bn2;  // beyond the matrix for the last bead; never read
  } else if (i1 + r < D) {
    const int i2 = i1 + r;
    // hess(j2,k2,j1,k1) = d grad(idof1) / d x(idof2)
    v = blocks[((size_t)bead * D + i2) * D + i1] / sqrt(massd[i1] * massd[i2]);
  }
  band[t] = v;
}

// dU_M/dtheta = sum over beads of dV/dtheta (water: theta = qH, De), one block per chain.
template <int NMOL>
__global__ void __launch_bounds__(kRingThreads) ring_param_grad_kernel(RingArgs A,
                                                                       const double* __restrict__ x,
                                                                       double* __restrict__ out) {
  __shared__ double red[kRingThreads];
  const int c = blockIdx.x, n = A.n[c];
  const double* xc = x + A.off[c];
  double p0 = 0.0, p1 = 0.0;
  for (int i = threadIdx.x; i < n; i += kRingThreads) {
    double xs[9 * NMOL], gs[9 * NMOL], pg[2];
#pragma unroll
    for (int d = 0; d < 9 * NMOL; ++d) xs[d] = xc[(size_t)d * n + i];
    (void)water_eval<NMOL>(A.pes, xs, gs, pg);
    p0 += pg[0];
    p1 += pg[1];
  }
  p0 = block_sum(p0, red);
  p1 = block_sum(p1, red);
  if (threadIdx.x == 0) {
    out[2 * c] = p0;
    out[2 * c + 1] = p1;
  }
}

}  // namespace

// ---------------------------------------------------------------- log-determinant
// Block LDL^T of the mass-weighted ring Hessian (the reference's band matrix):
//   A_i = (2/bn^2) I + M^-1/2 H_i M^-1/2,   coupling -(1/bn^2) I between beads (i, i+1)
//   for i >= 1 (0-based; bead 0 is not coupled to bead 1, instantonmod.f90:203).
//   S_0 = A_0,  S_i = A_i - (1/bn^4) S_{i-1}^-1 when (i-1, i) is coupled, else A_i.
// log|det| = sum log|det S_i|, inertia = sum inertia(S_i) (Haynsworth).  The S_i^-1 are
// kept for inverse iteration on the eigenvalue nearest zero.
namespace {
constexpr int kMaxD = 56;  // ndof bound of the log-determinant workspace (LDS)
constexpr int kPinP = 3;   // elements per thread whose step operands are read up front

// 1/x of a pivot: the hardware reciprocal estimate (about 26 correct bits) refined by two
// Newton-Raphson steps r <- r + r (1 - x r), each doubling the correct bits.  The full
// division adds scaling for denormal and near-overflow operands, which the pivots of these
// blocks are not, and its extra steps lie on the elimination's critical path.
__device__ __forceinline__ double pivot_recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}

// One wave per chain.  Lane l owns the elements e = l + 64 q (q < P) of every D x D block, so
// the Hessian block of the next bead is loaded into registers while the current one is
// eliminated, and the Schur update uses the previous inverse where it sits in LDS.  The
// Gauss-Jordan inverse runs without pivoting, double-buffered in LDS (one barrier per pair
// of eliminations): the pivot of elimination k is the k-th pivot of the matrix's LDL^T
// factorisation (a ratio of leading principal minors), so their logs sum to log|det| and
// their signs give the inertia (Sylvester).
// T threads per chain.  Thread t owns the elements e = t + T q (q < P) of every D x D block,
// so the Hessian block of the next bead is loaded into registers while the current one is
// eliminated, and the Schur update uses the previous inverse where it sits in LDS.  The
// Gauss-Jordan inverse runs without pivoting, double-buffered in LDS (one barrier per pair
// of eliminations): the pivot of elimination k is the k-th pivot of the matrix's LDL^T
// factorisation (a ratio of leading principal minors), so their logs sum to log|det| and
// their signs give the inertia (Sylvester).
//
// SYM: the lower triangle only.  The Gauss-Jordan iterates keep M[j][i] = s M[i][j], with
// s = -1 exactly when one of i, j has been swept (a pivot already eliminated) and +1
// otherwise, so thread t owns lower element t (r >= cc, row-major) and writes both it and its
// mirror: P elements per thread cover D (D + 1) / 2 instead of D^2, the reads are the same,
// and the LDS tile (PT elements per thread) still holds the whole matrix.  Threads beyond the
// triangle own the pad element (D, D).
template <int P, int T, bool SYM = false, int PT = P>
__global__ void __launch_bounds__(T) ring_logdet_kernel(
    int D, const int* __restrict__ nn, const int64_t* __restrict__ boff,
    const double* __restrict__ betan, const double* __restrict__ massd,
    const double* __restrict__ blocks, const double* __restrict__ shift, double* __restrict__ sinv,
    double* __restrict__ logdet, int* __restrict__ nneg, const int* __restrict__ chain_of) {
  // every LDS element t + T q exists, and so does every r D + k <= t + T q + D - 1 read for
  // it (and k D + cc < D^2 <= T P); the elements beyond D*D hold values nobody reads back,
  // which keeps every load and store unconditional (no branch, so the loads of the next
  // block stay in flight).  Sized by T P, so several workgroups share a CU's LDS.
  constexpr int kTile = T * PT + kMaxD;
  __shared__ double Sb[2][kTile];
  // one task per workgroup: chain c (task t of a chain_of map, else chain t) factored at
  // the shift of task t; outputs per task
  const int t = blockIdx.x, c = chain_of ? chain_of[t] : t;
  const int n = nn[c], tid = threadIdx.x, DD = D * D;
  const double bn2 = betan[c] * betan[c], c2 = 1.0 / (bn2 * bn2);
  const double sig = shift ? shift[t] : 0.0;  // factor H - sig I
  int er[P], ec[P], eg[P], ei[P], mi[P];  // row, column, global index, LDS index, mirror
  double wgt[P], diag[P], nxt[P];
#pragma unroll
  for (int q = 0; q < P; ++q) {
    const int e = tid + T * q;
    if constexpr (SYM) {
      int r = 0;
      while ((r + 1) * (r + 2) / 2 <= e && r < D) ++r;  // row of lower element e
      const bool pad = r >= D;
      er[q] = pad ? D : r;
      ec[q] = pad ? D : e - r * (r + 1) / 2;
    } else {
      er[q] = e / D;
      ec[q] = e - er[q] * D;
    }
    ei[q] = SYM ? er[q] * D + ec[q] : e;
    mi[q] = ec[q] * D + er[q];
    eg[q] = min(ei[q], DD - 1);  // clamped index of the global block loads
    const int rm = min(er[q], D - 1), cm = min(ec[q], D - 1);
    wgt[q] = 1.0 / sqrt(massd[rm] * massd[cm]);
    diag[q] = er[q] == ec[q] ? 2.0 / bn2 - sig : 0.0;
  }
  const double* base = blocks + (size_t)boff[c] * DD;
#pragma unroll
  for (int q = 0; q < P; ++q) nxt[q] = base[eg[q]];
  double lg_tot = 0.0;
  int neg_tot = 0, cur = 0;
  for (int i = 0; i < n; ++i) {
    // A_i(r, cc) = 2/bn^2 delta + H(cc -> r) / sqrt(m_r m_cc), minus c2 S_{i-1}^-1 when the
    // pair (i-1, i) is coupled (i >= 2; bead 0 is not coupled to bead 1)
    const bool coupled = i >= 2;
    double* M = Sb[cur];
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int e = ei[q];
      const double v = nxt[q] * wgt[q] + diag[q];
      const double prev = M[e];
      const double val = coupled ? v - c2 * prev : v;
      M[e] = val;
      if constexpr (SYM) M[mi[q]] = val;  // A_i and S_{i-1}^-1 are symmetric
    }
    // the next bead's block (a valid address for the last bead too), in flight during the
    // elimination
    const double* nb = base + (size_t)min(i + 1, n - 1) * DD;
#pragma unroll
    for (int q = 0; q < P; ++q) nxt[q] = nb[eg[q]];
    lds_sync<T>();
    // eliminations two pivots at a time (a 2 x 2 pivot block, inverted explicitly, is two
    // successive Gauss-Jordan steps), so a bead takes ceil(D/2) barriers; the pivots of the
    // pair are a and det/a, the same two LDL^T pivots as one at a time
    // thread k < D keeps pivot k as pnum / pden (the second pivot of a pair is det / a), divided
    // once per bead instead of in the step, where one thread's division holds up the barrier
    double pnum = 1.0, pden = 1.0;
    for (int k = 0; k < D; k += 2) {
      const double* A = Sb[cur];
      double* B = Sb[cur ^ 1];
      if (k + 1 < D) {
        const int l = k + 1;
        // every element's operands are read before the pivot block's inverse, and the empty
        // asm below needs them all: otherwise the compiler moves each read into the branch
        // that uses it, and a step waits on one LDS round trip per element and operand.
        // Pinned up to kPinP elements per thread: beyond that the held operands cost
        // occupancy (P = 6 would go from 3 to 2 waves per SIMD)
        double u0s[P], u1s[P], w0s[P], w1s[P], aes[P];
        auto operands = [&](int q) {
          const int r = er[q], cc = ec[q];
          u0s[q] = A[r * D + k];
          u1s[q] = A[r * D + l];
          w0s[q] = A[k * D + cc];
          w1s[q] = A[l * D + cc];
          aes[q] = A[ei[q]];
        };
        if constexpr (P <= kPinP) {
#pragma unroll
          for (int q = 0; q < P; ++q) operands(q);
        }
        const double a = A[k * D + k], b = A[k * D + l], cq = A[l * D + k], d = A[l * D + l];
        const double det = fma(a, d, -b * cq), id = pivot_recip(det);
        const double p00 = d * id, p01 = -b * id, p10 = -cq * id, p11 = a * id;  // block^-1
        if (tid < 64) {  // pivot lanes are in wave 0 (D <= kMaxD < 64)
          pnum = tid == k ? a : tid == l ? det : pnum;
          pden = tid == k ? 1.0 : tid == l ? a : pden;
        }
        if constexpr (P <= kPinP) {
#pragma unroll
          for (int q = 0; q < P; ++q)
            asm volatile("" ::"v"(u0s[q]), "v"(u1s[q]), "v"(w0s[q]), "v"(w1s[q]), "v"(aes[q]));
        }
#pragma unroll
        for (int q = 0; q < P; ++q) {
          if constexpr (P > kPinP) operands(q);
          const int e = ei[q], r = er[q], cc = ec[q];
          // every case is a' - u~ P w~ with P = block^-1: a pivot row (r in K) takes
          // u~ = -e_r (row r of P A[K][cc]), a pivot column w~ = e_cc (-A[r][K] P, or P
          // itself for both), and a' = 0 for either; elsewhere u~, w~, a' are the element's
          // row, column and value, A[r][cc] - A[r][K] P A[K][cc]
          const bool rk = (r >> 1) == (k >> 1), ck = (cc >> 1) == (k >> 1);  // k even
          const double u0 = rk ? ((r & 1) ? 0.0 : -1.0) : u0s[q];
          const double u1 = rk ? ((r & 1) ? -1.0 : 0.0) : u1s[q];
          const double w0 = ck ? ((cc & 1) ? 0.0 : 1.0) : w0s[q];
          const double w1 = ck ? ((cc & 1) ? 1.0 : 0.0) : w1s[q];
          const double ae = (rk || ck) ? 0.0 : aes[q];
          const double pw0 = p00 * w0 + p01 * w1, pw1 = p10 * w0 + p11 * w1;  // P A[K][cc]
          const double v = ae - (u0 * pw0 + u1 * pw1);
          B[e] = v;
          if constexpr (SYM) B[mi[q]] = (r <= l) != (cc <= l) ? -v : v;  // swept: 0..l
        }
      } else {
        double mrks[P], mkcs[P], aes[P];  // read first, as in the pair step
        auto operands = [&](int q) {
          mrks[q] = A[er[q] * D + k];
          mkcs[q] = A[k * D + ec[q]];
          aes[q] = A[ei[q]];
        };
        if constexpr (P <= kPinP) {
#pragma unroll
          for (int q = 0; q < P; ++q) operands(q);
        }
        const double p = A[k * D + k], ip = pivot_recip(p);
        if (tid < 64) {
          pnum = tid == k ? p : pnum;
          pden = tid == k ? 1.0 : pden;
        }
        if constexpr (P <= kPinP) {
#pragma unroll
          for (int q = 0; q < P; ++q) asm volatile("" ::"v"(mrks[q]), "v"(mkcs[q]), "v"(aes[q]));
        }
#pragma unroll
        for (int q = 0; q < P; ++q) {
          if constexpr (P > kPinP) operands(q);
          const int e = ei[q], r = er[q], cc = ec[q];
          const double mrk = mrks[q], mkc = mkcs[q], a = aes[q];
          const double vr = cc == k ? ip : mkc * ip;
          const double vo = cc == k ? -mrk * ip : a - mrk * mkc * ip;
          const double v = r == k ? vr : vo;
          B[e] = v;
          if constexpr (SYM) B[mi[q]] = (r <= k) != (cc <= k) ? -v : v;  // swept: 0..k
        }
      }
      cur ^= 1;
      lds_sync<T>();
    }
    // pivots: log|p| and the sign count; threads k < D <= 56 of wave 0 hold pivot k
    if (tid < 64) {
      const double mypiv = pnum / pden;
      double l = tid < D ? log(fabs(mypiv)) : 0.0;
      int ng = (tid < D && mypiv < 0.0) ? 1 : 0;
      for (int o = 32; o > 0; o >>= 1) {
        l += __shfl_xor(l, o);
        ng += __shfl_xor(ng, o);
      }
      lg_tot += l;
      neg_tot += ng;
    }
    if (sinv) {  // kept for inverse iteration; inertia-only calls pass null
      double* out = sinv + (size_t)(boff[c] + i) * DD;
      const double* Minv = Sb[cur];
      // each thread stores the elements it overwrites first in the next bead (no barrier
      // between): under SYM its lower element and the mirror, else its own elements
      if constexpr (SYM) {
#pragma unroll
        for (int q = 0; q < P; ++q)
          if (ei[q] < DD) {
            out[ei[q]] = Minv[ei[q]];
            out[mi[q]] = Minv[mi[q]];
          }
      } else {
#pragma unroll
        for (int q = 0; q < P; ++q)
          if (tid + T * q < DD) out[tid + T * q] = Minv[tid + T * q];
      }
    }
  }
  if (tid == 0) {
    logdet[t] = lg_tot;
    nneg[t] = neg_tot;
  }
}

// Gershgorin bounds of every chain's mass-weighted ring Hessian: over rows (bead i, dof r),
// the minimum and maximum of A_i(r, r) -/+ [sum_{c != r} |A_i(r, c)| + (coupled neighbours of
// bead i)/bn^2].  lohi[2c] = lower, lohi[2c+1] = upper.
__global__ void __launch_bounds__(kRingThreads) ring_gersh_kernel(
    int D, const int* __restrict__ nn, const int64_t* __restrict__ boff,
    const double* __restrict__ betan, const double* __restrict__ massd,
    const double* __restrict__ blocks, double* __restrict__ lohi) {
  __shared__ double rlo[kRingThreads], rhi[kRingThreads];
  const int c = blockIdx.x, n = nn[c];
  const double ibn2 = 1.0 / (betan[c] * betan[c]);
  const double* base = blocks + (size_t)boff[c] * D * D;
  double vlo = INFINITY, vhi = -INFINITY;
  for (int64_t e = threadIdx.x; e < (int64_t)n * D; e += kRingThreads) {
    const int i = (int)(e / D), r = (int)(e - (int64_t)i * D);
    const double* row = base + ((size_t)i * D + r) * D;
    double rad = 0.0;
    for (int k = 0; k < D; ++k)
      if (k != r) rad += fabs(row[k]) / sqrt(massd[r] * massd[k]);
    const int nb = (i >= 2 ? 1 : 0) + (i >= 1 && i + 1 < n ? 1 : 0);
    rad += nb * ibn2;
    const double ctr = 2.0 * ibn2 + row[r] / massd[r];
    vlo = fmin(vlo, ctr - rad);
    vhi = fmax(vhi, ctr + rad);
  }
  rlo[threadIdx.x] = vlo;
  rhi[threadIdx.x] = vhi;
  __syncthreads();
  for (int s = kRingThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      rlo[threadIdx.x] = fmin(rlo[threadIdx.x], rlo[threadIdx.x + s]);
      rhi[threadIdx.x] = fmax(rhi[threadIdx.x], rhi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    lohi[2 * c] = rlo[0];
    lohi[2 * c + 1] = rhi[0];
  }
}

// threads and per-thread element count of ring_logdet_kernel for D (D*D <= T P)
template <typename F>
int with_logdet_kernel(int D, F&& f) {
  if (D * D <= 64) return f(ring_logdet_kernel<1, 64>, 64);
  if (D * D <= 256) return f(ring_logdet_kernel<1, 256>, 256);
  if (D * D <= 768) return f(ring_logdet_kernel<2, 256, true, 3>, 256);  // D (D + 1) / 2 <= 512
  if (D * D <= 1536) return f(ring_logdet_kernel<4, 256, true, 6>, 256);  // D (D + 1) / 2 <= 1024
  return f(ring_logdet_kernel<7, 256, true, 13>, 256);  // D (D + 1) / 2 <= 1792
}

// The block solve y = H^-1 v through the stored factors S_i^-1 (the off-diagonal blocks are
// -c I and bead 0 is coupled to no other bead), in two sweeps whose serial path never waits
// for global memory:
//   forward  (m = 0 .. n-1): t_m = S_m^-1 z_m, with z_0 = v_0, z_1 = v_1 and
//                            z_{m+1} = v_{m+1} + c t_m for m >= 1;
//   backward (i = n-2 .. 1): y_i = t_i + c S_i^-1 y_{i+1}, with y_{n-1} = t_{n-1}, y_0 = t_0.
// The running vector (z or y, D values) sits in LDS, double-buffered; what a step reads from
// global memory (its block, and v_{m+1} or t_i) does not depend on it, so it is loaded P steps
// ahead into registers, and the per-step barrier orders LDS only, which leaves those loads in
// flight across it.  Lanes: Q per block row (row r = tid / Q, lane q), each holding the KPL
// columns k = q + Q j of its row; a row's sum is a DPP butterfly inside its Q lanes.  The LDS
// vector is zero beyond D, so clamped column loads add nothing.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  return __builtin_amdgcn_update_dpp(0.0, v, CTRL, 0xF, 0xF, false);
}

// the sum over each aligned group of Q (2, 4 or 8) lanes, in every lane of the group
template <int Q>
__device__ __forceinline__ double row_sum(double s) {
  s += dpp_mov<0xB1>(s);                         // quad_perm [1,0,3,2]: lane ^ 1
  if constexpr (Q >= 4) s += dpp_mov<0x4E>(s);   // quad_perm [2,3,0,1]: lane ^ 2
  if constexpr (Q >= 8) s += dpp_mov<0x141>(s);  // row_half_mirror: lane 7 - l of 8
  return s;
}

// A raw buffer over [p, p + bytes): its loads take the lane's byte offset in a VGPR and a
// uniform byte offset in an SGPR, so a step's loads need no 64-bit address per lane.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_buffer(const double* p, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double buffer_ld(__amdgpu_buffer_rsrc_t b, int voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(b, voff, soff, 0));
}

// One sweep; returns the LDS buffer that holds its last vector (t_{n-1} after the forward
// sweep, which is y_{n-1}).  src: v (forward) or the t the forward sweep left in dst
// (backward).  Byte offsets within a chain fit 31 bits (the host checks n D^2 8 < 2^31).
template <int Q, int KPL, int P, bool FWD>
__device__ __forceinline__ int solve_sweep(int D, int n, double c, const double* __restrict__ S,
                                           const double* src, double* dst, double* zb, int cur) {
  constexpr int QK = Q * KPL;
  const int tid = threadIdx.x, r = tid / Q, q = tid - r * Q;
  const int rr = min(r, D - 1);
  const bool wr = q == 0 && r < D;
  const int K = FWD ? n : n - 2, DD8 = D * D * 8, D8 = D * 8;
  const __amdgpu_buffer_rsrc_t Sb = raw_buffer(S, (int64_t)n * DD8);
  const __amdgpu_buffer_rsrc_t Eb = raw_buffer(src, (int64_t)n * D8);
  int ko[KPL];
#pragma unroll
  for (int j = 0; j < KPL; ++j) ko[j] = (rr * D + min(q + Q * j, D - 1)) * 8;
  const int eo = rr * 8;
  // bead of step k, and the bead of the row value it reads (clamped past the end)
  auto bead = [&](int k) { return FWD ? min(k, n - 1) : max(n - 2 - k, 0); };
  auto ebead = [&](int k) { return FWD ? min(k + 1, n - 1) : max(n - 2 - k, 0); };
  double a[P][KPL], e[P];
#pragma unroll
  for (int s = 0; s < P; ++s) {
    const int so = bead(s) * DD8;
#pragma unroll
    for (int j = 0; j < KPL; ++j) a[s][j] = buffer_ld(Sb, ko[j], so);
    e[s] = buffer_ld(Eb, eo, ebead(s) * D8);
  }
  for (int k0 = 0; k0 < K; k0 += P) {
#pragma unroll
    for (int s = 0; s < P; ++s) {
      const int k = k0 + s;
      if (k < K) {  // block-uniform
        const double* u = zb + cur * QK;
        double t0 = 0.0, t1 = 0.0;
#pragma unroll
        for (int j = 0; j < KPL; j += 2) {
          t0 += a[s][j] * u[q + Q * j];
          if (j + 1 < KPL) t1 += a[s][j + 1] * u[q + Q * (j + 1)];
        }
        const double t = row_sum<Q>(t0 + t1), ev = e[s];
        const int so = bead(k + P) * DD8;
#pragma unroll
        for (int j = 0; j < KPL; ++j) a[s][j] = buffer_ld(Sb, ko[j], so);
        e[s] = buffer_ld(Eb, eo, ebead(k + P) * D8);
        double* nz = zb + (cur ^ 1) * QK;
        if (FWD) {
          if (wr) {
            dst[(int64_t)k * D + r] = t;
            nz[r] = k + 1 < n ? (k >= 1 ? ev + c * t : ev) : t;
          }
        } else {
          const double y = ev + c * t;
          if (wr) {
            dst[(int64_t)(n - 2 - k) * D + r] = y;
            nz[r] = y;
          }
        }
        cur ^= 1;
        lds_sync<kRingThreads>();
      }
    }
  }
  return cur;
}

// Inverse iteration for the eigenvalue of H nearest the factorisation shift sig:
// v <- (H - sig)^-1 v / |.|, eta = sig + 1 / (v . (H - sig)^-1 v) for the normalised v.
// work holds [v | y] per chain (2 n D); fresh = 1 starts from a fixed generic vector,
// otherwise the previous v is continued.  Stops after iters iterations, or earlier once eta
// changes by at most tol relative; chains with act[c] == 0 are left as they are.
// D <= Q KPL.
template <int Q, int KPL, int P>
__global__ void __launch_bounds__(kRingThreads) ring_eta1_kernel(
    int D, const int* __restrict__ nn, const int64_t* __restrict__ boff,
    const double* __restrict__ betan, const double* __restrict__ sinv, double* __restrict__ work,
    const double* __restrict__ shift, int iters, int fresh, double tol,
    const int* __restrict__ act, double* __restrict__ eta1) {
  static_assert(Q * (kRingThreads / Q) == kRingThreads && (Q == 2 || Q == 4 || Q == 8), "lanes per row");
  __shared__ double red[kRingThreads];
  __shared__ double zb[2 * Q * KPL];
  const int cix = blockIdx.x, n = nn[cix], tid = threadIdx.x;
  if (act && !act[cix]) return;  // block-uniform
  const int64_t tot = (int64_t)n * D;
  const double c = 1.0 / (betan[cix] * betan[cix]);
  double* v = work + 2 * boff[cix] * D;
  double* y = v + tot;
  const double* S = sinv + (size_t)boff[cix] * D * D;
  const double sig = shift ? shift[cix] : 0.0;
  if (tid < 2 * Q * KPL) zb[tid] = 0.0;
  double nrm = 0.0;
  for (int64_t e = tid; e < tot; e += blockDim.x) {
    const double t = fresh ? 1.0 + 0.37 * (double)((e * 7919) % 101) / 101.0 : v[e];
    v[e] = t;
    nrm += t * t;
  }
  nrm = sqrt(block_sum(nrm, red));
  for (int64_t e = tid; e < tot; e += blockDim.x) v[e] /= nrm;
  __syncthreads();
  double eta = 0.0, prev = 0.0;
  for (int it = 0; it < iters; ++it) {
    if (tid < D) zb[tid] = v[tid];  // z_0 = v_0 in buffer 0
    __syncthreads();
    int cur = solve_sweep<Q, KPL, P, true>(D, n, c, S, v, y, zb, 0);
    __syncthreads();  // the t of every lane, in y, before the backward sweep reads them
    solve_sweep<Q, KPL, P, false>(D, n, c, S, y, y, zb, cur);
    __syncthreads();
    double mu = 0.0, yy = 0.0;
    for (int64_t e = tid; e < tot; e += blockDim.x) {
      mu += v[e] * y[e];
      yy += y[e] * y[e];
    }
    mu = block_sum(mu, red);
    yy = sqrt(block_sum(yy, red));
    for (int64_t e = tid; e < tot; e += blockDim.x) v[e] = y[e] / yy;
    __syncthreads();
    eta = sig + 1.0 / mu;
    if (it > 2 && fabs(eta - prev) <= tol * fabs(eta)) break;  // block-uniform
    prev = eta;
  }
  if (tid == 0) eta1[cix] = eta;
}

}  // namespace

// ---------------------------------------------------------------- host
struct rp_ring {
  rp_pes* pes = nullptr;
  int nchain = 0, D = 0, ndim = 0, natom = 0, fixedends = 1;
  std::vector<int> n;
  std::vector<double> betan;
  std::vector<int64_t> off, boff;
  int64_t tot = 0, nbeads = 0;
  int nmax = 0;  // the longest chain
  hipStream_t s = nullptr;
  DevBuf<double> d_x, d_g, d_v, d_hv, d_e, d_ab, d_betan, d_massd, d_blocks, d_sinv, d_work, d_ld,
      d_eta;
  DevBuf<int> d_n, d_nneg;
  DevBuf<int64_t> d_off, d_boff;
  // a potential that lives on the host: bead energies on the device, host staging
  DevBuf<double> d_pe;
  std::vector<double> hx, hg, he;
  // the instanton search's optional preconditioner (rp_ring_set_preconditioner)
  std::vector<double> hmass;  // per atom
  double pc_sigma = 0.0;
  DevBuf<double> d_pctab;
  DevBuf<int64_t> d_pcoff;
  // the EVB U_M kernel's bead tiles (chain, first bead) and its persistent grid
  DevBuf<int> d_umtask;
  int n_umtask = 0, um_grid = 0;
};

static RingArgs ring_args(rp_ring* r) {
  RingArgs a;
  a.pes = r->pes->dev();
  a.nchain = r->nchain;
  a.D = r->D;
  a.ndim = r->ndim;
  a.natom = r->natom;
  a.fixedends = r->fixedends;
  a.n = r->d_n.p;
  a.off = r->d_off.p;
  a.boff = r->d_boff.p;
  a.betan = r->d_betan.p;
  a.ab = r->d_ab.p;
  a.massd = r->d_massd.p;
  return a;
}

namespace rp {
// The potential of every bead from a host plugin: one copy of the chains to the host, the
// plugin's potforce per chain, the gradients and bead energies back (synchronous).
static int ring_host_potential(rp_ring* r, bool want_grad) {
  int rc;
  if (r->d_pe.n < (size_t)r->nbeads && (rc = r->d_pe.alloc(r->nbeads))) return rc;
  r->hx.resize(r->tot);
  r->hg.resize(r->tot);
  r->he.resize(r->nbeads);
  if ((rc = r->d_x.download(r->hx.data(), r->tot, r->s))) return rc;
  RP_HIP(hipStreamSynchronize(r->s));
  const rp_pes* P = r->pes;
  for (int c = 0; c < r->nchain; ++c)
    if (P->host_pf(P->host_user, r->n[c], r->hx.data() + r->off[c], r->hg.data() + r->off[c],
                   r->he.data() + r->boff[c]) != 0)
      return set_error(RP_ERR_ARG, "the host potential's potforce callback reported an error");
  for (double& e : r->he) e -= P->v0;
  if (want_grad && (rc = r->d_g.upload(r->hg.data(), r->tot, r->s))) return rc;
  if ((rc = r->d_pe.upload(r->he.data(), r->nbeads, r->s))) return rc;
  RP_HIP(hipStreamSynchronize(r->s));
  return RP_OK;
}

int ring_eval(rp_ring* r, bool want_grad) {
  RingArgs a = ring_args(r);
  RP_HIP(hipMemsetAsync(r->d_e.p, 0, sizeof(double) * r->nchain, r->s));
  const dim3 grid((unsigned)((r->nmax + kRingTile - 1) / kRingTile), (unsigned)r->nchain);
  if (r->pes->kind == PES_HOST) {
    int rc = ring_host_potential(r, want_grad);
    if (rc) return rc;
    ring_spring_kernel<<<grid, kRingTile, 0, r->s>>>(a, r->d_x.p, r->d_pe.p,
                                                     want_grad ? r->d_g.p : nullptr, r->d_e.p);
    RP_HIP(hipGetLastError());
    return RP_OK;
  }
  return dispatch_pes(r->pes, [&]<int K, int M>() -> int {
    if constexpr (K == PES_EVB) {  // the EVB's two GEMMs per bead tile on the matrix cores
      constexpr int TB = kUmWaves * 16;
      if (!r->d_umtask.p) {  // the bead tiles of every chain, once per ring
        std::vector<int> t;
        for (int c = 0; c < r->nchain; ++c)
          for (int b0 = 0; b0 < r->n[c]; b0 += TB) {
            t.push_back(c);
            t.push_back(b0);
          }
        int rc;
        if ((rc = r->d_umtask.alloc(t.size())) || (rc = r->d_umtask.upload(t.data(), t.size(), r->s)))
          return rc;
        RP_HIP(hipStreamSynchronize(r->s));  // t is released at the end of this block
        r->n_umtask = (int)(t.size() / 2);
        int ncu = 0, dev = 0, per_cu = 0;
        RP_HIP(hipGetDevice(&dev));
        RP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        RP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ring_um_evb_kernel<M>,
                                                            kUmWaves * 64, 0));
        r->um_grid = std::min(r->n_umtask, std::max(per_cu, 1) * std::max(ncu, 1));
      }
      ring_um_evb_kernel<M><<<r->um_grid, kUmWaves * 64, 0, r->s>>>(
          a, r->d_x.p, want_grad ? r->d_g.p : nullptr, r->d_e.p, r->d_umtask.p, r->n_umtask);
    } else {
      ring_um_kernel<K, M><<<grid, kRingTile, 0, r->s>>>(a, r->d_x.p,
                                                                want_grad ? r->d_g.p : nullptr, r->d_e.p);
    }
    RP_HIP(hipGetLastError());
    return RP_OK;
  });
}
int ring_tot(const rp_ring* r) { return (int)r->tot; }
int ring_nchain(const rp_ring* r) { return r->nchain; }
const std::vector<int>& ring_n(const rp_ring* r) { return r->n; }
const std::vector<int64_t>& ring_off(const rp_ring* r) { return r->off; }
int ring_D(const rp_ring* r) { return r->D; }
int ring_upload_x(rp_ring* r, const double* x) { return r->d_x.upload(x, r->tot, r->s); }
RingPc ring_pc(rp_ring* r) {
  return RingPc{r->pc_sigma > 0.0 ? r->d_pctab.p : nullptr, r->d_pcoff.p, r->ndim};
}
RingBufs ring_bufs(rp_ring* r) {
  return RingBufs{r->d_x.p, r->d_g.p, r->d_e.p, r->d_off.p, r->d_n.p, r->D, r->s};
}
int ring_download(rp_ring* r, double* e, double* g) {
  int rc;
  if (e && (rc = r->d_e.download(e, r->nchain, r->s))) return rc;
  if (g && (rc = r->d_g.download(g, r->tot, r->s))) return rc;
  RP_HIP(hipStreamSynchronize(r->s));
  return RP_OK;
}
}  // namespace rp

extern "C" int rp_ring_create(rp_pes* pes, const double* mass, int nchain, const int* n,
                              const double* betan, const double* a, const double* b,
                              int fixedends, rp_ring** out) {
  RP_REQUIRE(out, "rp_ring_create: out is NULL");
  *out = nullptr;
  RP_REQUIRE(pes && mass && n && betan && nchain > 0, "rp_ring_create: bad arguments");
  RP_REQUIRE(pes->ndof <= 56, "rp_ring_create: ndof %d > 56", pes->ndof);
  int rc = ensure_init();
  if (rc) return rc;
  rp_ring* r = new rp_ring();
  r->pes = pes;
  r->nchain = nchain;
  r->D = pes->ndof;
  r->ndim = pes->ndim;
  r->natom = pes->natom;
  r->fixedends = fixedends ? 1 : 0;
  const int D = r->D;
  r->n.assign(n, n + nchain);
  r->betan.assign(betan, betan + nchain);
  std::vector<double> ab((size_t)nchain * 2 * D, 0.0);
  for (int c = 0; c < nchain; ++c) {
    if (n[c] < 2) {
      delete r;
      return set_error(RP_ERR_ARG, "rp_ring_create: chain %d has n=%d < 2", c, n[c]);
    }
    r->off.push_back(r->tot);
    r->boff.push_back(r->nbeads);
    r->tot += (int64_t)n[c] * D;
    r->nbeads += n[c];
    r->nmax = std::max(r->nmax, n[c]);
    if (a) memcpy(ab.data() + (size_t)c * 2 * D, a + (size_t)c * D, D * sizeof(double));
    if (b) memcpy(ab.data() + (size_t)c * 2 * D + D, b + (size_t)c * D, D * sizeof(double));
  }
  r->hmass.assign(mass, mass + r->natom);
  std::vector<double> massd(D);
  for (int at = 0; at < r->natom; ++at)
    for (int j = 0; j < r->ndim; ++j) massd[at * r->ndim + j] = mass[at];
  if ((rc = [&]() -> int {
        RP_HIP(hipStreamCreateWithFlags(&r->s, hipStreamNonBlocking));
        int q;
        if ((q = r->d_x.alloc(r->tot)) || (q = r->d_g.alloc(r->tot)) || (q = r->d_v.alloc(r->tot)) ||
            (q = r->d_hv.alloc(r->tot)) || (q = r->d_e.alloc(nchain)))
          return q;
        RP_HIP(hipMemsetAsync(r->d_x.p, 0, r->tot * sizeof(double), r->s));
        if ((q = r->d_ab.alloc(ab.size())) || (q = r->d_ab.upload(ab.data(), ab.size(), r->s))) return q;
        if ((q = r->d_betan.alloc(nchain)) || (q = r->d_betan.upload(betan, nchain, r->s))) return q;
        if ((q = r->d_massd.alloc(D)) || (q = r->d_massd.upload(massd.data(), D, r->s))) return q;
        if ((q = r->d_n.alloc(nchain)) || (q = r->d_n.upload(n, nchain, r->s))) return q;
        if ((q = r->d_off.alloc(nchain)) || (q = r->d_off.upload(r->off.data(), nchain, r->s))) return q;
        if ((q = r->d_boff.alloc(nchain)) || (q = r->d_boff.upload(r->boff.data(), nchain, r->s))) return q;
        RP_HIP(hipStreamSynchronize(r->s));
        return RP_OK;
      }())) {
    rp_ring_destroy(r);
    return rc;
  }
  *out = r;
  return RP_OK;
}

// The spring part of the U_M Hessian per dof column, P = k tridiag(-1, 2, -1) + sigma I with
// k = m / betan^2 (instantonmod.f90:50-89; free ends: one-sided end rows, diagonal k), as
// Thomas-algorithm coefficients per (chain, atom, bead): {1/denominator, -off/denominator}.
extern "C" int rp_ring_set_preconditioner(rp_ring* r, double sigma) {
  RP_REQUIRE(r, "rp_ring_set_preconditioner: NULL ring");
  RP_REQUIRE(sigma == sigma, "rp_ring_set_preconditioner: sigma is NaN");
  r->pc_sigma = sigma > 0.0 ? sigma : 0.0;
  if (r->pc_sigma == 0.0) return RP_OK;
  std::vector<int64_t> pcoff(r->nchain);
  int64_t tot = 0;
  for (int c = 0; c < r->nchain; ++c) {
    pcoff[c] = tot;
    tot += (int64_t)r->natom * r->n[c] * 2;
  }
  std::vector<double> tab(tot);
  for (int c = 0; c < r->nchain; ++c) {
    const int n = r->n[c];
    for (int at = 0; at < r->natom; ++at) {
      const double k = r->hmass[at] / (r->betan[c] * r->betan[c]), off = -k;
      double* t = tab.data() + pcoff[c] + (int64_t)at * n * 2;
      double cp = 0.0;  // c'_{i-1}
      for (int i = 0; i < n; ++i) {
        const bool end = !r->fixedends && (i == 0 || i == n - 1);
        const double diag = (end ? k : 2.0 * k) + r->pc_sigma;
        const double inv = 1.0 / (diag - off * cp);
        cp = off * inv;
        t[2 * i] = inv;
        t[2 * i + 1] = -cp;
      }
    }
  }
  int rc;
  if ((rc = r->d_pctab.alloc(tot)) || (rc = r->d_pctab.upload(tab.data(), tot, r->s)) ||
      (rc = r->d_pcoff.alloc(r->nchain)) || (rc = r->d_pcoff.upload(pcoff.data(), r->nchain, r->s)))
    return rc;
  RP_HIP(hipStreamSynchronize(r->s));
  return RP_OK;
}

extern "C" int rp_ring_destroy(rp_ring* r) {
  if (!r) return RP_OK;
  if (r->s) (void)hipStreamSynchronize(r->s);
  for (DevBuf<double>* b : {&r->d_x, &r->d_g, &r->d_v, &r->d_hv, &r->d_e, &r->d_ab, &r->d_betan,
                            &r->d_massd, &r->d_blocks, &r->d_sinv, &r->d_work, &r->d_ld, &r->d_eta,
                            &r->d_pe, &r->d_pctab})
    b->free();
  r->d_n.free();
  r->d_nneg.free();
  r->d_umtask.free();
  r->d_off.free();
  r->d_boff.free();
  r->d_pcoff.free();
  if (r->s) (void)hipStreamDestroy(r->s);
  delete r;
  return RP_OK;
}

extern "C" int rp_ring_set_x(rp_ring* r, const double* x) {
  RP_REQUIRE(r && x, "rp_ring_set_x: bad arguments");
  int rc = r->d_x.upload(x, r->tot, r->s);
  if (rc) return rc;
  RP_HIP(hipStreamSynchronize(r->s));
  return RP_OK;
}

extern "C" int rp_ring_get_x(rp_ring* r, double* x) {
  RP_REQUIRE(r && x, "rp_ring_get_x: bad arguments");
  int rc = r->d_x.download(x, r->tot, r->s);
  if (rc) return rc;
  RP_HIP(hipStreamSynchronize(r->s));
  return RP_OK;
}

extern "C" int rp_ring_um(rp_ring* r, double* energy, double* grad) {
  RP_REQUIRE(r && energy, "rp_ring_um: bad arguments");
  int rc = ring_eval(r, grad != nullptr);
  if (rc) return rc;
  return ring_download(r, energy, grad);
}

static int ring_blocks(rp_ring* r, int singlewell);

// d_hv = H d_v at d_x (queued on the ring's stream)
static int ring_hvp(rp_ring* r) {
  int rc;
  RingArgs a = ring_args(r);
  const dim3 grid((unsigned)((r->nmax + kRingTile - 1) / kRingTile), (unsigned)r->nchain);
  if (r->pes->kind == PES_HOST) {  // the plugin's Hessian blocks, then the block product
    if ((rc = ring_blocks(r, 0))) return rc;
    ring_hvp_blocks_kernel<<<grid, kRingTile, 0, r->s>>>(a, r->d_blocks.p, r->d_v.p, r->d_hv.p);
    RP_HIP(hipGetLastError());
    return RP_OK;
  }
  return dispatch_pes(r->pes, [&]<int K, int M>() -> int {
    ring_hvp_kernel<K, M><<<grid, kRingTile, 0, r->s>>>(a, r->d_x.p, r->d_v.p, r->d_hv.p);
    RP_HIP(hipGetLastError());
    return RP_OK;
  });
}

extern "C" int rp_ring_hvp(rp_ring* r, const double* v, double* hv) {
  RP_REQUIRE(r && v && hv, "rp_ring_hvp: bad arguments");
  int rc;
  if ((rc = r->d_v.upload(v, r->tot, r->s))) return rc;
  if ((rc = ring_hvp(r))) return rc;
  if ((rc = r->d_hv.download(hv, r->tot, r->s))) return rc;
  RP_HIP(hipStreamSynchronize(r->s));
  return RP_OK;
}

// Device-memory twins of rp_ring_set_x / rp_ring_um / rp_ring_hvp: device-to-device copies
// on the ring's stream, complete on return.
extern "C" int rp_ring_set_x_device(rp_ring* r, const double* x) {
  RP_REQUIRE(r && x, "rp_ring_set_x_device: bad arguments");
  RP_HIP(hipMemcpyAsync(r->d_x.p, x, sizeof(double) * r->tot, hipMemcpyDeviceToDevice, r->s));
  RP_HIP(hipStreamSynchronize(r->s));
  return RP_OK;
}

extern "C" int rp_ring_um_device(rp_ring* r, double* energy, double* grad) {
  RP_REQUIRE(r && energy, "rp_ring_um_device: bad arguments");
  int rc = ring_eval(r, grad != nullptr);
  if (rc) return rc;
  if (grad)
    RP_HIP(hipMemcpyAsync(grad, r->d_g.p, sizeof(double) * r->tot, hipMemcpyDeviceToDevice, r->s));
  return ring_download(r, energy, nullptr);
}

extern "C" int rp_ring_hvp_device(rp_ring* r, const double* v, double* hv) {
  RP_REQUIRE(r && v && hv, "rp_ring_hvp_device: bad arguments");
  int rc;
  RP_HIP(hipMemcpyAsync(r->d_v.p, v, sizeof(double) * r->tot, hipMemcpyDeviceToDevice, r->s));
  if ((rc = ring_hvp(r))) return rc;
  RP_HIP(hipMemcpyAsync(hv, r->d_hv.p, sizeof(double) * r->tot, hipMemcpyDeviceToDevice, r->s));
  RP_HIP(hipStreamSynchronize(r->s));
  return RP_OK;
}

static int ring_blocks(rp_ring* r, int singlewell) {
  int rc;
  const size_t nb = (size_t)r->nbeads * r->D * r->D;
  if (r->d_blocks.n < nb && (rc = r->d_blocks.alloc(nb))) return rc;
  if (r->pes->kind == PES_HOST) {
    // the plugin's Vdoubleprime (or differenced gradients) per bead; singlewell reuses
    // bead 1's block for every bead (instantonmod.f90:183-185)
    const int D = r->D;
    std::vector<double> hx(r->tot), hb(nb), x1(D);
    if ((rc = r->d_x.download(hx.data(), r->tot, r->s))) return rc;
    RP_HIP(hipStreamSynchronize(r->s));
    for (int c = 0; c < r->nchain; ++c) {
      const int n = r->n[c];
      double* out = hb.data() + (size_t)r->boff[c] * D * D;
      if (singlewell) {
        for (int d = 0; d < D; ++d) x1[d] = hx[r->off[c] + (size_t)d * n];
        if ((rc = host_pes_hessian(r->pes, 1, x1.data(), out))) return rc;
        for (int i = 1; i < n; ++i) memcpy(out + (size_t)i * D * D, out, (size_t)D * D * sizeof(double));
      } else if ((rc = host_pes_hessian(r->pes, n, hx.data() + r->off[c], out))) {
        return rc;
      }
    }
    if ((rc = r->d_blocks.upload(hb.data(), nb, r->s))) return rc;
    RP_HIP(hipStreamSynchronize(r->s));
    return RP_OK;
  }
  RingArgs a = ring_args(r);
  int maxn = 0;
  for (int c = 0; c < r->nchain; ++c) maxn = std::max(maxn, r->n[c]);
  dim3 grid((maxn * r->D + 127) / 128, r->nchain);
  return dispatch_pes(r->pes, [&]<int K, int M>() -> int {
    ring_hblock_kernel<K, M><<<grid, 128, 0, r->s>>>(a, r->d_x.p, singlewell, r->d_blocks.p);
    RP_HIP(hipGetLastError());
    return RP_OK;
  });
}

extern "C" int rp_ring_hessian_band(rp_ring* r, int chain, int singlewell, double* band) {
  RP_REQUIRE(r && band && chain >= 0 && chain < r->nchain, "rp_ring_hessian_band: bad arguments");
  int rc;
  if ((rc = ring_blocks(r, singlewell))) return rc;
  const int n = r->n[chain], D = r->D;
  const int64_t tot = (int64_t)n * D, nband = (D + 1) * tot;
  DevBuf<double> d_band;
  if ((rc = d_band.alloc(nband))) return rc;
  ring_band_kernel<<<(unsigned)((nband + 255) / 256), 256, 0, r->s>>>(
      n, D, r->betan[chain] * r->betan[chain], r->d_massd.p,
      r->d_blocks.p + (size_t)r->boff[chain] * D * D, d_band.p);
  rc = hipGetLastError() == hipSuccess ? d_band.download(band, nband, r->s) : set_error(RP_ERR_HIP, "band kernel");
  if (!rc) RP_HIP(hipStreamSynchronize(r->s));
  d_band.free();
  return rc;
}

namespace rp {
// Launch order of factorisation tasks (task t factors chain chain_of[t]): longest chain first.
// A task's time is its chain's length, and the workgroups a launch holds beyond what the
// device runs at once start in block order as slots free; with the long chains first the
// short ones fill the tail, so a launch of several tasks per CU takes about as long as its
// longest chain instead of up to twice that.
static std::vector<int> longest_first(const rp_ring* r, const std::vector<int>& chain_of) {
  std::vector<int> ord(chain_of.size());
  for (size_t t = 0; t < ord.size(); ++t) ord[t] = (int)t;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int u, int v) { return r->n[chain_of[u]] > r->n[chain_of[v]]; });
  return ord;
}

template <class V>
static void permute(V& v, const std::vector<int>& ord) {
  V w(v.size());
  for (size_t t = 0; t < ord.size(); ++t) w[t] = v[ord[t]];
  v.swap(w);
}

// Factor every task t (chain chain_of[t], or chain t when chain_of is null) at shift[t]:
// log|det|, the number of negative pivots (eigenvalues below the shift) and, when keep, the
// S_i^-1 of the chain for inverse iteration.
static int ring_factor(rp_ring* r, int ntask, const double* shift, const int* chain_of,
                       bool keep, double* ld, int* nneg) {
  return with_logdet_kernel(r->D, [&](auto kern, int nt) -> int {
    kern<<<ntask, nt, 0, r->s>>>(r->D, r->d_n.p, r->d_boff.p, r->d_betan.p, r->d_massd.p,
                                r->d_blocks.p, shift, keep ? r->d_sinv.p : nullptr, ld, nneg,
                                chain_of);
    RP_HIP(hipGetLastError());
    return RP_OK;
  });
}

// eta[c] = the lowest eigenvalue of chain c's Hessian when it has none below zero (ng[c] == 0),
// else the eigenvalue nearest zero (which the caller keeps only when it is the one negative
// eigenvalue).  d_sinv holds the unshifted factors on entry; it is left holding shifted ones.
//
// The bottom of an instanton's spectrum can be a cluster (the springs' modes lie a few per
// cent apart at large N), where inverse iteration from zero converges slowly and a Rayleigh
// quotient shift can land on any member.  So for ng[c] == 0 the search keeps a bracket on
// Sylvester's inertia: lo with no eigenvalue below it (0 to start), hi with at least one (the
// Rayleigh quotient of a few unshifted inverse iterations, an upper bound on eta_1), and h1,
// the largest shift seen with exactly one below it (so eta_2 >= h1).  Rounds of multisection
// (three shifts per chain in one launch) narrow it until inverse iteration shifted to lo,
// which converges to the eigenvalue nearest lo, i.e. eta_1, contracts by
// (hi - lo) / (h1 - lo) <= 1/4 per iteration; that iteration then gives eta_1 to the last
// bits.  A chain whose bracket closes first takes its midpoint.
int ring_lowest_eigen(rp_ring* r, const std::vector<int>& ng, double* eta) {
  const int C = r->nchain;
  int rc;
  if ((int64_t)r->nmax * r->D * r->D * 8 >= ((int64_t)1 << 31))
    return set_error(RP_ERR_ARG, "ring log-determinant: a chain's block factors exceed 2 GiB");
  if (r->d_work.n < 2 * (size_t)r->tot && (rc = r->d_work.alloc(2 * (size_t)r->tot))) return rc;
  constexpr int M = 7;  // shifts per chain and multisection round
  DevBuf<double> d_shift, d_tld;
  DevBuf<int> d_act, d_map, d_cnt;
  auto cleanup = [&]() { d_shift.free(); d_tld.free(); d_act.free(); d_map.free(); d_cnt.free(); };
  if ((rc = d_shift.alloc(M * (size_t)C)) || (rc = d_tld.alloc(M * (size_t)C)) ||
      (rc = d_act.alloc(C)) || (rc = d_map.alloc(M * (size_t)C)) ||
      (rc = d_cnt.alloc(M * (size_t)C))) {
    cleanup();
    return rc;
  }
  // lanes per block row and prefetch depth of the block solve: 8 x 4 columns, 12 steps ahead
  // for D <= 32; 4 x 16 columns, 4 steps ahead up to kMaxD (the registers of a deeper ring)
  auto eta1 = [&](const double* shift, int iters, int fresh, double tol, const int* act) -> int {
    auto kern = r->D <= 32 ? ring_eta1_kernel<8, 4, 12> : ring_eta1_kernel<4, 16, 4>;
    kern<<<C, kRingThreads, 0, r->s>>>(r->D, r->d_n.p, r->d_boff.p, r->d_betan.p, r->d_sinv.p,
                                       r->d_work.p, shift, iters, fresh, tol, act, r->d_eta.p);
    RP_HIP(hipGetLastError());
    return RP_OK;
  };
  auto fetch_eta = [&](std::vector<double>& e) -> int {
    int q = r->d_eta.download(e.data(), C, r->s);
    if (!q && hipStreamSynchronize(r->s) != hipSuccess)
      q = set_error(RP_ERR_HIP, "inverse iteration");
    return q;
  };
  std::vector<double> mu(C), sh(C, 0.0);
  std::vector<int> act(C);
  // inverse iteration on the unshifted factors, from a fixed generic start
  if ((rc = eta1(nullptr, 6, 1, 0.0, nullptr)) || (rc = fetch_eta(mu))) return cleanup(), rc;
  // chains with negative eigenvalues: Rayleigh-quotient refinement toward the one nearest zero
  bool any = false;
  for (int c = 0; c < C; ++c) any |= (act[c] = ng[c] != 0);
  for (int round = 0; any && round < 3; ++round) {
    std::vector<int> map;
    for (int c = 0; c < C; ++c) {
      sh[c] = mu[c] * (1.0 - 1e-9);  // just off the estimate, so H - sh I stays factorable
      if (act[c]) map.push_back(c);
    }
    permute(map, longest_first(r, map));
    std::vector<double> tsh;
    for (int c : map) tsh.push_back(sh[c]);
    if ((rc = d_map.upload(map.data(), map.size(), r->s)) ||
        (rc = d_tld.upload(tsh.data(), tsh.size(), r->s)) ||
        (rc = ring_factor(r, (int)map.size(), d_tld.p, d_map.p, true, d_shift.p + C, d_cnt.p)) ||
        (rc = d_shift.upload(sh.data(), C, r->s)) || (rc = d_act.upload(act.data(), C, r->s)) ||
        (rc = eta1(d_shift.p, 4, 0, 1e-15, d_act.p)) || (rc = fetch_eta(mu)))
      return cleanup(), rc;
  }
  // chains with none: the lowest eigenvalue on an inertia bracket
  std::vector<double> lo(C, 0.0), hi(C), h1(C, -1.0);
  std::vector<char> open(C, 0);
  for (int c = 0; c < C; ++c) {
    eta[c] = mu[c];
    if (ng[c] != 0) continue;
    if (!(mu[c] > 0.0))
      return cleanup(), set_error(RP_ERR_CONVERGE,
                                  "chain %d: inverse iteration gave eta = %g with no negative "
                                  "eigenvalue", c, mu[c]);
    hi[c] = mu[c] * (1.0 + 1e-13);
    open[c] = 1;
  }
  auto needs_more = [&](int c) {
    if (hi[c] - lo[c] <= 4e-16 * hi[c]) return false;  // closed to a few ulps
    return !(h1[c] > lo[c] && hi[c] - lo[c] <= 0.25 * (h1[c] - lo[c]));
  };
  for (int round = 0; round < 60; ++round) {
    std::vector<int> map;
    std::vector<double> tsh;
    for (int c = 0; c < C; ++c) {
      if (!open[c] || !needs_more(c)) continue;
      for (int k = 1; k <= M; ++k) {
        map.push_back(c);
        // the first round's shifts crowd toward hi (eta_1 is usually within a few per cent
        // of the Rayleigh quotient), later rounds split the bracket evenly
        const double f = round == 0 ? 1.0 - std::ldexp(1.0, -k) : (double)k / (M + 1);
        tsh.push_back(lo[c] + (hi[c] - lo[c]) * f);
      }
    }
    if (map.empty()) break;
    const std::vector<int> ord = longest_first(r, map);
    permute(map, ord);
    permute(tsh, ord);
    std::vector<int> cnt(map.size());
    if ((rc = d_map.upload(map.data(), map.size(), r->s)) ||
        (rc = d_tld.upload(tsh.data(), tsh.size(), r->s)) ||
        (rc = ring_factor(r, (int)map.size(), d_tld.p, d_map.p, false, d_shift.p, d_cnt.p)) ||
        (rc = d_cnt.download(cnt.data(), cnt.size(), r->s)))
      return cleanup(), rc;
    if (hipStreamSynchronize(r->s) != hipSuccess)
      return cleanup(), set_error(RP_ERR_HIP, "inertia count");
    // the counts rise with the shift: the bracket closes onto the first rise
    for (size_t t = 0; t < map.size(); ++t) {
      const int c = map[t];
      const double x = tsh[t];
      if (cnt[t] == 0) lo[c] = std::max(lo[c], x);
      else hi[c] = std::min(hi[c], x);
      if (cnt[t] == 1) h1[c] = std::max(h1[c], x);
    }
  }
  // inverse iteration shifted to lo: every eigenvalue lies above it, so it converges to eta_1
  std::vector<int> map;
  std::vector<double> tsh;
  for (int c = 0; c < C; ++c) {
    act[c] = 0;
    if (!open[c]) continue;
    if (hi[c] - lo[c] <= 4e-16 * hi[c]) {
      eta[c] = 0.5 * (lo[c] + hi[c]);
      continue;
    }
    act[c] = 1;
    sh[c] = lo[c];
    map.push_back(c);
    tsh.push_back(lo[c]);
  }
  if (!map.empty()) {
    const std::vector<int> ord = longest_first(r, map);
    permute(map, ord);
    permute(tsh, ord);
    std::vector<double> e(C);
    if ((rc = d_map.upload(map.data(), map.size(), r->s)) ||
        (rc = d_tld.upload(tsh.data(), tsh.size(), r->s)) ||
        (rc = ring_factor(r, (int)map.size(), d_tld.p, d_map.p, true, d_shift.p + C, d_cnt.p)) ||
        (rc = d_shift.upload(sh.data(), C, r->s)) || (rc = d_act.upload(act.data(), C, r->s)) ||
        (rc = eta1(d_shift.p, 30, 0, 1e-14, d_act.p)) || (rc = fetch_eta(e)))
      return cleanup(), rc;
    for (int c : map) {
      if (!(e[c] > lo[c] && e[c] <= hi[c] * (1.0 + 1e-12)))
        return cleanup(), set_error(RP_ERR_CONVERGE,
                                    "chain %d: eta_1 = %.17g left its inertia bracket "
                                    "[%.17g, %.17g]", c, e[c], lo[c], hi[c]);
      eta[c] = e[c];
    }
  }
  cleanup();
  return RP_OK;
}
}  // namespace rp

namespace rp {
// Gershgorin bounds of every chain's band matrix (blocks already formed): lohi[2c], lohi[2c+1].
static int ring_gersh(rp_ring* r, std::vector<double>& lohi) {
  DevBuf<double> d_lohi;
  int rc;
  if ((rc = d_lohi.alloc(2 * (size_t)r->nchain))) return rc;
  ring_gersh_kernel<<<r->nchain, kRingThreads, 0, r->s>>>(r->D, r->d_n.p, r->d_boff.p, r->d_betan.p,
                                                          r->d_massd.p, r->d_blocks.p, d_lohi.p);
  lohi.resize(2 * (size_t)r->nchain);
  if (hipGetLastError() != hipSuccess) rc = set_error(RP_ERR_HIP, "Gershgorin kernel launch");
  if (!rc) rc = d_lohi.download(lohi.data(), lohi.size(), r->s);
  if (!rc && hipStreamSynchronize(r->s) != hipSuccess) rc = set_error(RP_ERR_HIP, "Gershgorin bounds");
  d_lohi.free();
  return rc;
}

// Eigenvalues by bisection on Sylvester's inertia: the number of eigenvalues below sigma is
// the number of negative pivots of the block LDL^T of H - sigma I.  Task t brackets the
// j[t]-th smallest eigenvalue (0-based) of chain c[t] in [a[t], b[t]], with at most j[t]
// eigenvalues below a[t] and more than j[t] below b[t]; one launch factors every task at its
// midpoint.  eig[t] gets the midpoint of the final bracket, within a few ulps of it.
static int ring_bisect(rp_ring* r, const std::vector<int>& tchain, const std::vector<int>& tj,
                       std::vector<double>& a, std::vector<double>& b, std::vector<double>& eig) {
  const int nt = (int)tchain.size();
  eig.assign(nt, 0.0);
  if (nt == 0) return RP_OK;
  int rc;
  DevBuf<double> d_mid, d_ldt;
  DevBuf<int> d_map, d_cnt;
  auto cleanup = [&]() { d_mid.free(); d_ldt.free(); d_map.free(); d_cnt.free(); };
  // launch slot u runs task ord[u]
  const std::vector<int> ord = longest_first(r, tchain);
  std::vector<int> lchain(tchain);
  permute(lchain, ord);
  if ((rc = d_mid.alloc(nt)) || (rc = d_ldt.alloc(nt)) || (rc = d_map.alloc(nt)) ||
      (rc = d_cnt.alloc(nt)) || (rc = d_map.upload(lchain.data(), nt, r->s))) {
    cleanup();
    return rc;
  }
  std::vector<double> mid(nt), lmid(nt);
  std::vector<int> k(nt);
  // 200 halvings take any bracket of doubles to adjacent values; the loop ends sooner, when
  // every bracket is within a few ulps
  for (int it = 0; !rc && it < 200; ++it) {
    bool open = false;
    for (int t = 0; t < nt; ++t) {
      mid[t] = 0.5 * (a[t] + b[t]);
      open |= (b[t] - a[t]) > 4e-16 * std::max(std::fabs(a[t]), std::fabs(b[t]));
    }
    if (!open) break;
    for (int u = 0; u < nt; ++u) lmid[u] = mid[ord[u]];
    if ((rc = d_mid.upload(lmid.data(), nt, r->s))) break;
    rc = with_logdet_kernel(r->D, [&](auto kern, int nth) -> int {
      kern<<<nt, nth, 0, r->s>>>(r->D, r->d_n.p, r->d_boff.p, r->d_betan.p, r->d_massd.p,
                                 r->d_blocks.p, d_mid.p, nullptr, d_ldt.p, d_cnt.p, d_map.p);
      RP_HIP(hipGetLastError());
      return RP_OK;
    });
    if (rc || (rc = d_cnt.download(k.data(), nt, r->s))) break;
    if (hipStreamSynchronize(r->s) != hipSuccess) {
      rc = set_error(RP_ERR_HIP, "inertia count");
      break;
    }
    for (int u = 0; u < nt; ++u) {
      const int t = ord[u];
      (k[u] > tj[t] ? b[t] : a[t]) = mid[t];
    }
  }
  for (int t = 0; t < nt; ++t) eig[t] = 0.5 * (a[t] + b[t]);
  cleanup();
  return rc;
}

// The negative eigenvalues of the chains listed (cnt[c] of them for chain c), each bracketed
// between the chain's Gershgorin lower bound and 0.  sumlog[c] gets sum ln|eta| over them,
// lowest[c] the smallest one.
int ring_negative_eigs(rp_ring* r, const std::vector<int>& chains, const std::vector<int>& cnt,
                       std::vector<double>& sumlog, std::vector<double>& lowest) {
  if (chains.empty()) return RP_OK;
  int rc;
  std::vector<double> lohi;
  if ((rc = ring_gersh(r, lohi))) return rc;
  std::vector<int> tchain, tj;
  std::vector<double> a, b, eig;
  for (int c : chains)
    for (int j = 0; j < cnt[c]; ++j) {
      tchain.push_back(c);
      tj.push_back(j);
      a.push_back(lohi[2 * c] - 1e-12 * std::fabs(lohi[2 * c]) - 1e-300);
      b.push_back(0.0);
    }
  if ((rc = ring_bisect(r, tchain, tj, a, b, eig))) return rc;
  for (size_t t = 0; t < tchain.size(); ++t) {
    const int c = tchain[t];
    if (!(eig[t] < 0.0))
      return set_error(RP_ERR_CONVERGE, "chain %d: the inertia bisection lost eigenvalue %d", c, tj[t]);
    sumlog[c] += std::log(std::fabs(eig[t]));
    lowest[c] = std::min(lowest[c], eig[t]);
  }
  return RP_OK;
}
}  // namespace rp

extern "C" int rp_ring_lowest_eigs(rp_ring* r, int singlewell, int k, double* eigs) {
  RP_REQUIRE(r && eigs && k >= 1, "rp_ring_lowest_eigs: bad arguments");
  for (int c = 0; c < r->nchain; ++c)
    RP_REQUIRE(k <= r->n[c] * r->D, "rp_ring_lowest_eigs: chain %d has only %d eigenvalues", c,
               r->n[c] * r->D);
  int rc;
  if ((rc = ring_blocks(r, singlewell))) return rc;
  std::vector<double> lohi;
  if ((rc = ring_gersh(r, lohi))) return rc;
  std::vector<int> tchain, tj;
  std::vector<double> a, b, eig;
  for (int c = 0; c < r->nchain; ++c)
    for (int j = 0; j < k; ++j) {
      tchain.push_back(c);
      tj.push_back(j);
      const double lo = lohi[2 * c], hi = lohi[2 * c + 1];
      a.push_back(lo - 1e-12 * std::fabs(lo) - 1e-300);
      b.push_back(hi + 1e-12 * std::fabs(hi) + 1e-300);
    }
  if ((rc = ring_bisect(r, tchain, tj, a, b, eig))) return rc;
  for (size_t t = 0; t < eig.size(); ++t) eigs[t] = eig[t];
  return RP_OK;
}

extern "C" int rp_ring_logdet(rp_ring* r, int singlewell, int skip_lowest, double* logdet_pos,
                              int* nneg, double* eta1) {
  RP_REQUIRE(r && logdet_pos, "rp_ring_logdet: bad arguments");
  int rc;
  const int D = r->D;
  if ((rc = ring_blocks(r, singlewell))) return rc;
  const size_t nb = (size_t)r->nbeads * D * D;
  if (r->d_sinv.n < nb && (rc = r->d_sinv.alloc(nb))) return rc;
  if ((rc = r->d_ld.alloc(r->nchain)) || (rc = r->d_nneg.alloc(r->nchain)) || (rc = r->d_eta.alloc(r->nchain)))
    return rc;
  if ((rc = with_logdet_kernel(D, [&](auto kern, int nt) -> int {
        kern<<<r->nchain, nt, 0, r->s>>>(D, r->d_n.p, r->d_boff.p, r->d_betan.p, r->d_massd.p,
                                        r->d_blocks.p, nullptr, r->d_sinv.p, r->d_ld.p, r->d_nneg.p,
                                        nullptr);
        RP_HIP(hipGetLastError());
        return RP_OK;
      })))
    return rc;
  std::vector<double> ld(r->nchain), et(r->nchain, 0.0);
  std::vector<int> ng(r->nchain);
  if ((rc = r->d_ld.download(ld.data(), r->nchain, r->s)) ||
      (rc = r->d_nneg.download(ng.data(), r->nchain, r->s)))
    return rc;
  RP_HIP(hipStreamSynchronize(r->s));
  const bool want_near = skip_lowest || eta1;
  if (want_near) {
    if ((rc = ring_lowest_eigen(r, ng, et.data()))) return rc;
  }
  // log|det| = sum ln|eta| over every eigenvalue, so the sum over the positive ones is log|det|
  // less ln|eta| over the negative ones.  With one negative eigenvalue that is the eigenvalue
  // nearest zero when inverse iteration found a negative one; every other negative spectrum
  // is bisected on the inertia.
  std::vector<int> bis;
  std::vector<double> neglog(r->nchain, 0.0), lowest(r->nchain, INFINITY);
  for (int c = 0; c < r->nchain; ++c) {
    if (ng[c] == 1 && want_near && et[c] < 0.0) {
      neglog[c] = std::log(-et[c]);
      lowest[c] = et[c];
    } else if (ng[c] >= 1) {
      bis.push_back(c);
    }
  }
  if ((rc = ring_negative_eigs(r, bis, ng, neglog, lowest))) return rc;
  for (int c = 0; c < r->nchain; ++c) {
    double v = ld[c] - neglog[c];  // sum of ln(eta) over eta > 0 (rpi_ser.f90:227-236)
    // lndetJ (rpi_ser.f90:355-365) also skips eta_1, the lowest eigenvalue: with every
    // eigenvalue positive that is the one nearest zero
    if (skip_lowest && ng[c] == 0) v -= std::log(std::fabs(et[c]));
    logdet_pos[c] = v;
    if (nneg) nneg[c] = ng[c];
    if (eta1) eta1[c] = ng[c] == 0 ? et[c] : lowest[c];
  }
  return RP_OK;
}

extern "C" int rp_ring_param_grad(rp_ring* r, double* out, int* nparam) {
  RP_REQUIRE(r && out, "rp_ring_param_grad: bad arguments");
  if (r->pes->kind != PES_WATER)
    return set_error(RP_ERR_ARG, "potential kind %d has no parameter sensitivities", r->pes->kind);
  int rc;
  DevBuf<double> d_out;
  if ((rc = d_out.alloc(2 * (size_t)r->nchain))) return rc;
  RingArgs a = ring_args(r);
  switch (r->natom) {
    case 3: ring_param_grad_kernel<1><<<r->nchain, kRingThreads, 0, r->s>>>(a, r->d_x.p, d_out.p); break;
    case 6: ring_param_grad_kernel<2><<<r->nchain, kRingThreads, 0, r->s>>>(a, r->d_x.p, d_out.p); break;
    default: ring_param_grad_kernel<6><<<r->nchain, kRingThreads, 0, r->s>>>(a, r->d_x.p, d_out.p); break;
  }
  if (hipGetLastError() != hipSuccess) rc = set_error(RP_ERR_HIP, "param grad kernel launch");
  if (!rc) rc = d_out.download(out, 2 * (size_t)r->nchain, r->s);
  if (!rc) RP_HIP(hipStreamSynchronize(r->s));
  d_out.free();
  if (nparam) *nparam = 2;
  return rc;
}

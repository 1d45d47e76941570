// rp_batch.h -- device side of the batched engine (one workgroup per polymer).
//
// A workgroup of 4 wavefronts owns one polymer for a block of K steps.  The state
// x, p lives in LDS as S[bead][col] (col < DC: x of dof col; DC <= col < 2 DC: p)
// with an odd row stride, and every step runs fused in the workgroup:
//   B   : each thread evaluates the potential of one bead from S, kicks p in S
//   fwd : [q | p~] = T [x | p]          FP64 MFMA, results stay in registers
//   NM  : q -= beadvec; A(dt/2) O A(dt/2) per (mode, dof) in registers -> S
//   bwd : [x | p] = T [q + beadvec | p~]  FP64 MFMA -> S
//   TI  : sum m (-x_N) db/dxi (verletmodule.f90:397-410)
// MFMA work is split in units of (two 16-row tiles) x (16 dof columns of x and of p):
// four 16x16 accumulators per unit, A fragments (rows of T) shared by the x and p
// MFMAs, B fragments (S rows) shared by the two row tiles.
//
// Every normal mode of the reference's chain has the same frequency 1/tau
// (beadmass = m (lam tau)^2, verletmodule.f90:315, 519), so the exact rotation by h
// and the O-step factor into uniform scalars and per-element m*lam_i^2:
//   q' = c q + s_q p~ / (m lam^2),  p~' = -s_p m lam^2 q + c p~
//   p~ <- c1^2 p~ + sqrt(m) lam_i kappa z
// with (c, s_q, s_p) = (cos(h/tau), sin(h/tau)/tau, tau sin(h/tau)) or the Cayley
// map's counterparts, c1 = exp(-gamma dt / tau), kappa = tau c2 sqrt((1+c1^2)/betan).
#pragma once
#include "rp_dispatch.h"
#include "rp_nm.h"
#include "rp_philox.h"

namespace rp {

// Per-polymer 64-bit fields (BatchArgs::pi64[poly * kPI64 + field])
enum : int { PI_STATE_OFF = 0, PI_CURSOR, PI_TICNT, PI_NREINIT, PI_KCURSOR, PI_NOISE_OFF,
             PI_KICK_OFF, PI_ACOUNT, PI_RKICK, kPI64 };
// Per-polymer doubles (BatchArgs::pf64[poly * kPF64 + field])
enum : int { PF_TISUM = 0, PF_XI, kPF64 };
// Per-mode family table (BatchArgs::modes[i * kMODE + field])
enum : int { MD_MU = 0, MD_IMU, MD_U, MD_W, MD_LAM, kMODE };

struct BatchArgs {
  PesDev pes;
  int N, NPAD, D, ndim, natom, LD;
  int first;          // first polymer (batch order) of this launch's family
  int nsteps;
  int64_t step0, imin;
  const int64_t* step0p;        // step counter in device memory (graph replays), else nullptr
  double dt, limit, stdev0;
  double c, sq, sp, c1sq, kappa;  // uniform rotation / O-step scalars (see above)
  double tau;
  int64_t noutput;
  uint64_t seed;
  double* state;                // per polymer [x: D*N][p: D*N]
  int64_t* pi64;                // per polymer 64-bit fields
  double* pf64;                 // per polymer doubles
  const uint64_t* stream;       // per polymer Philox stream id
  const double* T;              // NPAD*NPAD, zero padded
  const double* Tf;             // folded halves [odd|even][HPAD][HPAD] (even N), else null
  int fold;                     // 1: transforms through the folded halves (N even)
  const double* modes;          // per mode [mu, 1/mu, u, w, lam]
  const double* pdata;          // per polymer [a | b | dbdl], DC each, zero padded
  const double* massd;          // per dof mass, DC, padded with 1
  const double* noise;          // injected draws or nullptr
  int64_t noise_cap;
  const int32_t* kicks;         // injected Poisson draws or nullptr
  int64_t kick_cap;
  int waves;          // waves per workgroup of the step kernel (4 or 8)
  int npath;
  const double* pathd;          // [lampath (npath) | path (D*npath) | y2 (D*npath)]
  int* flags;                   // [0] NaN, [1] injected draws exhausted
};

int launch_batch_simple(const rp_pes* pes, const BatchArgs& a, int thermostat, int npoly,
                        hipStream_t s);
int launch_batch_evb(const rp_pes* pes, const BatchArgs& a, int thermostat, int npoly,
                     hipStream_t s);
int launch_batch_water(const rp_pes* pes, const BatchArgs& a, int thermostat, int npoly,
                       hipStream_t s);

inline int batch_dtiles(const rp_pes* pes) {
  const int D = pes->ndof;
  if (pes->kind == PES_WATER) return pes->natom == 3 ? 1 : pes->natom == 6 ? 2 : 4;
  if (pes->kind == PES_COUPLED2D || D <= 16) return 1;
  return 2;
}

// The EVB block as the fused kernel's MFMA kick reads it from LDS, for DC = 16 DT padded
// dofs: [kappa, 0 | cA | cB | lam | cub (DC each) | WA | WB (DC rows of stride DC + 1)],
// zero beyond the potential's own stride, so no operand load needs a clamp or a select.
__host__ __device__ constexpr int evb_kick_lds_size(int DC) { return 2 + 4 * DC + 2 * DC * (DC + 1); }

// Doubles of potential parameters staged in LDS by the fused kernel (the EVB matrices;
// other potentials read their few parameters from global memory).
inline int batch_pes_lds(const rp_pes* pes, int N) {
  (void)N;
  if (pes->kind != PES_EVB) return 0;
  return evb_kick_lds_size(16 * batch_dtiles(pes));
}

inline size_t batch_lds_bytes(int npad, int dt_tiles, int pes_lds) {
  return ((size_t)npad * (2 * dt_tiles * 16 + 1) + (size_t)npad * kMODE + (size_t)pes_lds) *
         sizeof(double);
}

}  // namespace rp

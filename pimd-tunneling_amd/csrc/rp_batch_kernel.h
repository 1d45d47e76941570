// rp_batch_kernel.h -- the fused persistent step kernel of the batched engine.
// Included by the per-potential instantiation units (rp_batch_k_*.hip).
#pragma once
#include "rp_batch.h"

namespace rp {


// f(integral_constant<int, 0..N-1>) expanded at compile time: the accumulator indices
// stay constants even where a loop body is too large for the unroller, so the
// accumulators never fall back to scratch memory.
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double batch_draw(const BatchArgs& A, int64_t noff, uint64_t strm,
                                             int64_t idx) {
  if (A.noise) return A.noise[noff + idx];
  return philox_normal(A.seed, strm, idx);
}

// Number of unit slots of this wave that hold a real unit (wave-uniform): slots u < NU.
template <int UPW, int W>
__device__ __forceinline__ int active_units(int wave, int units) {
  return max(0, min(UPW, (units - wave + W - 1) / W));
}

// f(integral_constant<int, nu>) for the wave-uniform nu in [0, UPW]: the GEMMs below are
// instantiated per active-unit count, so a wave issues no MFMA for an empty slot.
template <int UPW, typename F>
__device__ __forceinline__ void with_units(int nu, F&& f) {
  static_for<UPW + 1>([&](auto NU) __attribute__((always_inline)) {
    if (nu == decltype(NU)::value) f(NU);
  });
}

// acc = T S for the first NU units of this wave (see rp_batch.h for the unit layout);
// accumulators of the other slots are zero and are dropped by the consumers.  Measured on
// the 27-dof EVB at N = 192 (12 units in 16 slots of 8 waves): 8.64 -> 7.70 ms per 8-step
// launch against recomputing the last unit in the empty slots.
template <int DT, int UPW, int W, int NU>
__device__ __forceinline__ void gemm_T_n(const double* __restrict__ T, int NPAD, const double* S,
                                         int LD, int wave, int m, int g, d4 (&acc)[UPW][4]) {
#pragma unroll
  for (int u = 0; u < UPW; ++u)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[u][c] = d4{0.0, 0.0, 0.0, 0.0};
  if constexpr (NU > 0) {
    const int nk = NPAD >> 2;  // a multiple of 8: NPAD is a multiple of 32
    // Two register sets in ping-pong: the fragments of K-chunk kc+1 are loaded into one
    // set while the MFMAs of chunk kc read the other, so a load has a whole chunk of
    // MFMAs (16 x NU x 4 cycles x 16) to arrive and no register copy sits between them.
    const int units = (NPAD >> 5) * DT;
    int toff[NU], soff[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int unit = min(wave + u * W, units - 1);
      const int rp = unit / DT, dtile = unit - rp * DT;
      toff[u] = g * NPAD + rp * 32 + m;
      soff[u] = g * LD + dtile * 16 + m;
    }
    double a0[2][NU], a1[2][NU], bx[2][NU], bp[2][NU];
    auto load = [&](int set, int kc) __attribute__((always_inline)) {
      const double* Tk = T + (size_t)(4 * kc) * NPAD;  // wave-uniform chunk base
      const double* Sk = S + 4 * kc * LD;
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        a0[set][u] = Tk[toff[u]];
        a1[set][u] = Tk[toff[u] + 16];
        bx[set][u] = Sk[soff[u]];
        bp[set][u] = Sk[soff[u] + DT * 16];
      }
    };
    auto mma = [&](int set) __attribute__((always_inline)) {
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        acc[u][0] = mfma_f64(a0[set][u], bx[set][u], acc[u][0]);
        acc[u][1] = mfma_f64(a0[set][u], bp[set][u], acc[u][1]);
        acc[u][2] = mfma_f64(a1[set][u], bx[set][u], acc[u][2]);
        acc[u][3] = mfma_f64(a1[set][u], bp[set][u], acc[u][3]);
      }
    };
    load(0, 0);
    for (int kc = 0; kc < nk; kc += 2) {
      load(1, kc + 1);
      __builtin_amdgcn_sched_barrier(0);
      mma(0);
      __builtin_amdgcn_sched_barrier(0);
      load(0, kc + 2 < nk ? kc + 2 : kc + 1);  // the last prefetch re-reads a valid chunk
      __builtin_amdgcn_sched_barrier(0);
      mma(1);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

template <int DT, int UPW, int W>
__device__ __forceinline__ void gemm_T(const double* __restrict__ T, int NPAD, const double* S,
                                       int LD, int wave, int m, int g, d4 (&acc)[UPW][4]) {
  if constexpr (UPW == 1) {
    // one slot: a wave past the last unit recomputes the last unit (its result is dropped
    // by the consumers); the branch-free form keeps the one-unit kernels' registers
    gemm_T_n<DT, 1, W, 1>(T, NPAD, S, LD, wave, m, g, acc);
  } else {
    const int units = (NPAD >> 5) * DT;
    with_units<UPW>(active_units<UPW, W>(wave, units), [&](auto NU) __attribute__((always_inline)) {
      gemm_T_n<DT, UPW, W, decltype(NU)::value>(T, NPAD, S, LD, wave, m, g, acc);
    });
  }
}

// The folded transforms (even N, H = N/2, HP = H padded to 16).  A unit is a block of 16
// rows r of both halves and one dof tile; acc[u][0..1] hold the x / p columns of the odd
// half, acc[u][2..3] those of the even half.
//   forward:  odd r  (mode 2r+1) = sum_k Todd[r][k]  (x_k + x_{N-1-k})
//             even r (mode 2r+2) = sum_k Teven[r][k] (x_k - x_{N-1-k})
//             the sum and difference are formed from the two S rows as they are loaded;
//   backward: P_l = sum_r Todd[r][l] y_{2r},  Q_l = sum_r Teven[r][l] y_{2r+1}   (S rows = modes)
//             x_l = P_l + Q_l, x_{N-1-l} = P_l - Q_l                       (store_fold)
// The loop is the ping-pong of gemm_T over HP/4 K-chunks: half of gemm_T's MFMAs.
template <int DT, int UPW, int W, bool BWD, int NU>
__device__ __forceinline__ void gemm_fold_n(const double* __restrict__ Tf, int N, const double* S,
                                            int LD, int wave, int m, int g, d4 (&acc)[UPW][4]) {
  const int HP = ((N >> 1) + 15) / 16 * 16;
  const size_t Q = (size_t)HP * HP;
  const double* To = Tf + (BWD ? 2 * Q : 0);
  const double* Te = To + Q;
#pragma unroll
  for (int u = 0; u < UPW; ++u)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[u][c] = d4{0.0, 0.0, 0.0, 0.0};
  if constexpr (NU > 0) {
    const int nk = HP >> 2;  // a multiple of 4
    const int units = (HP >> 4) * DT;
    int toff[NU], soff[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int unit = min(wave + u * W, units - 1);
      const int rp = unit / DT, dtile = unit - rp * DT;
      toff[u] = g * HP + rp * 16 + m;
      soff[u] = dtile * 16 + m;
    }
    // T fragments (global, long latency) in ping-pong; S fragments (LDS) loaded just before
    // their MFMAs, which keeps the operand registers at 8 per unit
    double ao[2][NU], ae[2][NU];
    auto loadA = [&](int set, int kc) __attribute__((always_inline)) {
      const size_t tb = (size_t)(4 * kc) * HP;  // wave-uniform chunk base
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        ao[set][u] = To[tb + toff[u]];
        ae[set][u] = Te[tb + toff[u]];
      }
    };
    auto mma = [&](int set, int kc) __attribute__((always_inline)) {
      const int k = 4 * kc + g;
      const int ro = BWD ? 2 * k : k;                       // odd-half operand row
      const int re = BWD ? 2 * k + 1 : max(N - 1 - k, 0);   // even-half operand row
      double xo[NU], po[NU], xe[NU], pe[NU];
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        xo[u] = S[ro * LD + soff[u]];
        po[u] = S[ro * LD + DT * 16 + soff[u]];
        xe[u] = S[re * LD + soff[u]];
        pe[u] = S[re * LD + DT * 16 + soff[u]];
        if constexpr (!BWD) {
          // the fold in registers: rows k and N-1-k -> their sum (odd half) and difference
          // (even half).  Rows k >= H meet zero columns of the tables.
          const double sx = xo[u] + xe[u], dx = xo[u] - xe[u];
          const double sp = po[u] + pe[u], dp = po[u] - pe[u];
          xo[u] = sx;
          xe[u] = dx;
          po[u] = sp;
          pe[u] = dp;
        }
      }
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        acc[u][0] = mfma_f64(ao[set][u], xo[u], acc[u][0]);
        acc[u][1] = mfma_f64(ao[set][u], po[u], acc[u][1]);
        acc[u][2] = mfma_f64(ae[set][u], xe[u], acc[u][2]);
        acc[u][3] = mfma_f64(ae[set][u], pe[u], acc[u][3]);
      }
    };
    loadA(0, 0);
    for (int kc = 0; kc < nk; kc += 2) {
      loadA(1, kc + 1);
      __builtin_amdgcn_sched_barrier(0);
      mma(0, kc);
      __builtin_amdgcn_sched_barrier(0);
      loadA(0, kc + 2 < nk ? kc + 2 : kc + 1);  // the last prefetch re-reads a valid chunk
      __builtin_amdgcn_sched_barrier(0);
      mma(1, kc + 1);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

template <int DT, int UPW, int W, bool BWD>
__device__ __forceinline__ void gemm_fold(const double* __restrict__ Tf, int N, const double* S,
                                          int LD, int wave, int m, int g, d4 (&acc)[UPW][4]) {
  if constexpr (UPW == 1) {
    gemm_fold_n<DT, 1, W, BWD, 1>(Tf, N, S, LD, wave, m, g, acc);  // as in gemm_T
  } else {
    const int HP = ((N >> 1) + 15) / 16 * 16;
    const int units = (HP >> 4) * DT;
    with_units<UPW>(active_units<UPW, W>(wave, units), [&](auto NU) __attribute__((always_inline)) {
      gemm_fold_n<DT, UPW, W, BWD, decltype(NU)::value>(Tf, N, S, LD, wave, m, g, acc);
    });
  }
}

// S <- the backward folded accumulators: x_l = P + Q, x_{N-1-l} = P - Q (x and/or p).
template <int DT, int UPW, int W>
__device__ __forceinline__ void store_fold(double* S, int LD, int N, int wave, int m, int g,
                                           const d4 (&acc)[UPW][4], bool store_x) {
  const int H = N >> 1, HP = (H + 15) / 16 * 16;
  const int units = (HP >> 4) * DT;
  constexpr int DC = DT * 16;
#pragma unroll
  for (int u = 0; u < UPW; ++u) {
    const int unit = wave + u * W;
    if (unit < units) {
      const int rp = unit / DT, dtile = unit - rp * DT;
      const int col = dtile * 16 + m;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int l = rp * 16 + g + 4 * j;
        if (l < H) {
          if (store_x) {
            S[l * LD + col] = acc[u][0][j] + acc[u][2][j];
            S[(N - 1 - l) * LD + col] = acc[u][0][j] - acc[u][2][j];
          }
          S[l * LD + DC + col] = acc[u][1][j] + acc[u][3][j];
          S[(N - 1 - l) * LD + DC + col] = acc[u][1][j] - acc[u][3][j];
        }
      }
    }
  }
}

// S <- acc (x and/or p columns)
template <int DT, int UPW, int W>
__device__ __forceinline__ void store_acc(double* S, int LD, int NPAD, int wave, int m, int g,
                                          const d4 (&acc)[UPW][4], bool store_x) {
  const int units = (NPAD >> 5) * DT;
#pragma unroll
  for (int u = 0; u < UPW; ++u) {
    const int unit = wave + u * W;
    if (unit < units) {
      const int rp = unit / DT, dtile = unit - rp * DT;
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int i = rp * 32 + h * 16 + g + 4 * j;
          if (store_x) S[i * LD + dtile * 16 + m] = acc[u][2 * h][j];
          S[i * LD + DT * 16 + dtile * 16 + m] = acc[u][2 * h + 1][j];
        }
    }
  }
}

// The normals of one unit's two row halves: z[h][j] for lane group g at modes
// i = 32 rp + 16 h + g + 4 j of one dof (folded units: i = 2 (16 rp + g + 4 j) + h), draw
// index base + i.  When N is even the draws of
// modes (i, i+1) share one Philox block, so lane groups (2k, 2k+1) split the blocks of
// their pairs and swap halves with the partner lane 16 away: half the Philox/Box-Muller
// work.  The four blocks of a lane are independent chains, made side by side.
template <bool FOLD>
__device__ __forceinline__ void noise8(const BatchArgs& A, int64_t noff, uint64_t strm,
                                       int64_t base, int rp, int g, bool real_dof,
                                       double (&z)[2][4]) {
  const int N = A.N;
  if (A.noise || (N & 1)) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = FOLD ? 2 * (rp * 16 + g + 4 * j) + h : rp * 32 + 16 * h + g + 4 * j;
        z[h][j] = (real_dof && i < N) ? batch_draw(A, noff, strm, base + i) : 0.0;
      }
    return;
  }
  // base is even (the cursor and dof*N are multiples of N): pair (i even, i+1) = one block
  if constexpr (FOLD) {
    // folded units hold modes 2r and 2r+1 in the same lane: one block gives both
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i0 = 2 * (rp * 16 + g + 4 * j);
      double z0 = 0.0, z1 = 0.0;
      if (real_dof && i0 < N) philox_normal2(A.seed, strm, (uint64_t)(base + i0) >> 1, z0, z1);
      z[0][j] = z0;
      z[1][j] = z1;
    }
    return;
  }
  const int ibase = rp * 32;
  const bool odd = g & 1;
  double mine[2][2], other[2][2];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      // even lane: j = 0, 1; odd lane: j = 2, 3
      const int ie = ibase + 16 * h + (g & ~1) + 4 * (odd ? 2 + t : t);
      double z0 = 0.0, z1 = 0.0;
      if (real_dof && ie < N) philox_normal2(A.seed, strm, (uint64_t)(base + ie) >> 1, z0, z1);
      mine[h][t] = odd ? z1 : z0;
      other[h][t] = odd ? z0 : z1;
    }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const double r0 = __shfl_xor(other[h][0], 16), r1 = __shfl_xor(other[h][1], 16);
    if (!odd) {
      z[h][0] = mine[h][0]; z[h][1] = mine[h][1]; z[h][2] = r0; z[h][3] = r1;
    } else {
      z[h][0] = r0; z[h][1] = r1; z[h][2] = mine[h][0]; z[h][3] = mine[h][1];
    }
  }
}

// Normal-mode phase on the forward-transform accumulators: q = acc_x - beadvec,
// A(dt/2) [O A(dt/2)] per element, then S <- [q + beadvec | p~].
template <int DT, int UPW, int W, bool FOLD>
__device__ __forceinline__ void nm_phase(const BatchArgs& A, double* S, const double* MT,
                                         int wave, int m, int g, d4 (&acc)[UPW][4],
                                         const double* __restrict__ pa, int64_t noff,
                                         uint64_t strm, int64_t cur, bool with_o, bool& nan) {
  const int N = A.N, D = A.D, LD = A.LD;
  const int units = (A.NPAD >> 5) * DT;
  const double c = A.c, sq = A.sq, sp = A.sp;
  static_for<UPW>([&](auto U) __attribute__((always_inline)) {
    constexpr int u = decltype(U)::value;
    const int unit = wave + u * W;
    if (unit < units) {
      const int rp = unit / DT, dtile = unit - rp * DT;
      const int dof = dtile * 16 + m;
      const bool dreal = dof < D;
      const double mc = A.massd[dof], imc = 1.0 / mc, sqm = sqrt(mc);
      const double av = pa[dof], bvv = pa[DT * 16 + dof];
      double zz[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
      if (with_o) noise8<FOLD>(A, noff, strm, cur + (int64_t)dof * N, rp, g, dreal, zz);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double* z = zz[h];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int i = FOLD ? 2 * (rp * 16 + g + 4 * j) + h : rp * 32 + h * 16 + g + 4 * j;
          const double* md = MT + i * kMODE;
          const double wm = mc * md[MD_MU], iwm = imc * md[MD_IMU];
          const double bv = av * md[MD_U] + bvv * md[MD_W];
          double q = acc[u][2 * h][j] - bv;
          double pt = acc[u][2 * h + 1][j];
          double q1 = c * q + (sq * iwm) * pt;
          double p1 = -(sp * wm) * q + c * pt;
          nan |= (p1 != p1);
          if (with_o) {
            const double p2 = A.c1sq * p1 + (sqm * md[MD_LAM] * A.kappa) * z[j];
            q = c * q1 + (sq * iwm) * p2;
            pt = -(sp * wm) * q1 + c * p2;
            nan |= (pt != pt);
          } else {
            q = q1;
            pt = p1;
          }
          S[i * LD + dof] = q + bv;
          S[i * LD + DT * 16 + dof] = pt;
        }
      }
    }
  });
}

// B for the water cluster with one (bead, atom) per thread: the atom's gradient from its own
// molecule's Morse bonds and HOH bend and from every atom of the other molecules (charges,
// and Lennard-Jones between oxygens), the positions read from LDS.  Each pair term is made
// by both of its atoms (twice the pair arithmetic of one evaluation per bead), and a thread
// holds one atom's three components instead of a bead's 3 x 18: nothing spills.  The terms
// are those of water_eval (rp_pes.h).
template <int NMOL, int DT, int W>
__device__ __forceinline__ void water_kick_atoms(const BatchArgs& A, double* S, int tid, bool& nan) {
  const double* dp = A.pes.dp;
  const double De = dp[0], al = dp[1], re = dp[2], kb = dp[3], th0 = dp[4], eps = dp[5],
               sig = dp[6], qH = dp[7];
  constexpr int NA = 3 * NMOL;
  const int N = A.N, LD = A.LD;
  for (int it = tid; it < N * NA; it += W * 64) {
    const int b = it / NA, i = it - b * NA;
    const int m = i / 3, ai = i - 3 * m;  // molecule, and 0 = O, 1 / 2 = H
    const double* xb = S + b * LD;
    const double* xm = xb + 9 * m;
    // own molecule: bonds u_h = H_h - O and the bend
    double u[2][3], r[2], dEdr[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      double rr = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        u[h][d] = xm[3 * (h + 1) + d] - xm[d];
        rr += u[h][d] * u[h][d];
      }
      r[h] = sqrt(rr);
      const double e = exp(-al * (r[h] - re)), om = 1.0 - e;
      dEdr[h] = (2.0 * De * al) * om * e / r[h];
    }
    const double dot = u[0][0] * u[1][0] + u[0][1] * u[1][1] + u[0][2] * u[1][2];
    const double ir01 = 1.0 / (r[0] * r[1]);
    const double cs = dot * ir01;
    const double dth = acos(cs) - th0;
    const double dEdc = -kb * dth / sqrt(1.0 - cs * cs);
    const double i00 = cs / (r[0] * r[0]), i11 = cs / (r[1] * r[1]);
    double g[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double t0 = dEdr[0] * u[0][d], t1 = dEdr[1] * u[1][d];
      const double d0 = dEdc * (u[1][d] * ir01 - i00 * u[0][d]);
      const double d1 = dEdc * (u[0][d] * ir01 - i11 * u[1][d]);
      g[d] = ai == 0 ? -(t0 + t1) - (d0 + d1) : (ai == 1 ? t0 + d0 : t1 + d1);
    }
    // the other molecules
    const double si = ai == 0 ? -2.0 : 1.0;
    const double xi0 = xb[3 * i], xi1 = xb[3 * i + 1], xi2 = xb[3 * i + 2];
    for (int j = 0; j < NA; ++j) {
      if (j / 3 == m) continue;
      const double dx0 = xi0 - xb[3 * j], dx1 = xi1 - xb[3 * j + 1], dx2 = xi2 - xb[3 * j + 2];
      const double rr = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
      const double ir2 = 1.0 / rr, ir = sqrt(ir2);
      const double sij = si * ((j % 3 == 0) ? -2.0 : 1.0);
      double dEr = (-qH * qH * sij) * ir2 * ir;  // (dE/dr) / r
      if (ai == 0 && j % 3 == 0) {
        const double s2 = (sig * sig) * ir2, s6 = s2 * s2 * s2;
        dEr += (4.0 * eps) * (-12.0 * s6 * s6 + 6.0 * s6) * ir2;
      }
      g[0] += dEr * dx0;
      g[1] += dEr * dx1;
      g[2] += dEr * dx2;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double* pp = S + b * LD + DT * 16 + 3 * i + d;
      const double np = *pp - g[d] * A.dt;
      nan |= (np != np);
      *pp = np;
    }
  }
}

// B: p -= dt grad V(x) for every bead (verletmodule.f90:565-585)
template <int KIND, int MAXD, int DT, int W>
__device__ __forceinline__ void pes_kick(const BatchArgs& A, double* S, const double* PL, int tid,
                                         bool& nan) {
  static_assert(KIND != PES_EVB, "the EVB kick runs on the matrix cores (evb_kick_mfma)");
  (void)PL;
  if constexpr (KIND == PES_WATER) {
    water_kick_atoms<MAXD / 9, DT, W>(A, S, tid, nan);
    return;
  }
  const int N = A.N, D = A.D, LD = A.LD;
  for (int b = tid; b < N; b += W * 64) {
    double xs[MAXD], gs[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) xs[d] = (d < D) ? S[b * LD + d] : 0.0;
    (void)eval_kind<KIND, MAXD>(A.pes, xs, gs);
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
      if (d < D) {
        double np = S[b * LD + DT * 16 + d] - gs[d] * A.dt;
        nan |= (np != np);
        S[b * LD + DT * 16 + d] = np;
      }
    }
  }
}

// The EVB parameter block (evb_eval_at's layout, stride M) -> the kick's LDS layout
// (evb_kick_lds_size): vectors and matrices zero-padded to DC = 16 DT, matrix rows of
// stride DC + 1 (the y GEMM reads a column of 16 rows per lane group: an odd stride puts
// them in distinct banks).
template <int M, int DT, int W>
__device__ __forceinline__ void stage_evb_kick(const double* __restrict__ dp, double* PL, int tid) {
  constexpr int DC = DT * 16, LW = DC + 1, NV = 2 + 4 * DC;
  const double* WA = dp + 1;
  const double* cA = WA + M * M;
  const double* WB = cA + M;
  const double* cB = WB + M * M;
  const double* lam = cB + M;
  const double* cub = lam + M;
  for (int e = tid; e < evb_kick_lds_size(DC); e += W * 64) {
    double v = 0.0;
    if (e == 0) {
      v = dp[0];
    } else if (e >= 2 && e < NV) {
      const int k = (e - 2) / DC, i = (e - 2) - k * DC;
      const double* vec = k == 0 ? cA : k == 1 ? cB : k == 2 ? lam : cub;
      if (i < M) v = vec[i];
    } else if (e >= NV) {
      const int f = e - NV, w = f / (DC * LW), r = f - w * (DC * LW);
      const int i = r / LW, j = r - i * LW;
      if (i < M && j < M) v = (w ? WB : WA)[i * M + j];
    }
    PL[e] = v;
  }
}

// B for the EVB on the matrix cores.  For a tile of 16 beads, both wells' coordinates
// y = Wt x - cw are one GEMM (rows: 2 x 16*RT well rows; K: the dofs) and the mixed
// gradient wa WtA^T hA + wb WtB^T hB is a second GEMM (rows: dofs; K: both wells' rows)
// whose B operand is the first GEMM's accumulators, scaled in place: item j of row tile
// rt holds row 16 rt + g + 4 j of bead m, exactly the operand of K-chunk 4 rt + j.
// The per-bead well energies are reduced over the four lane groups with two xor
// shuffles.  Tiles are dealt to waves round-robin.  The matrices come zero-padded from
// stage_evb_kick, so padded rows give y = h = 0 and padded dofs no gradient.
template <int M, int DT, int W>
__device__ __forceinline__ void evb_kick_mfma(const BatchArgs& A, double* S, const double* PL,
                                              int wave, int m, int g, bool& nan) {
  constexpr int RT = DT;             // row tiles per well
  constexpr int KY = 4 * RT;         // K-chunks of the y GEMM (dofs, padded to 16 RT)
  constexpr int KG = 8 * RT;         // K-chunks of the gradient GEMM (both wells' rows)
  constexpr int DC = DT * 16, LW = DC + 1;
  const int N = A.N, D = A.D, LD = A.LD;
  const double kappa = PL[0];
  const double* cA = PL + 2;
  const double* cB = cA + DC;
  const double* lam = cB + DC;
  const double* cub = lam + DC;
  const double* WA = cub + DC;
  const double* WB = WA + DC * LW;
  const int ntile = A.NPAD >> 4;
  for (int t = wave; t < ntile; t += W) {
    const int b0 = t * 16;
    d4 ay[2 * RT];
#pragma unroll
    for (int r = 0; r < 2 * RT; ++r) ay[r] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kc = 0; kc < KY; ++kc) {
      // keep at most two K-chunks of operands in flight: the kick's registers stay below the
      // budget of two workgroups per CU
      if (kc % 2 == 0) __builtin_amdgcn_sched_barrier(0);
      const int j = 4 * kc + g;
      const double bx = S[(b0 + m) * LD + j];
#pragma unroll
      for (int r = 0; r < 2 * RT; ++r) {
        const double* Wt = r < RT ? WA : WB;
        ay[r] = mfma_f64(Wt[((r % RT) * 16 + m) * LW + j], bx, ay[r]);
      }
    }
    // h = lam y + 3 c y^2 and the well energies, per bead m
    double va = 0.0, vb = 0.0;
#pragma unroll
    for (int r = 0; r < 2 * RT; ++r)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const int i = (r % RT) * 16 + g + 4 * jj;
        const double li = lam[i], ci = cub[i];
        const double y = ay[r][jj] - (r < RT ? cA[i] : cB[i]), y2 = y * y;
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
    // gradient: G[dof][bead] = sum over both wells' rows i of Wt[i][dof] h_i
    d4 ag[DT];
#pragma unroll
    for (int d = 0; d < DT; ++d) ag[d] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kc = 0; kc < KG; ++kc) {
      if (kc % 4 == 0) __builtin_amdgcn_sched_barrier(0);
      const int r = kc / 4, jj = kc % 4;
      const int i = (r % RT) * 16 + 4 * jj + g;
      const double* Wt = r < RT ? WA : WB;
      const double bh = ay[r][jj];
#pragma unroll
      for (int d = 0; d < DT; ++d) ag[d] = mfma_f64(Wt[i * LW + d * 16 + m], bh, ag[d]);
    }
    const int bead = b0 + m;
#pragma unroll
    for (int d = 0; d < DT; ++d)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const int dof = d * 16 + g + 4 * jj;
        if (bead < N && dof < D) {
          const double np = S[bead * LD + DC + dof] - ag[d][jj] * A.dt;
          nan |= (np != np);
          S[bead * LD + DC + dof] = np;
        }
      }
  }
}

// NM-space momenta of a resample / init_path (verletmodule.f90:102-109, 211-219):
// v_i = sqrt(1/betan) z sqrt(beadmass_i), draws in (dim outer, atom inner) blocks of N.
template <int DT, int W>
__device__ __forceinline__ void write_resample(const BatchArgs& A, double* S, const double* MT,
                                               int tid, int64_t noff, uint64_t strm,
                                               int64_t cur) {
  constexpr int DC = DT * 16;
  const int N = A.N, NPAD = A.NPAD, D = A.D, LD = A.LD, natom = A.natom, ndim = A.ndim;
  for (int e = tid; e < NPAD * DC; e += W * 64) {
    const int c = e / NPAD, i = e - c * NPAD;
    double v = 0.0;
    if (i < N && c < D) {
      const int atom = c / ndim, dim = c - atom * ndim;
      const double z = batch_draw(A, noff, strm, cur + (int64_t)(dim * natom + atom) * N + i);
      v = (A.stdev0 * z) * (sqrt(A.massd[c]) * MT[i * kMODE + MD_LAM] * A.tau);
    }
    S[i * LD + DC + c] = v;
  }
}

// Phases of one step.  Every phase is: pre-op, sync, acc = T S, sync, post-op, sync.
//   Langevin (time_step_pile):  [kick] T [NM: A O A] T [store]
//   Andersen (time_step_nm):    ([resample] T [store p]) T [NM: A] T [store] [kick]
//                               T [NM: A] T [store]
//   re-initialisation:          [resample] T [store p, path spline -> x]
enum : int { PH_FWD_PILE = 0, PH_FWD_A = 1, PH_BWD = 2, PH_RESAMPLE = 3, PH_REINIT = 4 };

// One-unit 4-wave kernels (small bead counts, whose workgroups need little LDS) are held to
// two waves per SIMD so that two polymers share a CU.
template <int KIND, int MAXD, int DT, int UPW, int W, bool LANGEVIN, bool FOLD>
__global__ void __launch_bounds__(W * 64, (W <= 4 && UPW == 1) ? 2 : 1)
    batch_kernel(BatchArgs A) {
  extern __shared__ double S[];
  __shared__ int s_flag;
  constexpr int DC = DT * 16;
  const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4;
  (void)m;  // lane coordinates of the MFMA fragments: unused by the opaque-copy instances
  (void)g;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform (SGPR)
  const int poly = A.first + blockIdx.x;
  const int N = A.N, NPAD = A.NPAD, D = A.D, LD = A.LD;
  double* MT = S + NPAD * LD;
  double* PL = MT + NPAD * kMODE;  // EVB block (stage_evb_kick's layout), KIND == PES_EVB only
  int64_t* pi = A.pi64 + (size_t)poly * kPI64;
  double* st = A.state + pi[PI_STATE_OFF];
  const double* pa = A.pdata + (size_t)poly * 3 * DC;
  const uint64_t strm = A.stream[poly];
  const int64_t noff = pi[PI_NOISE_OFF];
  // the thermostat is a compile-time choice: each kernel carries one step sequence only
  constexpr bool langevin = LANGEVIN;
  int64_t cur = pi[PI_CURSOR];
  bool nan = false;
  // state -> LDS (zero padding); per-mode table -> LDS
  for (int e = tid; e < NPAD * 2 * DC; e += W * 64) {
    const int c = e / NPAD, i = e - c * NPAD;
    const int d = c < DC ? c : c - DC;
    double v = 0.0;
    if (i < N && d < D) v = st[(c < DC ? 0 : (size_t)D * N) + (size_t)d * N + i];
    S[i * LD + c] = v;
  }
  for (int e = tid; e < NPAD * kMODE; e += W * 64) MT[e] = A.modes[e];
  if constexpr (KIND == PES_EVB) stage_evb_kick<MAXD, DT, W>(A.pes.dp, PL, tid);
  if (tid == 0) s_flag = 0;
  double tisum = 0.0;
  int64_t ticnt = 0, nre = 0;
  int64_t acount = 0, rkick = 0, kcur = 0;
  if (!langevin) {
    acount = pi[PI_ACOUNT];
    rkick = pi[PI_RKICK];
    kcur = pi[PI_KCURSOR];
  }
  const int64_t koff = pi[PI_KICK_OFF];
  const int64_t need = (int64_t)N * D;  // draws per O-step / resample / re-initialisation
  __syncthreads();
  bool stop = false;
  const int64_t step0 = A.step0p ? *A.step0p : A.step0;
  for (int s = 0; s < A.nsteps; ++s) {
    const int64_t ii = step0 + s + 1;  // the reference's step counter
    // the step's phase list, 4 bits per phase in one register (an indexed array would live
    // in scratch memory): at most 6 phases
    uint32_t ph = 0;
    int nph = 0;
    auto push = [&](int p) __attribute__((always_inline)) { ph |= (uint32_t)p << (4 * nph++); };
    bool resample = false;
    if (langevin) {
      if (A.noise && cur + need > A.noise_cap) { stop = true; break; }
      push(PH_FWD_PILE);
      push(PH_BWD);
    } else {
      // propagate_pimd_nm (verletmodule.f90:202-234): Poisson-timed resample first
      if (rkick < 0) {
        if (A.kicks && kcur >= A.kick_cap) { stop = true; break; }
        rkick = A.kicks ? A.kicks[koff + kcur]
                        : philox_poisson(A.seed, strm, kcur, (double)A.noutput);
        ++kcur;
      }
      acount = acount + 1;
      if (acount >= rkick) {
        if ((A.noise && cur + need > A.noise_cap) || (A.kicks && kcur >= A.kick_cap)) {
          stop = true;
          break;
        }
        resample = true;
        push(PH_RESAMPLE);
      }
      push(PH_FWD_A);
      push(PH_BWD);
      push(PH_FWD_A);
      push(PH_BWD);
    }
    int nbwd = 0;
    bool ti_done = false;
#pragma unroll 1
    for (int k = 0; k < nph; ++k) {
      const int p = (int)((ph >> (4 * k)) & 15u);
      // With two or more units per wave, the lane indices are re-derived from an opaque
      // copy in every phase: the per-lane LDS addresses are then recomputed where they are
      // used instead of being hoisted out of the step loop and spilled (measured: 16 % less
      // time per launch at N = 256, and under 256 VGPRs for two workgroups per CU at N = 96).
      // One-unit waves keep the hoisted addresses: the opaque copy removes their few spill
      // slots (254 -> 237 VGPRs on the 27-dof EVB) but costs 4-5 % per launch at N = 32-128.
      int lane_k = lane;
      if constexpr (UPW >= 2 || !LANGEVIN) asm volatile("" : "+v"(lane_k));
      const int m = lane_k & 15, g = lane_k >> 4, tid = wave * 64 + lane_k;
      if (p == PH_FWD_PILE || (p == PH_FWD_A && nbwd == 1)) {
        if constexpr (KIND == PES_EVB) {  // B(dt) on the matrix cores
          evb_kick_mfma<MAXD, DT, W>(A, S, PL, wave, m, g, nan);
        } else {
          pes_kick<KIND, MAXD, DT, W>(A, S, PL, tid, nan);
        }
        __syncthreads();
      } else if (p == PH_RESAMPLE || p == PH_REINIT) {
        write_resample<DT, W>(A, S, MT, tid, noff, strm, cur);
        __syncthreads();
      }
      d4 acc[UPW][4];
      const bool fwd = p == PH_FWD_PILE || p == PH_FWD_A;
      if constexpr (FOLD) {
        if (fwd) {
          gemm_fold<DT, UPW, W, false>(A.Tf, N, S, LD, wave, m, g, acc);
        } else {
          gemm_fold<DT, UPW, W, true>(A.Tf, N, S, LD, wave, m, g, acc);
        }
      } else {
        gemm_T<DT, UPW, W>(A.T, NPAD, S, LD, wave, m, g, acc);
      }
      __syncthreads();
      if (p == PH_FWD_PILE || p == PH_FWD_A) {
        nm_phase<DT, UPW, W, FOLD>(A, S, MT, wave, m, g, acc, pa, noff, strm, cur, p == PH_FWD_PILE, nan);
        if (p == PH_FWD_PILE) cur += need;
      } else {
        if constexpr (FOLD)
          store_fold<DT, UPW, W>(S, LD, N, wave, m, g, acc, p == PH_BWD);
        else
          store_acc<DT, UPW, W>(S, LD, NPAD, wave, m, g, acc, p == PH_BWD);
        if (p == PH_BWD) ++nbwd;
        if (p == PH_REINIT) {
          // init_path positions: x_k = splint(path, (k-1) xi/(n-1)) (verletmodule.f90:40-48)
          const double xi = A.pf64[(size_t)poly * kPF64 + PF_XI];
          const double* lampath = A.pathd;
          for (int e = tid; e < N * D; e += W * 64) {
            const int cc = e / N, i = e - cc * N;
            const double xe = (double)i * xi / (double)(N - 1);
            S[i * LD + cc] = splint_dev(lampath, lampath + A.npath + (size_t)cc * A.npath,
                                        lampath + A.npath * (1 + D) + (size_t)cc * A.npath,
                                        A.npath, xe);
          }
        }
        if (p == PH_RESAMPLE || p == PH_REINIT) cur += need;
      }
      __syncthreads();
      // TI estimator after the step (verletmodule.f90:397-410, 236-244)
      if (k == nph - 1 && !ti_done && ii > A.imin) {
        ti_done = true;
        if (wave == 0) {
          // sum over dofs of m (-x_N) db/dxi: one dof per lane, then an xor-shuffle reduction
          // (the reference's dim-outer order differs only in rounding)
          double contr = 0.0;
          for (int d = lane; d < D; d += 64) contr += A.massd[d] * (-S[(N - 1) * LD + d]) * pa[2 * DC + d];
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) contr += __shfl_xor(contr, o);
          const bool ok = !langevin || fabs(contr) < A.limit || A.limit < 0.0;
          if (ok) {
            tisum = tisum + contr;
            ++ticnt;
          }
          if (lane == 0) s_flag = ok ? 0 : 1;
        }
        if (langevin && A.limit >= 0.0) {
          __syncthreads();
          if (s_flag) {
            if (A.noise && cur + need > A.noise_cap) {
              stop = true;
            } else {
              push(PH_REINIT);  // one more phase: re-initialise the path
              ++nre;
            }
          }
        }
      }
    }
    if (stop) break;
    if (resample) {
      rkick = A.kicks ? A.kicks[koff + kcur] : philox_poisson(A.seed, strm, kcur, (double)A.noutput);
      ++kcur;
      acount = 0;
    }
  }
  if (stop && tid == 0) atomicOr(&A.flags[1], 1);
  __syncthreads();
  // LDS -> state
  for (int e = tid; e < N * 2 * D; e += W * 64) {
    const int c = e / N, i = e - c * N;
    const int d = c < D ? c : c - D;
    const int sc = c < D ? d : DC + d;
    st[(c < D ? 0 : (size_t)D * N) + (size_t)d * N + i] = S[i * LD + sc];
  }
  if (nan) atomicOr(&A.flags[0], 1);
  if (tid == 0) {
    A.pf64[(size_t)poly * kPF64 + PF_TISUM] += tisum;
    pi[PI_TICNT] += ticnt;
    pi[PI_NREINIT] += nre;
    pi[PI_CURSOR] = cur;
    if (!langevin) {
      pi[PI_ACOUNT] = acount;
      pi[PI_RKICK] = rkick;
      pi[PI_KCURSOR] = kcur;
    }
  }
}

template <int KIND, int MAXD, int DT, bool LANG, bool FOLD>
int launch_batch_thermo(const BatchArgs& a, int npoly, hipStream_t s) {
  const int units = (a.NPAD >> 5) * DT;
  const size_t lds = batch_lds_bytes(a.NPAD, DT, KIND == PES_EVB ? evb_kick_lds_size(DT * 16) : 0);
  auto go = [&](auto kern, int waves) -> int {
    RP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    kern<<<npoly, waves * 64, lds, s>>>(a);
    RP_HIP(hipGetLastError());
    return RP_OK;
  };
  if constexpr (DT <= 2) {
    // one unit per wave, four 2-wave workgroups per CU (at most 256 VGPRs each): small
    // polymers whose 2 units would leave half of a 4-wave workgroup idle
    if (a.waves == 2 && units <= 2) return go(batch_kernel<KIND, MAXD, DT, 1, 2, LANG, FOLD>, 2);
  }
  if (a.waves == 8) {  // two waves per SIMD: units spread over 8 waves, at most 2 each
    const int upw = (units + 7) / 8;
    if (upw <= 1) return go(batch_kernel<KIND, MAXD, DT, 1, 8, LANG, FOLD>, 8);
    if (upw <= 2) return go(batch_kernel<KIND, MAXD, DT, 2, 8, LANG, FOLD>, 8);
    return set_error(RP_ERR_ARG, "batch: %d MFMA units are too many for 8 waves", units);
  }
  const int upw = (units + 3) / 4;
  if (upw <= 1) return go(batch_kernel<KIND, MAXD, DT, 1, 4, LANG, FOLD>, 4);
  if (upw <= 2) return go(batch_kernel<KIND, MAXD, DT, 2, 4, LANG, FOLD>, 4);
  if (upw <= 3) return go(batch_kernel<KIND, MAXD, DT, 3, 4, LANG, FOLD>, 4);
  if (upw <= 4) return go(batch_kernel<KIND, MAXD, DT, 4, 4, LANG, FOLD>, 4);
  return set_error(RP_ERR_ARG, "batch: %d MFMA units per wave is more than supported", upw);
}

// ALLOW_FOLD = false keeps a potential on the full transforms (fewer kernel instances).
template <int KIND, int MAXD, int DT, bool ALLOW_FOLD = true>
int launch_batch_kind(const BatchArgs& a, int thermostat, int npoly, hipStream_t s) {
  const bool lang = thermostat == RP_THERMO_LANGEVIN;
  if constexpr (ALLOW_FOLD) {
    if (a.fold) {
      if (lang) return launch_batch_thermo<KIND, MAXD, DT, true, true>(a, npoly, s);
      return launch_batch_thermo<KIND, MAXD, DT, false, true>(a, npoly, s);
    }
  }
  if (lang) return launch_batch_thermo<KIND, MAXD, DT, true, false>(a, npoly, s);
  return launch_batch_thermo<KIND, MAXD, DT, false, false>(a, npoly, s);
}

}  // namespace rp

// rp_philox.h -- counter-based Philox4x32-10 normals and Poisson draws.
//
// Written from the published description (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC11).  A trajectory owns one stream id; its
// k-th N(0,1) draw is a pure function of (seed, stream, k), so any kernel can make
// the draw the reference's sequential VSL stream would make at that position
// (verletmodule.f90:55,106,215,641) without carrying generator state.
//   block b = k >> 1 -> Philox(ctr = {b_lo, b_hi, stream_lo, stream_hi},
//                              key = {seed_lo, seed_hi}) -> two 53-bit uniforms
//   Box-Muller: draw 2b uses r cos(2 pi u1), draw 2b+1 uses r sin(2 pi u1).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace rp {

__host__ __device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
  }
}

__host__ __device__ __forceinline__ void philox_uniform2(uint64_t seed, uint64_t stream,
                                                         uint64_t block, double& u0, double& u1) {
  uint32_t c[4] = {(uint32_t)block, (uint32_t)(block >> 32), (uint32_t)stream,
                   (uint32_t)(stream >> 32)};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  uint64_t a = ((uint64_t)c[1] << 32) | c[0], b = ((uint64_t)c[3] << 32) | c[2];
  u0 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
  u1 = ((double)(b >> 11) + 0.5) * 0x1.0p-53;
}

// Both normals of one block (draws 2*block and 2*block+1).  The angle 2 pi u1 goes
// through sincospi(2 u1), whose argument reduction is exact.
__device__ __forceinline__ void philox_normal2(uint64_t seed, uint64_t stream, uint64_t block,
                                               double& z0, double& z1) {
  double u0, u1;
  philox_uniform2(seed, stream, block, u0, u1);
  double r = sqrt(-2.0 * log(u0));
  double s, c;
  sincospi(2.0 * u1, &s, &c);
  z0 = r * c;
  z1 = r * s;
}

__device__ __forceinline__ double philox_normal(uint64_t seed, uint64_t stream, int64_t index) {
  double z0, z1;
  philox_normal2(seed, stream, (uint64_t)index >> 1, z0, z1);
  return (index & 1) ? z1 : z0;
}

// Poisson(mean) on the stream's disjoint high half: inversion below mean 16, a rounded
// normal approximation above (VSL's POISNORM, which the reference selects at
// verletmodule.f90:362, is likewise a normal approximation for large means).
__host__ __device__ __forceinline__ int32_t philox_poisson(uint64_t seed, uint64_t stream,
                                                           int64_t index, double mean) {
  double u0, u1;
  philox_uniform2(seed, stream ^ 0x8000000000000000ull, (uint64_t)index, u0, u1);
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

}  // namespace rp

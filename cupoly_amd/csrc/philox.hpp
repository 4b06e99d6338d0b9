// Philox4x32-10 draws and the three distributions, shared by the sampler
// kernels (sampler.hip) and the fused sample-and-multiply kernel (dgt.hip).
// Counter layout and distributions: see sampler.hip.
#pragma once
#include "internal.hpp"
#include "cdt.inc"

// The discrete-Gaussian cumulative table travels by value in the kernel
// arguments (304 bytes), so every translation unit sees the same data.
struct cp_cdt {
  uint64_t v[2 * CUPOLY_GAUSS_BOUND];
};

static inline cp_cdt cp_make_cdt() {
  cp_cdt c;
  for (int i = 0; i < 2 * CUPOLY_GAUSS_BOUND; i++) c.v[i] = cupoly_gauss_cdt[i];
  return c;
}

__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t (&o)[4]) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  o[0] = c0;
  o[1] = c1;
  o[2] = c2;
  o[3] = c3;
}

__device__ __forceinline__ void draw(uint64_t seed, uint64_t stream, uint32_t coef, uint32_t tag,
                                     uint32_t rid, uint32_t attempt, uint32_t (&o)[4]) {
  philox4x32_10(coef, (tag << 24) | (rid << 16) | attempt, (uint32_t)stream,
                (uint32_t)(stream >> 32), (uint32_t)seed, (uint32_t)(seed >> 32), o);
}

__device__ __forceinline__ uint64_t uniform_coef(uint64_t seed, uint64_t stream, uint32_t coef,
                                                 uint32_t rid, uint64_t q) {
  const uint64_t lim = (~0ULL / q) * q;
  for (uint32_t att = 0;; att++) {  // leaves with probability 1 - 2^-60 per attempt
    uint32_t w[4];
    draw(seed, stream, coef, 1, rid, att, w);
    const uint64_t x0 = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
    if (x0 < lim) return x0 % q;
    const uint64_t x1 = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
    if (x1 < lim) return x1 % q;
  }
}

__device__ __forceinline__ int ternary_coef(uint64_t seed, uint64_t stream, uint32_t coef) {
  for (uint32_t att = 0;; att++) {
    uint32_t w[4];
    draw(seed, stream, coef, 2, 0, att, w);
#pragma unroll
    for (int h = 0; h < 4; h++)
      if (w[h] < 4294967295u) return (int)(w[h] % 3u) - 1;
  }
}

__device__ __forceinline__ int gaussian_coef(uint64_t seed, uint64_t stream, uint32_t coef,
                                             const cp_cdt &cdt) {
  uint32_t w[4];
  draw(seed, stream, coef, 3, 0, 0, w);
  const uint64_t u = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
  int idx = 0;
#pragma unroll
  for (int i = 0; i < 2 * CUPOLY_GAUSS_BOUND; i++) idx += (u >= cdt.v[i]);
  return idx - CUPOLY_GAUSS_BOUND;
}

__device__ __forceinline__ int narrow_coef(uint64_t seed, uint64_t stream, uint32_t coef) {
  for (uint32_t att = 0;; att++) {
    uint32_t w[4];
    draw(seed, stream, coef, 4, 0, att, w);
#pragma unroll
    for (int h = 0; h < 4; h++)
      if (w[h] < 4294967295u) return (int)(w[h] % 3u);
  }
}

__device__ __forceinline__ int binary_coef(uint64_t seed, uint64_t stream, uint32_t coef) {
  uint32_t w[4];
  draw(seed, stream, coef, 5, 0, 0, w);
  return (int)(w[0] & 1u);
}

// Fixed-weight sampling key of coefficient coef (< 2^16): 48 random bits
// above the index, so keys are distinct; neg = sign of a selected coefficient.
__device__ __forceinline__ uint64_t fixed_weight_key(uint64_t seed, uint64_t stream, uint32_t coef,
                                                     bool &neg) {
  uint32_t w[4];
  draw(seed, stream, coef, 6, 0, 0, w);
  neg = (w[2] & 1u) != 0;
  return ((((uint64_t)w[0] << 16) | (w[1] >> 16)) << 16) | (coef & 0xFFFFu);
}

// small signed value -> residue mod q
__device__ __forceinline__ uint64_t small_mod(int e, uint64_t q) {
  return e < 0 ? q - (uint64_t)(-e) : (uint64_t)e;
}


// Counter-based sampling on Philox4x32-10.
//
// Reference: Sampler (src/cuda/sampler.cu, include/cuPoly/cuda/sampler.h)
// keeps one cuRAND XORWOW state per coefficient (48 bytes each, seeded with
// 42) and fills polynomials that are then moved to the DGT domain.  Here the
// generator is stateless: the draw for coefficient j of stream s under seed
// k is Philox4x32-10 (Salmon et al., SC'11) of
//     counter = (j, tag << 24 | residue << 16 | attempt, s_lo, s_hi),
//     key     = (k_lo, k_hi),
// so any launch shape gives the same polynomial and a CPU can recompute it.
//   uniform mod q_i  (tag 1): 64-bit words, rejection above floor(2^64/q)q;
//   ternary {-1,0,1} (tag 2): 32-bit words, rejection of 2^32 - 1, w mod 3 - 1;
//   Gaussian         (tag 3): sigma 3.2 truncated at |e| <= 19 by inversion of
//                              a 64-bit cumulative table (generated with
//                              60-digit arithmetic).
// Small values are written to every residue of the base as e mod q_i, like the
// reference's samplers (sampler.cu:94-97, 485-487).
#include "internal.hpp"
#include "cdt.inc"

__constant__ uint64_t c_gauss_cdt[2 * CUPOLY_GAUSS_BOUND];

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

__device__ __forceinline__ int gaussian_coef(uint64_t seed, uint64_t stream, uint32_t coef) {
  uint32_t w[4];
  draw(seed, stream, coef, 3, 0, 0, w);
  const uint64_t u = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
  int idx = 0;
#pragma unroll
  for (int i = 0; i < 2 * CUPOLY_GAUSS_BOUND; i++) idx += (u >= c_gauss_cdt[i]);
  return idx - CUPOLY_GAUSS_BOUND;
}

// one thread per (slot, residue): coefficients slot and slot + ng
__global__ void __launch_bounds__(256)
    k_sample_uniform(cp_gi *__restrict__ out, int ng, int nres, int roff,
                     const cp_prime *__restrict__ primes, uint64_t seed, uint64_t stream) {
  const long total = (long)ng * nres;
  for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < total;
       s += (long)gridDim.x * blockDim.x) {
    const int r = (int)(s / ng), slot = (int)(s % ng);
    const int rid = roff + r;
    const uint64_t q = primes[rid].p;
    cp_gi v;
    v.re = uniform_coef(seed, stream, (uint32_t)slot, (uint32_t)rid, q);
    v.imag = uniform_coef(seed, stream, (uint32_t)(slot + ng), (uint32_t)rid, q);
    out[s] = v;
  }
}

template <int KIND>
__global__ void __launch_bounds__(256)
    k_sample_small(cp_gi *__restrict__ out, int ng, int nres, int roff,
                   const cp_prime *__restrict__ primes, uint64_t seed, uint64_t stream) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= ng) return;
  int e0, e1;
  if constexpr (KIND == CUPOLY_TERNARY) {
    e0 = ternary_coef(seed, stream, (uint32_t)slot);
    e1 = ternary_coef(seed, stream, (uint32_t)(slot + ng));
  } else {
    e0 = gaussian_coef(seed, stream, (uint32_t)slot);
    e1 = gaussian_coef(seed, stream, (uint32_t)(slot + ng));
  }
  for (int r = 0; r < nres; r++) {
    const uint64_t q = primes[roff + r].p;
    cp_gi v;
    v.re = e0 < 0 ? q - (uint64_t)(-e0) : (uint64_t)e0;
    v.imag = e1 < 0 ? q - (uint64_t)(-e1) : (uint64_t)e1;
    out[(size_t)r * ng + slot] = v;
  }
}

static bool g_cdt_ready_dev[64];

static int ensure_cdt(int dev) {
  if (dev >= 0 && dev < 64 && g_cdt_ready_dev[dev]) return CUPOLY_OK;
  CP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gauss_cdt), cupoly_gauss_cdt, sizeof(cupoly_gauss_cdt)));
  if (dev >= 0 && dev < 64) g_cdt_ready_dev[dev] = true;
  return CUPOLY_OK;
}

extern "C" int cupoly_sample_coeff(cupoly_ctx *c, cupoly_gi *out, int kind, uint64_t seed,
                                   uint64_t stream, int base) {
  if (!c || !out) return cp_fail(CUPOLY_ERR_ARG, "sample: null argument");
  int off, cnt;
  if (!cp_base_range(c->e, base, &off, &cnt)) return cp_fail(CUPOLY_ERR_UNSUPPORTED, "Unknown base");
  const int ng = c->e->ng;
  switch (kind) {
    case CUPOLY_UNIFORM: {
      long total = (long)ng * cnt;
      int grid = (int)((total + 255) / 256);
      hipLaunchKernelGGL(k_sample_uniform, dim3(grid), dim3(256), 0, c->stream, (cp_gi *)out, ng,
                         cnt, off, c->e->d_primes, seed, stream);
      break;
    }
    case CUPOLY_TERNARY:
      hipLaunchKernelGGL(k_sample_small<CUPOLY_TERNARY>, dim3((ng + 255) / 256), dim3(256), 0,
                         c->stream, (cp_gi *)out, ng, cnt, off, c->e->d_primes, seed, stream);
      break;
    case CUPOLY_DISCRETE_GAUSSIAN: {
      int rc = ensure_cdt(c->e->device);
      if (rc) return rc;
      hipLaunchKernelGGL(k_sample_small<CUPOLY_DISCRETE_GAUSSIAN>, dim3((ng + 255) / 256),
                         dim3(256), 0, c->stream, (cp_gi *)out, ng, cnt, off, c->e->d_primes, seed,
                         stream);
      break;
    }
    default: return cp_fail(CUPOLY_ERR_ARG, "Invalid sampler: " + std::to_string(kind));
  }
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

extern "C" int cupoly_sample(cupoly_ctx *c, cupoly_gi *out, int kind, uint64_t seed,
                             uint64_t stream, int base) {
  int rc = cupoly_sample_coeff(c, out, kind, seed, stream, base);
  if (rc) return rc;
  return cupoly_dgt(c, out, base, CUPOLY_FORWARD);
}

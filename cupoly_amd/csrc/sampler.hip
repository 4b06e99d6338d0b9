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
//                              60-digit arithmetic);
//   narrow {0,1,2}   (tag 4): as ternary, value w mod 3 (the reference's BFV-
//                              mode narrow sampler, sampler.cu:72-103);
//   binary {0,1}     (tag 5): bit 0 of the first word (sampler.cu:158-182);
//   ZO / Hamming weight (tag 6): exactly h coefficients +-1 (h = n/2 for ZO,
//                              h = 64 for HAMMINGWEIGHT, settings.h:43), at the
//                              positions of the h smallest keys (48 random
//                              bits above the 16-bit index, so all distinct),
//                              sign from bit 0 of the third word.  The
//                              reference draws these on the host with rand()
//                              and std::shuffle (sampler.cu:222-278, 401-446);
//                              here one workgroup radix-selects the h-th
//                              smallest key in eight 8-bit passes over LDS
//                              histograms.
// Small values are written to every residue of the base as e mod q_i, like the
// reference's samplers (sampler.cu:94-97, 485-487).
#include "philox.hpp"

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
                   const cp_prime *__restrict__ primes, uint64_t seed, uint64_t stream,
                   cp_cdt cdt) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= ng) return;
  int e0, e1;
  if constexpr (KIND == CUPOLY_TERNARY) {
    e0 = ternary_coef(seed, stream, (uint32_t)slot);
    e1 = ternary_coef(seed, stream, (uint32_t)(slot + ng));
  } else if constexpr (KIND == CUPOLY_NARROW) {
    e0 = narrow_coef(seed, stream, (uint32_t)slot);
    e1 = narrow_coef(seed, stream, (uint32_t)(slot + ng));
  } else if constexpr (KIND == CUPOLY_BINARY) {
    e0 = binary_coef(seed, stream, (uint32_t)slot);
    e1 = binary_coef(seed, stream, (uint32_t)(slot + ng));
  } else {
    e0 = gaussian_coef(seed, stream, (uint32_t)slot, cdt);
    e1 = gaussian_coef(seed, stream, (uint32_t)(slot + ng), cdt);
  }
  for (int r = 0; r < nres; r++) {
    const uint64_t q = primes[roff + r].p;
    cp_gi v;
    v.re = small_mod(e0, q);
    v.imag = small_mod(e1, q);
    out[(size_t)r * ng + slot] = v;
  }
}

// One workgroup: the h coefficients (of n = 2 ng) with the smallest keys
// become +-1, all others 0; every residue gets the value mod q_r.  The h-th
// smallest key is found by radix selection, 8 bits per pass from the top.
// Selection stops as soon as every key sharing the resolved top bits with
// the h-th one is itself selected (its bucket holds exactly the remaining
// rank, e.g. a bucket of one key): then "top bits <= the resolved prefix"
// picks the same h keys as a full 64-bit comparison, usually after 2 or 3
// passes instead of 8.  With KPT > 0 each thread keeps its KPT keys (and
// signs) in registers, so Philox runs once per coefficient; KPT = 0 (n above
// 16384) recomputes them in each pass.
#define FW_THREADS 1024
template <int KPT>
__global__ void __launch_bounds__(FW_THREADS)
    k_sample_fixed_weight(cp_gi *__restrict__ out, int ng, int nres, int roff,
                          const cp_prime *__restrict__ primes, uint64_t seed, uint64_t stream,
                          int h) {
  __shared__ unsigned hist[256];
  __shared__ uint64_t s_prefix;
  __shared__ int s_need, s_done;
  const int n = 2 * ng, tid = (int)threadIdx.x;
  uint64_t keys[KPT > 0 ? KPT : 1];
  uint32_t negs = 0;
  if constexpr (KPT > 0) {
#pragma unroll
    for (int u = 0; u < KPT; u++) {
      const int j = tid + u * FW_THREADS;
      bool neg = false;
      keys[u] = j < n ? fixed_weight_key(seed, stream, (uint32_t)j, neg) : ~0ull;
      negs |= (uint32_t)neg << u;
    }
  }
  uint64_t prefix = 0;  // top bits of the h-th smallest key, 8 more per pass
  int need = h;         // its rank among the keys that share those bits
  int shift = 0;        // low bits not resolved when selection stopped
  if (h > 0 && h < n) {
    for (int pass = 0; pass < 8; pass++) {
      shift = 56 - 8 * pass;
      for (int i = tid; i < 256; i += FW_THREADS) hist[i] = 0;
      __syncthreads();
      if constexpr (KPT > 0) {
#pragma unroll
        for (int u = 0; u < KPT; u++) {
          const uint64_t key = keys[u];
          if (tid + u * FW_THREADS < n &&
              (pass == 0 || (key >> (shift + 8)) == (prefix >> (shift + 8))))
            atomicAdd(&hist[(key >> shift) & 255], 1u);
        }
      } else {
        for (int j = tid; j < n; j += FW_THREADS) {
          bool neg;
          const uint64_t key = fixed_weight_key(seed, stream, (uint32_t)j, neg);
          if (pass == 0 || (key >> (shift + 8)) == (prefix >> (shift + 8)))
            atomicAdd(&hist[(key >> shift) & 255], 1u);
        }
      }
      __syncthreads();
      if (tid == 0) {
        int cum = 0, d = 0;
        for (; d < 255; d++) {
          if (cum + (int)hist[d] >= need) break;
          cum += (int)hist[d];
        }
        s_need = need - cum;
        s_done = (need - cum) == (int)hist[d];
        s_prefix = prefix | ((uint64_t)d << shift);
      }
      __syncthreads();
      prefix = s_prefix;
      need = s_need;
      const bool done = s_done != 0;
      __syncthreads();  // hist and the broadcast slots are reused by the next pass
      if (done) break;  // uniform
    }
  }
  const uint64_t bound = prefix >> shift;
  for (int u = 0, j = tid; j < n; u++, j += FW_THREADS) {
    bool neg;
    uint64_t key;
    if constexpr (KPT > 0) {
      key = keys[u < KPT ? u : 0];
      neg = (negs >> u) & 1;
    } else {
      key = fixed_weight_key(seed, stream, (uint32_t)j, neg);
    }
    const bool sel = h >= n || (h > 0 && (key >> shift) <= bound);
    const int e = sel ? (neg ? -1 : 1) : 0;
    const int slot = j < ng ? j : j - ng;
    for (int r = 0; r < nres; r++) {
      uint64_t *dst = j < ng ? &out[(size_t)r * ng + slot].re : &out[(size_t)r * ng + slot].imag;
      *dst = small_mod(e, primes[roff + r].p);
    }
  }
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
                         c->stream, (cp_gi *)out, ng, cnt, off, c->e->d_primes, seed, stream,
                         cp_make_cdt());
      break;
    case CUPOLY_DISCRETE_GAUSSIAN:
      hipLaunchKernelGGL(k_sample_small<CUPOLY_DISCRETE_GAUSSIAN>, dim3((ng + 255) / 256),
                         dim3(256), 0, c->stream, (cp_gi *)out, ng, cnt, off, c->e->d_primes, seed,
                         stream, cp_make_cdt());
      break;
    case CUPOLY_NARROW:
      hipLaunchKernelGGL(k_sample_small<CUPOLY_NARROW>, dim3((ng + 255) / 256), dim3(256), 0,
                         c->stream, (cp_gi *)out, ng, cnt, off, c->e->d_primes, seed, stream,
                         cp_make_cdt());
      break;
    case CUPOLY_BINARY:
      hipLaunchKernelGGL(k_sample_small<CUPOLY_BINARY>, dim3((ng + 255) / 256), dim3(256), 0,
                         c->stream, (cp_gi *)out, ng, cnt, off, c->e->d_primes, seed, stream,
                         cp_make_cdt());
      break;
    case CUPOLY_ZO:
    case CUPOLY_HAMMINGWEIGHT: {
      if (2 * ng > 65536) return cp_fail(CUPOLY_ERR_UNSUPPORTED, "fixed-weight sampler: n > 65536");
      const int h = kind == CUPOLY_ZO ? ng : (2 * ng < CUPOLY_HAMMINGWEIGHT_H ? 2 * ng
                                                                           : CUPOLY_HAMMINGWEIGHT_H);
      if (2 * ng <= 16 * FW_THREADS)
        hipLaunchKernelGGL(k_sample_fixed_weight<16>, dim3(1), dim3(FW_THREADS), 0, c->stream,
                           (cp_gi *)out, ng, cnt, off, c->e->d_primes, seed, stream, h);
      else
        hipLaunchKernelGGL(k_sample_fixed_weight<0>, dim3(1), dim3(FW_THREADS), 0, c->stream,
                           (cp_gi *)out, ng, cnt, off, c->e->d_primes, seed, stream, h);
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

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

// Fixed weight (ZO, Hamming weight): the h coefficients (of n = 2 ng) with
// the smallest keys become +-1, all others 0; every residue gets the value
// mod q_r.  Three launches:
//   1. k_fw_keys:   every coefficient's key (one Philox draw), in parallel,
//                   into the context's scratch;
//   2. k_fw_select: one workgroup finds the h-th smallest key by radix
//                   selection, 8 bits per pass from the top, each pass a
//                   256-bucket histogram and a parallel scan of it.
//                   Selection stops as soon as every key sharing the
//                   resolved top bits with the h-th one is itself selected
//                   (its bucket holds exactly the remaining rank, e.g. a
//                   bucket of one key): then "top bits <= the resolved
//                   prefix" picks the same h keys as a full 64-bit
//                   comparison, usually after 2 or 3 passes instead of 8.
//                   Writes (bound, shift);
//   3. k_fw_write:  every coefficient redraws its key and sign and writes its
//                   value to every residue, in parallel.
// The keys are distinct (the coefficient index is in their low bits), so the
// selected set is the same for any launch shape.
#define FW_THREADS 1024

__global__ void __launch_bounds__(256)
    k_fw_keys(uint64_t *__restrict__ keys, int n, uint64_t seed, uint64_t stream) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  bool neg;
  keys[j] = fixed_weight_key(seed, stream, (uint32_t)j, neg);
}

// sel[0] = bound (resolved top bits of the h-th key), sel[1] = shift; h in
// (0, n).  With KPT > 0 (n <= KPT * FW_THREADS) every thread loads its KPT
// keys once, all loads in flight together, and keeps them in registers for
// every pass; KPT = 0 reads them again in each pass, 16 loads at a time.
template <int KPT>
__global__ void __launch_bounds__(FW_THREADS)
    k_fw_select(uint64_t *__restrict__ sel, const uint64_t *__restrict__ keys, int n, int h) {
  __shared__ unsigned hist[256];
  __shared__ unsigned scan[256];
  __shared__ uint64_t s_prefix;
  __shared__ int s_need, s_done;
  const int tid = (int)threadIdx.x;
  constexpr int CH = KPT > 0 ? KPT : 16;
  uint64_t kr[CH];
  if constexpr (KPT > 0) {
#pragma unroll
    for (int u = 0; u < KPT; u++) {
      const int j = tid + u * FW_THREADS;
      kr[u] = j < n ? keys[j] : 0;
    }
  }
  uint64_t prefix = 0;  // top bits of the h-th smallest key, 8 more per pass
  int need = h;         // its rank among the keys that share those bits
  int shift = 0;        // low bits not resolved when selection stopped
  for (int pass = 0; pass < 8; pass++) {
    shift = 56 - 8 * pass;
    if (tid < 256) hist[tid] = 0;
    __syncthreads();
    for (int base = 0; base < n; base += CH * FW_THREADS) {
      if constexpr (KPT == 0) {
#pragma unroll
        for (int u = 0; u < CH; u++) {
          const int j = base + tid + u * FW_THREADS;
          kr[u] = j < n ? keys[j] : 0;
        }
      }
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int j = base + tid + u * FW_THREADS;
        const uint64_t key = kr[u];
        if (j < n && (pass == 0 || (key >> (shift + 8)) == (prefix >> (shift + 8))))
          atomicAdd(&hist[(key >> shift) & 255], 1u);
      }
      if constexpr (KPT > 0) break;  // every key is in registers
    }
    __syncthreads();
    // inclusive scan of the 256 buckets (Hillis-Steele, one thread per bucket)
    unsigned v = tid < 256 ? hist[tid] : 0;
    for (int d = 1; d < 256; d <<= 1) {
      if (tid < 256) scan[tid] = v;
      __syncthreads();
      if (tid < 256 && tid >= d) v += scan[tid - d];
      __syncthreads();
    }
    // the bucket holding rank `need`: cumulative count before it < need <= through it
    if (tid < 256) {
      const int incl = (int)v, excl = incl - (int)hist[tid];
      if (excl < need && need <= incl) {
        s_need = need - excl;
        s_done = (need - excl) == (int)hist[tid];
        s_prefix = prefix | ((uint64_t)tid << shift);
      }
    }
    __syncthreads();
    prefix = s_prefix;
    need = s_need;
    const bool done = s_done != 0;
    __syncthreads();  // hist, scan and the broadcast slots are reused by the next pass
    if (done) break;  // uniform
  }
  if (tid == 0) {
    sel[0] = prefix >> shift;
    sel[1] = (uint64_t)shift;
  }
}

// sel == nullptr: h = 0 (no coefficient selected) or h >= n (all selected)
__global__ void __launch_bounds__(256)
    k_fw_write(cp_gi *__restrict__ out, int ng, int nres, int roff,
               const cp_prime *__restrict__ primes, uint64_t seed, uint64_t stream,
               const uint64_t *__restrict__ sel, bool all) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= 2 * ng) return;
  bool neg;
  const uint64_t key = fixed_weight_key(seed, stream, (uint32_t)j, neg);
  const bool on = sel ? (key >> sel[1]) <= sel[0] : all;
  const int e = on ? (neg ? -1 : 1) : 0;
  const int slot = j < ng ? j : j - ng;
  for (int r = 0; r < nres; r++) {
    uint64_t *dst = j < ng ? &out[(size_t)r * ng + slot].re : &out[(size_t)r * ng + slot].imag;
    *dst = small_mod(e, primes[roff + r].p);
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
      const int n = 2 * ng;
      if (n > 65536) return cp_fail(CUPOLY_ERR_UNSUPPORTED, "fixed-weight sampler: n > 65536");
      const int h = kind == CUPOLY_ZO ? ng : (n < CUPOLY_HAMMINGWEIGHT_H ? n
                                                                        : CUPOLY_HAMMINGWEIGHT_H);
      const dim3 g((n + 255) / 256);
      if (h > 0 && h < n) {
        // keys (n words) and the selection (2 words) in the context's scratch
        // row d_tmp, which holds nres * ng >= n Gaussian slots
        uint64_t *keys = (uint64_t *)c->d_tmp, *sel = keys + n;
        hipLaunchKernelGGL(k_fw_keys, g, dim3(256), 0, c->stream, keys, n, seed, stream);
        if (n <= 16 * FW_THREADS)
          hipLaunchKernelGGL(k_fw_select<16>, dim3(1), dim3(FW_THREADS), 0, c->stream, sel,
                             (const uint64_t *)keys, n, h);
        else
          hipLaunchKernelGGL(k_fw_select<0>, dim3(1), dim3(FW_THREADS), 0, c->stream, sel,
                             (const uint64_t *)keys, n, h);
        hipLaunchKernelGGL(k_fw_write, g, dim3(256), 0, c->stream, (cp_gi *)out, ng, cnt, off,
                           c->e->d_primes, seed, stream, (const uint64_t *)sel, false);
      } else {
        hipLaunchKernelGGL(k_fw_write, g, dim3(256), 0, c->stream, (cp_gi *)out, ng, cnt, off,
                           c->e->d_primes, seed, stream, (const uint64_t *)nullptr, h >= n);
      }
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

// Coefficient edges on the device: multi-word integers <-> RNS residues.
//
// Reference: poly_crt (src/arithmetic/polynomial.cu:190-219) folds NTL ZZ
// coefficients into the residue layout (coefficient j -> re of slot j,
// coefficient j + n/2 -> imag of slot j, each reduced mod every prime), and
// poly_icrt (polynomial.cu:221-291) rebuilds
//     x = sum_i (Q/q_i) [x_i (Q/q_i)^-1]_{q_i}  mod Q
// with host big integers.  Here both run on little-endian 64-bit limbs,
// crt with one thread per (slot, residue), icrt with one per coefficient:
//   crt:  x mod p by Horner from the top limb, r <- r (2^64 mod p) + w, with a
//         Shoup product by the per-prime constant 2^64 mod p;
//   icrt: the sum in k+1 limbs (128-bit multiply-accumulate against the limbs
//         of Q/q_i), then one subtraction of m Q where m = floor(sum_i y_i/q_i)
//         is estimated in FP64 and corrected exactly (the estimate is off by at
//         most one when the sum is within rounding of an integer).
#include "internal.hpp"

// w mod q for one 64-bit word: 2^64 < 3q for q > 2^63 - 2^30.
__device__ __forceinline__ uint64_t word_mod(uint64_t w, uint64_t q) {
  w = w >= q ? w - q : w;
  return w >= q ? w - q : w;
}

__device__ __forceinline__ uint64_t limbs_mod(const uint64_t *__restrict__ x, int limbs, uint64_t q,
                                              uint64_t c, uint64_t cp) {
  uint64_t r = 0;
  for (int l = limbs - 1; l >= 0; l--) r = cp_add(cp_shoup(r, c, cp, q), word_mod(x[l], q), q);
  return r;
}

// coefs: [ncoefs][limbs]; out: [nres][ng] GIs of residues roff..roff+nres-1;
// one thread per (slot, residue r = blockIdx.y)
__global__ void __launch_bounds__(256)
    k_crt(cp_gi *__restrict__ out, const uint64_t *__restrict__ coefs, int ncoefs, int limbs,
          int ng, int roff, const cp_crt *__restrict__ C, const cp_prime *__restrict__ primes) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= ng) return;
  const int r = blockIdx.y;
  const bool has0 = slot < ncoefs, has1 = slot + ng < ncoefs;
  const uint64_t *x0 = coefs + (size_t)slot * limbs, *x1 = coefs + (size_t)(slot + ng) * limbs;
  const uint64_t q = primes[roff + r].p;
  const uint64_t c = C->two64[roff + r][0], cp = C->two64[roff + r][1];
  cp_gi v;
  v.re = has0 ? limbs_mod(x0, limbs, q, c, cp) : 0;
  v.imag = has1 ? limbs_mod(x1, limbs, q, c, cp) : 0;
  out[(size_t)r * ng + slot] = v;
}

// in: [K][ng] GIs of base Q; out: [2 ng][K] limbs of x in [0, Q).
template <int K>
__global__ void __launch_bounds__(256)
    k_icrt(uint64_t *__restrict__ out, const cp_gi *__restrict__ in, int ng,
           const cp_hps *__restrict__ H, const cp_crt *__restrict__ C) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= 2 * ng) return;
  const int slot = c < ng ? c : c - ng;
  uint64_t acc[K + 1];
#pragma unroll
  for (int l = 0; l <= K; l++) acc[l] = 0;
  double v = 0.0;
#pragma unroll
  for (int i = 0; i < K; i++) {
    const cp_gi g = in[(size_t)i * ng + slot];
    const uint64_t q = H->q[i].p;
    const uint64_t y = cp_shoup(c < ng ? g.re : g.imag, H->invQqi[i][0], H->invQqi[i][1], q);
    v += (double)y * C->rcp_q[i];
    uint64_t carry = 0;
#pragma unroll
    for (int l = 0; l < K; l++) {
      const cp_u128 t = (cp_u128)y * C->M[i][l] + acc[l] + carry;
      acc[l] = (uint64_t)t;
      carry = (uint64_t)(t >> 64);
    }
    acc[K] += carry;
  }
  // acc - m Q with m = floor(v) in [0, K]
  const uint64_t m = (uint64_t)v;
  uint64_t borrow = 0, mcarry = 0;
#pragma unroll
  for (int l = 0; l <= K; l++) {
    const cp_u128 mq = (cp_u128)m * C->Q[l] + mcarry;
    mcarry = (uint64_t)(mq >> 64);
    const uint64_t s = (uint64_t)mq;
    const uint64_t d = acc[l] - s - borrow;
    borrow = (acc[l] < s || (acc[l] == s && borrow)) ? 1 : 0;
    acc[l] = d;
  }
  if (borrow) {  // m was one too large: add Q back
    uint64_t carry = 0;
#pragma unroll
    for (int l = 0; l <= K; l++) {
      const cp_u128 t = (cp_u128)acc[l] + C->Q[l] + carry;
      acc[l] = (uint64_t)t;
      carry = (uint64_t)(t >> 64);
    }
  } else {  // m was one too small: subtract Q once more if acc >= Q
    bool ge = true;
#pragma unroll
    for (int l = K; l >= 0; l--) {
      if (acc[l] != C->Q[l]) {
        ge = acc[l] > C->Q[l];
        break;
      }
    }
    if (ge) {
      uint64_t b2 = 0;
#pragma unroll
      for (int l = 0; l <= K; l++) {
        const uint64_t s = C->Q[l];
        const uint64_t d = acc[l] - s - b2;
        b2 = (acc[l] < s || (acc[l] == s && b2)) ? 1 : 0;
        acc[l] = d;
      }
    }
  }
#pragma unroll
  for (int l = 0; l < K; l++) out[(size_t)c * K + l] = acc[l];
}

extern "C" int cupoly_crt(cupoly_ctx *c, cupoly_gi *out, const uint64_t *coefs, int ncoefs,
                          int limbs, int base) {
  if (!c || !out || (!coefs && ncoefs > 0) || ncoefs < 0 || limbs < 1)
    return cp_fail(CUPOLY_ERR_ARG, "crt: bad argument");
  if (base != CUPOLY_QBASE && base != CUPOLY_QBBASE && base != CUPOLY_BBASE)
    return cp_fail(CUPOLY_ERR_UNSUPPORTED, "crt: base must be Q, B or QB");
  const int ng = c->e->ng;
  if (ncoefs > 2 * ng) return cp_fail(CUPOLY_ERR_ARG, "crt: more coefficients than the ring degree");
  int off, cnt;
  if (!cp_base_range(c->e, base, &off, &cnt)) return cp_fail(CUPOLY_ERR_UNSUPPORTED, "Unknown base");
  hipLaunchKernelGGL(k_crt, dim3((ng + 255) / 256, cnt), dim3(256), 0, c->stream, (cp_gi *)out,
                     coefs, ncoefs, limbs, ng, off, c->e->d_crt, c->e->d_primes);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

template <int K>
static void launch_icrt(cupoly_ctx *c, uint64_t *out, const cupoly_gi *in) {
  const int ng = c->e->ng;
  hipLaunchKernelGGL(k_icrt<K>, dim3((2 * ng + 255) / 256), dim3(256), 0, c->stream, out,
                     (const cp_gi *)in, ng, c->e->d_hps, c->e->d_crt);
}

extern "C" int cupoly_icrt(cupoly_ctx *c, uint64_t *out, const cupoly_gi *in) {
  if (!c || !out || !in) return cp_fail(CUPOLY_ERR_ARG, "icrt: null argument");
  switch (c->e->k) {
    case 1: launch_icrt<1>(c, out, in); break;
    case 2: launch_icrt<2>(c, out, in); break;
    case 3: launch_icrt<3>(c, out, in); break;
    case 4: launch_icrt<4>(c, out, in); break;
    case 5: launch_icrt<5>(c, out, in); break;
    case 6: launch_icrt<6>(c, out, in); break;
    case 7: launch_icrt<7>(c, out, in); break;
    case 8: launch_icrt<8>(c, out, in); break;
    case 9: launch_icrt<9>(c, out, in); break;
    case 10: launch_icrt<10>(c, out, in); break;
    case 11: launch_icrt<11>(c, out, in); break;
    case 12: launch_icrt<12>(c, out, in); break;
    case 13: launch_icrt<13>(c, out, in); break;
    case 14: launch_icrt<14>(c, out, in); break;
    default: return cp_fail(CUPOLY_ERR_UNSUPPORTED, "icrt: k outside 1..14");
  }
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

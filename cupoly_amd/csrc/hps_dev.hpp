// Per-slot HPS steps shared by the single-polynomial kernels (hps.hip) and
// the batched homomorphic-multiply kernels (homomul.hip).  A thread owns one
// Gaussian slot c: coefficients c and c + n/2 of every residue.
//
// Floating point is reproduced operation for operation from the reference
// (src/cuda/cudaengine.cu:382-440, 513-534, 688-703): see hps.hip.
#pragma once
#include <math.h>

#include "internal.hpp"

// Two FP64 fused multiply-adds rounded toward -infinity (CUDA's __fma_rd, the
// reference's rounding of the HPS fractional sums): vr = RD(xr r + vr),
// vi = RD(xi r + vi).  The MODE register's FP64 rounding field (bits 3:2) is
// set to round-down for the two instructions and back to round-to-nearest in
// the same asm block, so no other floating-point instruction runs under the
// directed mode (the compiler's own mode tracking does not look inside asm).
// A VALU instruction may follow a MODE write at once; a MODE write after a
// VALU instruction needs one wait state, hence the s_nop.
__device__ __forceinline__ void cp_fma_rd2(double xr, double xi, double r, double &vr,
                                           double &vi) {
  asm volatile(
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 2\n\t"
      "v_fma_f64 %0, %2, %4, %0\n\t"
      "v_fma_f64 %1, %3, %4, %1\n\t"
      "s_nop 0\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0"
      : "+&v"(vr), "+&v"(vi)
      : "v"(xr), "v"(xi), "v"(r));
}

// The reference's addmod / submod macros (include/cuPoly/cuda/dgt.h:73-74):
// one conditional correction on 64-bit wrap-around words, kept verbatim for
// the additions that may see unreduced operands in the complex scaling.
__device__ __forceinline__ uint64_t addmod_ref(uint64_t a, uint64_t b, uint64_t p) {
  const uint64_t s = a + b;
  return s >= p ? s - p : s;
}
__device__ __forceinline__ uint64_t submod_ref(uint64_t a, uint64_t b, uint64_t p) {
  return a < b ? a + (p - b) : a - b;
}

// |value - llrint(value)| <= 1e-9 ? llrint(value) : floor(value)  (cudaengine.cu:530-531)
__device__ __forceinline__ uint64_t snap(double value) {
  const long long rn = llrint(value);
  const double diff = fabs(value - (double)rn);
  return diff <= 1e-9 ? (uint64_t)rn : (uint64_t)floor(value);
}

__device__ __forceinline__ uint64_t shoup2(uint64_t a, const uint64_t (&w)[2], uint64_t p) {
  return cp_shoup(a, w[0], w[1], p);
}

// floor(a * b / 2^64) for a 128-bit a = (alo, ahi): (lo, hi) words
// (the reference's __umul128hi, cudaengine.cu:313-320).
__device__ __forceinline__ void umul128hi(uint64_t alo, uint64_t ahi, uint64_t b, uint64_t &lo,
                                          uint64_t &hi) {
  const cp_u128 x = (cp_u128)alo * b;
  const cp_u128 y = (cp_u128)ahi * b + (uint64_t)(x >> 64);
  lo = (uint64_t)y;
  hi = (uint64_t)(y >> 64);
}

// Q -> B for slot c: in[j*qs + c] (j < k) -> out[i*bs + c] (i < kl).  All
// inputs are read before any output is written, so in == out is allowed.
__device__ __forceinline__ void hps_q_to_b_slot(const cp_gi *in, size_t qs, cp_gi *out, size_t bs,
                                                int c, const cp_hps *__restrict__ H) {
  const int k = H->k, kl = H->kl;
  uint64_t ar[CP_MAX_BASE], ai[CP_MAX_BASE];
  double vr = 0.0, vi = 0.0;
#pragma unroll
  for (int j = 0; j < CP_MAX_BASE; j++) {
    if (j < k) {
      const uint64_t q = H->q[j].p;
      const cp_gi x = ld_gi_nt((const CP_GLOBAL cp_gi *)in + (size_t)j * qs + c);
      ar[j] = shoup2(x.re, H->invQqi[j], q);
      ai[j] = shoup2(x.imag, H->invQqi[j], q);
    }
  }
  // v = RD(aux_j * RD(1/q_j) + v) in ascending j (cudaengine.cu:526-527)
#pragma unroll
  for (int j = 0; j < CP_MAX_BASE; j++)
    if (j < k) cp_fma_rd2((double)ar[j], (double)ai[j], H->rcp_q[j], vr, vi);
  const uint64_t v_re = snap(vr), v_im = snap(vi);
  for (int i = 0; i < kl; i++) {
    // (sum_j a_j (Q/q_j mod b_i) - v (Q mod b_i)) mod b_i: every term is
    // reduced, so the exact sum reduced once is the reference's chain of
    // addmod / submod
    const uint64_t b = H->b[i].p, qb = H->Qbi[i][0];
    const uint64_t nqb = qb ? b - qb : 0;  // -Q mod b_i
    cp_acc sr, si;
    cp_acc_init(sr);
    cp_acc_init(si);
#pragma unroll
    for (int j = 0; j < CP_MAX_BASE; j++)
      if (j < k) cp_acc_mac2(sr, si, ar[j], ai[j], H->Qqibi[i][j][0]);
    cp_acc_mac2(sr, si, v_re, v_im, nqb);
    cp_gi o;
    o.re = cp_acc_reduce(sr, b);
    o.imag = cp_acc_reduce(si, b);
    st_gi_nt((CP_GLOBAL cp_gi *)out + (size_t)i * bs + c, o);
  }
}

// Complex scaling t/Q, residue i of both coefficients of one slot
// (cudaengine.cu:401-433): the reference forms
//     y_i = sum_j (x_j w_ij mod b_i  (+)  h_j)  (+)  round  (+)  lambda_i xB_i mod b_i
// with its single-correction addmod (+), where h_j = floor(x_j omega_frac_j
// / 2^128) < q_j and round = llrint(sum_j RN(l_j 2^-64)) <= k.  When every
// h_j < b_i, every operand of the chain is below b_i, the chain is plain
// modular addition, and the exact sum reduced once (a 160-bit accumulation,
// hsum = sum_j h_j + round added as its starting value) is the same word.
// An h_j >= b_i needs x_j within 2^30 of q_j (the primes lie within 2^30 of
// 2^63): about one coefficient in 2^33.  Such a slot takes the reference's
// chain as written (`slow`), so the result is the same word in every case.
__device__ __forceinline__ void scale_exact2(const cp_acc &hsum_r, const cp_acc &hsum_i,
                                             const uint64_t *xr, const uint64_t *xi, int kq,
                                             const uint64_t (*wint)[2], uint64_t xbr,
                                             uint64_t xbi, const uint64_t (&lam)[2], uint64_t b,
                                             uint64_t &rr, uint64_t &ri) {
  cp_acc ar = hsum_r, ai = hsum_i;
#pragma unroll
  for (int j = 0; j < kq; j++) cp_acc_mac2(ar, ai, xr[j], xi[j], wint[j][0]);
  cp_acc_mac2(ar, ai, xbr, xbi, lam[0]);
  rr = cp_acc_reduce(ar, b);
  ri = cp_acc_reduce(ai, b);
}

// hsum = sum_j h_j + round as a 160-bit accumulator value
__device__ __forceinline__ void hsum_add(cp_acc &s, uint64_t v) {
  unsigned c = 0;
  s.lo = cp_addc64(s.lo, v, c);
  s.hi += c;
}

// The reference's chain for residue i of one coefficient, operation for
// operation (cudaengine.cu:401-433): x[j] the Q words, h[j] their fraction
// products' high words, round = llrint of the FP64 sum of the low words.
__device__ __forceinline__ uint64_t scale_chain_word(const uint64_t *x, const uint64_t *h, int kq,
                                                     const uint64_t (*wint)[2], uint64_t round,
                                                     uint64_t xb, const uint64_t (&lam)[2],
                                                     uint64_t b) {
  uint64_t r = 0;
#pragma unroll
  for (int j = 0; j < kq; j++) r = addmod_ref(r, addmod_ref(shoup2(x[j], wint[j], b), h[j], b), b);
  r = addmod_ref(r, round, b);
  return addmod_ref(r, shoup2(xb, lam, b), b);
}

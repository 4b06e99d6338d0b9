// Per-slot HPS steps shared by the single-polynomial kernels (hps.hip) and
// the batched homomorphic-multiply kernels (homomul.hip).  A thread owns one
// Gaussian slot c: coefficients c and c + n/2 of every residue.
//
// Floating point is reproduced operation for operation from the reference
// (src/cuda/cudaengine.cu:382-440, 513-534, 688-703): see hps.hip.
#pragma once
#include <math.h>

#include "internal.hpp"

extern "C" __device__ double __ocml_fma_rtn_f64(double, double, double);

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
      const cp_gi x = in[(size_t)j * qs + c];
      ar[j] = shoup2(x.re, H->invQqi[j], q);
      ai[j] = shoup2(x.imag, H->invQqi[j], q);
    }
  }
  // v = RD(aux_j * RD(1/q_j) + v) in ascending j (cudaengine.cu:526-527)
#pragma unroll
  for (int j = 0; j < CP_MAX_BASE; j++) {
    if (j < k) {
      const double r = H->rcp_q[j];
      vr = __ocml_fma_rtn_f64((double)ar[j], r, vr);
      vi = __ocml_fma_rtn_f64((double)ai[j], r, vi);
    }
  }
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
    for (int j = 0; j < CP_MAX_BASE; j++) {
      if (j < k) {
        cp_acc_mac(sr, ar[j], H->Qqibi[i][j][0]);
        cp_acc_mac(si, ai[j], H->Qqibi[i][j][0]);
      }
    }
    cp_acc_mac(sr, v_re, nqb);
    cp_acc_mac(si, v_im, nqb);
    cp_gi o;
    o.re = cp_acc_reduce(sr, b);
    o.imag = cp_acc_reduce(si, b);
    out[(size_t)i * bs + c] = o;
  }
}

// Complex scaling t/Q for slot c: aq[j*qs + c] (j < k), ab[i*bs + c] (i < kl)
// -> yr[i], yi[i] (residues mod b_i of the re / imag coefficient).
// omega_frac_j does not depend on the output residue i (it is t (QB/q_j)^-1
// mod q_j B / q_j, reduced to its fraction), so the 192-bit products x_j
// omega_frac_j and the FP64 sum of their low words are formed once per slot
// and reused for every i; the sum runs over j in the same order as the
// reference's per-(i, slot) thread, so it is the same double.
__device__ __forceinline__ void hps_scale_slot(const cp_gi *__restrict__ aq, size_t qs,
                                               const cp_gi *__restrict__ ab, size_t bs, int c,
                                               const cp_hps *__restrict__ H,
                                               uint64_t (&yr)[CP_MAX_BASE],
                                               uint64_t (&yi)[CP_MAX_BASE]) {
  const int k = H->k, kl = H->kl;
  cp_gi x[CP_MAX_BASE];
  uint64_t hr[CP_MAX_BASE], hi[CP_MAX_BASE];
  const double two_m64 = 0x1p-64;  // __drcp_rn(2^64)
  double acc_r = 0.0, acc_i = 0.0;
#pragma unroll
  for (int j = 0; j < CP_MAX_BASE; j++) {
    if (j < k) {
      x[j] = aq[(size_t)j * qs + c];
      uint64_t lr, li;
      umul128hi(H->wfrac[j][0], H->wfrac[j][1], x[j].re, lr, hr[j]);
      umul128hi(H->wfrac[j][0], H->wfrac[j][1], x[j].imag, li, hi[j]);
      acc_r = fma((double)lr, two_m64, acc_r);  // __fma_rn (cudaengine.cu:413-414)
      acc_i = fma((double)li, two_m64, acc_i);
    }
  }
  const uint64_t round_r = (uint64_t)llrint(acc_r), round_i = (uint64_t)llrint(acc_i);
#pragma unroll
  for (int i = 0; i < CP_MAX_BASE; i++) {
    if (i < kl) {
      const uint64_t b = H->b[i].p;
      uint64_t rr = 0, ri = 0;
#pragma unroll
      for (int j = 0; j < CP_MAX_BASE; j++) {
        if (j < k) {
          const uint64_t int_r = shoup2(x[j].re, H->wint[i][j], b);
          const uint64_t int_i = shoup2(x[j].imag, H->wint[i][j], b);
          rr = addmod_ref(rr, addmod_ref(int_r, hr[j], b), b);
          ri = addmod_ref(ri, addmod_ref(int_i, hi[j], b), b);
        }
      }
      rr = addmod_ref(rr, round_r, b);
      ri = addmod_ref(ri, round_i, b);
      const cp_gi y = ab[(size_t)i * bs + c];
      yr[i] = addmod_ref(rr, shoup2(y.re, H->lam[i], b), b);
      yi[i] = addmod_ref(ri, shoup2(y.imag, H->lam[i], b), b);
    }
  }
}

// B -> Q for slot c from residues xr[i], xi[i] mod b_i -> out[j*qs + c] (j < k).
__device__ __forceinline__ void hps_b_to_q_slot(const uint64_t (&xr)[CP_MAX_BASE],
                                                const uint64_t (&xi)[CP_MAX_BASE], cp_gi *out,
                                                size_t qs, int c, const cp_hps *__restrict__ H) {
  const int k = H->k, kl = H->kl;
  uint64_t ar[CP_MAX_BASE], ai[CP_MAX_BASE];
  double vr = 0.0, vi = 0.0;
  // v = RN(v + RN(aux_i) / RN(b_i)) in ascending i (cudaengine.cu:699-700)
#pragma unroll
  for (int i = 0; i < CP_MAX_BASE; i++) {
    if (i < kl) {
      const uint64_t b = H->b[i].p;
      ar[i] = shoup2(xr[i], H->invBbi[i], b);
      ai[i] = shoup2(xi[i], H->invBbi[i], b);
      const double bd = (double)b;
      vr = vr + (double)ar[i] / bd;
      vi = vi + (double)ai[i] / bd;
    }
  }
  const uint64_t v_re = snap(vr), v_im = snap(vi);
  for (int j = 0; j < k; j++) {
    // (sum_i a_i (B/b_i mod q_j) - v (B mod q_j)) mod q_j, reduced once
    const uint64_t q = H->q[j].p, bq = H->Bqi[j][0];
    const uint64_t nbq = bq ? q - bq : 0;  // -B mod q_j
    cp_acc sr, si;
    cp_acc_init(sr);
    cp_acc_init(si);
#pragma unroll
    for (int i = 0; i < CP_MAX_BASE; i++) {
      if (i < kl) {
        cp_acc_mac(sr, ar[i], H->Bbiqi[j][i][0]);
        cp_acc_mac(si, ai[i], H->Bbiqi[j][i][0]);
      }
    }
    cp_acc_mac(sr, v_re, nbq);
    cp_acc_mac(si, v_im, nbq);
    cp_gi o;
    o.re = cp_acc_reduce(sr, q);
    o.imag = cp_acc_reduce(si, q);
    out[(size_t)j * qs + c] = o;
  }
}

/*
 * cupoly_oracle.c - CPU restatement of the cuPoly polynomial hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path (cupoly_amd) never calls it.
 *
 * Every function below restates one step of the reference's algorithm on the
 * CPU with exact integer arithmetic (unsigned __int128) and IEEE-754 FP64 with
 * explicit rounding modes (<fenv.h>).  The reference file:line each function
 * follows is given in its comment.  Pins (what checks each function against
 * something other than itself) are listed in tests/test_oracle_*.py:
 *   - or_mulmod/or_powmod      : Python big-int `%`; the reference's Barrett
 *                                mod() restated in a test agrees with it.
 *   - or_dgt_forward/inverse   : O(N^2) closed-form DFT over Z_p[i] (Python),
 *                                inverse o forward = id, DGT(const) = const.
 *   - or_polymul               : schoolbook negacyclic product (Python ints)
 *                                and the committed n = 2048, k = 2 golden vector.
 *   - or_basis_ext_q_to_b /
 *     or_basis_ext_b_to_q      : exact CRT value x mod b_i (Python ints) away
 *                                from the 1e-9 snap window; i^2 mod t KAT.
 *   - or_complex_scaling       : exact round(t*x/Q) mod b_i within +-1, as the
 *                                reference's own test allows.
 *   - or_simple_scaling /
 *     or_scale_residue         : exact round(t*x/Q) mod t (Python ints), and
 *                                scaled messages recovered for t = 65537.
 *   - or_pointwise             : Gaussian-integer arithmetic in Python ints.
 *   - or_dgt_poly              : the reference's hierarchical kernels restated
 *                                in Python; full bit reversal of the placement.
 *   - or_negacyclic_schoolbook : Python big-integer negacyclic products.
 *   - or_fma_rd / or_rcp_rd    : exact rational rounding (Python fractions).
 *   - or_philox4x32_10         : Random123 known-answer vectors.
 *   - samplers                 : chi-square against the ideal distributions;
 *                                the fixed-weight sampler also by its exact
 *                                weight, uniform positions and signs, and a
 *                                brute-force smallest-key selection.
 * The word decomposition (xi, rho, dot) and the CRT are restated in
 * oracle.py and pinned there by sum_i rho(a)_i xi(b)_i = a*b and by the CRT
 * identity against Python integers.
 *
 * Layout conventions (same as the reference, src/arithmetic/polynomial.cu:190-221):
 *   a polynomial of ring degree n = 2*Ng over nres residues is stored as
 *   Gaussian integers [rid][Ng] {re, imag}; coefficient j goes to re[j] and
 *   coefficient j + Ng to imag[j].  Here a GI array is a uint64 array of
 *   length 2*Ng per residue, interleaved re, imag.
 */
#include <fenv.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* Modular core: dgt.cu:394-466 (mod, mulmod), dgt.h:73-82 (addmod/submod),  */
/* dgt.cu:508-518 (fast_pow), dgt.cu:547-570 (GIMul, Karatsuba).           */
/* ------------------------------------------------------------------------ */
uint64_t or_mulmod(uint64_t a, uint64_t b, uint64_t p) {
  return (uint64_t)(((u128)a * b) % p);
}

/* addmod as the reference writes it: one conditional subtraction on 64-bit
 * wrap-around arithmetic (dgt.h:73-74).  Used verbatim where the reference
 * feeds it values that need not be reduced (complex scaling). */
static inline uint64_t addmod_ref(uint64_t a, uint64_t b, uint64_t p) {
  uint64_t s = a + b;
  return s >= p ? s - p : s;
}
static inline uint64_t submod_ref(uint64_t a, uint64_t b, uint64_t p) {
  return a < b ? a + (p - b) : a - b;
}

uint64_t or_powmod(uint64_t a, uint64_t e, uint64_t p) {
  uint64_t r = 1 % p, s = a % p;
  while (e) {
    if (e & 1) r = or_mulmod(r, s, p);
    s = or_mulmod(s, s, p);
    e >>= 1;
  }
  return r;
}

uint64_t or_invmod(uint64_t a, uint64_t p) { return or_powmod(a, p - 2, p); }

static inline void gi_mul(uint64_t *cr, uint64_t *ci, uint64_t ar, uint64_t ai,
                          uint64_t br, uint64_t bi, uint64_t p) {
  /* (ar + i ai)(br + i bi) mod p; exact, canonical result. */
  uint64_t re = submod_ref(or_mulmod(ar, br, p), or_mulmod(ai, bi, p), p);
  uint64_t im = addmod_ref(or_mulmod(ar, bi, p), or_mulmod(ai, br, p), p);
  *cr = re;
  *ci = im;
}

/* ------------------------------------------------------------------------ */
/* DGT: dgt.cu:681-1067 (transforms), dgt.cu:1509-1692 (gjk, nthroot),      */
/* dgt.cu:1879-1960 (HDGT bit-reversal and gN/ginvN).                        */
/*                                                                          */
/* The reference's hierarchical Na x Nb DGT places X_k at                   */
/* bitrev_Na(k / Nb) + Na * bitrev_Nb(k % Nb) and its single-kernel DGT at   */
/* bitrev_Ng(k); both equal bitrev_Ng(k) because reversing the bits of      */
/* (k1 : k2) gives (rev k2 : rev k1).  So one in-place radix-2 DIF            */
/* (Gentleman-Sande, as doDGT at dgt.cu:755-777) after the twist xi^n        */
/* restates both.  omega = PROOT^((p-1)/Ng) (dgt.cu:1526-1538, 1900-1904).   */
/* ------------------------------------------------------------------------ */
void or_dgt_forward(uint64_t *x, int Ng, uint64_t p, const uint64_t xi[2],
                    uint64_t omega) {
  /* twist: x[n] *= xi^n  (hdgt_tr<Nb> first overload, dgt.cu:819-862) */
  uint64_t tr = 1, ti = 0;
  for (int n = 0; n < Ng; n++) {
    gi_mul(&x[2 * n], &x[2 * n + 1], x[2 * n], x[2 * n + 1], tr, ti, p);
    gi_mul(&tr, &ti, tr, ti, xi[0], xi[1], p);
  }
  /* DIF stages: stride s has half-size m = Ng >> (s+1), twiddle omega^(j*2^s) */
  for (int m = Ng >> 1, s = 0; m >= 1; m >>= 1, s++) {
    uint64_t ws = or_powmod(omega, (uint64_t)1 << s, p);
    for (int g = 0; g < Ng; g += 2 * m) {
      uint64_t w = 1;
      for (int j = 0; j < m; j++) {
        uint64_t *u = &x[2 * (g + j)], *v = &x[2 * (g + j + m)];
        uint64_t sr = addmod_ref(u[0], v[0], p), si = addmod_ref(u[1], v[1], p);
        uint64_t dr = submod_ref(u[0], v[0], p), di = submod_ref(u[1], v[1], p);
        u[0] = sr;
        u[1] = si;
        v[0] = or_mulmod(dr, w, p);
        v[1] = or_mulmod(di, w, p);
        w = or_mulmod(w, ws, p);
      }
    }
  }
}

/* Inverse: Cooley-Tukey DIT from the last stride to the first with
 * omega^(-j*2^s) (idgttransform, dgt.cu:716-753; doIDGT :781-806), then the
 * untwist xi^(-n) / Ng (the reference folds 1/Ng into d_ginvN or
 * d_invnthroot, dgt.cu:1654-1662, 1933-1939). */
void or_dgt_inverse(uint64_t *x, int Ng, uint64_t p, const uint64_t xiinv[2],
                    uint64_t omega) {
  uint64_t oinv = or_invmod(omega, p);
  int L = 0;
  while ((1 << L) < Ng) L++;
  for (int s = L - 1; s >= 0; s--) {
    int m = Ng >> (s + 1);
    uint64_t ws = or_powmod(oinv, (uint64_t)1 << s, p);
    for (int g = 0; g < Ng; g += 2 * m) {
      uint64_t w = 1;
      for (int j = 0; j < m; j++) {
        uint64_t *u = &x[2 * (g + j)], *v = &x[2 * (g + j + m)];
        uint64_t tr = or_mulmod(v[0], w, p), ti = or_mulmod(v[1], w, p);
        uint64_t ar = addmod_ref(u[0], tr, p), ai = addmod_ref(u[1], ti, p);
        uint64_t br = submod_ref(u[0], tr, p), bi = submod_ref(u[1], ti, p);
        u[0] = ar;
        u[1] = ai;
        v[0] = br;
        v[1] = bi;
        w = or_mulmod(w, ws, p);
      }
    }
  }
  uint64_t ninv = or_invmod((uint64_t)Ng, p);
  uint64_t tr = ninv, ti = 0; /* xi^-n / Ng */
  for (int n = 0; n < Ng; n++) {
    gi_mul(&x[2 * n], &x[2 * n + 1], x[2 * n], x[2 * n + 1], tr, ti, p);
    gi_mul(&tr, &ti, tr, ti, xiinv[0], xiinv[1], p);
  }
}

/* Whole polynomials: nres residues, residue r uses primes[r], xi[2r..2r+1],
 * omega[r].  direction 0 = forward, 1 = inverse (dgt.cu:1069-1094). */
void or_dgt_poly(uint64_t *x, int Ng, int nres, const uint64_t *primes,
                 const uint64_t *xi, const uint64_t *omega, int direction) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < nres; r++) {
    if (direction == 0)
      or_dgt_forward(x + (size_t)2 * Ng * r, Ng, primes[r], xi + 2 * r, omega[r]);
    else
      or_dgt_inverse(x + (size_t)2 * Ng * r, Ng, primes[r], xi + 2 * r, omega[r]);
  }
}

/* ------------------------------------------------------------------------ */
/* Pointwise ops in the DGT domain: dgt.cu:1096-1499.                        */
/* op: 0 add, 1 sub, 2 mul (GIMul), 3 muladd d = a*b + c                    */
/* ------------------------------------------------------------------------ */
void or_pointwise(uint64_t *out, const uint64_t *a, const uint64_t *b,
                  const uint64_t *c, int Ng, int nres, const uint64_t *primes,
                  int op) {
  for (int r = 0; r < nres; r++) {
    uint64_t p = primes[r];
    for (int i = 0; i < Ng; i++) {
      size_t o = (size_t)2 * (r * (size_t)Ng + i);
      uint64_t re, im;
      switch (op) {
        case 0:
          re = addmod_ref(a[o], b[o], p);
          im = addmod_ref(a[o + 1], b[o + 1], p);
          break;
        case 1:
          re = submod_ref(a[o], b[o], p);
          im = submod_ref(a[o + 1], b[o + 1], p);
          break;
        case 2:
          gi_mul(&re, &im, a[o], a[o + 1], b[o], b[o + 1], p);
          break;
        default:
          gi_mul(&re, &im, a[o], a[o + 1], b[o], b[o + 1], p);
          re = addmod_ref(re, c[o], p);
          im = addmod_ref(im, c[o + 1], p);
          break;
      }
      out[o] = re;
      out[o + 1] = im;
    }
  }
}

/* Coefficient-domain negacyclic product through the DGT, for `batch`
 * independent products of nres residues each (poly_mul at
 * polynomial.cu:892-914 followed by the inverse in poly_copy_to_host). */
void or_polymul_batch(uint64_t *c, const uint64_t *a, const uint64_t *b,
                      int batch, int Ng, int nres, const uint64_t *primes,
                      const uint64_t *xi, const uint64_t *xiinv,
                      const uint64_t *omega) {
  size_t rs = (size_t)2 * Ng;
#pragma omp parallel
  {
    uint64_t *ta = (uint64_t *)malloc(rs * sizeof(uint64_t));
    uint64_t *tb = (uint64_t *)malloc(rs * sizeof(uint64_t));
#pragma omp for schedule(dynamic)
    for (long w = 0; w < (long)batch * nres; w++) {
      int r = (int)(w % nres);
      size_t off = (size_t)w * rs;
      uint64_t p = primes[r];
      memcpy(ta, a + off, rs * sizeof(uint64_t));
      memcpy(tb, b + off, rs * sizeof(uint64_t));
      or_dgt_forward(ta, Ng, p, xi + 2 * r, omega[r]);
      or_dgt_forward(tb, Ng, p, xi + 2 * r, omega[r]);
      for (int i = 0; i < Ng; i++)
        gi_mul(&ta[2 * i], &ta[2 * i + 1], ta[2 * i], ta[2 * i + 1], tb[2 * i],
               tb[2 * i + 1], p);
      or_dgt_inverse(ta, Ng, p, xiinv + 2 * r, omega[r]);
      memcpy(c + off, ta, rs * sizeof(uint64_t));
    }
    free(ta);
    free(tb);
  }
}

/* Schoolbook negacyclic product in Z_p[X]/(X^n + 1) on flat coefficient
 * vectors: the ring product NTL's ZZ_pX computes in the reference's Mul test
 * (tests/unit_tests/test_bfv_mode.cu:1180-1244). */
void or_negacyclic_schoolbook(uint64_t *c, const uint64_t *a, const uint64_t *b,
                              int n, uint64_t p) {
  for (int i = 0; i < n; i++) c[i] = 0;
  for (int i = 0; i < n; i++) {
    if (!a[i]) continue;
    for (int j = 0; j < n; j++) {
      uint64_t t = or_mulmod(a[i], b[j], p);
      int k = i + j;
      if (k < n)
        c[k] = addmod_ref(c[k], t, p);
      else
        c[k - n] = submod_ref(c[k - n], t, p);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* FP64 with explicit rounding                                              */
/* ------------------------------------------------------------------------ */
/* RD(a*b + c), one rounding (the reference's __fma_rd). */
double or_fma_rd(double a, double b, double c) {
  int old = fegetround();
  fesetround(FE_DOWNWARD);
  volatile double r = fma(a, b, c);
  fesetround(old);
  return r;
}

/* RD(1 / RN((double)q)) (the reference's __drcp_rd((double)q)). */
double or_rcp_rd(uint64_t q) {
  volatile double qd = (double)q; /* RN conversion, default mode */
  int old = fegetround();
  fesetround(FE_DOWNWARD);
  volatile double r = 1.0 / qd;
  fesetround(old);
  return r;
}

/* The "snap" the reference applies to v (cudaengine.cu:530-531, :702):
 * nearest integer when within FPERROR = 1e-9, else floor. */
static inline uint64_t snap_v(double value) {
  long long rn = llrint(value); /* round half even, default mode */
  double diff = fabs(value - (double)rn);
  return diff <= 1e-9 ? (uint64_t)rn : (uint64_t)floor(value);
}

/* ------------------------------------------------------------------------ */
/* HPS basis extension Q -> B: cudaengine.cu:483-649.                       */
/* in : [k][Ng] GI, coefficient domain, residues mod q_j                    */
/* out: [kl][Ng] GI, residues mod b_i                                       */
/* invQqi[j] = (Q/q_j)^-1 mod q_j; Qqibi[i*k + j] = (Q/q_j) mod b_i;         */
/* Qbi[i] = Q mod b_i; rcp[j] = or_rcp_rd(q_j).                              */
/* ------------------------------------------------------------------------ */
void or_basis_ext_q_to_b(uint64_t *out, const uint64_t *in, int Ng, int k, int kl,
                         const uint64_t *q, const uint64_t *b,
                         const uint64_t *invQqi, const uint64_t *Qqibi,
                         const uint64_t *Qbi, const double *rcp, uint64_t *v_out) {
  int n = 2 * Ng;
  uint64_t *aux = (uint64_t *)malloc((size_t)k * n * sizeof(uint64_t));
  double *auxd = (double *)malloc((size_t)k * sizeof(double));
  /* polynomial_basis_ext_Q_aux: aux[j][c] = x_j * invQqi[j] mod q_j, split
   * into the re half (c < Ng) and imag half (c >= Ng). */
  for (int j = 0; j < k; j++)
    for (int c = 0; c < Ng; c++) {
      aux[(size_t)j * n + c] = or_mulmod(in[2 * ((size_t)j * Ng + c)], invQqi[j], q[j]);
      aux[(size_t)j * n + c + Ng] =
          or_mulmod(in[2 * ((size_t)j * Ng + c) + 1], invQqi[j], q[j]);
    }
  for (int c = 0; c < n; c++) {
    /* polynomial_basis_ext_Q_compute_v: value = fma_rd(aux, rcp_rd(q), value)
     * in ascending j, then the snap. */
    for (int j = 0; j < k; j++) auxd[j] = (double)aux[(size_t)j * n + c];
    double value = 0;
    for (int j = 0; j < k; j++) value = or_fma_rd(auxd[j], rcp[j], value);
    uint64_t v = snap_v(value);
    if (v_out) v_out[c] = v;
    /* polynomial_basis_ext_Q_to_B */
    for (int i = 0; i < kl; i++) {
      uint64_t bi = b[i], acc = 0;
      for (int j = 0; j < k; j++)
        acc = addmod_ref(acc, or_mulmod(aux[(size_t)j * n + c], Qqibi[i * k + j], bi), bi);
      acc = submod_ref(acc, or_mulmod(v, Qbi[i], bi), bi);
      if (c < Ng)
        out[2 * ((size_t)i * Ng + c)] = acc;
      else
        out[2 * ((size_t)i * Ng + c - Ng) + 1] = acc;
    }
  }
  free(aux);
  free(auxd);
}

/* 128 x 64 -> upper 128 bits of the 192-bit product (cudaengine.cu:313-320). */
static inline void umul128hi(uint64_t alo, uint64_t ahi, uint64_t b, uint64_t *lo,
                             uint64_t *hi) {
  u128 x = (u128)alo * b;
  u128 y = (u128)ahi * b + (uint64_t)(x >> 64);
  *lo = (uint64_t)y;
  *hi = (uint64_t)(y >> 64);
}

/* ------------------------------------------------------------------------ */
/* HPS complex scaling t/Q: cudaengine.cu:382-440.                          */
/* aQ: [k][Ng] GI, aB: [kl][Ng] GI, both coefficient domain.                 */
/* out: uint64 [kl][2Ng] (re half, then imag half), as ctx->d_coefs_B.       */
/* wint[i*k + j], wfrac[2*(i*k + j)] = {lo, hi}, lambda[i].                  */
/* ------------------------------------------------------------------------ */
void or_complex_scaling(uint64_t *out, const uint64_t *aQ, const uint64_t *aB, int Ng,
                        int k, int kl, const uint64_t *b, const uint64_t *wint,
                        const uint64_t *wfrac, const uint64_t *lambda) {
  const double two_m64 = 0x1p-64; /* __drcp_rn(2^64) */
  for (int i = 0; i < kl; i++) {
    uint64_t bi = b[i];
    for (int c = 0; c < Ng; c++) {
      double acc_re = 0, acc_im = 0;
      uint64_t res_re = 0, res_im = 0;
      for (int j = 0; j < k; j++) {
        uint64_t vre = aQ[2 * ((size_t)j * Ng + c)], vim = aQ[2 * ((size_t)j * Ng + c) + 1];
        uint64_t w = wint[i * k + j];
        uint64_t int_re = or_mulmod(vre, w, bi), int_im = or_mulmod(vim, w, bi);
        uint64_t flo = wfrac[2 * (i * k + j)], fhi = wfrac[2 * (i * k + j) + 1];
        uint64_t lre, hre, lim, him;
        umul128hi(flo, fhi, vre, &lre, &hre);
        umul128hi(flo, fhi, vim, &lim, &him);
        acc_re = fma((double)lre, two_m64, acc_re);
        acc_im = fma((double)lim, two_m64, acc_im);
        res_re = addmod_ref(res_re, addmod_ref(int_re, hre, bi), bi);
        res_im = addmod_ref(res_im, addmod_ref(int_im, him, bi), bi);
      }
      res_re = addmod_ref(res_re, (uint64_t)llrint(acc_re), bi);
      res_im = addmod_ref(res_im, (uint64_t)llrint(acc_im), bi);
      uint64_t xre = aB[2 * ((size_t)i * Ng + c)], xim = aB[2 * ((size_t)i * Ng + c) + 1];
      res_re = addmod_ref(res_re, or_mulmod(xre, lambda[i], bi), bi);
      res_im = addmod_ref(res_im, or_mulmod(xim, lambda[i], bi), bi);
      out[(size_t)i * 2 * Ng + c] = res_re;
      out[(size_t)i * 2 * Ng + c + Ng] = res_im;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* HPS basis extension B -> Q: cudaengine.cu:659-751.                       */
/* in : uint64 [kl][2Ng] (complex-scaling layout).  out: [k][Ng] GI.         */
/* invBbi[i] = (B/b_i)^-1 mod b_i; Bbiqi[j*kl + i] = (B/b_i) mod q_j;        */
/* Bqi[j] = B mod q_j.                                                      */
/* ------------------------------------------------------------------------ */
void or_basis_ext_b_to_q(uint64_t *out, const uint64_t *in, int Ng, int k, int kl,
                         const uint64_t *q, const uint64_t *b,
                         const uint64_t *invBbi, const uint64_t *Bbiqi,
                         const uint64_t *Bqi, uint64_t *v_out) {
  int n = 2 * Ng;
  uint64_t *aux = (uint64_t *)malloc((size_t)kl * sizeof(uint64_t));
  for (int c = 0; c < n; c++) {
    double value = 0;
    for (int i = 0; i < kl; i++) {
      aux[i] = or_mulmod(in[(size_t)i * n + c], invBbi[i], b[i]);
      value += (double)aux[i] / (double)b[i];
    }
    uint64_t v = snap_v(value);
    if (v_out) v_out[c] = v;
    for (int j = 0; j < k; j++) {
      uint64_t qj = q[j], acc = 0;
      for (int i = 0; i < kl; i++)
        acc = addmod_ref(acc, or_mulmod(aux[i], Bbiqi[j * kl + i], qj), qj);
      acc = submod_ref(acc, or_mulmod(v, Bqi[j], qj), qj);
      if (c < Ng)
        out[2 * ((size_t)j * Ng + c)] = acc;
      else
        out[2 * ((size_t)j * Ng + c - Ng) + 1] = acc;
    }
  }
  free(aux);
}

/* ------------------------------------------------------------------------ */
/* HPS simple scaling Q -> t, power-of-two t: cudaengine.cu:323-363.         */
/* in: [k][Ng] GI coefficient domain; out: [Ng] GI residues mod t.           */
/* twint[j] = floor(t * (Q/q_j)^-1 / q_j), twfrac[2j..] its fraction * 2^128 */
/* ------------------------------------------------------------------------ */
/* cudaengine.cu:323-363 for t a power of two (masks); for other t < 2^32
 * the same sums with exact reductions mod t (x_j reduced before the
 * product, which the reference's masked form cannot do). */
void or_simple_scaling(uint64_t *out, const uint64_t *in, int Ng, int k, uint64_t t,
                       const uint64_t *twint, const uint64_t *twfrac) {
  const double two_m64 = 0x1p-64;
  const int pow2 = (t & (t - 1)) == 0;
  uint64_t mask = t - 1;
  for (int c = 0; c < Ng; c++) {
    double acc[2] = {0, 0};
    uint64_t res[2] = {0, 0};
    for (int j = 0; j < k; j++)
      for (int h = 0; h < 2; h++) {
        uint64_t v = in[2 * ((size_t)j * Ng + c) + h];
        uint64_t lo, hi;
        umul128hi(twfrac[2 * j], twfrac[2 * j + 1], v, &lo, &hi);
        acc[h] = fma((double)lo, two_m64, acc[h]);
        if (pow2)
          res[h] += (((v * twint[j]) & mask) + (hi & mask)) & mask;
        else
          res[h] += ((v % t) * twint[j] % t + hi % t) % t;
      }
    for (int h = 0; h < 2; h++) {
      uint64_t r = res[h] + (uint64_t)llrint(acc[h]);
      out[2 * (size_t)c + h] = pow2 ? (r & mask) : (r % t);
    }
  }
}

/* xi / rho of the BFV word decomposition (cudaengine.cu:992-1062,
 * mulint_dgt): one residue times a constant, both halves of every slot. */
void or_scale_residue(uint64_t *out, const uint64_t *in, int Ng, uint64_t p, uint64_t c) {
  for (size_t i = 0; i < (size_t)2 * Ng; i++) out[i] = or_mulmod(in[i], c, p);
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as  */
/* easy as 1, 2, 3", SC'11).                                                */
/* ------------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Sampler counter layout (this project's; the reference's cuRAND XORWOW
 * stream is not reproducible):
 *   key     = (seed & 0xffffffff, seed >> 32)
 *   counter = (coefficient index, tag << 24 | residue << 16 | attempt,
 *              stream & 0xffffffff, stream >> 32)
 * tag 1 = uniform mod q_i, 2 = ternary, 3 = discrete Gaussian.
 * Output: flat coefficient values, `n` per residue for uniform, `n` signed
 * values for the small distributions. */
static void philox_draw(uint64_t seed, uint64_t stream, uint32_t coef, uint32_t tag,
                        uint32_t rid, uint32_t attempt, uint32_t out[4]) {
  uint32_t ctr[4] = {coef, (tag << 24) | (rid << 16) | attempt, (uint32_t)stream,
                     (uint32_t)(stream >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  or_philox4x32_10(ctr, key, out);
}

uint64_t or_sample_uniform_coef(uint64_t seed, uint64_t stream, uint32_t coef,
                                uint32_t rid, uint64_t q) {
  uint64_t lim = (UINT64_MAX / q) * q; /* accept x < floor(2^64/q)*q ... */
  /* ... using the complement form so that lim never overflows: values in
   * [lim, 2^64) are rejected; UINT64_MAX/q*q <= 2^64 - 1. */
  for (uint32_t att = 0;; att++) {
    uint32_t w[4];
    philox_draw(seed, stream, coef, 1, rid, att, w);
    uint64_t x0 = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
    if (x0 < lim) return x0 % q;
    uint64_t x1 = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
    if (x1 < lim) return x1 % q;
  }
}

int or_sample_ternary_coef(uint64_t seed, uint64_t stream, uint32_t coef) {
  for (uint32_t att = 0;; att++) {
    uint32_t w[4];
    philox_draw(seed, stream, coef, 2, 0, att, w);
    for (int h = 0; h < 4; h++)
      if (w[h] < 4294967295u) return (int)(w[h] % 3u) - 1;
  }
}

/* cdt[0..2B] : cumulative table, cdt[i] = round(2^64 * P(e <= i - B)) with
 * cdt[2B] implicit 2^64 (passed as UINT64_MAX and compared with <=). */
int or_sample_gaussian_coef(uint64_t seed, uint64_t stream, uint32_t coef,
                            const uint64_t *cdt, int bound) {
  uint32_t w[4];
  philox_draw(seed, stream, coef, 3, 0, 0, w);
  uint64_t u = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
  int idx = 0;
  while (idx < 2 * bound && u >= cdt[idx]) idx++;
  return idx - bound;
}

/* narrow: uniform {0, 1, 2} (sampler.cu:72-103, BFV mode), tag 4, the
 * same exact rejection as the ternary draw. */
int or_sample_narrow_coef(uint64_t seed, uint64_t stream, uint32_t coef) {
  for (uint32_t att = 0;; att++) {
    uint32_t w[4];
    philox_draw(seed, stream, coef, 4, 0, att, w);
    for (int h = 0; h < 4; h++)
      if (w[h] < 4294967295u) return (int)(w[h] % 3u);
  }
}

/* binary: uniform {0, 1} (sampler.cu:158-182), tag 5. */
int or_sample_binary_coef(uint64_t seed, uint64_t stream, uint32_t coef) {
  uint32_t w[4];
  philox_draw(seed, stream, coef, 5, 0, 0, w);
  return (int)(w[0] & 1u);
}

/* Fixed weight h (ZO: h = n/2, sampler.cu:222-278; Hamming weight: h = 64,
 * sampler.cu:401-446): the nonzero positions are the h coefficients with
 * the smallest keys, key = 48 random bits (tag 6) above the 16-bit index
 * (unique), each nonzero +1 or -1 by one more random bit. */
uint64_t or_fixed_weight_key(uint64_t seed, uint64_t stream, uint32_t coef, int *neg) {
  uint32_t w[4];
  philox_draw(seed, stream, coef, 6, 0, 0, w);
  *neg = (int)(w[2] & 1u);
  return ((((uint64_t)w[0] << 16) | (w[1] >> 16)) << 16) | (coef & 0xFFFFu);
}

static int cmp_u64(const void *x, const void *y) {
  const uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
  return a < b ? -1 : a > b;
}

void or_sample_fixed_weight(int64_t *out, int n, int h, uint64_t seed, uint64_t stream) {
  uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
  for (int c = 0; c < n; c++) {
    int neg;
    keys[c] = or_fixed_weight_key(seed, stream, (uint32_t)c, &neg);
    out[c] = 0;
  }
  qsort(keys, (size_t)n, sizeof(uint64_t), cmp_u64);
  for (int i = 0; i < h && i < n; i++) {
    const uint32_t c = (uint32_t)(keys[i] & 0xFFFFu);
    int neg;
    (void)or_fixed_weight_key(seed, stream, c, &neg);
    out[c] = neg ? -1 : 1;
  }
  free(keys);
}

/* Fill flat coefficient arrays for a whole polynomial. */
void or_sample_uniform_poly(uint64_t *out, int n, int nres, const uint64_t *primes,
                            uint64_t seed, uint64_t stream) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < nres; r++)
    for (int c = 0; c < n; c++)
      out[(size_t)r * n + c] = or_sample_uniform_coef(seed, stream, c, r, primes[r]);
}

void or_sample_small_poly(int64_t *out, int n, int kind, uint64_t seed, uint64_t stream,
                          const uint64_t *cdt, int bound) {
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; c++)
    out[c] = kind == 2   ? or_sample_ternary_coef(seed, stream, c)
             : kind == 4 ? or_sample_narrow_coef(seed, stream, c)
             : kind == 5 ? or_sample_binary_coef(seed, stream, c)
                         : or_sample_gaussian_coef(seed, stream, c, cdt, bound);
}

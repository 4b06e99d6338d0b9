/*
 * cpu_port.c - the CPU baseline: an optimised CPU implementation of the
 * batched negacyclic product c = a * b mod (X^n + 1, q_r), the algorithm of
 * the reference's poly_mul (src/arithmetic/polynomial.cu:892-914): twist by
 * xi^n, forward DGT of a and b (dgt.cu:819-908), slot-wise Gaussian product
 * (dgt.cu:547-570), inverse DGT and untwist xi^-n / N (dgt.cu:917-1006).
 *
 * TEST INFRASTRUCTURE, like the rest of oracle/: only the benchmark's CPU
 * baseline leg and the tests load it; the product path never does.  It is
 * what a CPU port of the path looks like when written for speed (the
 * oracle in cupoly_oracle.c is written for clarity: a 128-bit division per
 * product, twiddles rebuilt inside the butterflies):
 *   - twiddles omega^e and the twist / untwist factors are precomputed once
 *     per plan with their Shoup quotients w' = floor(w 2^64 / p), so every
 *     constant product is one 64x64 high product and two low products;
 *   - each forward stage reads its twiddles from a contiguous table;
 *   - products of two variables fold the 127-bit product twice with
 *     2^63 = delta (mod p), p = 2^63 - delta (every prime of the 63-bit
 *     bucket has delta < 2^30);
 *   - no 128-bit division anywhere outside plan creation;
 *   - OpenMP over (product, residue) units, one working copy of a and b per
 *     unit in the thread's cache.
 * Results are exact residues, so they are bit-identical to the GPU's and
 * the oracle's (tests/test_oracle_pins.py checks the port against the
 * oracle; the benchmark checks it against the GPU result it times).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

typedef struct {
  int ng, logng, nres;
  uint64_t *p;      /* [nres] */
  uint64_t *tw;     /* [nres][ng - 1][2]: stage s (half size m = ng >> (s+1)) at
                       offset ng - 2m: omega^(j 2^s), j < m, with Shoup quotient */
  uint64_t *twi;    /* same for omega^-1 */
  uint64_t *ts;     /* [nres][ng][6]: xi^n as (c, c', d, d', c + d, (c + d)') */
  uint64_t *tsi;    /* [nres][ng][6]: xi^-n / ng */
} cpu_plan;

static uint64_t mulmod_slow(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)(((u128)a * b) % p); }
static uint64_t powmod_slow(uint64_t a, uint64_t e, uint64_t p) {
  uint64_t r = 1 % p;
  a %= p;
  while (e) {
    if (e & 1) r = mulmod_slow(r, a, p);
    a = mulmod_slow(a, a, p);
    e >>= 1;
  }
  return r;
}
static uint64_t shoup_q(uint64_t w, uint64_t p) { return (uint64_t)(((u128)w << 64) / p); }

static inline uint64_t red2(uint64_t a, uint64_t p) { return a >= p ? a - p : a; }
static inline uint64_t addm(uint64_t a, uint64_t b, uint64_t p) { return red2(a + b, p); }
static inline uint64_t subm(uint64_t a, uint64_t b, uint64_t p) { return a >= b ? a - b : a + p - b; }
/* a w mod p for any 64-bit a, w < p, wq = floor(w 2^64 / p) */
static inline uint64_t shoup(uint64_t a, uint64_t w, uint64_t wq, uint64_t p) {
  const uint64_t q = (uint64_t)(((u128)a * wq) >> 64);
  return red2(a * w - q * p, p);
}
/* a b mod p, a b < 2^127, p = 2^63 - delta */
static inline uint64_t mulm(uint64_t a, uint64_t b, uint64_t p) {
  const uint64_t delta = (1ULL << 63) - p;
  const u128 x = (u128)a * b;
  const uint64_t L = (uint64_t)x & ((1ULL << 63) - 1), H = (uint64_t)(x >> 63);
  const u128 y = (u128)H * delta + L;
  const uint64_t z = ((uint64_t)y & ((1ULL << 63) - 1)) + (uint64_t)(y >> 63) * delta;
  return red2(z, p);
}
/* (xr + i xi) * (c + i d) with the precomputed (c, c', d, d', c+d, (c+d)') */
static inline void gi_const(uint64_t *xr, uint64_t *xim, const uint64_t *k, uint64_t p) {
  const uint64_t s1 = shoup(*xr, k[0], k[1], p), s2 = shoup(*xim, k[2], k[3], p);
  const uint64_t s3 = shoup(*xr + *xim, k[4], k[5], p);
  *xr = subm(s1, s2, p);
  *xim = subm(s3, addm(s1, s2, p), p);
}

cpu_plan *cpu_plan_create(int ng, int nres, const uint64_t *primes, const uint64_t *xi,
                          const uint64_t *xiinv, const uint64_t *omega) {
  cpu_plan *pl = (cpu_plan *)calloc(1, sizeof(cpu_plan));
  if (!pl) return NULL;
  pl->ng = ng;
  pl->nres = nres;
  while ((1 << pl->logng) < ng) pl->logng++;
  pl->p = (uint64_t *)malloc(sizeof(uint64_t) * nres);
  pl->tw = (uint64_t *)malloc(sizeof(uint64_t) * 2 * (size_t)nres * ng);
  pl->twi = (uint64_t *)malloc(sizeof(uint64_t) * 2 * (size_t)nres * ng);
  pl->ts = (uint64_t *)malloc(sizeof(uint64_t) * 6 * (size_t)nres * ng);
  pl->tsi = (uint64_t *)malloc(sizeof(uint64_t) * 6 * (size_t)nres * ng);
  if (!pl->p || !pl->tw || !pl->twi || !pl->ts || !pl->tsi) return NULL;
  for (int r = 0; r < nres; r++) {
    const uint64_t p = primes[r];
    pl->p[r] = p;
    const uint64_t om = omega[r], omi = powmod_slow(om, p - 2, p);
    uint64_t *tw = pl->tw + 2 * (size_t)r * ng, *twi = pl->twi + 2 * (size_t)r * ng;
    for (int s = 0, m = ng >> 1; m >= 1; s++, m >>= 1) {
      const uint64_t ws = powmod_slow(om, (uint64_t)1 << s, p), wsi = powmod_slow(omi, (uint64_t)1 << s, p);
      uint64_t w = 1, wi = 1;
      for (int j = 0; j < m; j++) {
        const size_t o = 2 * ((size_t)(ng - 2 * m) + j);
        tw[o] = w;
        tw[o + 1] = shoup_q(w, p);
        twi[o] = wi;
        twi[o + 1] = shoup_q(wi, p);
        w = mulmod_slow(w, ws, p);
        wi = mulmod_slow(wi, wsi, p);
      }
    }
    const uint64_t ninv = powmod_slow((uint64_t)ng, p - 2, p);
    uint64_t fr = 1, fi = 0, gr = ninv, gi = 0;
    for (int g = 0; g < ng; g++) {
      uint64_t *t = pl->ts + 6 * ((size_t)r * ng + g), *u = pl->tsi + 6 * ((size_t)r * ng + g);
      t[0] = fr; t[1] = shoup_q(fr, p); t[2] = fi; t[3] = shoup_q(fi, p);
      t[4] = (fr + fi) % p; t[5] = shoup_q(t[4], p);
      u[0] = gr; u[1] = shoup_q(gr, p); u[2] = gi; u[3] = shoup_q(gi, p);
      u[4] = (gr + gi) % p; u[5] = shoup_q(u[4], p);
      /* f *= xi, g *= xi^-1 (Gaussian products) */
      const uint64_t a = xi[2 * r], b = xi[2 * r + 1], c = xiinv[2 * r], d = xiinv[2 * r + 1];
      const uint64_t nr = subm(mulmod_slow(fr, a, p), mulmod_slow(fi, b, p), p);
      const uint64_t ni = addm(mulmod_slow(fr, b, p), mulmod_slow(fi, a, p), p);
      const uint64_t mr = subm(mulmod_slow(gr, c, p), mulmod_slow(gi, d, p), p);
      const uint64_t mi = addm(mulmod_slow(gr, d, p), mulmod_slow(gi, c, p), p);
      fr = nr; fi = ni; gr = mr; gi = mi;
    }
  }
  return pl;
}

void cpu_plan_destroy(cpu_plan *pl) {
  if (!pl) return;
  free(pl->p); free(pl->tw); free(pl->twi); free(pl->ts); free(pl->tsi);
  free(pl);
}

/* x: ng Gaussian slots (interleaved re, im) of residue r */
static void fwd(const cpu_plan *pl, int r, uint64_t *x) {
  const int ng = pl->ng;
  const uint64_t p = pl->p[r];
  const uint64_t *ts = pl->ts + 6 * (size_t)r * ng, *tw = pl->tw + 2 * (size_t)r * ng;
  for (int g = 0; g < ng; g++) gi_const(&x[2 * g], &x[2 * g + 1], ts + 6 * (size_t)g, p);
  for (int m = ng >> 1; m >= 1; m >>= 1) {
    const uint64_t *w = tw + 2 * (size_t)(ng - 2 * m);
    for (int g = 0; g < ng; g += 2 * m) {
      uint64_t *u = x + 2 * (size_t)g, *v = u + 2 * (size_t)m;
      for (int j = 0; j < m; j++) {
        const uint64_t ur = u[2 * j], ui = u[2 * j + 1], vr = v[2 * j], vi = v[2 * j + 1];
        u[2 * j] = addm(ur, vr, p);
        u[2 * j + 1] = addm(ui, vi, p);
        v[2 * j] = shoup(ur + p - vr, w[2 * j], w[2 * j + 1], p);
        v[2 * j + 1] = shoup(ui + p - vi, w[2 * j], w[2 * j + 1], p);
      }
    }
  }
}

static void inv(const cpu_plan *pl, int r, uint64_t *x) {
  const int ng = pl->ng;
  const uint64_t p = pl->p[r];
  const uint64_t *tsi = pl->tsi + 6 * (size_t)r * ng, *twi = pl->twi + 2 * (size_t)r * ng;
  for (int m = 1; m < ng; m <<= 1) {
    const uint64_t *w = twi + 2 * (size_t)(ng - 2 * m);
    for (int g = 0; g < ng; g += 2 * m) {
      uint64_t *u = x + 2 * (size_t)g, *v = u + 2 * (size_t)m;
      for (int j = 0; j < m; j++) {
        const uint64_t tr = shoup(v[2 * j], w[2 * j], w[2 * j + 1], p);
        const uint64_t ti = shoup(v[2 * j + 1], w[2 * j], w[2 * j + 1], p);
        const uint64_t ur = u[2 * j], ui = u[2 * j + 1];
        u[2 * j] = addm(ur, tr, p);
        u[2 * j + 1] = addm(ui, ti, p);
        v[2 * j] = subm(ur, tr, p);
        v[2 * j + 1] = subm(ui, ti, p);
      }
    }
  }
  for (int g = 0; g < ng; g++) gi_const(&x[2 * g], &x[2 * g + 1], tsi + 6 * (size_t)g, p);
}

/* c[i] = a[i] * b[i] for `batch` products of nres residues of ng slots each
 * ([batch][nres][ng][2] words, coefficient domain); c may not alias a or b.
 * Returns 0, or -1 if a working buffer could not be allocated. */
int cpu_polymul_batch(const cpu_plan *pl, uint64_t *c, const uint64_t *a, const uint64_t *b,
                      int batch) {
  const int ng = pl->ng, nres = pl->nres;
  const long units = (long)batch * nres;
  int fail = 0;
#pragma omp parallel
  {
    uint64_t *xb = (uint64_t *)malloc(sizeof(uint64_t) * 2 * (size_t)ng);
    if (!xb) {
#pragma omp atomic write
      fail = 1;
    }
#pragma omp for schedule(static)
    for (long u = 0; u < units; u++) {
      if (!xb) continue;
      const int r = (int)(u % nres);
      const uint64_t p = pl->p[r];
      const size_t off = (size_t)u * 2 * ng;
      uint64_t *xa = c + off;  /* a's spectrum, then the product, in place in c */
      memcpy(xa, a + off, sizeof(uint64_t) * 2 * (size_t)ng);
      memcpy(xb, b + off, sizeof(uint64_t) * 2 * (size_t)ng);
      fwd(pl, r, xa);
      fwd(pl, r, xb);
      for (int g = 0; g < ng; g++) {
        const uint64_t ar = xa[2 * g], ai = xa[2 * g + 1], br = xb[2 * g], bi = xb[2 * g + 1];
        const uint64_t s1 = mulm(ar, br, p), s2 = mulm(ai, bi, p);
        const uint64_t s3 = mulm(addm(ar, ai, p), addm(br, bi, p), p);
        xa[2 * g] = subm(s1, s2, p);
        xa[2 * g + 1] = subm(s3, addm(s1, s2, p), p);
      }
      inv(pl, r, xa);
    }
    free(xb);
  }
  return fail ? -1 : 0;
}

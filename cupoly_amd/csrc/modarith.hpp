// Modular arithmetic on 63-bit RNS primes, shared by host and device code.
//
// The reference reduces every 64x64 product with a Barrett step
// (src/cuda/dgt.cu:394-466) and multiplies Gaussian integers with Karatsuba
// (dgt.cu:547-570).  Its results are exact residues in [0, p), so any exact
// reduction reproduces them bit for bit.  Here:
//   * products by a constant known in advance (twiddles, twists, RNS
//     constants) use Shoup's precomputed quotient w' = floor(w * 2^64 / p):
//     one high and two low 64-bit products, valid for any 64-bit input;
//   * products of two variables fold the 127-bit product twice with
//     2^63 = delta (mod p) (pseudo-Mersenne reduction), then one correction.
// Both use p = 2^63 - delta with delta < 2^30; every prime of the 63-bit
// bucket has delta < 2^29 (the engine checks the bound).  On gfx950 the
// 128-bit products lower to v_mad_u64_u32 / v_mul_hi_u32 chains on the
// integer pipes.
#pragma once
#include <stdint.h>

#ifndef CP_HD
#define CP_HD __host__ __device__ __forceinline__
#endif

typedef unsigned __int128 cp_u128;

struct alignas(16) cp_gi {  // GaussianInteger {re, imag}, 16 bytes (include/cuPoly/settings.h:70-80)
  uint64_t re, imag;
};

// One RNS prime, 2^63 - 2^30 < p < 2^63.
struct cp_prime {
  uint64_t p;
};

CP_HD uint64_t cp_mulhi(uint64_t a, uint64_t b) {
  return (uint64_t)(((cp_u128)a * b) >> 64);
}

// High words of two 64 x 64 products, hx = (x bx) >> 64 and hy = (y by) >> 64.
// The middle column of the schoolbook product, x1 bx0 + x0 bx1 + hi(x0 bx0),
// can reach 2^65.  Left to itself the compiler keeps it below 2^64 by
// splitting it into zero-extended 32-bit halves: four register moves and a
// 64-bit add per product.  Two cheaper forms:
//   * cp_mulhi2: the middle column as a 128-bit sum, which the compiler turns
//     into an add with carry into a zero word (seven instructions, no move);
//     valid for operands in SGPRs (uniform twiddles);
//   * cp_mulhi2v: the second multiply-add of the column takes the first's
//     full 64-bit result, and its carry-out (an SGPR lane mask) goes into the
//     top word with one add-with-carry: six instructions per product.  Its asm
//     takes every operand in VGPRs.  The two products share each asm block,
//     so at least two other VALU instructions separate a carry's write from
//     its read (gfx950 needs two wait states between a VALU write of a lane
//     mask and a VALU read of it as a carry).
CP_HD uint64_t cp_mulhi_col(uint64_t x, uint64_t b) {
  const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
  const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t m = (uint64_t)x0 * b1 + (uint64_t)(((uint64_t)x0 * b0) >> 32);  // < 2^64
  const cp_u128 s = (cp_u128)((uint64_t)x1 * b0) + m;                         // < 2^65
  return (uint64_t)x1 * b1 + (uint64_t)(s >> 32);
}
CP_HD void cp_mulhi2(uint64_t x, uint64_t bx, uint64_t y, uint64_t by, uint64_t &hx,
                     uint64_t &hy) {
  hx = cp_mulhi_col(x, bx);
  hy = cp_mulhi_col(y, by);
}
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ void cp_mulhi2v(uint64_t x, uint64_t bx, uint64_t y, uint64_t by,
                                           uint64_t &hx, uint64_t &hy) {
  const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
  const uint32_t y0 = (uint32_t)y, y1 = (uint32_t)(y >> 32);
  const uint32_t bx0 = (uint32_t)bx, bx1 = (uint32_t)(bx >> 32);
  const uint32_t by0 = (uint32_t)by, by1 = (uint32_t)(by >> 32);
  // x0 bx1 + hi(x0 bx0) < 2^64
  const uint64_t mx = (uint64_t)x0 * bx1 + __umulhi(x0, bx0);
  const uint64_t my = (uint64_t)y0 * by1 + __umulhi(y0, by0);
  uint64_t sx, sy, cx, cy;  // sx + cx 2^64 = x1 bx0 + mx
  asm("v_mad_u64_u32 %0, %2, %4, %5, %6\n\t"
      "v_mad_u64_u32 %1, %3, %7, %8, %9"
      : "=&v"(sx), "=&v"(sy), "=&s"(cx), "=&s"(cy)
      : "v"(x1), "v"(bx0), "v"(mx), "v"(y1), "v"(by0), "v"(my));
  const uint64_t tx = (uint64_t)x1 * bx1 + (sx >> 32);  // < 2^64 - 2^32
  const uint64_t ty = (uint64_t)y1 * by1 + (sy >> 32);
  uint32_t ux, uy;
  uint64_t jx, jy;
  asm("v_addc_co_u32_e64 %0, %2, %4, 0, %5\n\t"
      "v_addc_co_u32_e64 %1, %3, %6, 0, %7"
      : "=&v"(ux), "=&v"(uy), "=&s"(jx), "=&s"(jy)
      : "v"((uint32_t)(tx >> 32)), "s"(cx), "v"((uint32_t)(ty >> 32)), "s"(cy));
  hx = ((uint64_t)ux << 32) | (uint32_t)tx;
  hy = ((uint64_t)uy << 32) | (uint32_t)ty;
}
#else
CP_HD void cp_mulhi2v(uint64_t x, uint64_t bx, uint64_t y, uint64_t by, uint64_t &hx,
                      uint64_t &hy) {
  cp_mulhi2(x, bx, y, by, hx, hy);
}
#endif

CP_HD uint64_t cp_add(uint64_t a, uint64_t b, uint64_t p) {
  uint64_t s = a + b;  // a, b < p < 2^63: no wrap
  return s >= p ? s - p : s;
}

CP_HD uint64_t cp_sub(uint64_t a, uint64_t b, uint64_t p) {
  return a >= b ? a - b : a + (p - b);
}

// Reduce from [0, 2p) to [0, p), p < 2^63.  t = a - p is formed as a + (2^64 - p):
// for a >= p it is below p (top bit clear), for a < p it is at least
// 2^64 - p > 2^63 (top bit set), so the sign of t selects.  On the GPU this is
// one 64-bit add (v_lshl_add_u64), one compare and two selects; the empty asm
// hides 2^64 - p from the optimizer, which would otherwise turn the add back
// into a two-instruction subtract with borrow.
CP_HD uint64_t cp_red2(uint64_t a, uint64_t p) {
  uint64_t np = 0 - p;
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+v"(np));
#endif
  const uint64_t t = a + np;
  return (int64_t)t < 0 ? a : t;
}

// Shoup: a * w mod p in [0, 2p) for any a < 2^64, w < p, wp = floor(w 2^64 / p).
// The remainder a w - q p is formed mod 2^64 as a w + q (2^64 - p); with
// p = 2^63 - delta, delta < 2^30 (every prime of the 63-bit bucket), that is
// a w + q delta + (q mod 2) 2^63: two 32x32+64 multiply-adds for the low word
// and three 32-bit products for the high word, and no 64-bit subtraction.
CP_HD uint64_t cp_shoup_rem(uint64_t a, uint64_t w, uint64_t q, uint64_t p) {
  const uint32_t delta = (uint32_t)((1ULL << 63) - p);
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
  const uint32_t w0 = (uint32_t)w, w1 = (uint32_t)(w >> 32);
  const uint32_t q0 = (uint32_t)q, q1 = (uint32_t)(q >> 32);
  const uint64_t x = (uint64_t)a0 * w0 + (uint64_t)q0 * delta;
  const uint32_t hi = (uint32_t)(x >> 32) + a0 * w1 + a1 * w0 + q1 * delta + (q0 << 31);
  return ((uint64_t)hi << 32) | (uint32_t)x;
}

CP_HD uint64_t cp_shoup_lazy(uint64_t a, uint64_t w, uint64_t wp, uint64_t p) {
  return cp_shoup_rem(a, w, cp_mulhi(a, wp), p);
}

CP_HD uint64_t cp_shoup(uint64_t a, uint64_t w, uint64_t wp, uint64_t p) {
  return cp_red2(cp_shoup_lazy(a, w, wp, p), p);
}

// Two Shoup products, x w_x mod p and y w_y mod p in [0, p).  cp_shoup2v
// forms the quotients with cp_mulhi2v (every operand in VGPRs: twiddles and
// twists read per lane); cp_shoup2 with cp_mulhi2 (any operands).
CP_HD void cp_shoup2(uint64_t x, uint64_t wx, uint64_t wxp, uint64_t y, uint64_t wy,
                     uint64_t wyp, uint64_t p, uint64_t &rx, uint64_t &ry) {
  uint64_t qx, qy;
  cp_mulhi2(x, wxp, y, wyp, qx, qy);
  rx = cp_red2(cp_shoup_rem(x, wx, qx, p), p);
  ry = cp_red2(cp_shoup_rem(y, wy, qy, p), p);
}
CP_HD void cp_shoup2v(uint64_t x, uint64_t wx, uint64_t wxp, uint64_t y, uint64_t wy,
                      uint64_t wyp, uint64_t p, uint64_t &rx, uint64_t &ry) {
  uint64_t qx, qy;
  cp_mulhi2v(x, wxp, y, wyp, qx, qy);
  rx = cp_red2(cp_shoup_rem(x, wx, qx, p), p);
  ry = cp_red2(cp_shoup_rem(y, wy, qy, p), p);
}

// a * b mod p for a * b < 2^127 (a < 2^64, b < 2^63 or the reverse), with
// p = 2^63 - delta, delta < 2^30 (every prime of the 63-bit bucket has
// delta < 2^29).  Pseudo-Mersenne folding, since 2^63 = delta (mod p):
//   x = H 2^63 + L           ->  y = L + H delta  < 2^94,
//   y = H2 2^63 + L2         ->  z = L2 + H2 delta < 2^63 + 2^61 < 2p,
// then one conditional subtraction.  Four 32x32 products for x, two for
// H delta and one for H2 delta: about half the instructions of a division-
// based reduction of the full product.
CP_HD uint64_t cp_mulmod(uint64_t a, uint64_t b, const cp_prime& pr) {
  const uint64_t p = pr.p;
  const uint32_t delta = (uint32_t)((1ULL << 63) - p);
  const cp_u128 x = (cp_u128)a * b;
  const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
  const uint64_t L = lo & ((1ULL << 63) - 1);
  const uint64_t H = (hi << 1) | (lo >> 63);
  const cp_u128 y = (cp_u128)H * delta + L;
  const uint64_t L2 = (uint64_t)y & ((1ULL << 63) - 1);
  const uint32_t H2 = (uint32_t)(y >> 63);
  const uint64_t z = L2 + (uint64_t)H2 * delta;
  return cp_red2(z, p);
}

// The 127-bit product's pseudo-Mersenne reduction with its words given:
// lo + hi 2^64 = a b < 2^127.  The folds are written on 32-bit words so that
// each is one multiply-add: A = H0 delta + L (< 2^64), B = H1 delta + A.hi,
// y = A.lo + B 2^32; then z = (y mod 2^63) + (y >> 63) delta.
CP_HD uint64_t cp_fold127(uint64_t lo, uint64_t hi, uint32_t delta, uint64_t p) {
  const uint64_t L = lo & ((1ULL << 63) - 1);
  const uint64_t H = (hi << 1) | (lo >> 63);
  const uint64_t A = (uint64_t)(uint32_t)H * delta + L;
  const uint64_t B = (uint64_t)(uint32_t)(H >> 32) * delta + (A >> 32);
  const uint32_t H2 = (uint32_t)(B >> 31);  // y >> 63, B < 2^62 + 2^32
  const uint64_t L2 = ((uint64_t)((uint32_t)B & 0x7fffffffu) << 32) | (uint32_t)A;
  return cp_red2(L2 + (uint64_t)H2 * delta, p);
}

// Two products of variables, rx = x y mod p and ru = u v mod p, each
// product < 2^127 (one operand below 2^63).  The 128-bit products take the
// carry-out of the middle column's second multiply-add (as cp_mulhi2v), then
// cp_fold127.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ void cp_mulmod2(uint64_t x, uint64_t y, uint64_t u, uint64_t v,
                                           const cp_prime &pr, uint64_t &rx, uint64_t &ru) {
  const uint64_t p = pr.p;
  const uint32_t delta = (uint32_t)((1ULL << 63) - p);
  const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32), y0 = (uint32_t)y, y1 = (uint32_t)(y >> 32);
  const uint32_t u0 = (uint32_t)u, u1 = (uint32_t)(u >> 32), v0 = (uint32_t)v, v1 = (uint32_t)(v >> 32);
  const uint64_t px = (uint64_t)x0 * y0, pu = (uint64_t)u0 * v0;
  const uint64_t mx = (uint64_t)x0 * y1 + (px >> 32);  // < 2^64 (one operand < 2^63)
  const uint64_t mu = (uint64_t)u0 * v1 + (pu >> 32);
  uint64_t sx, su, cx, cu;  // sx + cx 2^64 = x1 y0 + mx
  asm("v_mad_u64_u32 %0, %2, %4, %5, %6\n\t"
      "v_mad_u64_u32 %1, %3, %7, %8, %9"
      : "=&v"(sx), "=&v"(su), "=&s"(cx), "=&s"(cu)
      : "v"(x1), "v"(y0), "v"(mx), "v"(u1), "v"(v0), "v"(mu));
  const uint64_t tx = (uint64_t)x1 * y1 + (sx >> 32);  // + cx 2^32 below: < 2^63 + 2^33
  const uint64_t tu = (uint64_t)u1 * v1 + (su >> 32);
  uint32_t hx, hu;
  uint64_t jx, ju;
  asm("v_addc_co_u32_e64 %0, %2, %4, 0, %5\n\t"
      "v_addc_co_u32_e64 %1, %3, %6, 0, %7"
      : "=&v"(hx), "=&v"(hu), "=&s"(jx), "=&s"(ju)
      : "v"((uint32_t)(tx >> 32)), "s"(cx), "v"((uint32_t)(tu >> 32)), "s"(cu));
  rx = cp_fold127(((uint64_t)(uint32_t)sx << 32) | (uint32_t)px,
                  ((uint64_t)hx << 32) | (uint32_t)tx, delta, p);
  ru = cp_fold127(((uint64_t)(uint32_t)su << 32) | (uint32_t)pu,
                  ((uint64_t)hu << 32) | (uint32_t)tu, delta, p);
}
#else
CP_HD void cp_mulmod2(uint64_t x, uint64_t y, uint64_t u, uint64_t v, const cp_prime &pr,
                      uint64_t &rx, uint64_t &ru) {
  const uint32_t delta = (uint32_t)((1ULL << 63) - pr.p);
  const cp_u128 a = (cp_u128)x * y, b = (cp_u128)u * v;
  rx = cp_fold127((uint64_t)a, (uint64_t)(a >> 64), delta, pr.p);
  ru = cp_fold127((uint64_t)b, (uint64_t)(b >> 64), delta, pr.p);
}
#endif

// Exact sums of products of 64-bit words, reduced mod p once at the end
// (p = 2^63 - delta, delta < 2^30).  The sum is kept in 160 bits,
// x = top 2^128 + hi 2^64 + lo (up to 2^32 products), and folded with
// 2^64 = 2 delta and 2^128 = 4 delta^2 (mod p):
//   y = lo + hi (2 delta) + top (4 delta^2) < 2^64 + 2^94 + 2^92 < 2^95,
//   z = (y mod 2^63) + (y >> 63) delta < 2^63 + 2^62 < 2p, then one red2.
// A product costs one 64x64 -> 128 multiply and a three-word add, where a
// modular chain costs a Shoup product and a conditional correction per term.
struct cp_acc {
  uint64_t lo, hi;
  uint32_t top;
};

CP_HD uint64_t cp_addc64(uint64_t a, uint64_t b, unsigned &carry) {
  const uint64_t s = a + b;
  const unsigned c1 = s < a;
  const uint64_t r = s + carry;
  carry = c1 | (unsigned)(r < s);
  return r;
}

CP_HD void cp_acc_init(cp_acc &s) {
  s.lo = 0;
  s.hi = 0;
  s.top = 0;
}

CP_HD void cp_acc_mac(cp_acc &s, uint64_t a, uint64_t b) {
  const cp_u128 x = (cp_u128)a * b;
  unsigned c = 0;
  s.lo = cp_addc64(s.lo, (uint64_t)x, c);
  s.hi = cp_addc64(s.hi, (uint64_t)(x >> 64), c);
  s.top += c;
}

// Two accumulations by one multiplier, sx += x b and sy += y b, b < 2^63.
// On the GPU the product's middle column and the accumulator add are chained
// through the carry-outs of v_mad_u64_u32 and add-with-carry instead of the
// 64-bit compares the compiler makes of `s < a`:
//   P  = x0 b0 + lo            (carry c0 at 2^64)
//   M  = x0 b1 + P.hi          (< 2^64)
//   M2 = x1 b0 + M             (carry c1 at 2^96)
//   T  = x1 b1 + M2.hi + c1 2^32  (< 2^63 + 2^33, as b1 < 2^31)
//   lo = (P.lo, M2.lo),  hi += T + c0,  top += carry.
// The x and y chains alternate, and each carry is read at least two
// instructions (or an s_nop) after its write: gfx950 needs two wait states
// between a VALU write of a lane mask and a VALU read of it as a carry.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ void cp_acc_mac2(cp_acc &sx, cp_acc &sy, uint64_t x, uint64_t y,
                                            uint64_t b) {
  const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
  const uint32_t y0 = (uint32_t)y, y1 = (uint32_t)(y >> 32);
  const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  uint64_t px, py, cx0, cy0;
  asm("v_mad_u64_u32 %0, %2, %4, %6, %7\n\t"
      "v_mad_u64_u32 %1, %3, %5, %6, %8"
      : "=&v"(px), "=&v"(py), "=&s"(cx0), "=&s"(cy0)
      : "v"(x0), "v"(y0), "v"(b0), "v"(sx.lo), "v"(sy.lo));
  const uint64_t mx = (uint64_t)x0 * b1 + (px >> 32);
  const uint64_t my = (uint64_t)y0 * b1 + (py >> 32);
  uint64_t m2x, m2y, cx1, cy1;
  asm("v_mad_u64_u32 %0, %2, %4, %6, %7\n\t"
      "v_mad_u64_u32 %1, %3, %5, %6, %8"
      : "=&v"(m2x), "=&v"(m2y), "=&s"(cx1), "=&s"(cy1)
      : "v"(x1), "v"(y1), "v"(b0), "v"(mx), "v"(my));
  const uint64_t tx = (uint64_t)x1 * b1 + (m2x >> 32);
  const uint64_t ty = (uint64_t)y1 * b1 + (m2y >> 32);
  uint32_t txl = (uint32_t)tx, txh = (uint32_t)(tx >> 32);
  uint32_t tyl = (uint32_t)ty, tyh = (uint32_t)(ty >> 32);
  uint32_t hxl = (uint32_t)sx.hi, hxh = (uint32_t)(sx.hi >> 32);
  uint32_t hyl = (uint32_t)sy.hi, hyh = (uint32_t)(sy.hi >> 32);
  uint32_t topx = sx.top, topy = sy.top;
  uint64_t j0, j1, e1x, e1y, e2x, e2y, j2, j3;
  asm("v_addc_co_u32_e64 %[txh], %[j0], %[txh], 0, %[cx1]\n\t"
      "v_addc_co_u32_e64 %[tyh], %[j1], %[tyh], 0, %[cy1]\n\t"
      "v_addc_co_u32_e64 %[hxl], %[e1x], %[hxl], %[txl], %[cx0]\n\t"
      "v_addc_co_u32_e64 %[hyl], %[e1y], %[hyl], %[tyl], %[cy0]\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32_e64 %[hxh], %[e2x], %[hxh], %[txh], %[e1x]\n\t"
      "v_addc_co_u32_e64 %[hyh], %[e2y], %[hyh], %[tyh], %[e1y]\n\t"
      "s_nop 0\n\t"
      "v_addc_co_u32_e64 %[topx], %[j2], %[topx], 0, %[e2x]\n\t"
      "v_addc_co_u32_e64 %[topy], %[j3], %[topy], 0, %[e2y]"
      : [txh] "+&v"(txh), [tyh] "+&v"(tyh), [hxl] "+&v"(hxl), [hyl] "+&v"(hyl),
        [hxh] "+&v"(hxh), [hyh] "+&v"(hyh), [topx] "+&v"(topx), [topy] "+&v"(topy),
        [j0] "=&s"(j0), [j1] "=&s"(j1), [e1x] "=&s"(e1x), [e1y] "=&s"(e1y), [e2x] "=&s"(e2x),
        [e2y] "=&s"(e2y), [j2] "=&s"(j2), [j3] "=&s"(j3)
      : [txl] "v"(txl), [tyl] "v"(tyl), [cx0] "s"(cx0), [cy0] "s"(cy0), [cx1] "s"(cx1),
        [cy1] "s"(cy1));
  sx.lo = ((uint64_t)(uint32_t)m2x << 32) | (uint32_t)px;
  sy.lo = ((uint64_t)(uint32_t)m2y << 32) | (uint32_t)py;
  sx.hi = ((uint64_t)hxh << 32) | hxl;
  sy.hi = ((uint64_t)hyh << 32) | hyl;
  sx.top = topx;
  sy.top = topy;
}
#else
CP_HD void cp_acc_mac2(cp_acc &sx, cp_acc &sy, uint64_t x, uint64_t y, uint64_t b) {
  cp_acc_mac(sx, x, b);
  cp_acc_mac(sy, y, b);
}
#endif

CP_HD uint64_t cp_acc_reduce(const cp_acc &s, uint64_t p) {
  const uint64_t delta = (1ULL << 63) - p;
  const cp_u128 y = (cp_u128)s.lo + (cp_u128)s.hi * (2 * delta) +
                    (cp_u128)s.top * (4 * delta * delta);
  const uint64_t z = ((uint64_t)y & ((1ULL << 63) - 1)) + (uint64_t)(y >> 63) * delta;
  return cp_red2(z, p);
}

// Gaussian integer product (a.re + i a.im)(b.re + i b.im) mod p, Karatsuba
// with three exact products as the reference's GIMul.
CP_HD cp_gi cp_gimul(cp_gi a, cp_gi b, const cp_prime& pr) {
  uint64_t p = pr.p;
  uint64_t s1 = cp_mulmod(a.re, b.re, pr);
  uint64_t s2 = cp_mulmod(a.imag, b.imag, pr);
  uint64_t s3 = cp_mulmod(cp_add(a.re, a.imag, p), cp_add(b.re, b.imag, p), pr);
  cp_gi c;
  c.re = cp_sub(s1, s2, p);
  c.imag = cp_sub(s3, cp_add(s1, s2, p), p);
  return c;
}

// Gaussian product by a constant (c + i d) given as three Shoup pairs:
// k[0..1] = (c, c'), k[2..3] = (d, d'), k[4..5] = (c+d mod p, (c+d)').
CP_HD cp_gi cp_gimul_const(cp_gi a, const uint64_t* k, uint64_t p) {
  uint64_t s1, s2;
  cp_shoup2v(a.re, k[0], k[1], a.imag, k[2], k[3], p, s1, s2);
  uint64_t s3 = cp_shoup(a.re + a.imag, k[4], k[5], p);  // a.re + a.imag < 2^64
  cp_gi c;
  c.re = cp_sub(s1, s2, p);
  c.imag = cp_sub(s3, cp_add(s1, s2, p), p);
  return c;
}

// Host-side helpers for table construction.
static inline uint64_t cp_shoup_factor(uint64_t w, uint64_t p) {
  return (uint64_t)(((cp_u128)w << 64) / p);
}

static inline cp_prime cp_make_prime(uint64_t p) {
  cp_prime pr;
  pr.p = p;
  return pr;
}

static inline uint64_t cp_powmod_host(uint64_t a, uint64_t e, uint64_t p) {
  uint64_t r = 1, s = a % p;
  while (e) {
    if (e & 1) r = (uint64_t)(((cp_u128)r * s) % p);
    s = (uint64_t)(((cp_u128)s * s) % p);
    e >>= 1;
  }
  return r;
}

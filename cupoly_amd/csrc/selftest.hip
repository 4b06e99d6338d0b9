// Device self-test of the arithmetic primitives whose gfx950 forms are inline
// asm: the carry-chained high products (cp_mulhi2v), the Shoup products built
// on them (cp_shoup2v), the paired products of variables (cp_mulmod2), the
// carry-chained 160-bit accumulation (cp_acc_mac2) and the round-down FMAs
// under the MODE switch (cp_fma_rd2).  The host compiles the portable forms of
// the same functions; the kernels run the asm.  A test feeds operands chosen
// to exercise every carry and compares the words with exact integer (and
// rational) arithmetic.
#include "hps_dev.hpp"

// One record per thread: in = (x, y, w, z); out = 12 words:
//   0, 1   hi(x y), hi(z y)                       cp_mulhi2v
//   2, 3   the same                               cp_mulhi2 (portable form)
//   4, 5   x w' mod p, z w' mod p (w' = w mod p)  cp_shoup2v
//   6, 7   x (y >> 1) mod p, (z mod p) y mod p    cp_mulmod2
//   8, 9   sum_k x b_k mod p, sum_k z b_k mod p   cp_acc_mac2, b_k = (y >> 1) ^ (k c), 6 terms
//   10, 11 bits of RD(a r + c) for the two lanes of cp_fma_rd2, with
//          a = x >> 11, r = (y >> 11) 2^-63, c = (z >> 11) 2^-52 (and the
//          second lane with x, z swapped)
__global__ void __launch_bounds__(256)
    k_arith_selftest(const uint64_t *__restrict__ in, uint64_t *__restrict__ out, int n,
                     uint64_t p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t x = in[4 * i], y = in[4 * i + 1], w = in[4 * i + 2] % p, z = in[4 * i + 3];
  const uint64_t wp = (uint64_t)(((cp_u128)w << 64) / p);
  uint64_t *o = out + 12 * (size_t)i;
  uint64_t h0, h1;
  cp_mulhi2v(x, y, z, y, h0, h1);
  o[0] = h0;
  o[1] = h1;
  cp_mulhi2(x, y, z, y, h0, h1);
  o[2] = h0;
  o[3] = h1;
  cp_shoup2v(x, w, wp, z, w, wp, p, h0, h1);
  o[4] = h0;
  o[5] = h1;
  const cp_prime pr{p};
  cp_mulmod2(x, y >> 1, z % p, y, pr, h0, h1);
  o[6] = h0;
  o[7] = h1;
  cp_acc sx, sz;
  cp_acc_init(sx);
  cp_acc_init(sz);
  for (int k = 0; k < 6; k++) {
    const uint64_t b = ((y >> 1) ^ ((uint64_t)k * 0x1D8E4E27C47D124FULL)) & ((1ULL << 63) - 1);
    cp_acc_mac2(sx, sz, x, z, b);
  }
  o[8] = cp_acc_reduce(sx, p);
  o[9] = cp_acc_reduce(sz, p);
  const double a0 = (double)(x >> 11), a1 = (double)(z >> 11);
  const double r = ldexp((double)(y >> 11), -63);
  double v0 = ldexp((double)(z >> 11), -52), v1 = ldexp((double)(x >> 11), -52);
  cp_fma_rd2(a0, a1, r, v0, v1);
  o[10] = (uint64_t)__double_as_longlong(v0);
  o[11] = (uint64_t)__double_as_longlong(v1);
}

extern "C" int cupoly_arith_selftest(cupoly_ctx *c, const uint64_t *in, uint64_t *out, int n,
                                     uint64_t p) {
  if (!c || !in || !out || n < 0) return cp_fail(CUPOLY_ERR_ARG, "arith_selftest: bad argument");
  if (p >= (1ULL << 63) || p <= (1ULL << 63) - (1ULL << 30))
    return cp_fail(CUPOLY_ERR_ARG, "arith_selftest: p must lie in (2^63 - 2^30, 2^63)");
  if (n == 0) return CUPOLY_OK;
  hipLaunchKernelGGL(k_arith_selftest, dim3((n + 255) / 256), dim3(256), 0, c->stream, in, out, n,
                     p);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

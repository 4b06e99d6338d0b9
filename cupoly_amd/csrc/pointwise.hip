// Slot-wise operations in the DGT domain and scalar operations.
//
// Reference: DGTEngine::execute_{add,sub,double_add,dr2,mul_dgt_gi,mul_dgt_u64,
// muladd}_dgt (src/cuda/dgt.cu:1096-1499) and CUDAEngine::execute_polynomial_
// {negate,op_by_int,double_op_by_int} (src/cuda/cudaengine.cu:115-302).
// These are HBM-streaming kernels: each thread handles whole 16-byte
// Gaussian slots (global_load_dwordx4) in a grid-stride loop sized to the
// device, and the residue of slot s is s / (n/2).  Results are canonical
// residues, equal to the reference's (its addmod/submod/GIMul are exact on
// canonical inputs).
//
// One deliberate difference: the reference's scalar kernels index primes
// without the base offset (cudaengine.cu:126, 181, 251), so on base B they
// reduce by Q's primes.  Here every op honours the base (identical results on
// bases Q and QB).
#include "internal.hpp"

enum { PW_ADD, PW_SUB, PW_MUL, PW_MULADD, PW_MULU64, PW_NEG, PW_ADDI, PW_SUBI, PW_MULI };

template <int OP>
__global__ void __launch_bounds__(256)
    k_pointwise(cp_gi *__restrict__ out, const cp_gi *__restrict__ a,
                const cp_gi *__restrict__ b, const cp_gi *__restrict__ c,
                const uint64_t *__restrict__ bu, uint64_t x, int ng, int nres, int roff,
                const cp_prime *__restrict__ primes) {
  const long total = (long)ng * nres;
  for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < total;
       s += (long)gridDim.x * blockDim.x) {
    const int rid = roff + (int)(s / ng);
    const cp_prime pr = primes[rid];
    const uint64_t p = pr.p;
    const cp_gi va = a[s];
    cp_gi r;
    if constexpr (OP == PW_ADD) {
      const cp_gi vb = b[s];
      r.re = cp_add(va.re, vb.re, p);
      r.imag = cp_add(va.imag, vb.imag, p);
    } else if constexpr (OP == PW_SUB) {
      const cp_gi vb = b[s];
      r.re = cp_sub(va.re, vb.re, p);
      r.imag = cp_sub(va.imag, vb.imag, p);
    } else if constexpr (OP == PW_MUL) {
      r = cp_gimul(va, b[s], pr);
    } else if constexpr (OP == PW_MULADD) {
      r = cp_gimul(va, b[s], pr);
      const cp_gi vc = c[s];
      r.re = cp_add(r.re, vc.re, p);
      r.imag = cp_add(r.imag, vc.imag, p);
    } else if constexpr (OP == PW_MULU64) {
      const uint64_t w = bu[s];
      r.re = cp_mulmod(va.re, w, pr);
      r.imag = cp_mulmod(va.imag, w, pr);
    } else if constexpr (OP == PW_NEG) {
      r.re = va.re == 0 ? 0 : p - va.re;
      r.imag = va.imag == 0 ? 0 : p - va.imag;
    } else if constexpr (OP == PW_ADDI) {
      r.re = cp_add(va.re, x % p, p);
      r.imag = va.imag;
    } else if constexpr (OP == PW_SUBI) {
      r.re = cp_sub(va.re, x % p, p);
      r.imag = va.imag;
    } else {  // PW_MULI: GIMul(a, {x, 0}) = (a.re x, a.imag x)
      const uint64_t xm = x % p;
      r.re = cp_mulmod(va.re, xm, pr);
      r.imag = cp_mulmod(va.imag, xm, pr);
    }
    out[s] = r;
  }
}

// ct21 = a0 b0, ct22 = a0 b1 + a1 b0, ct23 = a1 b1 (dgt.cu:1251-1279); ct22 by
// Karatsuba over the pair, (a0 + a1)(b0 + b1) - ct21 - ct23, exact mod p.
__global__ void __launch_bounds__(256)
    k_dr2(cp_gi *__restrict__ ct21, cp_gi *__restrict__ ct22, cp_gi *__restrict__ ct23,
          const cp_gi *__restrict__ ct01, const cp_gi *__restrict__ ct02,
          const cp_gi *__restrict__ ct11, const cp_gi *__restrict__ ct12, int ng, int nres,
          int roff, const cp_prime *__restrict__ primes) {
  const long total = (long)ng * nres;
  for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < total;
       s += (long)gridDim.x * blockDim.x) {
    const cp_prime pr = primes[roff + (int)(s / ng)];
    const uint64_t p = pr.p;
    const cp_gi a0 = ct01[s], a1 = ct02[s], b0 = ct11[s], b1 = ct12[s];
    const cp_gi r1 = cp_gimul(a0, b0, pr), r3 = cp_gimul(a1, b1, pr);
    cp_gi sa, sb;
    sa.re = cp_add(a0.re, a1.re, p);
    sa.imag = cp_add(a0.imag, a1.imag, p);
    sb.re = cp_add(b0.re, b1.re, p);
    sb.imag = cp_add(b0.imag, b1.imag, p);
    cp_gi r2 = cp_gimul(sa, sb, pr);
    r2.re = cp_sub(r2.re, cp_add(r1.re, r3.re, p), p);
    r2.imag = cp_sub(r2.imag, cp_add(r1.imag, r3.imag, p), p);
    ct21[s] = r1;
    ct22[s] = r2;
    ct23[s] = r3;
  }
}

static int grid_for(long total) {
  long g = (total + 255) / 256;
  if (g > 256 * 16) g = 256 * 16;  // 16 workgroups per CU, grid-stride beyond
  return (int)(g < 1 ? 1 : g);
}

template <int OP>
static int launch_pw(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, const cupoly_gi *b,
                     const cupoly_gi *add, const uint64_t *bu, uint64_t x, int base) {
  if (!c || !out || !a) return cp_fail(CUPOLY_ERR_ARG, "pointwise: null argument");
  int off, cnt;
  if (!cp_base_range(c->e, base, &off, &cnt)) return cp_fail(CUPOLY_ERR_UNSUPPORTED, "Unknown base");
  const long total = (long)c->e->ng * cnt;
  hipLaunchKernelGGL(k_pointwise<OP>, dim3(grid_for(total)), dim3(256), 0, c->stream,
                     (cp_gi *)out, (const cp_gi *)a, (const cp_gi *)b, (const cp_gi *)add, bu, x,
                     c->e->ng, cnt, off, c->e->d_primes);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

int cp_pointwise_mul_launch(cupoly_ctx *c, cp_gi *out, const cp_gi *a, const cp_gi *b, int nres) {
  const long total = (long)c->e->ng * nres;
  hipLaunchKernelGGL(k_pointwise<PW_MUL>, dim3(grid_for(total)), dim3(256), 0, c->stream, out, a, b,
                     (const cp_gi *)nullptr, (const uint64_t *)nullptr, (uint64_t)0, c->e->ng, nres,
                     0, c->e->d_primes);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

extern "C" int cupoly_add_dgt(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a,
                              const cupoly_gi *b, int base) {
  if (!b) return cp_fail(CUPOLY_ERR_ARG, "add_dgt: null b");
  return launch_pw<PW_ADD>(c, out, a, b, nullptr, nullptr, 0, base);
}
extern "C" int cupoly_sub_dgt(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a,
                              const cupoly_gi *b, int base) {
  if (!b) return cp_fail(CUPOLY_ERR_ARG, "sub_dgt: null b");
  return launch_pw<PW_SUB>(c, out, a, b, nullptr, nullptr, 0, base);
}
extern "C" int cupoly_mul_dgt_gi(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a,
                                 const cupoly_gi *b, int base) {
  if (!b) return cp_fail(CUPOLY_ERR_ARG, "mul_dgt_gi: null b");
  return launch_pw<PW_MUL>(c, out, a, b, nullptr, nullptr, 0, base);
}
extern "C" int cupoly_mul_dgt_u64(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a,
                                  const uint64_t *b, int base) {
  if (!b) return cp_fail(CUPOLY_ERR_ARG, "mul_dgt_u64: null b");
  return launch_pw<PW_MULU64>(c, out, a, nullptr, nullptr, b, 0, base);
}
extern "C" int cupoly_muladd_dgt(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a,
                                 const cupoly_gi *b, const cupoly_gi *add, int base) {
  if (!b || !add) return cp_fail(CUPOLY_ERR_ARG, "muladd_dgt: null argument");
  return launch_pw<PW_MULADD>(c, out, a, b, add, nullptr, 0, base);
}
extern "C" int cupoly_double_add_dgt(cupoly_ctx *c, cupoly_gi *c1, const cupoly_gi *a1,
                                     const cupoly_gi *b1, cupoly_gi *c2, const cupoly_gi *a2,
                                     const cupoly_gi *b2, int base) {
  int rc = cupoly_add_dgt(c, c1, a1, b1, base);
  return rc ? rc : cupoly_add_dgt(c, c2, a2, b2, base);
}
extern "C" int cupoly_dr2_dgt(cupoly_ctx *c, cupoly_gi *ct21, cupoly_gi *ct22, cupoly_gi *ct23,
                              const cupoly_gi *ct01, const cupoly_gi *ct02, const cupoly_gi *ct11,
                              const cupoly_gi *ct12, int base) {
  if (!c || !ct21 || !ct22 || !ct23 || !ct01 || !ct02 || !ct11 || !ct12)
    return cp_fail(CUPOLY_ERR_ARG, "dr2_dgt: null argument");
  int off, cnt;
  if (!cp_base_range(c->e, base, &off, &cnt)) return cp_fail(CUPOLY_ERR_UNSUPPORTED, "Unknown base");
  const long total = (long)c->e->ng * cnt;
  hipLaunchKernelGGL(k_dr2, dim3(grid_for(total)), dim3(256), 0, c->stream, (cp_gi *)ct21,
                     (cp_gi *)ct22, (cp_gi *)ct23, (const cp_gi *)ct01, (const cp_gi *)ct02,
                     (const cp_gi *)ct11, (const cp_gi *)ct12, c->e->ng, cnt, off, c->e->d_primes);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}
extern "C" int cupoly_negate(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, int base) {
  return launch_pw<PW_NEG>(c, out, a, nullptr, nullptr, nullptr, 0, base);
}
extern "C" int cupoly_op_by_int(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, uint64_t x,
                                int op, int base) {
  switch (op) {
    case CUPOLY_ADD: return launch_pw<PW_ADDI>(c, out, a, nullptr, nullptr, nullptr, x, base);
    case CUPOLY_SUB: return launch_pw<PW_SUBI>(c, out, a, nullptr, nullptr, nullptr, x, base);
    case CUPOLY_MUL: return launch_pw<PW_MULI>(c, out, a, nullptr, nullptr, nullptr, x, base);
    default: return cp_fail(CUPOLY_ERR_ARG, "op_by_int: Unknown operation");
  }
}
extern "C" int cupoly_double_op_by_int(cupoly_ctx *c, cupoly_gi *b1, const cupoly_gi *a1,
                                       uint64_t x1, cupoly_gi *b2, const cupoly_gi *a2, uint64_t x2,
                                       int op, int base) {
  int rc;
  switch (op) {
    case CUPOLY_MULMUL:
      rc = cupoly_op_by_int(c, b1, a1, x1, CUPOLY_MUL, base);
      return rc ? rc : cupoly_op_by_int(c, b2, a2, x2, CUPOLY_MUL, base);
    case CUPOLY_ADDADD:
      rc = cupoly_op_by_int(c, b1, a1, x1, CUPOLY_ADD, base);
      return rc ? rc : cupoly_op_by_int(c, b2, a2, x2, CUPOLY_ADD, base);
    default: return cp_fail(CUPOLY_ERR_ARG, "double_op_by_int: Unknown operation");
  }
}

// The reference's order for one product (polynomial.cu:892-914 plus the
// inverse in poly_copy_to_host): forward(a), forward(b), mul, inverse(c).
extern "C" int cupoly_polymul_plain(cupoly_ctx *c, cupoly_gi *out, cupoly_gi *a, cupoly_gi *b,
                                    int batch) {
  if (c && batch == 0) return CUPOLY_OK;  // empty batch: pointers may be null
  if (!c || !out || !a || !b || batch < 0) return cp_fail(CUPOLY_ERR_ARG, "polymul_plain: bad argument");
  const size_t stride = (size_t)c->e->k * c->e->ng;
  for (int i = 0; i < batch; i++) {
    int rc;
    if ((rc = cupoly_dgt(c, a + i * stride, CUPOLY_QBASE, CUPOLY_FORWARD)) ||
        (rc = cupoly_dgt(c, b + i * stride, CUPOLY_QBASE, CUPOLY_FORWARD)) ||
        (rc = cupoly_mul_dgt_gi(c, out + i * stride, a + i * stride, b + i * stride, CUPOLY_QBASE)) ||
        (rc = cupoly_dgt(c, out + i * stride, CUPOLY_QBASE, CUPOLY_INVERSE)))
      return rc;
  }
  return CUPOLY_OK;
}

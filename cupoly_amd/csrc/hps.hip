// HPS basis extension and scaling (Halevi-Polyakov-Shoup RNS BFV).
//
// Reference: CUDAEngine::execute_polynomial_{basis_ext_Q_to_B, complex_scaling,
// basis_ext_B_to_Q, simple_scaling} (src/cuda/cudaengine.cu:323-796) and the
// poly_* orchestration (src/arithmetic/polynomial.cu:310-457).
//
// The reference runs each extension as three launches (aux, v, extension)
// with one thread per coefficient and scratch arrays in HBM.  Here each
// extension is one kernel: a thread owns one Gaussian slot (coefficients j
// and j + n/2), keeps its k (or kl) residues in registers, and writes the
// other base directly, so the aux / v arrays never touch memory.  Integer
// steps are exact (Shoup products by the precomputed constants), so they
// equal the reference's canonical Barrett results.  The floating-point steps
// are reproduced operation for operation in IEEE-754 binary64:
//   Q->B v:   value = RD(RN(aux_j) * RD(1/RN(q_j)) + value), ascending j
//             (__fma_rd / __drcp_rd at cudaengine.cu:527); the reciprocal is a
//             per-prime constant and the FMA runs in round-toward-minus-
//             infinity mode (__ocml_fma_rtn_f64 sets the MODE register);
//   B->Q v:   value = RN(value + RN(RN(aux_i) / RN(b_i))), ascending i (:700);
//   scaling:  acc = RN(RN(lo) * 2^-64 + acc) (:413-414), then llrint;
//   snap:     |value - llrint(value)| <= 1e-9 ? llrint(value) : floor(value)
//             (:530-531, :702).
// The library is compiled with -ffp-contract=off so no other FP operation is
// fused.  The additions that may see unreduced operands in the complex scaling
// keep the reference's single conditional subtraction on 64-bit wrap-around
// words (dgt.h:73-74), so even those rare values are reproduced.
#include "hps_dev.hpp"

// --- Q -> B, in place: a holds [k][ng] Q residues in, [kl][ng] B residues out.
__global__ void __launch_bounds__(256)
    k_basis_ext_q_to_b(cp_gi *a, int ng, const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  hps_q_to_b_slot(a, ng, a, ng, c, H);
}

// --- complex scaling t/Q: out = uint64 [kl][2ng] (re half, imag half).
__global__ void __launch_bounds__(256)
    k_complex_scaling(uint64_t *__restrict__ out, const cp_gi *__restrict__ aq,
                      const cp_gi *__restrict__ ab, int ng, const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  uint64_t yr[CP_MAX_BASE], yi[CP_MAX_BASE];
  hps_scale_slot(aq, ng, ab, ng, c, H, yr, yi);
#pragma unroll
  for (int i = 0; i < CP_MAX_BASE; i++) {
    if (i < H->kl) {
      out[(size_t)i * 2 * ng + c] = yr[i];
      out[(size_t)i * 2 * ng + ng + c] = yi[i];
    }
  }
}

// --- B -> Q: in = uint64 [kl][2ng], out = [k][ng] GI.
__global__ void __launch_bounds__(256)
    k_basis_ext_b_to_q(cp_gi *__restrict__ out, const uint64_t *__restrict__ in, int ng,
                       const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  uint64_t xr[CP_MAX_BASE], xi[CP_MAX_BASE];
#pragma unroll
  for (int i = 0; i < CP_MAX_BASE; i++) {
    if (i < H->kl) {
      xr[i] = in[(size_t)i * 2 * ng + c];
      xi[i] = in[(size_t)i * 2 * ng + ng + c];
    }
  }
  hps_b_to_q_slot(xr, xi, out, ng, c, H);
}

// --- simple scaling Q -> t (cudaengine.cu:323-363): round(t x / Q) mod t per
// coefficient.  For each residue j, t (Q/q_j)^-1 mod q_j / q_j = w_int + w_frac
// (w_frac as a 128-bit fraction): x_j w_int + floor(x_j w_frac) is summed mod t
// and the fractional words are accumulated in FP64 and rounded once.  The
// reference reduces mod t with masks, which is exact only for t a power of two
// (for other t the product x_j w_int overflows, known-issues.md item 1); here
// x_j is reduced mod t first, so any t < 2^32 is exact, and for powers of two
// the result equals the masked one.
__global__ void __launch_bounds__(256)
    k_simple_scaling(cp_gi *__restrict__ a, int ng, const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int k = H->k;
  const uint64_t t = H->t;
  const bool pow2 = (t & (t - 1)) == 0;
  const uint64_t mask = t - 1;
  const double two_m64 = 0x1p-64;
  double acc_r = 0.0, acc_i = 0.0;
  uint64_t rr = 0, ri = 0;
  for (int j = 0; j < k; j++) {
    const cp_gi v = a[(size_t)j * ng + c];
    const uint64_t tw = H->twint[j];
    uint64_t lr, hr, li, hi;
    umul128hi(H->twfrac[j][0], H->twfrac[j][1], v.re, lr, hr);
    umul128hi(H->twfrac[j][0], H->twfrac[j][1], v.imag, li, hi);
    acc_r = fma((double)lr, two_m64, acc_r);
    acc_i = fma((double)li, two_m64, acc_i);
    if (pow2) {
      rr += ((v.re * tw & mask) + (hr & mask)) & mask;
      ri += ((v.imag * tw & mask) + (hi & mask)) & mask;
    } else {  // t < 2^32: (x mod t) w_int < 2^64
      rr += ((v.re % t) * tw % t + hr % t) % t;
      ri += ((v.imag % t) * tw % t + hi % t) % t;
    }
  }
  cp_gi o;
  if (pow2) {
    o.re = (rr + (uint64_t)llrint(acc_r)) & mask;
    o.imag = (ri + (uint64_t)llrint(acc_i)) & mask;
  } else {
    o.re = (rr + (uint64_t)llrint(acc_r)) % t;
    o.imag = (ri + (uint64_t)llrint(acc_i)) % t;
  }
  a[c] = o;
}

static inline dim3 slots_grid(int ng) { return dim3((ng + 255) / 256); }

extern "C" int cupoly_basis_ext_q_to_b(cupoly_ctx *c, cupoly_gi *a) {
  if (!c || !a) return cp_fail(CUPOLY_ERR_ARG, "basis_ext_q_to_b: null argument");
  hipLaunchKernelGGL(k_basis_ext_q_to_b, slots_grid(c->e->ng), dim3(256), 0, c->stream,
                     (cp_gi *)a, c->e->ng, c->e->d_hps);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

extern "C" int cupoly_complex_scaling(cupoly_ctx *c, uint64_t *b, const cupoly_gi *a_q,
                                      const cupoly_gi *a_b) {
  if (!c || !b || !a_q || !a_b) return cp_fail(CUPOLY_ERR_ARG, "complex_scaling: null argument");
  hipLaunchKernelGGL(k_complex_scaling, slots_grid(c->e->ng), dim3(256), 0, c->stream, b,
                     (const cp_gi *)a_q, (const cp_gi *)a_b, c->e->ng, c->e->d_hps);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

extern "C" int cupoly_basis_ext_b_to_q(cupoly_ctx *c, cupoly_gi *b, const uint64_t *a) {
  if (!c || !b || !a) return cp_fail(CUPOLY_ERR_ARG, "basis_ext_b_to_q: null argument");
  hipLaunchKernelGGL(k_basis_ext_b_to_q, slots_grid(c->e->ng), dim3(256), 0, c->stream,
                     (cp_gi *)b, a, c->e->ng, c->e->d_hps);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

extern "C" int cupoly_simple_scaling(cupoly_ctx *c, cupoly_gi *a) {
  if (!c || !a) return cp_fail(CUPOLY_ERR_ARG, "simple_scaling: null argument");
  if ((c->e->t & (c->e->t - 1)) && c->e->t >= (1ULL << 32))
    return cp_fail(CUPOLY_ERR_UNSUPPORTED,
                   "simple_scaling: t must be a power of 2 or below 2^32");
  hipLaunchKernelGGL(k_simple_scaling, slots_grid(c->e->ng), dim3(256), 0, c->stream, (cp_gi *)a,
                     c->e->ng, c->e->d_hps);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

// poly_basis_extension_Q_to_B: DGT-domain Q in, DGT-domain B out.
extern "C" int cupoly_poly_basis_ext_q_to_b(cupoly_ctx *c, cupoly_gi *out_b,
                                            const cupoly_gi *a_q) {
  if (!c || !out_b || !a_q) return cp_fail(CUPOLY_ERR_ARG, "poly_basis_ext_q_to_b: null argument");
  cupoly_engine *e = c->e;
  const size_t qbytes = (size_t)e->k * e->ng * sizeof(cp_gi);
  const size_t bbytes = (size_t)e->kl * e->ng * sizeof(cp_gi);
  int rc;
  CP_HIP(hipMemcpyAsync(c->d_tmp, a_q, qbytes, hipMemcpyDeviceToDevice, c->stream));
  if ((rc = cp_dgt_launch(c, c->d_tmp, 0, e->k, 1, CUPOLY_INVERSE))) return rc;
  if ((rc = cupoly_basis_ext_q_to_b(c, (cupoly_gi *)c->d_tmp))) return rc;
  if ((rc = cp_dgt_launch(c, c->d_tmp, e->k, e->kl, 1, CUPOLY_FORWARD))) return rc;
  CP_HIP(hipMemcpyAsync(out_b, c->d_tmp, bbytes, hipMemcpyDeviceToDevice, c->stream));
  return CUPOLY_OK;
}

// poly_complex_scaling_tDivQ: DGT-domain Q and B in, DGT-domain Q out.
extern "C" int cupoly_poly_complex_scaling(cupoly_ctx *c, cupoly_gi *out_q, const cupoly_gi *a_q,
                                           const cupoly_gi *a_b) {
  if (!c || !out_q || !a_q || !a_b)
    return cp_fail(CUPOLY_ERR_ARG, "poly_complex_scaling: null argument");
  cupoly_engine *e = c->e;
  const size_t qbytes = (size_t)e->k * e->ng * sizeof(cp_gi);
  const size_t bbytes = (size_t)e->kl * e->ng * sizeof(cp_gi);
  int rc;
  CP_HIP(hipMemcpyAsync(c->d_tmp, a_q, qbytes, hipMemcpyDeviceToDevice, c->stream));
  CP_HIP(hipMemcpyAsync(c->d_tmp2, a_b, bbytes, hipMemcpyDeviceToDevice, c->stream));
  if ((rc = cp_dgt_launch(c, c->d_tmp, 0, e->k, 1, CUPOLY_INVERSE))) return rc;
  if ((rc = cp_dgt_launch(c, c->d_tmp2, e->k, e->kl, 1, CUPOLY_INVERSE))) return rc;
  if ((rc = cupoly_complex_scaling(c, c->d_aux, (cupoly_gi *)c->d_tmp, (cupoly_gi *)c->d_tmp2)))
    return rc;
  if ((rc = cupoly_basis_ext_b_to_q(c, out_q, c->d_aux))) return rc;
  return cp_dgt_launch(c, (cp_gi *)out_q, 0, e->k, 1, CUPOLY_FORWARD);
}

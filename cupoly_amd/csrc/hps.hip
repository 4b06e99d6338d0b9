// HPS basis extension and scaling (Halevi-Polyakov-Shoup RNS BFV).
//
// Reference: CUDAEngine::execute_polynomial_{basis_ext_Q_to_B, complex_scaling,
// basis_ext_B_to_Q, simple_scaling} (src/cuda/cudaengine.cu:323-796) and the
// poly_* orchestration (src/arithmetic/polynomial.cu:310-457).
//
// The reference runs each extension as three launches (aux, v, extension)
// with one thread per coefficient and scratch arrays in HBM.  Here each
// extension is one kernel with one thread per (Gaussian slot, output
// residue): a thread reads the slot's input residues (coefficients j and
// j + n/2), forms the per-slot values in registers and writes its output
// residue directly, so the aux / v arrays never touch memory.  Integer
// steps are exact (Shoup products by the precomputed constants), so they
// equal the reference's canonical Barrett results.  The floating-point steps
// are reproduced operation for operation in IEEE-754 binary64:
//   Q->B v:   value = RD(RN(aux_j) * RD(1/RN(q_j)) + value), ascending j
//             (__fma_rd / __drcp_rd at cudaengine.cu:527); the reciprocal is a
//             per-prime constant and the FMA runs in round-toward-minus-
//             infinity mode (cp_fma_rd2 sets the MODE register);
//   B->Q v:   value = RN(value + RN(RN(aux_i) / RN(b_i))), ascending i (:700);
//   scaling:  acc = RN(RN(lo) * 2^-64 + acc) (:413-414), then llrint;
//   snap:     |value - llrint(value)| <= 1e-9 ? llrint(value) : floor(value)
//             (:530-531, :702).
// The library is compiled with -ffp-contract=off so no other FP operation is
// fused.  The additions that may see unreduced operands in the complex scaling
// keep the reference's single conditional subtraction on 64-bit wrap-around
// words (dgt.h:73-74), so even those rare values are reproduced.
#include "hps_dev.hpp"

// Each single-polynomial step gives every output residue of a slot its own
// thread: one thread per slot would put the kl (or k) output residues of a
// 2^14-coefficient polynomial on 32 workgroups, a few CUs each running a
// long serial chain.  The per-slot values every output residue needs (the
// words (Q/q_j)^-1 x_j and the FP64 sums) are recomputed by each thread with
// the same operations in the same order, so every output word is the same.

// --- Q -> B, in place: a holds [k][ng] Q residues in, [kl][ng] B residues out.
// blockDim = (64, kl): thread (x, y) is slot blockIdx.x * 64 + x, B residue y.
// Every thread reads its slot's Q words before the barrier and writes only
// after it, and all kl threads of a slot are in the workgroup, so the
// in-place overwrite cannot race with a read.
__global__ void __launch_bounds__(64 * CP_MAX_BASE)
    k_basis_ext_q_to_b(cp_gi *a, int ng, const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * 64 + threadIdx.x, i = threadIdx.y;
  const int k = H->k;
  const bool on = c < ng;
  uint64_t ar[CP_MAX_BASE], ai[CP_MAX_BASE];
  double vr = 0.0, vi = 0.0;
  if (on) {
#pragma unroll
    for (int j = 0; j < CP_MAX_BASE; j++) {
      if (j < k) {
        const uint64_t q = H->q[j].p;
        const cp_gi x = a[(size_t)j * ng + c];
        ar[j] = shoup2(x.re, H->invQqi[j], q);
        ai[j] = shoup2(x.imag, H->invQqi[j], q);
      }
    }
    // v = RD(aux_j * RD(1/q_j) + v), ascending j (cudaengine.cu:526-527)
#pragma unroll
    for (int j = 0; j < CP_MAX_BASE; j++)
      if (j < k) cp_fma_rd2((double)ar[j], (double)ai[j], H->rcp_q[j], vr, vi);
  }
  __syncthreads();
  if (!on) return;
  const uint64_t v_re = snap(vr), v_im = snap(vi);
  // exact sum, reduced once (as hps_q_to_b_slot)
  const uint64_t b = H->b[i].p, qb = H->Qbi[i][0];
  const uint64_t nqb = qb ? b - qb : 0;  // -Q mod b_i
  cp_acc sr, si;
  cp_acc_init(sr);
  cp_acc_init(si);
#pragma unroll
  for (int j = 0; j < CP_MAX_BASE; j++) {
    if (j < k) {
      cp_acc_mac2(sr, si, ar[j], ai[j], H->Qqibi[i][j][0]);
    }
  }
  cp_acc_mac2(sr, si, v_re, v_im, nqb);
  a[(size_t)i * ng + c] = cp_gi{cp_acc_reduce(sr, b), cp_acc_reduce(si, b)};
}

// --- complex scaling t/Q: out = uint64 [kl][2ng] (re half, imag half); one
// thread per (slot, B residue i = blockIdx.y), K = k.  The exact sum reduced
// once where the reference's chain sees reduced operands only, the chain
// itself elsewhere (or everywhere with chain = true): the same words.
template <int K>
__global__ void __launch_bounds__(256)
    k_complex_scaling(uint64_t *__restrict__ out, const cp_gi *__restrict__ aq,
                      const cp_gi *__restrict__ ab, int ng, const cp_hps *__restrict__ H,
                      bool chain) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int i = blockIdx.y;
  const uint64_t b = H->b[i].p;
  uint64_t xr[K], xi[K], fr[K], fi[K];
  const double two_m64 = 0x1p-64;  // __drcp_rn(2^64)
  double acc_r = 0.0, acc_i = 0.0;
  bool slow = chain;
  cp_acc hr, hi;
  cp_acc_init(hr);
  cp_acc_init(hi);
#pragma unroll
  for (int j = 0; j < K; j++) {
    const cp_gi x = aq[(size_t)j * ng + c];
    xr[j] = x.re;
    xi[j] = x.imag;
    uint64_t lr, li;
    umul128hi(H->wfrac[j][0], H->wfrac[j][1], x.re, lr, fr[j]);
    umul128hi(H->wfrac[j][0], H->wfrac[j][1], x.imag, li, fi[j]);
    acc_r = fma((double)lr, two_m64, acc_r);  // __fma_rn (cudaengine.cu:413-414)
    acc_i = fma((double)li, two_m64, acc_i);
    slow |= (fr[j] >= b) | (fi[j] >= b);
    hsum_add(hr, fr[j]);
    hsum_add(hi, fi[j]);
  }
  const uint64_t round_r = (uint64_t)llrint(acc_r), round_i = (uint64_t)llrint(acc_i);
  hsum_add(hr, round_r);
  hsum_add(hi, round_i);
  const cp_gi y = ab[(size_t)i * ng + c];
  uint64_t rr, ri;
  if (!slow) {
    scale_exact2(hr, hi, xr, xi, K, H->wint[i], y.re, y.imag, H->lam[i], b, rr, ri);
  } else {
    rr = scale_chain_word(xr, fr, K, H->wint[i], round_r, y.re, H->lam[i], b);
    ri = scale_chain_word(xi, fi, K, H->wint[i], round_i, y.imag, H->lam[i], b);
  }
  out[(size_t)i * 2 * ng + c] = rr;
  out[(size_t)i * 2 * ng + ng + c] = ri;
}

// --- B -> Q: in = uint64 [kl][2ng], out = [k][ng] GI; one thread per (slot,
// Q residue j = blockIdx.y).
__global__ void __launch_bounds__(256)
    k_basis_ext_b_to_q(cp_gi *__restrict__ out, const uint64_t *__restrict__ in, int ng,
                       const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int j = blockIdx.y, kl = H->kl;
  const uint64_t q = H->q[j].p;
  double vr = 0.0, vi = 0.0;
  cp_acc sr, si;  // exact sums (every term reduced), reduced once
  cp_acc_init(sr);
  cp_acc_init(si);
  for (int i = 0; i < kl; i++) {  // ascending i (cudaengine.cu:699-700)
    const uint64_t b = H->b[i].p;
    const uint64_t ar = shoup2(in[(size_t)i * 2 * ng + c], H->invBbi[i], b);
    const uint64_t ai = shoup2(in[(size_t)i * 2 * ng + ng + c], H->invBbi[i], b);
    const double bd = (double)b;
    vr = vr + (double)ar / bd;
    vi = vi + (double)ai / bd;
    cp_acc_mac2(sr, si, ar, ai, H->Bbiqi[j][i][0]);
  }
  const uint64_t v_re = snap(vr), v_im = snap(vi);
  const uint64_t bq = H->Bqi[j][0], nbq = bq ? q - bq : 0;  // -B mod q_j
  cp_acc_mac2(sr, si, v_re, v_im, nbq);
  out[(size_t)j * ng + c] = cp_gi{cp_acc_reduce(sr, q), cp_acc_reduce(si, q)};
}

// x mod t for t < 2^32 that is not a power of two, m = floor(2^64 / t):
// q = floor(x m / 2^64) is floor(x / t) or one less, so one correction.
__device__ __forceinline__ uint64_t mod_t(uint64_t x, uint64_t t, uint64_t m) {
  const uint64_t q = cp_mulhi(x, m);
  const uint64_t r = x - q * t;
  return r >= t ? r - t : r;
}

// --- simple scaling Q -> t (cudaengine.cu:323-363): round(t x / Q) mod t per
// coefficient, one thread per coefficient (c < ng: re of slot c, else imag of
// slot c - ng).  For each residue j, t (Q/q_j)^-1 mod q_j / q_j = w_int +
// w_frac (w_frac as a 128-bit fraction): x_j w_int + floor(x_j w_frac) is
// summed mod t and the fractional words are accumulated in FP64 and rounded
// once.  The reference reduces mod t with masks, which is exact only for t a
// power of two (for other t the product x_j w_int overflows,
// known-issues.md item 1); here x_j is reduced mod t first (a Barrett
// reduction by floor(2^64 / t)), so any t < 2^32 is exact, and for powers
// of two the result equals the masked one.
__global__ void __launch_bounds__(256)
    k_simple_scaling(cp_gi *__restrict__ a, int ng, const cp_hps *__restrict__ H,
                     uint64_t tmu) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= 2 * ng) return;
  const int slot = c < ng ? c : c - ng;
  const bool im = c >= ng;
  const int k = H->k;
  const uint64_t t = H->t;
  const bool pow2 = (t & (t - 1)) == 0;
  const uint64_t mask = t - 1;
  const double two_m64 = 0x1p-64;
  double acc = 0.0;
  uint64_t r = 0;
  // each thread reads and writes only its own 8-byte half of a slot (re or
  // imag), so writing residue 0 in place cannot race with its sibling thread
  const uint64_t *w = reinterpret_cast<const uint64_t *>(a) + (im ? 1 : 0);
  for (int j = 0; j < k; j++) {
    const uint64_t v = w[2 * ((size_t)j * ng + slot)];
    const uint64_t tw = H->twint[j];
    uint64_t lo, hi;
    umul128hi(H->twfrac[j][0], H->twfrac[j][1], v, lo, hi);
    acc = fma((double)lo, two_m64, acc);  // ascending j, as the reference
    if (pow2)
      r += ((v * tw & mask) + (hi & mask)) & mask;
    else  // t < 2^32: (x mod t) w_int < 2^64, and w_int < t
      r += mod_t(mod_t(mod_t(v, t, tmu) * tw, t, tmu) + mod_t(hi, t, tmu), t, tmu);
  }
  const uint64_t res = pow2 ? (r + (uint64_t)llrint(acc)) & mask
                            : mod_t(r + (uint64_t)llrint(acc), t, tmu);
  reinterpret_cast<uint64_t *>(a)[2 * (size_t)slot + (im ? 1 : 0)] = res;
}

static inline unsigned grid256(long threads) { return (unsigned)((threads + 255) / 256); }

extern "C" int cupoly_basis_ext_q_to_b(cupoly_ctx *c, cupoly_gi *a) {
  if (!c || !a) return cp_fail(CUPOLY_ERR_ARG, "basis_ext_q_to_b: null argument");
  const int ng = c->e->ng;
  hipLaunchKernelGGL(k_basis_ext_q_to_b, dim3((ng + 63) / 64), dim3(64, c->e->kl), 0, c->stream,
                     (cp_gi *)a, ng, c->e->d_hps);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

template <int K>
static void launch_complex_scaling(cupoly_ctx *c, uint64_t *b, const cp_gi *aq, const cp_gi *ab) {
  const int ng = c->e->ng;
  hipLaunchKernelGGL(k_complex_scaling<K>, dim3(grid256(ng), c->e->kl), dim3(256), 0, c->stream,
                     b, aq, ab, ng, c->e->d_hps, c->scale_chain);
}

extern "C" int cupoly_complex_scaling(cupoly_ctx *c, uint64_t *b, const cupoly_gi *a_q,
                                      const cupoly_gi *a_b) {
  if (!c || !b || !a_q || !a_b) return cp_fail(CUPOLY_ERR_ARG, "complex_scaling: null argument");
  const cp_gi *aq = (const cp_gi *)a_q, *ab = (const cp_gi *)a_b;
  switch (c->e->k) {
    case 1: launch_complex_scaling<1>(c, b, aq, ab); break;
    case 2: launch_complex_scaling<2>(c, b, aq, ab); break;
    case 3: launch_complex_scaling<3>(c, b, aq, ab); break;
    case 4: launch_complex_scaling<4>(c, b, aq, ab); break;
    case 5: launch_complex_scaling<5>(c, b, aq, ab); break;
    case 6: launch_complex_scaling<6>(c, b, aq, ab); break;
    case 7: launch_complex_scaling<7>(c, b, aq, ab); break;
    case 8: launch_complex_scaling<8>(c, b, aq, ab); break;
    case 9: launch_complex_scaling<9>(c, b, aq, ab); break;
    case 10: launch_complex_scaling<10>(c, b, aq, ab); break;
    case 11: launch_complex_scaling<11>(c, b, aq, ab); break;
    case 12: launch_complex_scaling<12>(c, b, aq, ab); break;
    case 13: launch_complex_scaling<13>(c, b, aq, ab); break;
    case 14: launch_complex_scaling<14>(c, b, aq, ab); break;
    default: return cp_fail(CUPOLY_ERR_UNSUPPORTED, "complex_scaling: k outside 1..14");
  }
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

extern "C" int cupoly_basis_ext_b_to_q(cupoly_ctx *c, cupoly_gi *b, const uint64_t *a) {
  if (!c || !b || !a) return cp_fail(CUPOLY_ERR_ARG, "basis_ext_b_to_q: null argument");
  const int ng = c->e->ng;
  hipLaunchKernelGGL(k_basis_ext_b_to_q, dim3(grid256(ng), c->e->k), dim3(256), 0, c->stream,
                     (cp_gi *)b, a, ng, c->e->d_hps);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

extern "C" int cupoly_simple_scaling(cupoly_ctx *c, cupoly_gi *a) {
  if (!c || !a) return cp_fail(CUPOLY_ERR_ARG, "simple_scaling: null argument");
  const uint64_t t = c->e->t;
  const bool pow2 = (t & (t - 1)) == 0;
  if (!pow2 && t >= (1ULL << 32))
    return cp_fail(CUPOLY_ERR_UNSUPPORTED,
                   "simple_scaling: t must be a power of 2 or below 2^32");
  const uint64_t tmu = pow2 ? 0 : (uint64_t)(((cp_u128)1 << 64) / t);  // floor(2^64 / t)
  const int ng = c->e->ng;
  hipLaunchKernelGGL(k_simple_scaling, dim3(grid256(2L * ng)), dim3(256), 0, c->stream,
                     (cp_gi *)a, ng, c->e->d_hps, tmu);
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

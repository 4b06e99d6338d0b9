// Batched BFV homomorphic multiplication core (tensor product + HPS rescale).
//
// For each pair of ciphertexts ct1 = (c0, c1), ct2 = (d0, d1) in base Q,
// coefficient domain, compute the three polynomials
//     y_m = round(t/Q * z_m) mod Q,  z_0 = c0 d0, z_1 = c0 d1 + c1 d0, z_2 = c1 d1,
// with the products taken in base QB (HPS, Halevi-Polyakov-Shoup).  This is
// the composition the reference builds from its pieces:
// poly_basis_extension_Q_to_B (src/arithmetic/polynomial.cu:407-457), DGT,
// poly_dr2 (polynomial.cu:951-999 -> dgt.cu:1251-1312), inverse DGT and
// poly_complex_scaling_tDivQ (polynomial.cu:350-404, cudaengine.cu:382-796).
//
// Pipeline, every stage batched over all ciphertext pairs (one launch each):
//   1. k_ext_q_to_qb:   each of the 4 inputs -> QB residues (Q copied, B by the
//                       HPS extension), one thread per Gaussian slot;
//   2. forward DGT of the 4*(k+kl) residue polynomials per pair;
//   3. k_tensor:        slot-wise c0 d0, c0 d1 + c1 d0, c1 d1 in every residue;
//   4. inverse DGT of the 3*(k+kl) results;
//   5. k_scale_to_q:    complex scaling t/Q (QB -> B) and B -> Q fused in
//                       registers, one thread per slot, writing base Q.
// Scratch lives in a grow-only workspace owned by the context.
#include "hps_dev.hpp"

// Where ciphertext i lives: base + i * stride, or table[i] (gathered
// requests of a serving batch).
struct cp_cts {
  cp_gi *base;
  size_t stride;
  cp_gi *const *table;
  __device__ __forceinline__ cp_gi *at(int i) const { return table ? table[i] : base + i * stride; }
};

// blockIdx.y = pair * 4 + which (0: c0, 1: c1, 2: d0, 3: d1)
__global__ void __launch_bounds__(256)
    k_ext_q_to_qb(cp_gi *__restrict__ X, cp_cts ct1, cp_cts ct2, int ng,
                  const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int k = H->k, kl = H->kl;
  const int pair = blockIdx.y >> 2, which = blockIdx.y & 3;
  const cp_gi *in = (which < 2 ? ct1.at(pair) : ct2.at(pair)) + (size_t)(which & 1) * k * ng;
  cp_gi *out = X + (size_t)blockIdx.y * (k + kl) * ng;
  for (int j = 0; j < k; j++) out[(size_t)j * ng + c] = in[(size_t)j * ng + c];
  hps_q_to_b_slot(in, ng, out + (size_t)k * ng, ng, c, H);
}

// Y[pair][m][r][slot] from X[pair][0..3][r][slot]; nres = k + kl residues.
__global__ void __launch_bounds__(256)
    k_tensor(cp_gi *__restrict__ Y, const cp_gi *__restrict__ X, int ng, int nres, long total,
             const cp_prime *__restrict__ primes) {
  const long per_pair = (long)nres * ng;
  for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < total;
       s += (long)gridDim.x * blockDim.x) {
    const long pair = s / per_pair, rs = s % per_pair;
    const cp_prime pr = primes[(int)(rs / ng)];
    const uint64_t p = pr.p;
    const cp_gi *x = X + pair * 4 * per_pair + rs;
    const cp_gi a0 = x[0], a1 = x[per_pair], b0 = x[2 * per_pair], b1 = x[3 * per_pair];
    const cp_gi r0 = cp_gimul(a0, b0, pr), r2 = cp_gimul(a1, b1, pr);
    const cp_gi u = cp_gimul(a0, b1, pr), v = cp_gimul(a1, b0, pr);
    cp_gi r1;
    r1.re = cp_add(u.re, v.re, p);
    r1.imag = cp_add(u.imag, v.imag, p);
    cp_gi *y = Y + pair * 3 * per_pair + rs;
    y[0] = r0;
    y[per_pair] = r1;
    y[2 * per_pair] = r2;
  }
}

// blockIdx.y = output polynomial index (pair * 3 + m)
__global__ void __launch_bounds__(256)
    k_scale_to_q(cp_cts out, const cp_gi *__restrict__ Y, int ng, const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int k = H->k, kl = H->kl;
  const int pair = blockIdx.y / 3, m = blockIdx.y % 3;
  const cp_gi *y = Y + (size_t)blockIdx.y * (k + kl) * ng;
  uint64_t yr[CP_MAX_BASE], yi[CP_MAX_BASE];
  hps_scale_slot(y, ng, y + (size_t)k * ng, ng, c, H, yr, yi);
  hps_b_to_q_slot(yr, yi, out.at(pair) + (size_t)m * k * ng, ng, c, H);
}

int cp_workspace(cupoly_ctx *c, size_t bytes, void **out) {
  if (c->ws_bytes < bytes) {
    if (c->d_ws) {
      CP_HIP(hipStreamSynchronize(c->stream));
      CP_HIP(hipFree(c->d_ws));
      c->d_ws = nullptr;
      c->ws_bytes = 0;
    }
    CP_HIP(hipMalloc(&c->d_ws, bytes));
    c->ws_bytes = bytes;
  }
  *out = c->d_ws;
  return CUPOLY_OK;
}

static int hom_mul_run(cupoly_ctx *c, cp_cts out, cp_cts ct1, cp_cts ct2, int batch) {
  cupoly_engine *e = c->e;
  const int k = e->k, kl = e->kl, nres = k + kl, ng = e->ng;
  if (batch > 65535 / 4) return cp_fail(CUPOLY_ERR_ARG, "hom_mul: batch > 16383");
  const size_t xs = (size_t)batch * 4 * nres * ng, ys = (size_t)batch * 3 * nres * ng;
  void *ws;
  int rc = cp_workspace(c, (xs + ys) * sizeof(cp_gi), &ws);
  if (rc) return rc;
  cp_gi *X = (cp_gi *)ws, *Y = X + xs;
  const dim3 sg((ng + 255) / 256);
  hipLaunchKernelGGL(k_ext_q_to_qb, dim3(sg.x, batch * 4), dim3(256), 0, c->stream, X, ct1, ct2,
                     ng, e->d_hps);
  CP_LAUNCH_CHECK();
  if ((rc = cp_dgt_launch(c, X, 0, nres, batch * 4, CUPOLY_FORWARD))) return rc;
  const long total = (long)batch * nres * ng;
  long g = (total + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_tensor, dim3((unsigned)g), dim3(256), 0, c->stream, Y, X, ng, nres, total,
                     e->d_primes);
  CP_LAUNCH_CHECK();
  if ((rc = cp_dgt_launch(c, Y, 0, nres, batch * 3, CUPOLY_INVERSE))) return rc;
  hipLaunchKernelGGL(k_scale_to_q, dim3(sg.x, batch * 3), dim3(256), 0, c->stream, out, Y, ng,
                     e->d_hps);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

extern "C" int cupoly_hom_mul_batch(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *ct1,
                                    const cupoly_gi *ct2, int batch) {
  if (!c || !out || !ct1 || !ct2 || batch < 0)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_batch: bad argument");
  if (batch == 0) return CUPOLY_OK;
  const size_t k = c->e->k, ng = c->e->ng;
  cp_cts o{(cp_gi *)out, 3 * k * ng, nullptr}, a{(cp_gi *)ct1, 2 * k * ng, nullptr},
      b{(cp_gi *)ct2, 2 * k * ng, nullptr};
  return hom_mul_run(c, o, a, b, batch);
}

extern "C" int cupoly_hom_mul_gather(cupoly_ctx *c, cupoly_gi *const *out,
                                     const cupoly_gi *const *ct1, const cupoly_gi *const *ct2,
                                     int batch) {
  if (!c || !out || !ct1 || !ct2 || batch < 0)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_gather: bad argument");
  if (batch == 0) return CUPOLY_OK;
  for (int i = 0; i < batch; i++)
    if (!out[i] || !ct1[i] || !ct2[i]) return cp_fail(CUPOLY_ERR_ARG, "hom_mul_gather: null entry");
  // pointer tables to the device on the context's stream, from a pinned
  // staging slot that is not rewritten until its previous copy has finished
  const size_t need = (size_t)batch * 3 * sizeof(void *);
  if (c->ptr_bytes < need) {
    if (c->d_ptrs) {
      CP_HIP(hipStreamSynchronize(c->stream));
      CP_HIP(hipFree(c->d_ptrs));
      c->d_ptrs = nullptr;
      c->ptr_bytes = 0;
    }
    CP_HIP(hipMalloc((void **)&c->d_ptrs, need));
    c->ptr_bytes = need;
  }
  if (c->ring_slot_bytes < need) {
    if (c->h_ring) {
      CP_HIP(hipStreamSynchronize(c->stream));
      CP_HIP(hipHostFree(c->h_ring));
      c->h_ring = nullptr;
    }
    CP_HIP(hipHostMalloc(&c->h_ring, need * cupoly_ctx::kRing, hipHostMallocDefault));
    c->ring_slot_bytes = need;
    for (int i = 0; i < cupoly_ctx::kRing; i++) c->ring_used[i] = false;
  }
  const int slot = c->ring_next;
  c->ring_next = (slot + 1) % cupoly_ctx::kRing;
  if (!c->ring_ev[slot]) CP_HIP(hipEventCreateWithFlags(&c->ring_ev[slot], hipEventDisableTiming));
  if (c->ring_used[slot]) CP_HIP(hipEventSynchronize(c->ring_ev[slot]));
  const void **h = (const void **)((char *)c->h_ring + (size_t)slot * c->ring_slot_bytes);
  for (int i = 0; i < batch; i++) {
    h[i] = out[i];
    h[batch + i] = ct1[i];
    h[2 * batch + i] = ct2[i];
  }
  CP_HIP(hipMemcpyAsync(c->d_ptrs, h, need, hipMemcpyHostToDevice, c->stream));
  CP_HIP(hipEventRecord(c->ring_ev[slot], c->stream));
  c->ring_used[slot] = true;
  cp_gi **d = (cp_gi **)c->d_ptrs;
  cp_cts o{nullptr, 0, d}, a{nullptr, 0, d + batch}, b{nullptr, 0, d + 2 * batch};
  return hom_mul_run(c, o, a, b, batch);
}

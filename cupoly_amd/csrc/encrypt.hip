// Fused sampling + multiply-add: out = a * s + e in the DGT domain.
//
// The encryption-shaped step of RLWE schemes: the reference samples s and e
// into poly_t objects (Sampler::sample, src/cuda/sampler.cu:576-635, each
// followed by a forward DGT) and then calls poly_mul_add (polynomial.cu:916-948
// -> muladd_dgt, dgt.cu:1439-1499): five launches and four polynomials through
// HBM.  Here two launches:
//   1. k_sample_se draws s (ternary) and e (discrete Gaussian) of every
//      instance once from their Philox counters (they are the same integers
//      in every residue) and packs each Gaussian slot's four small values
//      (s re, s imag, e re, e imag) as signed bytes: 4 bytes per slot;
//   2. one workgroup per (instance, residue) reads its slots' packed values,
//      reduces them mod q_r, twists and transforms e and then s in LDS, and
//      writes a * s + e slot by slot as it reads a (once, DGT domain).  With e
//      transformed first, its spectrum is the only value held across the
//      transform of s, and a is read only at the end: no register spill at
//      128 VGPRs (the order s, a * s, e held a * s across the transform of e
//      and spilled 92 bytes per lane).
// HBM traffic: a in and out out, 16 k n bytes per instance, plus the packed
// samples (2 n bytes written once, read k times).
//
// Instance i uses Philox stream stream0 + 2i for s and stream0 + 2i + 1 for e,
// so the small polynomials are the ones cupoly_sample_coeff draws for those
// streams (same values in every residue).
#include "dgt_dev.hpp"
#include "philox.hpp"

// one thread per (instance, Gaussian slot g): s and e coefficients g and g + ng
__global__ void __launch_bounds__(256)
    k_sample_se(uint32_t *__restrict__ se, int ng, long total, uint64_t seed, uint64_t stream0,
                cp_cdt cdt) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const uint64_t inst = (uint64_t)(idx / ng);
    const uint32_t g = (uint32_t)(idx % ng);
    const int s0 = ternary_coef(seed, stream0 + 2 * inst, g);
    const int s1 = ternary_coef(seed, stream0 + 2 * inst, g + (uint32_t)ng);
    const int e0 = gaussian_coef(seed, stream0 + 2 * inst + 1, g, cdt);
    const int e1 = gaussian_coef(seed, stream0 + 2 * inst + 1, g + (uint32_t)ng, cdt);
    se[idx] = (uint32_t)(uint8_t)s0 | ((uint32_t)(uint8_t)s1 << 8) |
              ((uint32_t)(uint8_t)e0 << 16) | ((uint32_t)(uint8_t)e1 << 24);
  }
}

__device__ __forceinline__ int se_byte(uint32_t w, int k) { return (int)(int8_t)(w >> (8 * k)); }

template <int LOGN, int ROWS>
__global__ void __launch_bounds__(((1 << LOGN) / EPT) * ROWS, fused_min_waves<LOGN>())
    k_sample_muladd(cp_gi *__restrict__ out, const cp_gi *__restrict__ a, int nunits, int nres,
                    const cp_prime *__restrict__ primes, const uint64_t *__restrict__ twf_all,
                    const uint64_t *__restrict__ tsf_all, const uint32_t *__restrict__ se_all) {
  constexpr int N = 1 << LOGN, T = N / EPT;
  using LP = LastPass<LOGN, 0, LOGN>;
  extern __shared__ cp_gi lds_all[];
  const int row = ROWS == 1 ? 0 : (int)(threadIdx.x / T);
  const int t = ROWS == 1 ? (int)threadIdx.x : (int)(threadIdx.x % T);
  cp_gi *lds = lds_all + row * lds_row<LOGN>();
  int unit = blockIdx.x * ROWS + row;
  const bool active = unit < nunits;
  if (!active) unit = nunits - 1;
  int rid = unit % nres;
  if (ROWS == 1) rid = __builtin_amdgcn_readfirstlane(rid);
  const uint64_t inst = (uint64_t)(unit / nres);
  const cp_prime pr = primes[rid];
  const uint64_t p = pr.p;
  const uint64_t *twf = twf_all + (size_t)rid * N;
  const uint64_t *tsf = tsf_all + (size_t)rid * N * 6;
  const uint32_t *se = se_all + inst * N;
  const cp_gi *ap = a + (size_t)unit * N;
  cp_gi *op = out + (size_t)unit * N;
  MapIdentity id;
  cp_gi x[EPT], y[EPT];
  // e first: its spectrum is the only value held across the transform of s
  const int te = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int g = first_index<LOGN, 0, LOGN>(te, i);
    const uint32_t w = se[g];
    y[i].re = small_mod(se_byte(w, 2), p);
    y[i].imag = small_mod(se_byte(w, 3), p);
    twist(y[i], tsf, g, p);
    __builtin_amdgcn_sched_barrier(0);
  }
  dif_passes<LOGN, LOGN, 0, LOGN, 0>(y, lds, t, twf, p, id);
  const int ts = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int g = first_index<LOGN, 0, LOGN>(ts, i);
    const uint32_t w = se[g];
    x[i].re = small_mod(se_byte(w, 0), p);
    x[i].imag = small_mod(se_byte(w, 1), p);
    twist(x[i], tsf, g, p);
    __builtin_amdgcn_sched_barrier(0);
  }
  dif_passes<LOGN, LOGN, 0, LOGN, 0>(x, lds, t, twf, p, id);
  const int to = cp_opaque(t);
  if (active) {
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      const int j = pass_index<LOGN, LP::LS0, LP::RL>(to, i);
      const cp_gi m = cp_gimul(ap[j], x[i], pr);
      cp_gi r;
      r.re = cp_add(m.re, y[i].re, p);
      r.imag = cp_add(m.imag, y[i].imag, p);
      op[j] = r;
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

template <int LOGN>
static int sample_muladd_t(cupoly_ctx *c, cp_gi *out, const cp_gi *a, int batch, uint64_t seed,
                           uint64_t stream0) {
  cupoly_engine *e = c->e;
  const int nres = e->k, nunits = nres * batch;
  constexpr int R = rows_for<LOGN>(), T = (1 << LOGN) / EPT;
  const size_t lds = (size_t)R * lds_row<LOGN>() * sizeof(cp_gi);
  // packed samples of every instance: 4 bytes per Gaussian slot
  const long total = (long)batch * e->ng;
  void *ws;
  int rc = cp_workspace(c, (size_t)total * sizeof(uint32_t), &ws);
  if (rc) return rc;
  long g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_sample_se, dim3((unsigned)g), dim3(256), 0, c->stream, (uint32_t *)ws,
                     e->ng, total, seed, stream0, cp_make_cdt());
  CP_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_sample_muladd<LOGN, R>), dim3((nunits + R - 1) / R), dim3(T * R), lds,
                     c->stream, out, a, nunits, nres, e->d_primes, e->d_tw_fwd, e->d_twist_fwd,
                     (const uint32_t *)ws);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

// Dynamic LDS limit of k_sample_muladd, set once per engine on the engine's
// device (engine creation).
template <int LOGN>
static int encrypt_setup_t() {
  constexpr int R = rows_for<LOGN>();
  const size_t lds = (size_t)R * lds_row<LOGN>() * sizeof(cp_gi);
  if (lds > 64 * 1024)
    CP_HIP(hipFuncSetAttribute((const void *)k_sample_muladd<LOGN, R>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return CUPOLY_OK;
}

int cp_encrypt_setup(cupoly_engine *e) {
  switch (e->logng) {
    case 4: return encrypt_setup_t<4>();
    case 5: return encrypt_setup_t<5>();
    case 6: return encrypt_setup_t<6>();
    case 7: return encrypt_setup_t<7>();
    case 8: return encrypt_setup_t<8>();
    case 9: return encrypt_setup_t<9>();
    case 10: return encrypt_setup_t<10>();
    case 11: return encrypt_setup_t<11>();
    case 12: return encrypt_setup_t<12>();
    case 13: return encrypt_setup_t<13>();
    default: return CUPOLY_OK;  // larger rings compose the generic kernels
  }
}

extern "C" int cupoly_sample_muladd_batch(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a,
                                          int batch, uint64_t seed, uint64_t stream0) {
  if (c && batch == 0) return CUPOLY_OK;  // empty batch: pointers may be null
  if (!c || !out || !a || batch < 0) return cp_fail(CUPOLY_ERR_ARG, "sample_muladd: bad argument");
  if ((long long)batch * c->e->k >= (1LL << 31))
    return cp_fail(CUPOLY_ERR_ARG, "sample_muladd: batch * k must stay below 2^31");
  switch (c->e->logng) {
    case 4: return sample_muladd_t<4>(c, (cp_gi *)out, (const cp_gi *)a, batch, seed, stream0);
    case 5: return sample_muladd_t<5>(c, (cp_gi *)out, (const cp_gi *)a, batch, seed, stream0);
    case 6: return sample_muladd_t<6>(c, (cp_gi *)out, (const cp_gi *)a, batch, seed, stream0);
    case 7: return sample_muladd_t<7>(c, (cp_gi *)out, (const cp_gi *)a, batch, seed, stream0);
    case 8: return sample_muladd_t<8>(c, (cp_gi *)out, (const cp_gi *)a, batch, seed, stream0);
    case 9: return sample_muladd_t<9>(c, (cp_gi *)out, (const cp_gi *)a, batch, seed, stream0);
    case 10: return sample_muladd_t<10>(c, (cp_gi *)out, (const cp_gi *)a, batch, seed, stream0);
    case 11: return sample_muladd_t<11>(c, (cp_gi *)out, (const cp_gi *)a, batch, seed, stream0);
    case 12: return sample_muladd_t<12>(c, (cp_gi *)out, (const cp_gi *)a, batch, seed, stream0);
    case 13: return sample_muladd_t<13>(c, (cp_gi *)out, (const cp_gi *)a, batch, seed, stream0);
    default: break;
  }
  // n > 16384: sample, transform and multiply-add as separate launches
  cupoly_engine *e = c->e;
  const size_t stride = (size_t)e->k * e->ng;
  void *ws;
  int rc = cp_workspace(c, 2 * stride * sizeof(cp_gi), &ws);
  if (rc) return rc;
  cupoly_gi *s = (cupoly_gi *)ws, *ee = s + stride;
  for (int i = 0; i < batch; i++) {
    if ((rc = cupoly_sample(c, s, CUPOLY_TERNARY, seed, stream0 + 2 * (uint64_t)i, CUPOLY_QBASE)) ||
        (rc = cupoly_sample(c, ee, CUPOLY_DISCRETE_GAUSSIAN, seed, stream0 + 2 * (uint64_t)i + 1,
                           CUPOLY_QBASE)) ||
        (rc = cupoly_muladd_dgt(c, out + i * stride, a + i * stride, s, ee, CUPOLY_QBASE)))
      return rc;
  }
  return CUPOLY_OK;
}

// RNS word decomposition of BFV key switching, and the dot product of
// polynomial vectors.
//
// Reference: CUDAEngine::execute_xi_rns / execute_rho_bfv_rns
// (src/cuda/cudaengine.cu:992-1062) behind poly_xi_bfv / poly_rho_bfv
// (src/arithmetic/polynomial.cu:610-693), and poly_dot
// (polynomial.cu:1439-1451).
//   xi(a)[i]:  residue i of output i is a_i (Q/q_i)^-1 mod q_i,
//   rho(a)[i]: residue i of output i is a_i (Q/q_i) mod q_i,
// for i < k (Q residues only, also when a is in base QB), in whichever
// domain a is (a product by a constant commutes with the DGT).  As in the
// reference only that residue of each output is written; the others keep
// their contents (zero for freshly initialised polynomials).  Then
// sum_i rho(a)[i] * xi(b)[i] = a * b, the identity the reference tests
// (test_bfv_mode.cu:492-546).
// dot(a, b) = sum_i a[i] * b[i] slot-wise in the DGT domain; the reference
// builds it from one poly_mul and k - 1 poly_mul_add (2k - 1 launches and k
// round trips of the accumulator); here one streaming kernel reads each
// operand once.
#include "internal.hpp"

struct cp_ptrs {
  cp_gi *p[CP_MAX_BASE];
};
struct cp_cptrs {
  const cp_gi *p[CP_MAX_BASE];
};

template <bool XI>
__global__ void __launch_bounds__(256)
    k_xi_rho(cp_ptrs out, const cp_gi *__restrict__ a, int ng, int k,
             const cp_hps *__restrict__ H) {
  const long total = (long)ng * k;
  for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < total;
       s += (long)gridDim.x * blockDim.x) {
    const int rid = (int)(s / ng);
    const uint64_t q = H->q[rid].p;
    const uint64_t w = XI ? H->invQqi[rid][0] : H->Qqi[rid][0];
    const uint64_t wp = XI ? H->invQqi[rid][1] : H->Qqi[rid][1];
    cp_gi v = a[s];
    v.re = cp_shoup(v.re, w, wp, q);
    v.imag = cp_shoup(v.imag, w, wp, q);
    out.p[rid][s] = v;
  }
}

// out (+)= sum_{i < cnt} a[i] * b[i] over nres residues from roff.
__global__ void __launch_bounds__(256)
    k_dot(cp_gi *__restrict__ out, cp_cptrs a, cp_cptrs b, int cnt, bool accumulate, int ng,
          int nres, int roff, const cp_prime *__restrict__ primes) {
  const long total = (long)ng * nres;
  for (long s = blockIdx.x * (long)blockDim.x + threadIdx.x; s < total;
       s += (long)gridDim.x * blockDim.x) {
    const cp_prime pr = primes[roff + (int)(s / ng)];
    const uint64_t p = pr.p;
    cp_gi acc = accumulate ? out[s] : cp_gi{0, 0};
    for (int i = 0; i < cnt; i++) {
      const cp_gi m = cp_gimul(a.p[i][s], b.p[i][s], pr);
      acc.re = cp_add(acc.re, m.re, p);
      acc.imag = cp_add(acc.imag, m.imag, p);
    }
    out[s] = acc;
  }
}

static int grid_for(long total) {
  long g = (total + 255) / 256;
  if (g > 256 * 16) g = 256 * 16;
  return (int)(g < 1 ? 1 : g);
}

static int xi_rho(cupoly_ctx *c, cupoly_gi *const *out, const cupoly_gi *a, bool xi) {
  if (!c || !out || !a) return cp_fail(CUPOLY_ERR_ARG, xi ? "xi_bfv: null argument" : "rho_bfv: null argument");
  const int k = c->e->k, ng = c->e->ng;
  cp_ptrs o{};
  for (int i = 0; i < k; i++) {
    if (!out[i]) return cp_fail(CUPOLY_ERR_ARG, "xi/rho: null output " + std::to_string(i));
    o.p[i] = (cp_gi *)out[i];
  }
  const long total = (long)ng * k;
  if (xi)
    hipLaunchKernelGGL(k_xi_rho<true>, dim3(grid_for(total)), dim3(256), 0, c->stream, o,
                       (const cp_gi *)a, ng, k, c->e->d_hps);
  else
    hipLaunchKernelGGL(k_xi_rho<false>, dim3(grid_for(total)), dim3(256), 0, c->stream, o,
                       (const cp_gi *)a, ng, k, c->e->d_hps);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

extern "C" int cupoly_xi_bfv(cupoly_ctx *c, cupoly_gi *const *out, const cupoly_gi *a) {
  return xi_rho(c, out, a, true);
}

extern "C" int cupoly_rho_bfv(cupoly_ctx *c, cupoly_gi *const *out, const cupoly_gi *a, int base) {
  if (base != CUPOLY_QBASE && base != CUPOLY_QBBASE)
    return cp_fail(CUPOLY_ERR_UNSUPPORTED, "rho_bfv: base must be Q or QB");
  return xi_rho(c, out, a, false);
}

extern "C" int cupoly_dot(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *const *a,
                          const cupoly_gi *const *b, int count, int base) {
  if (!c || !out || !a || !b || count < 1) return cp_fail(CUPOLY_ERR_ARG, "dot: bad argument");
  int off, cnt;
  if (!cp_base_range(c->e, base, &off, &cnt)) return cp_fail(CUPOLY_ERR_UNSUPPORTED, "Unknown base");
  for (int i = 0; i < count; i++)
    if (!a[i] || !b[i]) return cp_fail(CUPOLY_ERR_ARG, "dot: null operand " + std::to_string(i));
  const long total = (long)c->e->ng * cnt;
  for (int i0 = 0; i0 < count; i0 += CP_MAX_BASE) {  // operands in chunks of CP_MAX_BASE
    const int n = count - i0 < CP_MAX_BASE ? count - i0 : CP_MAX_BASE;
    cp_cptrs pa{}, pb{};
    for (int i = 0; i < n; i++) {
      pa.p[i] = (const cp_gi *)a[i0 + i];
      pb.p[i] = (const cp_gi *)b[i0 + i];
    }
    hipLaunchKernelGGL(k_dot, dim3(grid_for(total)), dim3(256), 0, c->stream, (cp_gi *)out, pa,
                       pb, n, i0 > 0, c->e->ng, cnt, off, c->e->d_primes);
    CP_LAUNCH_CHECK();
  }
  return CUPOLY_OK;
}

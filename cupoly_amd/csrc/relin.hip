// BFV relinearization with a device-resident key, and the full BFV
// multiplication (tensor product + HPS rescale + relinearization).
//
// The reference library provides the pieces of key switching and leaves the
// scheme to its callers: the RNS word decomposition poly_xi_bfv /
// poly_rho_bfv (src/arithmetic/polynomial.cu:610-693 ->
// src/cuda/cudaengine.cu:992-1062) and the dot product poly_dot
// (polynomial.cu:1439-1451).  This file composes them into a
// relinearization that decrypts correctly:
//   key     rlk_i = (b_i, a_i), i < k: a_i uniform mod Q, e_i discrete
//           Gaussian, b_i = -(a_i s + e_i) + rho(s^2)_i, where
//           rho(x)_i = x (Q/q_i) mod Q is zero in every residue but i (the
//           reference's rho).  Held in the DGT domain of base Q,
//           [k][2][k][n/2] Gaussian slots, for the key's lifetime.
//   digits  D_i = [y2_i (Q/q_i)^-1]_{q_i} (the reference's xi, residue i) is
//           an integer below q_i; here it is written to EVERY residue j as
//           D_i mod q_j.  (The reference's xi writes residue i only, which
//           satisfies sum_i rho(a)_i xi(b)_i = a b, but leaves the digit out
//           of the other residues, where the key's a_i s + e_i part needs it.)
//   output  y0' = y0 + sum_i D_i b_i,  y1' = y1 + sum_i D_i a_i  (mod Q), so
//           y0' + y1' s = y0 + y1 s + s^2 sum_i D_i (Q/q_i) - sum_i D_i e_i
//                       = y0 + y1 s + y2 s^2 - sum_i D_i e_i   (mod Q),
//           the relinearized ciphertext with the added noise sum_i D_i e_i.
// Two forms:
//   * n = 2^11 .. 2^14 (k_relin_fti): one workgroup per (pair, residue j)
//     forms the k digits of residue j, transforms each, and accumulates both
//     key products (the b-sum in registers, the a-sum parked thread-major in
//     scratch), then runs the two inverse transforms and adds y0, y1;
//   * other sizes: digits for all residues, batched forward transforms, one
//     streaming dot-product kernel, batched inverse transforms, one add.
#include <string>

// one radix-2 stage's twiddle loads in flight at a time (dgt_dev.hpp): 1.4%
// faster here than the compiler's schedule across stages
#define CP_STAGE_BARRIER 1
#include "dgt_dev.hpp"

struct cupoly_relin_key {
  cupoly_engine *e = nullptr;
  cp_gi *d = nullptr;   // [k][2][k][ng], DGT domain: [i][0] = b_i, [i][1] = a_i
  cp_gi *ds = nullptr;  // the same in the split basis (dgt_dev.hpp), for k_relin_fti
};

// Where polynomial-set i lives: base + i * stride, or table[i].
struct cp_rcts {
  cp_gi *base;
  size_t stride;
  cp_gi *const *table;
  __device__ __forceinline__ cp_gi *at(int i) const { return table ? table[i] : base + i * stride; }
};

// ---- key generation ---------------------------------------------------------
// b_i = -(a_i s + e_i) + [j == i] s^2 (Q/q_i mod q_i), slot-wise in the DGT
// domain; blockIdx.y = i * k + j.
__global__ void __launch_bounds__(256)
    k_relin_key_b(cp_gi *__restrict__ key, const cp_gi *__restrict__ e_all,
                  const cp_gi *__restrict__ s, const cp_gi *__restrict__ s2, int ng,
                  const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int k = H->k, i = blockIdx.y / k, j = blockIdx.y % k;
  const cp_prime pr = H->q[j];
  const uint64_t p = pr.p;
  const size_t row = (size_t)j * ng + c;
  const cp_gi a = key[((size_t)(2 * i + 1) * k) * ng + row];
  const cp_gi as = cp_gimul(a, s[row], pr), e = e_all[(size_t)i * k * ng + row];
  cp_gi b;
  b.re = cp_sub(0, cp_add(as.re, e.re, p), p);
  b.imag = cp_sub(0, cp_add(as.imag, e.imag, p), p);
  if (i == j) {
    const cp_gi v = s2[row];
    b.re = cp_add(b.re, cp_shoup(v.re, H->Qqi[i][0], H->Qqi[i][1], p), p);
    b.imag = cp_add(b.imag, cp_shoup(v.imag, H->Qqi[i][0], H->Qqi[i][1], p), p);
  }
  key[((size_t)(2 * i) * k) * ng + row] = b;
}

// The key in the split basis of dgt_dev.hpp: a DGT-domain slot X of residue j
// becomes (X.re + J X.im, X.re - J X.im) mod q_j.  a + ib -> a +- Jb is a ring
// homomorphism Z_p[i] -> Z_p, so this is the split-basis spectrum of the same
// polynomial, position for position.  blockIdx.y = row of the key (i, m, j),
// j = blockIdx.y % k.
__global__ void __launch_bounds__(256)
    k_relin_key_split(cp_gi *__restrict__ ds, const cp_gi *__restrict__ d, int ng,
                      const cp_hps *__restrict__ H, const uint64_t *__restrict__ sk_all) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int j = blockIdx.y % H->k;
  const uint64_t p = H->q[j].p, *sk = sk_all + (size_t)j * 8;
  const size_t at = (size_t)blockIdx.y * ng + c;
  const cp_gi x = d[at];
  const uint64_t t = cp_shoup(x.imag, sk[0], sk[1], p);
  // times h = 1/2N (sk[2], sk[3]): every product with the key then carries the
  // inverse transform's scale, and the conversion back is split_back_h2
  cp_gi v;
  cp_shoup2(cp_add(x.re, t, p), sk[2], sk[3], cp_sub(x.re, t, p), sk[2], sk[3], p, v.re, v.imag);
  ds[at] = v;
}

// Digit i of coefficient-domain word x (residue i of y2), in residue j.
__device__ __forceinline__ uint64_t relin_digit(uint64_t x, const cp_hps *__restrict__ H, int i,
                                                uint64_t qj) {
  const uint64_t d = cp_shoup(x, H->invQqi[i][0], H->invQqi[i][1], H->q[i].p);
  return cp_red2(d, qj);  // d < q_i < 2 q_j: the bucket primes lie within 2^30 of 2^63
}

// ---- generic form ------------------------------------------------------------
// Dg[pair][i][j][c] = D_i mod q_j; blockIdx.y = pair * k + i.
__global__ void __launch_bounds__(256)
    k_relin_digits(cp_gi *__restrict__ Dg, const cp_gi *__restrict__ T, int ng,
                   const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int k = H->k, pair = blockIdx.y / k, i = blockIdx.y % k;
  const cp_gi x = T[(((size_t)pair * 3 + 2) * k + i) * ng + c];
  cp_gi *d = Dg + (size_t)blockIdx.y * k * ng + c;
  for (int j = 0; j < k; j++) {
    const uint64_t qj = H->q[j].p;
    d[(size_t)j * ng] = cp_gi{relin_digit(x.re, H, i, qj), relin_digit(x.imag, H, i, qj)};
  }
}

// ACC[pair][m][j][c] = sum_i Dg^[pair][i][j][c] key[i][m][j][c]; blockIdx.y = pair * k + j.
__global__ void __launch_bounds__(256)
    k_relin_dot(cp_gi *__restrict__ ACC, const cp_gi *__restrict__ Dg,
                const cp_gi *__restrict__ key, int ng, const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int k = H->k, pair = blockIdx.y / k, j = blockIdx.y % k;
  const cp_prime pr = H->q[j];
  const uint64_t p = pr.p;
  cp_gi a0{0, 0}, a1{0, 0};
  for (int i = 0; i < k; i++) {
    const cp_gi d = Dg[(((size_t)pair * k + i) * k + j) * ng + c];
    const cp_gi m0 = cp_gimul(d, key[(((size_t)2 * i) * k + j) * ng + c], pr);
    const cp_gi m1 = cp_gimul(d, key[(((size_t)2 * i + 1) * k + j) * ng + c], pr);
    a0.re = cp_add(a0.re, m0.re, p);
    a0.imag = cp_add(a0.imag, m0.imag, p);
    a1.re = cp_add(a1.re, m1.re, p);
    a1.imag = cp_add(a1.imag, m1.imag, p);
  }
  ACC[(((size_t)pair * 2) * k + j) * ng + c] = a0;
  ACC[(((size_t)pair * 2 + 1) * k + j) * ng + c] = a1;
}

// out[pair][m][j][c] = T[pair][m][j][c] + ACC[pair][m][j][c]; blockIdx.y = (pair * 2 + m) * k + j.
__global__ void __launch_bounds__(256)
    k_relin_add(cp_rcts out, const cp_gi *__restrict__ T, const cp_gi *__restrict__ ACC, int ng,
                const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int k = H->k, j = blockIdx.y % k, pm = blockIdx.y / k, pair = pm / 2, m = pm % 2;
  const uint64_t p = H->q[j].p;
  const cp_gi y = T[(((size_t)pair * 3 + m) * k + j) * ng + c];
  const cp_gi a = ACC[(size_t)blockIdx.y * ng + c];
  out.at(pair)[((size_t)m * k + j) * ng + c] = cp_gi{cp_add(y.re, a.re, p), cp_add(y.imag, a.imag, p)};
}

// ---- fused form, n = 2^11 .. 2^14 ---------------------------------------------
// In the split basis of dgt_dev.hpp (digits converted and twisted, products
// with the split-basis key, split-tree inverses): nothing of the transform
// domain leaves the workgroup but the parked a-sum, which only it reads.
template <int LOGN, int ROWS>
__global__ void __launch_bounds__(((1 << LOGN) / EPT) * ROWS, fused_min_waves<LOGN>())
    k_relin_fti(cp_rcts out, const cp_gi *__restrict__ T, const cp_gi *__restrict__ key,
                cp_gi *__restrict__ park_all, int batch, const cp_hps *__restrict__ H,
                const uint64_t *__restrict__ twf_all, const uint64_t *__restrict__ root_all,
                const uint64_t *__restrict__ stw_all, const uint64_t *__restrict__ sk_all) {
  constexpr int N = 1 << LOGN, TT = N / EPT;
  using LP = LastPass<LOGN, 0, LOGN>;
  extern __shared__ cp_gi lds_all[];
  const int row = ROWS == 1 ? 0 : (int)(threadIdx.x / TT);
  const int t = ROWS == 1 ? (int)threadIdx.x : (int)(threadIdx.x % TT);
  cp_gi *lds = lds_all + row * lds_row<LOGN>();
  const int k = H->k;
  const int nunits = batch * k;
  int u = blockIdx.x * ROWS + row;
  const bool active = u < nunits;
  if (!active) u = nunits - 1;
  int j = u / batch;  // residue-major: consecutive workgroups share residue j's tables
  if (ROWS == 1) j = __builtin_amdgcn_readfirstlane(j);
  const int pair = u - j * batch;
  const cp_prime pr = H->q[j];
  const uint64_t p = pr.p;
  const uint64_t *twf = twf_all + (size_t)j * N;
  const CP_GLOBAL uint64_t *root = (const CP_GLOBAL uint64_t *)(root_all + (size_t)j * N * 4);
  const CP_GLOBAL uint64_t *stw = (const CP_GLOBAL uint64_t *)(stw_all + (size_t)j * N * 4);
  const uint64_t *sk = sk_all + (size_t)j * 8;
  const uint64_t J = sk[0], Jp = sk[1];
  const cp_gi *y = T + (size_t)pair * 3 * k * N;  // y0, y1, y2 of this pair, [3][k][N]
  cp_gi *park = park_all + (size_t)u * N;          // the a-sum, thread-major
  MapIdentity id;
  cp_gi x[EPT], acc[EPT];
#pragma nounroll
  for (int i = 0; i < k; i++) {
    const int tl = cp_opaque(t);
    const cp_gi *y2i = y + ((size_t)2 * k + i) * N;
#pragma unroll
    for (int r = 0; r < EPT; r += 2) {
      const int g0 = first_index<LOGN, 0, LOGN>(tl, r), g1 = first_index<LOGN, 0, LOGN>(tl, r + 1);
      const cp_gi v0 = y2i[g0], v1 = y2i[g1];
      x[r] = cp_gi{relin_digit(v0.re, H, i, p), relin_digit(v0.imag, H, i, p)};
      x[r + 1] = cp_gi{relin_digit(v1.re, H, i, p), relin_digit(v1.imag, H, i, p)};
      split_twist2(x[r], x[r + 1], stw, g0, g1, J, Jp, p);
      __builtin_amdgcn_sched_barrier(0);
    }
    dif_passes<LOGN, LOGN, 0, LOGN, 0>(x, lds, t, twf, p, id);
    const cp_gi *kb = key + ((size_t)(2 * i) * k + j) * N, *ka = key + ((size_t)(2 * i + 1) * k + j) * N;
    const int tk = cp_opaque(t);
#pragma unroll
    for (int r = 0; r < EPT; r++) {
      const int pos = pass_index<LOGN, LP::LS0, LP::RL>(tk, r);
      const cp_gi m0 = split_mul(x[r], kb[pos], pr), m1 = split_mul(x[r], ka[pos], pr);
      if (i == 0) {
        acc[r] = m0;
        if (active) park[r * TT + tk] = m1;
      } else {
        acc[r].re = cp_add(acc[r].re, m0.re, p);
        acc[r].imag = cp_add(acc[r].imag, m0.imag, p);
        const cp_gi q = park[r * TT + tk];
        if (active) park[r * TT + tk] = cp_gi{cp_add(q.re, m1.re, p), cp_add(q.imag, m1.imag, p)};
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // the output address may come from a pointer table: typed as global memory,
  // so its stores are global stores (a generic pointer's flat stores also
  // count against the LDS wait counter of the exchanges)
  typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
  typedef __attribute__((address_space(1))) u64x2 global_gi;
  global_gi *o = (global_gi *)out.at(pair);
#pragma unroll
  for (int m = 0; m < 2; m++) {
    if (m == 1) {
      const int tl = cp_opaque(t);
#pragma unroll
      for (int r = 0; r < EPT; r++) acc[r] = park[r * TT + tl];
    }
    gsi_passes<LOGN, LOGN, 0, LOGN, 0>(acc, lds, t, root, p, id);
    const int ts = cp_opaque(t);
    if (active) {
      const cp_gi *ym = y + ((size_t)m * k + j) * N;
      global_gi *om = o + ((size_t)m * k + j) * N;
#pragma unroll
      for (int r = 0; r < EPT; r++) {
        if (r % 2 == 0) split_back_h2(acc[r], acc[r + 1], J, Jp, p);
        const int g = first_index<LOGN, 0, LOGN>(ts, r);
        const cp_gi v = ld_gi_nt((const CP_GLOBAL cp_gi *)ym + g);
        __builtin_nontemporal_store(
            u64x2{cp_add(v.re, acc[r].re, p), cp_add(v.imag, acc[r].imag, p)}, om + g);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
}

template <int LOGN>
static int launch_relin_fti(cupoly_ctx *c, cp_rcts out, const cp_gi *T, const cp_gi *key,
                            cp_gi *park, int batch) {  // key: the split-basis copy
  constexpr int R = rows_for<LOGN>(), THREADS = ((1 << LOGN) / EPT) * R;
  const size_t lds = (size_t)R * lds_row<LOGN>() * sizeof(cp_gi);
  cupoly_engine *e = c->e;
  const int nunits = batch * e->k;
  hipLaunchKernelGGL((k_relin_fti<LOGN, R>), dim3((nunits + R - 1) / R), dim3(THREADS), lds,
                     c->stream, out, T, key, park, batch, e->d_hps, e->d_tw_fwd,
                     e->d_split_root, e->d_split_tw, e->d_split_k);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

template <int LOGN>
static int relin_setup_t() {
  constexpr int R = rows_for<LOGN>();
  const int lds = (int)((size_t)R * lds_row<LOGN>() * sizeof(cp_gi));
  CP_HIP(hipFuncSetAttribute((const void *)k_relin_fti<LOGN, R>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  return CUPOLY_OK;
}

int cp_relin_setup(cupoly_engine *e) {
  switch (e->logng) {
    case 10: return relin_setup_t<10>();
    case 11: return relin_setup_t<11>();
    case 12: return relin_setup_t<12>();
    case 13: return relin_setup_t<13>();
    default: return CUPOLY_OK;
  }
}

// The fused form once the batch has a (pair, residue) unit for every CU;
// smaller batches take the composed form, whose transforms spread each
// residue polynomial over several workgroups (k transforms in series in one
// workgroup would set the latency of a batch that leaves most CUs idle).
static bool relin_fused_sizes(const cupoly_engine *e) { return e->logng >= 10 && e->logng <= 13; }
static int relin_fused_min_batch(const cupoly_engine *e) { return (e->ncu + e->k - 1) / e->k; }
static bool relin_uses_fused(const cupoly_engine *e, int batch) {
  return relin_fused_sizes(e) && batch >= relin_fused_min_batch(e);
}

// Scratch of relin_run beyond its input for `batch` pairs in one form: the
// parked a-sums (k residue polynomials per pair) or the digits and both sums.
static size_t relin_ws_form(const cupoly_engine *e, int batch, bool fused) {
  const size_t polys = fused ? (size_t)e->k : (size_t)e->k * e->k + 2 * (size_t)e->k;
  return (size_t)batch * polys * e->ng * sizeof(cp_gi);
}

// Scratch that covers every batch of up to `batch` pairs.
size_t cp_relin_ws_bytes(const cupoly_engine *e, int batch) {
  if (!relin_fused_sizes(e)) return relin_ws_form(e, batch, false);
  const int b0 = relin_fused_min_batch(e);
  const size_t f = batch >= b0 ? relin_ws_form(e, batch, true) : 0;
  const size_t g = relin_ws_form(e, batch < b0 ? batch : b0 - 1, false);
  return f > g ? f : g;
}

// y = T [batch][3][k][ng] (coefficient domain, base Q) -> out[pair] = [2][k][ng];
// ws holds at least cp_relin_ws_bytes(e, batch).
static int relin_run(cupoly_ctx *c, const cupoly_relin_key *key, cp_rcts out, const cp_gi *T,
                     int batch, void *ws) {
  cupoly_engine *e = c->e;
  const int k = e->k, ng = e->ng;
  const dim3 sg((ng + 255) / 256);
  if (relin_uses_fused(e, batch)) {
    cp_gi *park = (cp_gi *)ws;
    switch (e->logng) {
      case 10: return launch_relin_fti<10>(c, out, T, key->ds, park, batch);
      case 11: return launch_relin_fti<11>(c, out, T, key->ds, park, batch);
      case 12: return launch_relin_fti<12>(c, out, T, key->ds, park, batch);
      default: return launch_relin_fti<13>(c, out, T, key->ds, park, batch);
    }
  }
  cp_gi *Dg = (cp_gi *)ws, *ACC = Dg + (size_t)batch * k * k * ng;
  hipLaunchKernelGGL(k_relin_digits, dim3(sg.x, batch * k), dim3(256), 0, c->stream, Dg, T, ng,
                     e->d_hps);
  CP_LAUNCH_CHECK();
  int rc = cp_dgt_launch(c, Dg, 0, k, batch * k, CUPOLY_FORWARD);
  if (rc) return rc;
  hipLaunchKernelGGL(k_relin_dot, dim3(sg.x, batch * k), dim3(256), 0, c->stream, ACC, Dg, key->d,
                     ng, e->d_hps);
  CP_LAUNCH_CHECK();
  if ((rc = cp_dgt_launch(c, ACC, 0, k, batch * 2, CUPOLY_INVERSE))) return rc;
  hipLaunchKernelGGL(k_relin_add, dim3(sg.x, batch * 2 * k), dim3(256), 0, c->stream, out, T, ACC,
                     ng, e->d_hps);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

// Gathered output table: the device pointer table of the hom-mul gather form
// holds out, ct1, ct2; relin_gather_out views its first `batch` entries.
int cp_hom_mul_relin_run(cupoly_ctx *c, const cupoly_relin_key *key, cupoly_gi *const *d_out_table,
                         cupoly_gi *out_base, const cp_gi *T, int batch, void *ws) {
  const size_t k = c->e->k, ng = c->e->ng;
  cp_rcts o{(cp_gi *)out_base, 2 * k * ng, (cp_gi *const *)d_out_table};
  return relin_run(c, key, o, (const cp_gi *)T, batch, ws);
}

const cupoly_engine *cp_relin_key_engine(const cupoly_relin_key *key) { return key->e; }

// ---- C ABI ---------------------------------------------------------------------
extern "C" int cupoly_relin_key_create(cupoly_ctx *c, const cupoly_gi *s, uint64_t seed,
                                       uint64_t stream0, cupoly_relin_key **out) {
  if (!out) return cp_fail(CUPOLY_ERR_ARG, "relin_key_create: null out");
  *out = nullptr;
  if (!c || !s) return cp_fail(CUPOLY_ERR_ARG, "relin_key_create: null argument");
  cupoly_engine *e = c->e;
  const int k = e->k, ng = e->ng;
  const size_t poly = (size_t)k * ng;  // one base-Q polynomial, in slots
  cupoly_relin_key *key = new cupoly_relin_key;
  key->e = e;
  cp_gi *tmp = nullptr;  // s^ [k][ng], s^2 [k][ng], e_i [k][k][ng]
  int rc = CUPOLY_OK;
  hipError_t err = hipMalloc((void **)&key->d, 2 * k * poly * sizeof(cp_gi));
  if (err == hipSuccess) err = hipMalloc((void **)&tmp, (2 + k) * poly * sizeof(cp_gi));
  if (err != hipSuccess) rc = cp_check_hip(err, "relin_key_create: hipMalloc");
  cp_gi *sh = tmp, *s2 = tmp + poly, *ee = tmp + 2 * poly;
  if (!rc) {
    err = hipMemcpyAsync(sh, s, poly * sizeof(cp_gi), hipMemcpyDeviceToDevice, c->stream);
    if (err != hipSuccess) rc = cp_check_hip(err, "relin_key_create: copy of s");
  }
  if (!rc) rc = cupoly_dgt(c, (cupoly_gi *)sh, CUPOLY_QBASE, CUPOLY_FORWARD);
  if (!rc) rc = cupoly_mul_dgt_gi(c, (cupoly_gi *)s2, (cupoly_gi *)sh, (cupoly_gi *)sh, CUPOLY_QBASE);
  for (int i = 0; i < k && !rc; i++) {
    rc = cupoly_sample(c, (cupoly_gi *)(key->d + (size_t)(2 * i + 1) * poly), CUPOLY_UNIFORM, seed,
                       stream0 + 2 * (uint64_t)i, CUPOLY_QBASE);
    if (!rc)
      rc = cupoly_sample(c, (cupoly_gi *)(ee + (size_t)i * poly), CUPOLY_DISCRETE_GAUSSIAN, seed,
                         stream0 + 2 * (uint64_t)i + 1, CUPOLY_QBASE);
  }
  if (!rc && relin_fused_sizes(e)) {
    err = hipMalloc((void **)&key->ds, 2 * k * poly * sizeof(cp_gi));
    if (err != hipSuccess) rc = cp_check_hip(err, "relin_key_create: hipMalloc");
  }
  if (!rc) {
    hipLaunchKernelGGL(k_relin_key_b, dim3((ng + 255) / 256, k * k), dim3(256), 0, c->stream,
                       key->d, ee, sh, s2, ng, e->d_hps);
    if (key->ds)
      hipLaunchKernelGGL(k_relin_key_split, dim3((ng + 255) / 256, 2 * k * k), dim3(256), 0,
                         c->stream, key->ds, key->d, ng, e->d_hps, e->d_split_k);
    err = hipGetLastError();
    if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
    if (err != hipSuccess) rc = cp_check_hip(err, "relin_key_create: key kernel");
  } else if (tmp) {
    (void)hipStreamSynchronize(c->stream);  // nothing may still use tmp when it is freed
  }
  if (tmp) (void)hipFree(tmp);
  if (rc) {
    if (key->d) (void)hipFree(key->d);
    if (key->ds) (void)hipFree(key->ds);
    delete key;
    return rc;
  }
  *out = key;
  return CUPOLY_OK;
}

extern "C" int cupoly_relin_key_destroy(cupoly_relin_key *key) {
  if (!key) return CUPOLY_OK;
  hipError_t err = key->d ? hipFree(key->d) : hipSuccess;
  const hipError_t err2 = key->ds ? hipFree(key->ds) : hipSuccess;
  if (err == hipSuccess) err = err2;
  delete key;
  return err == hipSuccess ? CUPOLY_OK : cp_check_hip(err, "relin_key_destroy");
}

extern "C" int cupoly_relin_key_data(const cupoly_relin_key *key, const cupoly_gi **data) {
  if (!key || !data) return cp_fail(CUPOLY_ERR_ARG, "relin_key_data: null argument");
  *data = (const cupoly_gi *)key->d;
  return CUPOLY_OK;
}

extern "C" int cupoly_relin_key_copy(cupoly_ctx *c, const cupoly_relin_key *key, cupoly_gi *dst) {
  if (!c || !key || !dst) return cp_fail(CUPOLY_ERR_ARG, "relin_key_copy: null argument");
  if (key->e != c->e) return cp_fail(CUPOLY_ERR_STATE, "relin_key_copy: key of another engine");
  const size_t bytes = (size_t)2 * key->e->k * key->e->k * key->e->ng * sizeof(cp_gi);
  CP_HIP(hipMemcpyAsync(dst, key->d, bytes, hipMemcpyDeviceToDevice, c->stream));
  return CUPOLY_OK;
}

extern "C" int cupoly_relinearize(cupoly_ctx *c, const cupoly_relin_key *key, cupoly_gi *out,
                                  const cupoly_gi *in, int batch) {
  if (c && key && batch == 0) return CUPOLY_OK;
  if (!c || !key || !out || !in || batch < 0)
    return cp_fail(CUPOLY_ERR_ARG, "relinearize: bad argument");
  if (key->e != c->e) return cp_fail(CUPOLY_ERR_STATE, "relinearize: key of another engine");
  if (batch > 65535 / 4) return cp_fail(CUPOLY_ERR_ARG, "relinearize: batch > 16383");
  if ((const void *)out == (const void *)in)
    return cp_fail(CUPOLY_ERR_ARG, "relinearize: out must not be in");
  void *ws;
  int rc = cp_workspace(c, cp_relin_ws_bytes(c->e, batch), &ws);
  if (rc) return rc;
  return cp_hom_mul_relin_run(c, key, nullptr, out, (const cp_gi *)in, batch, ws);
}

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
//   1. k_ext_q_to_qb:   each of the 4 inputs -> its B residues by the HPS
//                       extension, one thread per Gaussian slot (the Q rows
//                       are read from the input by the transforms, or copied
//                       for the generic sizes);
//   2. n = 2^11 .. 2^14, k_hom_fti: one workgroup per (pair, residue) runs
//                       the 4 forward transforms, the slot-wise tensor
//                       product and the 3 inverse transforms;
//      n = 2^15: forward transforms (k_dgt_fwd_tm16), then the tensor
//                       product fused into the inverse (k_tensor_idgt16);
//   3. k_scale_to_q:    complex scaling t/Q (QB -> B) and B -> Q fused in
//                       registers, one thread per slot, writing base Q.
// Scratch lives in a grow-only workspace owned by the context.
#include <mutex>
#include <string>

#include "dgt_dev.hpp"
#include "hps_dev.hpp"

// Where ciphertext i lives: base + i * stride, or table[i] (gathered
// requests of a serving batch).
struct cp_cts {
  cp_gi *base;
  size_t stride;
  cp_gi *const *table;
  __device__ __forceinline__ cp_gi *at(int i) const { return table ? table[i] : base + i * stride; }
};

// blockIdx.y = pair * 4 + which (0: c0, 1: c1, 2: d0, 3: d1).  X holds
// `xrows` residue polynomials per input; the B residues go to rows
// boff .. boff + kl - 1, the Q residues (copy_q) to rows 0 .. k - 1.
__global__ void __launch_bounds__(256)
    k_ext_q_to_qb(cp_gi *__restrict__ X, cp_cts ct1, cp_cts ct2, int ng,
                  const cp_hps *__restrict__ H, int xrows, int boff, bool copy_q) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int k = H->k;
  const int pair = blockIdx.y >> 2, which = blockIdx.y & 3;
  const cp_gi *in = (which < 2 ? ct1.at(pair) : ct2.at(pair)) + (size_t)(which & 1) * k * ng;
  cp_gi *out = X + (size_t)blockIdx.y * xrows * ng;
  if (copy_q)  // else the forward transform reads the Q rows from the input itself
    for (int j = 0; j < k; j++) out[(size_t)j * ng + c] = in[(size_t)j * ng + c];
  hps_q_to_b_slot(in, ng, out + (size_t)boff * ng, ng, c, H);
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

// Complex scaling t/Q (QB -> B) and B -> Q in registers, K = k known at
// compile time.  blockIdx.y = output polynomial index (pair * 3 + m).  The
// fraction products x_j omega_frac_j and their FP64 sum do not depend on the
// B residue i and are formed once; the B residues are produced one at a time
// and folded straight into the B -> Q sums (ascending i, as the reference's
// loops).  B -> Q keeps one exact 160-bit sum per output residue for K <= 4
// and one modular word for larger K (K 160-bit sums would stay live across
// the whole B-residue loop: 272 VGPRs at K = 8); both give the exact residue,
// the reference's chain of reduced terms.
template <int K>
__global__ void __launch_bounds__(256)
    k_scale_to_q(cp_cts out, const cp_gi *__restrict__ Y, int ng, const cp_hps *__restrict__ H,
                 bool chain) {
  constexpr bool kAcc = K <= 4;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int kl = H->kl;
  const int pair = blockIdx.y / 3, m = blockIdx.y % 3;
  const cp_gi *y = Y + (size_t)blockIdx.y * (K + kl) * ng;
  const cp_gi *yb = y + (size_t)K * ng;
  uint64_t xr[K], xi[K];
  uint64_t fr[K], fi[K];
  uint64_t bmin = H->b[0].p;
  for (int i = 1; i < kl; i++) bmin = H->b[i].p < bmin ? H->b[i].p : bmin;
  const double two_m64 = 0x1p-64;
  double acc_r = 0.0, acc_i = 0.0;
  bool slow = chain;
  cp_acc hr, hi;
  cp_acc_init(hr);
  cp_acc_init(hi);
#pragma unroll
  for (int j = 0; j < K; j++) {
    const cp_gi v = ld_gi_nt((const CP_GLOBAL cp_gi *)y + (size_t)j * ng + c);
    xr[j] = v.re;
    xi[j] = v.imag;
    uint64_t lr, li;
    umul128hi(H->wfrac[j][0], H->wfrac[j][1], xr[j], lr, fr[j]);
    umul128hi(H->wfrac[j][0], H->wfrac[j][1], xi[j], li, fi[j]);
    acc_r = fma((double)lr, two_m64, acc_r);  // __fma_rn, ascending j
    acc_i = fma((double)li, two_m64, acc_i);
    slow |= (fr[j] >= bmin) | (fi[j] >= bmin);
    hsum_add(hr, fr[j]);
    hsum_add(hi, fi[j]);
  }
  const uint64_t round_r = (uint64_t)llrint(acc_r), round_i = (uint64_t)llrint(acc_i);
  hsum_add(hr, round_r);
  hsum_add(hi, round_i);
  cp_acc ar_acc[kAcc ? K : 1], ai_acc[kAcc ? K : 1];
  uint64_t sr[kAcc ? 1 : K], si[kAcc ? 1 : K];
#pragma unroll
  for (int j = 0; j < K; j++) {
    if constexpr (kAcc) {
      cp_acc_init(ar_acc[j]);
      cp_acc_init(ai_acc[j]);
    } else {
      sr[j] = si[j] = 0;
    }
  }
  double vr = 0.0, vi = 0.0;
  for (int i = 0; i < kl; i++) {
    const uint64_t b = H->b[i].p;
    const cp_gi yv = ld_gi_nt((const CP_GLOBAL cp_gi *)yb + (size_t)i * ng + c);
    uint64_t rr, ri;
    if (!slow) {
      scale_exact2(hr, hi, xr, xi, K, H->wint[i], yv.re, yv.imag, H->lam[i], b, rr, ri);
    } else {  // the reference's chain, operation for operation
      rr = scale_chain_word(xr, fr, K, H->wint[i], round_r, yv.re, H->lam[i], b);
      ri = scale_chain_word(xi, fi, K, H->wint[i], round_i, yv.imag, H->lam[i], b);
    }
    // B -> Q, term i (cudaengine.cu:688-703)
    const uint64_t ar = shoup2(rr, H->invBbi[i], b), ai = shoup2(ri, H->invBbi[i], b);
    const double bd = (double)b;
    vr = vr + (double)ar / bd;
    vi = vi + (double)ai / bd;
#pragma unroll
    for (int j = 0; j < K; j++) {
      if constexpr (kAcc) {
        cp_acc_mac2(ar_acc[j], ai_acc[j], ar, ai, H->Bbiqi[j][i][0]);
      } else {
        const uint64_t q = H->q[j].p;
        sr[j] = addmod_ref(sr[j], shoup2(ar, H->Bbiqi[j][i], q), q);
        si[j] = addmod_ref(si[j], shoup2(ai, H->Bbiqi[j][i], q), q);
      }
    }
  }
  const uint64_t v_re = snap(vr), v_im = snap(vi);
  cp_gi *o = out.at(pair) + (size_t)m * K * ng;
#pragma unroll
  for (int j = 0; j < K; j++) {
    const uint64_t q = H->q[j].p;
    cp_gi r;
    if constexpr (kAcc) {
      const uint64_t bq = H->Bqi[j][0], nbq = bq ? q - bq : 0;  // -B mod q_j
      cp_acc_mac2(ar_acc[j], ai_acc[j], v_re, v_im, nbq);
      r.re = cp_acc_reduce(ar_acc[j], q);
      r.imag = cp_acc_reduce(ai_acc[j], q);
    } else {
      r.re = submod_ref(sr[j], shoup2(v_re, H->Bqi[j], q), q);
      r.imag = submod_ref(si[j], shoup2(v_im, H->Bqi[j], q), q);
    }
    st_gi_nt((CP_GLOBAL cp_gi *)o + (size_t)j * ng + c, r);
  }
}

// ---- n <= 2^14: every transform of one (pair, residue) in one workgroup ----
// The four forward transforms (c0, c1, d0, d1), the slot-wise tensor product
// and the three inverse transforms of residue r of one ciphertext pair run in
// one workgroup, so the spectra never make a full trip through HBM and each
// workgroup's load and store phases are spread over seven transforms.  The
// registers hold one spectrum beside the one being transformed, as in the
// fused product kernel; the two further spectra the product needs wait in
// the pair's own y1 and y2 output rows, addressed thread-major (slot i of
// thread t at i * T + t: coalesced, and read back only by the thread that
// wrote it):
//   S_c0 = F(c0)                      held
//   S_d0 = F(d0)         -> y1 park
//   y0 = I(S_c0 S_d0)
//   S_d1 = F(d1);  S_c0 S_d1 -> y2 park;  S_d1 held
//   S_c1 = F(c1)
//   y1 = I(S_c0 S_d1 + S_c1 S_d0),  y2 = I(S_c1 S_d1)
// The final y1 and y2 stores overwrite the parks in another layout; every
// thread has read its parks before the first workgroup barrier of the y1
// inverse (each inverse pass plan of 2^10 .. 2^13 slots has at least one), so
// no store can overtake a park read.  Workgroups run residue-major (all pairs
// of residue 0, then residue 1, ...), so consecutive workgroups share one
// residue's twiddle and twist tables in L2.
// Loads and stores of the seven-transform kernel, in the split basis of
// dgt_dev.hpp: the spectra never leave the workgroup except through its own
// parks, so the basis does not show in Y.
template <int LOGN>
__device__ __forceinline__ void fti_raw(cp_gi (&x)[EPT], const cp_gi *__restrict__ src, int t) {
  t = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i++)
    x[i] = ld_gi_nt((const CP_GLOBAL cp_gi *)src + first_index<LOGN, 0, LOGN>(t, i));
}
template <int LOGN>
__device__ __forceinline__ void fti_twist(cp_gi (&x)[EPT], int t,
                                          const CP_GLOBAL uint64_t *__restrict__ stw, uint64_t J,
                                          uint64_t Jp, uint64_t p) {
  t = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i += 2) {
    split_twist2(x[i], x[i + 1], stw, first_index<LOGN, 0, LOGN>(t, i),
                 first_index<LOGN, 0, LOGN>(t, i + 1), J, Jp, p);
    __builtin_amdgcn_sched_barrier(0);
  }
}
// (the d operands were twisted with the table that carries 1/2N, so every
// product row holds it: split_back_h2)
template <int LOGN>
__device__ __forceinline__ void fti_store(cp_gi *__restrict__ dst, cp_gi (&x)[EPT], int t,
                                          const uint64_t *__restrict__ sk, uint64_t p,
                                          bool active) {
  t = cp_opaque(t);
  const uint64_t J = sk[0], Jp = sk[1];
  if (active) {
#pragma unroll
    for (int i = 0; i < EPT; i += 2) {
      split_back_h2(x[i], x[i + 1], J, Jp, p);
      st_gi_nt((CP_GLOBAL cp_gi *)dst + first_index<LOGN, 0, LOGN>(t, i), x[i]);
      st_gi_nt((CP_GLOBAL cp_gi *)dst + first_index<LOGN, 0, LOGN>(t, i + 1), x[i + 1]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

template <int LOGN, int ROWS>
__global__ void __launch_bounds__(((1 << LOGN) / EPT) * ROWS, fused_min_waves<LOGN>())
    k_hom_fti(cp_gi *__restrict__ Y, const cp_gi *__restrict__ XB, int batch, int nres, int k,
              cp_cts ct1, cp_cts ct2, const cp_prime *__restrict__ primes,
              const uint64_t *__restrict__ twf_all, const uint64_t *__restrict__ root_all,
              const uint64_t *__restrict__ stw_all, const uint64_t *__restrict__ stwh_all,
              const uint64_t *__restrict__ sk_all) {
  constexpr int N = 1 << LOGN, T = N / EPT;
  extern __shared__ cp_gi lds_all[];
  const int row = ROWS == 1 ? 0 : (int)(threadIdx.x / T);
  const int t = ROWS == 1 ? (int)threadIdx.x : (int)(threadIdx.x % T);
  cp_gi *lds = lds_all + row * lds_row<LOGN>();
  const int nunits = batch * nres;
  int u = blockIdx.x * ROWS + row;
  const bool active = u < nunits;
  if (!active) u = nunits - 1;
  int rid = u / batch;
  if (ROWS == 1) rid = __builtin_amdgcn_readfirstlane(rid);
  const int pair = u - rid * batch;
  const cp_prime pr = primes[rid];
  const uint64_t p = pr.p;
  const uint64_t *twf = twf_all + (size_t)rid * N;
  const CP_GLOBAL uint64_t *root = (const CP_GLOBAL uint64_t *)(root_all + (size_t)rid * N * 4);
  const CP_GLOBAL uint64_t *stw = (const CP_GLOBAL uint64_t *)(stw_all + (size_t)rid * N * 4);
  const CP_GLOBAL uint64_t *stwh = (const CP_GLOBAL uint64_t *)(stwh_all + (size_t)rid * N * 4);
  const uint64_t *sk = sk_all + (size_t)rid * 8;
  const uint64_t J = sk[0], Jp = sk[1];
  // a Q residue comes straight from its ciphertext, a B residue from the
  // extension (XB = [pair][which][kl][N])
  const cp_gi *c0, *c1, *d0, *d1;
  if (rid < k) {
    c0 = ct1.at(pair) + (size_t)rid * N;
    c1 = c0 + (size_t)k * N;
    d0 = ct2.at(pair) + (size_t)rid * N;
    d1 = d0 + (size_t)k * N;
  } else {
    const size_t kn = (size_t)(nres - k) * N;
    c0 = XB + (size_t)pair * 4 * kn + (size_t)(rid - k) * N;
    c1 = c0 + kn;
    d0 = c1 + kn;
    d1 = d0 + kn;
  }
  cp_gi *y0 = Y + ((size_t)pair * 3 * nres + rid) * N;
  cp_gi *y1 = y0 + (size_t)nres * N, *y2 = y1 + (size_t)nres * N;
  if (ROWS == 1) {
    c0 = uniform_ptr(c0);
    c1 = uniform_ptr(c1);
    d0 = uniform_ptr(d0);
    d1 = uniform_ptr(d1);
    y0 = uniform_ptr(y0);
    y1 = uniform_ptr(y1);
    y2 = uniform_ptr(y2);
  }
  MapIdentity id;
  cp_gi xa[EPT], xb[EPT];
  // Operand loads overlap register passes: d0 goes into registers during
  // S_c0's last pass; d1 and c1 are copied into LDS (lds_copy_issue, each
  // wave into its own area) during the last pass of the transform before
  // theirs, once a barrier has freed the array.
  static_assert(T % 64 == 0, "whole waves per row");
  cp_gi *area = lds + __builtin_amdgcn_readfirstlane(t >> 6) * (EPT * 64);
  const int lane = t & 63;
  fti_raw<LOGN>(xa, c0, t);
  fti_twist<LOGN>(xa, t, stw, J, Jp, p);
  dif_passes<LOGN, LOGN, 0, LOGN, 0>(xa, lds, t, twf, p, id,  // S_c0, held
                                     [&] { fti_raw<LOGN>(xb, d0, t); });
  fti_twist<LOGN>(xb, t, stwh, J, Jp, p);
  dif_passes<LOGN, LOGN, 0, LOGN, 0>(xb, lds, t, twf, p, id);  // S_d0
  if (active) {
#pragma unroll
    for (int i = 0; i < EPT; i++) y1[i * T + cp_opaque(t)] = xb[i];
  }
#pragma unroll
  for (int i = 0; i < EPT; i++) xb[i] = split_mul(xa[i], xb[i], pr);
  gsi_passes<LOGN, LOGN, 0, LOGN, 0, true>(xb, lds, t, root, p, id, [&] {  // right after S_d0's passes
    __syncthreads();
    lds_copy_issue<LOGN>(area, (const CP_GLOBAL cp_gi *)d1, t);
  });
  fti_store<LOGN>(y0, xb, t, sk, p, active);
  lds_copy_wait();
  lds_copy_read(xb, area, lane);
  fti_twist<LOGN>(xb, t, stwh, J, Jp, p);
  // S_d1, held; the first exchange's barrier orders every wave's read of its
  // copy before the array is overwritten
  dif_passes<LOGN, LOGN, 0, LOGN, 0>(xb, lds, t, twf, p, id, [&] {
    __syncthreads();
    lds_copy_issue<LOGN>(area, (const CP_GLOBAL cp_gi *)c1, t);
  });
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const cp_gi w = split_mul(xa[i], xb[i], pr);
    if (active) y2[i * T + cp_opaque(t)] = w;
    __builtin_amdgcn_sched_barrier(0);
  }
  lds_copy_wait();
  lds_copy_read(xa, area, lane);
  fti_twist<LOGN>(xa, t, stw, J, Jp, p);
  dif_passes<LOGN, LOGN, 0, LOGN, 0>(xa, lds, t, twf, p, id);  // S_c1
  const int tp = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    // one slot's park loads in flight at a time (all 16 hoisted would spill)
    const cp_gi pd0 = y1[i * T + tp], pc0d1 = y2[i * T + tp];
    const cp_gi cc = xa[i];
    xb[i] = split_mul(cc, xb[i], pr);
    const cp_gi w = split_mul(cc, pd0, pr);
    xa[i].re = cp_add(w.re, pc0d1.re, p);
    xa[i].imag = cp_add(w.imag, pc0d1.imag, p);
    __builtin_amdgcn_sched_barrier(0);
  }
  gsi_passes<LOGN, LOGN, 0, LOGN, 0, true>(xa, lds, t, root, p, id);  // right after S_c1's passes
  fti_store<LOGN>(y1, xa, t, sk, p, active);
  gsi_passes<LOGN, LOGN, 0, LOGN, 0>(xb, lds, t, root, p, id);
  fti_store<LOGN>(y2, xb, t, sk, p, active);
}

// ---- small batches: one thread per (slot, output residue) ------------------
// A batch of one or two pairs gives the per-slot kernels above only a few
// waves per SIMD, each running a long serial chain over all residues.  These
// forms give each output residue its own thread; the per-slot values every
// residue needs (a_j = x_j (Q/q_j)^-1, the FP64 sums) are recomputed per
// thread by the same operations in the same order, so every residue is the
// same as in the per-slot kernels.

// Q -> B, output residue i = blockIdx.z (blockIdx.y = pair * 4 + which)
__global__ void __launch_bounds__(256)
    k_ext_q_to_b_wide(cp_gi *__restrict__ X, cp_cts ct1, cp_cts ct2, int ng,
                      const cp_hps *__restrict__ H, int xrows, int boff, bool copy_q) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int k = H->k, i = blockIdx.z;
  const int pair = blockIdx.y >> 2, which = blockIdx.y & 3;
  const cp_gi *in = (which < 2 ? ct1.at(pair) : ct2.at(pair)) + (size_t)(which & 1) * k * ng;
  cp_gi *out = X + (size_t)blockIdx.y * xrows * ng;
  if (copy_q && i == 0)
    for (int j = 0; j < k; j++) out[(size_t)j * ng + c] = in[(size_t)j * ng + c];
  uint64_t ar[CP_MAX_BASE], ai[CP_MAX_BASE];
  double vr = 0.0, vi = 0.0;
#pragma unroll
  for (int j = 0; j < CP_MAX_BASE; j++) {
    if (j < k) {
      const uint64_t q = H->q[j].p;
      const cp_gi x = in[(size_t)j * ng + c];
      ar[j] = shoup2(x.re, H->invQqi[j], q);
      ai[j] = shoup2(x.imag, H->invQqi[j], q);
    }
  }
#pragma unroll
  for (int j = 0; j < CP_MAX_BASE; j++)
    if (j < k) cp_fma_rd2((double)ar[j], (double)ai[j], H->rcp_q[j], vr, vi);
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
  out[(size_t)(boff + i) * ng + c] = cp_gi{cp_acc_reduce(sr, b), cp_acc_reduce(si, b)};
}

// t/Q scaling, B residue i = blockIdx.z of output polynomial blockIdx.y, and
// the first step of B -> Q: Z[pm][i][c] = y_i (B/b_i)^-1 mod b_i.
template <int K>
__global__ void __launch_bounds__(256)
    k_scale_b_wide(cp_gi *__restrict__ Z, const cp_gi *__restrict__ Y, int ng,
                   const cp_hps *__restrict__ H, bool chain) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int kl = H->kl, i = blockIdx.z, pm = blockIdx.y;
  const cp_gi *y = Y + (size_t)pm * (K + kl) * ng;
  const uint64_t b = H->b[i].p;
  uint64_t xr[K], xi[K];
  uint64_t fr[K], fi[K];
  const double two_m64 = 0x1p-64;
  double acc_r = 0.0, acc_i = 0.0;
  bool slow = chain;
  cp_acc hr, hi;
  cp_acc_init(hr);
  cp_acc_init(hi);
#pragma unroll
  for (int j = 0; j < K; j++) {
    const cp_gi v = y[(size_t)j * ng + c];
    xr[j] = v.re;
    xi[j] = v.imag;
    uint64_t lr, li;
    umul128hi(H->wfrac[j][0], H->wfrac[j][1], xr[j], lr, fr[j]);
    umul128hi(H->wfrac[j][0], H->wfrac[j][1], xi[j], li, fi[j]);
    acc_r = fma((double)lr, two_m64, acc_r);
    acc_i = fma((double)li, two_m64, acc_i);
    slow |= (fr[j] >= b) | (fi[j] >= b);
    hsum_add(hr, fr[j]);
    hsum_add(hi, fi[j]);
  }
  const uint64_t round_r = (uint64_t)llrint(acc_r), round_i = (uint64_t)llrint(acc_i);
  hsum_add(hr, round_r);
  hsum_add(hi, round_i);
  const cp_gi yv = y[(size_t)(K + i) * ng + c];
  uint64_t rr, ri;
  if (!slow) {  // same word as the reference's chain (scale_exact)
    scale_exact2(hr, hi, xr, xi, K, H->wint[i], yv.re, yv.imag, H->lam[i], b, rr, ri);
  } else {
    rr = scale_chain_word(xr, fr, K, H->wint[i], round_r, yv.re, H->lam[i], b);
    ri = scale_chain_word(xi, fi, K, H->wint[i], round_i, yv.imag, H->lam[i], b);
  }
  Z[((size_t)pm * kl + i) * ng + c] = cp_gi{shoup2(rr, H->invBbi[i], b), shoup2(ri, H->invBbi[i], b)};
}

// B -> Q, output residue j = blockIdx.z of output polynomial blockIdx.y
__global__ void __launch_bounds__(256)
    k_b_to_q_wide(cp_cts out, const cp_gi *__restrict__ Z, int ng, int k,
                  const cp_hps *__restrict__ H) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ng) return;
  const int kl = H->kl, j = blockIdx.z, pm = blockIdx.y;
  const int pair = pm / 3, m = pm % 3;
  const uint64_t q = H->q[j].p;
  const cp_gi *z = Z + (size_t)pm * kl * ng;
  double vr = 0.0, vi = 0.0;
  cp_acc sr, si;  // exact sums (every term reduced), reduced once
  cp_acc_init(sr);
  cp_acc_init(si);
  for (int i = 0; i < kl; i++) {  // ascending i, as the per-slot form
    const cp_gi a = z[(size_t)i * ng + c];
    const double bd = (double)H->b[i].p;
    vr = vr + (double)a.re / bd;
    vi = vi + (double)a.imag / bd;
    cp_acc_mac2(sr, si, a.re, a.imag, H->Bbiqi[j][i][0]);
  }
  const uint64_t v_re = snap(vr), v_im = snap(vi);
  const uint64_t bq = H->Bqi[j][0], nbq = bq ? q - bq : 0;  // -B mod q_j
  cp_acc_mac2(sr, si, v_re, v_im, nbq);
  cp_gi r;
  r.re = cp_acc_reduce(sr, q);
  r.imag = cp_acc_reduce(si, q);
  out.at(pair)[((size_t)m * k + j) * ng + c] = r;
}

// ---- n = 2^15: the two kernels above for 16384-slot residue polynomials --
// One 1024-thread workgroup per residue polynomial, as k_polymul_16k: stage 0
// pairs slot g with g + 8192 in registers, and the 8192-slot halves then take
// stages 1..13 in turn through the LDS.  Spectra are thread-major per half,
// X[unit][h 8192 + i 1024 + t].  While one half is in the passes the other
// waits in its own final place: the slots a thread stores to there are slots
// only that thread has read (first_index<13>(t, i) is t modulo 1024), and it
// read them before.
__global__ void __launch_bounds__(1024)
    k_dgt_fwd_tm16(cp_gi *__restrict__ X, int nunits, int nres, int k, cp_cts ct1, cp_cts ct2,
                   const cp_prime *__restrict__ primes, const uint64_t *__restrict__ twf_all,
                   const uint64_t *__restrict__ stw_all, const uint64_t *__restrict__ stwh_all,
                   const uint64_t *__restrict__ sk_all) {
  constexpr int N = 16384, H = 8192;
  extern __shared__ cp_gi lds[];
  const int t = (int)threadIdx.x;
  // residue-major workgroup order, as k_dgt_fwd_tm
  const int per_r = nunits / nres, u = (int)blockIdx.x;
  const int rid = __builtin_amdgcn_readfirstlane(u / per_r);
  const int unit = (u - rid * per_r) * nres + rid;
  const uint64_t p = primes[rid].p;
  const CP_GLOBAL uint64_t *twf = (const CP_GLOBAL uint64_t *)(twf_all + (size_t)rid * N);
  const int pw = unit / nres, pair = pw >> 2, which = pw & 3;  // 0 c0, 1 c1, 2 d0, 3 d1
  // the d operands take the twists that carry 1/2N (split_back_h2 at the end)
  const CP_GLOBAL uint64_t *stw =
      (const CP_GLOBAL uint64_t *)((which >= 2 ? stwh_all : stw_all) + (size_t)rid * N * 4);
  const uint64_t J = sk_all[(size_t)rid * 8], Jp = sk_all[(size_t)rid * 8 + 1];
  CP_GLOBAL cp_gi *poly = (CP_GLOBAL cp_gi *)(X + (size_t)unit * N);
  const CP_GLOBAL cp_gi *src = poly;
  if (rid < k) {  // a Q residue comes straight from its ciphertext
    src = (const CP_GLOBAL cp_gi *)((which < 2 ? ct1.at(pair) : ct2.at(pair)) +
                                     (size_t)((which & 1) * k + rid) * N);
  }
  cp_gi x[EPT], y[EPT];
  stage0_16k(x, y, src, t, twf, stw, J, Jp, p);
  park_st(poly, 1, t, y);
  dif_passes<14, 13, 0, 13, 1>(x, lds, t, twf, p, MapBlock{0});
  park_st(poly, 0, t, x);
  park_ld(poly, 1, t, x);
  dif_passes<14, 13, 0, 13, 1>(x, lds, t, twf, p, MapBlock{H});
  park_st(poly, 1, t, x);
}

__global__ void __launch_bounds__(1024)
    k_tensor_idgt16(cp_gi *__restrict__ Y, const cp_gi *__restrict__ X, int nunits, int nres,
                    const cp_prime *__restrict__ primes, const uint64_t *__restrict__ root_all,
                    const uint64_t *__restrict__ sk_all) {
  constexpr int N = 16384, H = 8192;
  extern __shared__ cp_gi lds[];
  const int t = (int)threadIdx.x;
  const int per_r = nunits / nres, u = (int)blockIdx.x;
  const int rid = __builtin_amdgcn_readfirstlane(u / per_r);
  const int pm = u - rid * per_r, unit = pm * nres + rid;  // (pair * 3 + m) * nres + r
  const int pair = pm / 3, m = pm % 3;
  const cp_prime pr = primes[rid];
  const uint64_t p = pr.p;
  const CP_GLOBAL uint64_t *root = (const CP_GLOBAL uint64_t *)(root_all + (size_t)rid * N * 4);
  const uint64_t *sk = sk_all + (size_t)rid * 8;
  const size_t per_pair = (size_t)nres * N;
  const CP_GLOBAL cp_gi *x = (const CP_GLOBAL cp_gi *)(X + (size_t)pair * 4 * per_pair + (size_t)rid * N);
  const CP_GLOBAL cp_gi *xa = x + (m == 2 ? per_pair : 0), *xb = x + (m == 0 ? 2 : 3) * per_pair;
  CP_GLOBAL cp_gi *y = (CP_GLOBAL cp_gi *)(Y + (size_t)unit * N);
  cp_gi v[EPT], w[EPT];
#pragma nounroll
  for (int h = 0; h < 2; h++) {
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      const int g = h * H + i * 1024 + t;  // thread-major
      v[i] = split_mul(ld_gi(xa + g), ld_gi(xb + g), pr);
      if (m == 1) {  // + c1 d0
        const cp_gi z = split_mul(ld_gi(x + per_pair + g), ld_gi(x + 2 * per_pair + g), pr);
        v[i].re = cp_add(v[i].re, z.re, p);
        v[i].imag = cp_add(v[i].imag, z.imag, p);
      }
    }
    gsi_passes<14, 13, 0, 13, 1>(v, lds, t, root, p, MapBlock{h * H});
    if (h == 0) park_st(y, 0, t, v);  // half 0 waits in its own final place
  }
#pragma unroll
  for (int i = 0; i < EPT; i++) w[i] = v[i];
  park_ld(y, 0, t, v);
  int to = t;
  asm volatile("" : "+v"(to));
  istage0_16k(y, v, w, to, root, sk, p);
}

template <int LOGN>
static int launch_fti(cupoly_ctx *c, cp_gi *Y, const cp_gi *XB, int batch, cp_cts ct1,
                      cp_cts ct2) {
  constexpr int R = rows_for<LOGN>(), THREADS = ((1 << LOGN) / EPT) * R;
  const size_t lds = (size_t)R * lds_row<LOGN>() * sizeof(cp_gi);
  cupoly_engine *e = c->e;
  const int nunits = batch * e->nres;
  hipLaunchKernelGGL((k_hom_fti<LOGN, R>), dim3((nunits + R - 1) / R), dim3(THREADS), lds,
                     c->stream, Y, XB, batch, e->nres, e->k, ct1, ct2, e->d_primes, e->d_tw_fwd,
                     e->d_split_root, e->d_split_tw, e->d_split_tw_h, e->d_split_k);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

static int fti(cupoly_ctx *c, cp_gi *Y, const cp_gi *XB, int batch, cp_cts ct1, cp_cts ct2) {
  switch (c->e->logng) {
    case 10: return launch_fti<10>(c, Y, XB, batch, ct1, ct2);
    case 11: return launch_fti<11>(c, Y, XB, batch, ct1, ct2);
    case 12: return launch_fti<12>(c, Y, XB, batch, ct1, ct2);
    case 13: return launch_fti<13>(c, Y, XB, batch, ct1, ct2);
    default: return cp_fail(CUPOLY_ERR_UNSUPPORTED, "hom_mul: ring degree");
  }
}

// Dynamic LDS limits of the pipeline's transform kernels, set once per engine
// on the engine's device (engine creation).
template <int LOGN>
static int homomul_setup_t() {
  constexpr int R = rows_for<LOGN>();
  const int lds = (int)((size_t)R * lds_row<LOGN>() * sizeof(cp_gi));
  CP_HIP(hipFuncSetAttribute((const void *)k_hom_fti<LOGN, R>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  return CUPOLY_OK;
}

// ---- scratch of the allocation-free entry point ------------------------------
// cupoly_hom_mul_batch_noalloc runs the pipeline with its scratch in this
// file-scope device array: CUPOLY_NOALLOC_SCRATCH_BYTES on every GPU the
// library runs on, held until the process ends.  In host code the name
// `scratch` denotes a host-side stand-in, not device memory, so host code
// never passes it to a kernel or a copy: the array's address on each device
// is looked up once, with that device current (engine creation), by
// hipGetSymbolAddress and kept in g_scratch_dev; every use goes through it.
static __device__ __align__(256) unsigned char scratch[CUPOLY_NOALLOC_SCRATCH_BYTES];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[kMaxDevices];

static int scratch_setup(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices)
    return cp_fail(CUPOLY_ERR_UNSUPPORTED, "hom_mul scratch: device index beyond 63");
  if (g_scratch_dev[device]) return CUPOLY_OK;
  void *ptr = nullptr;
  CP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return cp_fail(CUPOLY_ERR_HIP, "hom_mul scratch: no device address");
  g_scratch_dev[device] = ptr;
  return CUPOLY_OK;
}

static void *scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch_dev[device] : nullptr;
}

int cp_homomul_setup(cupoly_engine *e) {
  // runs with e's device current
  if (int rc = scratch_setup(e->device)) return rc;
  switch (e->logng) {
    case 10: return homomul_setup_t<10>();
    case 11: return homomul_setup_t<11>();
    case 12: return homomul_setup_t<12>();
    case 13: return homomul_setup_t<13>();
    case 14: {
      const int lds = (int)(lds_row<13>() * sizeof(cp_gi));
      CP_HIP(hipFuncSetAttribute((const void *)k_dgt_fwd_tm16,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, lds));
      CP_HIP(hipFuncSetAttribute((const void *)k_tensor_idgt16,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, lds));
      return CUPOLY_OK;
    }
    default: return CUPOLY_OK;  // other sizes use the generic transforms
  }
}

template <int K>
static void launch_scale(cupoly_ctx *c, cp_cts out, const cp_gi *Y, int batch) {
  const int ng = c->e->ng;
  hipLaunchKernelGGL(k_scale_to_q<K>, dim3((ng + 255) / 256, batch * 3), dim3(256), 0, c->stream,
                     out, Y, ng, c->e->d_hps, c->scale_chain);
}

template <int K>
static void launch_scale_wide(cupoly_ctx *c, cp_cts out, const cp_gi *Y, cp_gi *Z, int batch) {
  const int ng = c->e->ng, kl = c->e->kl;
  const dim3 g((ng + 255) / 256, batch * 3);
  hipLaunchKernelGGL(k_scale_b_wide<K>, dim3(g.x, g.y, kl), dim3(256), 0, c->stream, Z, Y, ng,
                     c->e->d_hps, c->scale_chain);
  hipLaunchKernelGGL(k_b_to_q_wide, dim3(g.x, g.y, K), dim3(256), 0, c->stream, out, Z, ng, K,
                     c->e->d_hps);
}

static int scale_to_q_wide(cupoly_ctx *c, cp_cts out, const cp_gi *Y, cp_gi *Z, int batch) {
  switch (c->e->k) {
    case 1: launch_scale_wide<1>(c, out, Y, Z, batch); break;
    case 2: launch_scale_wide<2>(c, out, Y, Z, batch); break;
    case 3: launch_scale_wide<3>(c, out, Y, Z, batch); break;
    case 4: launch_scale_wide<4>(c, out, Y, Z, batch); break;
    case 5: launch_scale_wide<5>(c, out, Y, Z, batch); break;
    case 6: launch_scale_wide<6>(c, out, Y, Z, batch); break;
    case 7: launch_scale_wide<7>(c, out, Y, Z, batch); break;
    case 8: launch_scale_wide<8>(c, out, Y, Z, batch); break;
    case 9: launch_scale_wide<9>(c, out, Y, Z, batch); break;
    case 10: launch_scale_wide<10>(c, out, Y, Z, batch); break;
    case 11: launch_scale_wide<11>(c, out, Y, Z, batch); break;
    case 12: launch_scale_wide<12>(c, out, Y, Z, batch); break;
    case 13: launch_scale_wide<13>(c, out, Y, Z, batch); break;
    case 14: launch_scale_wide<14>(c, out, Y, Z, batch); break;
    default: return cp_fail(CUPOLY_ERR_UNSUPPORTED, "hom_mul: k outside 1..14");
  }
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

static int scale_to_q(cupoly_ctx *c, cp_cts out, const cp_gi *Y, int batch) {
  switch (c->e->k) {
    case 1: launch_scale<1>(c, out, Y, batch); break;
    case 2: launch_scale<2>(c, out, Y, batch); break;
    case 3: launch_scale<3>(c, out, Y, batch); break;
    case 4: launch_scale<4>(c, out, Y, batch); break;
    case 5: launch_scale<5>(c, out, Y, batch); break;
    case 6: launch_scale<6>(c, out, Y, batch); break;
    case 7: launch_scale<7>(c, out, Y, batch); break;
    case 8: launch_scale<8>(c, out, Y, batch); break;
    case 9: launch_scale<9>(c, out, Y, batch); break;
    case 10: launch_scale<10>(c, out, Y, batch); break;
    case 11: launch_scale<11>(c, out, Y, batch); break;
    case 12: launch_scale<12>(c, out, Y, batch); break;
    case 13: launch_scale<13>(c, out, Y, batch); break;
    case 14: launch_scale<14>(c, out, Y, batch); break;
    default: return cp_fail(CUPOLY_ERR_UNSUPPORTED, "hom_mul: k outside 1..14");
  }
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
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

// Device pointer table and pinned staging ring of the gathered form.
static int grow_ptr_tables(cupoly_ctx *c, size_t need) {
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
  return CUPOLY_OK;
}

// (k + kl)-residue scratch of the pipeline: 4 extended inputs and 3 products per pair.
// n = 2^11 .. 2^14: one workgroup runs every transform of a (pair, residue)
// once the batch has a (pair, residue) unit for every CU.  Smaller batches
// take the composed transforms, which spread each residue polynomial over
// several workgroups: seven transforms in series in one workgroup would set
// the latency of a batch that leaves most CUs idle.
static bool hom_mul_fti_sizes(const cupoly_engine *e) { return e->logng >= 10 && e->logng <= 13; }
static int hom_mul_fti_min_batch(const cupoly_engine *e) {
  const int nres = e->k + e->kl;
  return (e->ncu + nres - 1) / nres;
}
static bool hom_mul_uses_fti(const cupoly_engine *e, int batch) {
  return hom_mul_fti_sizes(e) && batch >= hom_mul_fti_min_batch(e);
}

// Scratch of the pipeline for `batch` pairs in one launch form: the extended
// inputs X (4 polynomials of the B rows only when the fused transforms read
// the Q rows from the ciphertexts, of all k + kl rows otherwise) and the 3
// products Y (k + kl rows), N Gaussian slots each.
static size_t hom_mul_ws_form(const cupoly_engine *e, int batch, bool fti_form) {
  const size_t xrows = fti_form ? e->kl : e->k + e->kl;
  return (size_t)batch * (4 * xrows + 3 * (size_t)(e->k + e->kl)) * e->ng * sizeof(cp_gi);
}

// Scratch that covers every batch of up to `batch` pairs run as one pipeline
// (the two forms need different amounts per pair, so the figure is the
// larger over both ranges).
static size_t hom_mul_ws_single(const cupoly_engine *e, int batch) {
  if (!hom_mul_fti_sizes(e)) return hom_mul_ws_form(e, batch, false);
  const int b0 = hom_mul_fti_min_batch(e);
  const size_t f = batch >= b0 ? hom_mul_ws_form(e, batch, true) : 0;
  const size_t g = hom_mul_ws_form(e, batch < b0 ? batch : b0 - 1, false);
  return f > g ? f : g;
}

// The pipeline on the context's stream with scratch `ws` of at least
// hom_mul_ws_single(e, batch) bytes; allocates nothing.
static int hom_mul_run(cupoly_ctx *c, cp_cts out, cp_cts ct1, cp_cts ct2, int batch, void *ws) {
  cupoly_engine *e = c->e;
  const int k = e->k, kl = e->kl, nres = k + kl, ng = e->ng;
  const bool fti_sizes = hom_mul_uses_fti(e, batch);
  const int xrows = fti_sizes ? kl : nres;
  const size_t xs = (size_t)batch * 4 * xrows * ng;
  cp_gi *X = (cp_gi *)ws, *Y = X + xs;
  const dim3 sg((ng + 255) / 256);
  int rc;
  // n = 2^15 uses the one-workgroup-per-polynomial kernels once the batch has
  // forward residue polynomials for three quarters of the CUs; smaller
  // batches keep the split transforms, which put more workgroups on each
  // polynomial (measured at k = 8: 2 pairs split 0.24 ms against 0.26 ms,
  // 3 pairs 0.32 against 0.28 ms)
  const bool big16 = e->logng == 14 && 4 * batch * 4 * nres >= 3 * e->ncu;
  // the Q rows are copied into X only for the generic transforms; the fused
  // ones read them from the ciphertexts
  const bool copy_q = !(fti_sizes || big16);
  const int boff = fti_sizes ? 0 : k;
  // per-residue threads while the per-slot forms would leave the GPU mostly idle
  const bool wide = (long)batch * 3 * ng < c->wide_slots;
  if (wide)
    hipLaunchKernelGGL(k_ext_q_to_b_wide, dim3(sg.x, batch * 4, kl), dim3(256), 0, c->stream, X,
                       ct1, ct2, ng, e->d_hps, xrows, boff, copy_q);
  else
    hipLaunchKernelGGL(k_ext_q_to_qb, dim3(sg.x, batch * 4), dim3(256), 0, c->stream, X, ct1,
                       ct2, ng, e->d_hps, xrows, boff, copy_q);
  CP_LAUNCH_CHECK();
  if (fti_sizes) {
    if ((rc = fti(c, Y, X, batch, ct1, ct2))) return rc;
  } else if (big16) {
    const size_t lds = lds_row<13>() * sizeof(cp_gi);
    hipLaunchKernelGGL(k_dgt_fwd_tm16, dim3(batch * 4 * nres), dim3(1024), lds, c->stream, X,
                       batch * 4 * nres, nres, k, ct1, ct2, e->d_primes, e->d_tw_fwd,
                       e->d_split_tw, e->d_split_tw_h, e->d_split_k);
    CP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tensor_idgt16, dim3(batch * 3 * nres), dim3(1024), lds, c->stream, Y, X,
                       batch * 3 * nres, nres, e->d_primes, e->d_split_root, e->d_split_k);
    CP_LAUNCH_CHECK();
  } else {
    if ((rc = cp_dgt_launch(c, X, 0, nres, batch * 4, CUPOLY_FORWARD))) return rc;
    const long total = (long)batch * nres * ng;
    long g = (total + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_tensor, dim3((unsigned)g), dim3(256), 0, c->stream, Y, X, ng, nres,
                       total, e->d_primes);
    CP_LAUNCH_CHECK();
    if ((rc = cp_dgt_launch(c, Y, 0, nres, batch * 3, CUPOLY_INVERSE))) return rc;
  }
  // X is dead once the products are formed: it holds the B -> Q operands
  // (batch * 3 * kl <= batch * 4 * xrows residue polynomials)
  return wide ? scale_to_q_wide(c, out, Y, X, batch) : scale_to_q(c, out, Y, batch);
}

// The full multiplication (relin.hip relinearizes): the rescaled tensor
// T = [batch][3][k][n/2] words, then the scratch shared by the two steps.
static size_t hom_mul_relin_ws_single(const cupoly_engine *e, int batch) {
  const size_t t = (size_t)batch * 3 * e->k * e->ng * sizeof(cp_gi);
  const size_t h = hom_mul_ws_single(e, batch), r = cp_relin_ws_bytes(e, batch);
  return t + (h > r ? h : r);
}

// A batch whose halves each still have a (pair, residue) unit per CU runs as
// two pipelines on the context's two streams.  With one stream the last
// round of the one-workgroup-per-unit kernel leaves CUs idle (1152 units at
// n = 2^14, k = 4, 128 pairs are 4.5 rounds over 256 CUs) and the HPS steps
// wait for it; with two, one half's extension, scaling and relinearization
// fill the CUs the other half's transforms leave free (measured: 128 pairs
// 1.63 against 1.73 ms, 2.20 against 2.36 ms relinearized).
static bool hom_mul_splits(const cupoly_engine *e, int batch) {
  return hom_mul_fti_sizes(e) && batch >= 2 * hom_mul_fti_min_batch(e);
}

static size_t pipeline_ws(const cupoly_engine *e, int batch, bool relin) {
  return relin ? hom_mul_relin_ws_single(e, batch) : hom_mul_ws_single(e, batch);
}

// Scratch for every batch of up to `batch` pairs, split or not: the two
// halves' scratch areas lie one after the other.
static size_t hom_mul_cover_ws(const cupoly_engine *e, int batch, bool relin) {
  size_t w = pipeline_ws(e, batch, relin);
  if (hom_mul_splits(e, batch)) {
    const int a = (batch + 1) / 2;
    const size_t s = pipeline_ws(e, a, relin) + pipeline_ws(e, batch - a, relin);
    if (s > w) w = s;
  }
  return w;
}
static size_t hom_mul_ws_bytes(const cupoly_engine *e, int batch) {
  return hom_mul_cover_ws(e, batch, false);
}
static size_t hom_mul_relin_ws_bytes(const cupoly_engine *e, int batch) {
  return hom_mul_cover_ws(e, batch, true);
}

// Request i of a batch onwards.
static cp_cts cts_from(const cp_cts &x, int i) {
  return x.table ? cp_cts{nullptr, 0, x.table + i} : cp_cts{x.base + (size_t)i * x.stride, x.stride, nullptr};
}

// One pipeline on c's stream with scratch ws of pipeline_ws(e, batch, key):
// tensor product + rescale into out ([3][k][n/2] per pair), or with a key
// also relinearized (out [2][k][n/2] per pair; T and the scratch shared by
// the two steps in ws).
static int hom_mul_one(cupoly_ctx *c, const cupoly_relin_key *key, cp_cts out, cp_cts ct1,
                       cp_cts ct2, int batch, void *ws) {
  if (!key) return hom_mul_run(c, out, ct1, ct2, batch, ws);
  const size_t k = c->e->k, ng = c->e->ng;
  cp_gi *T = (cp_gi *)ws;
  void *shared = T + (size_t)batch * 3 * k * ng;
  cp_cts t{T, 3 * k * ng, nullptr};
  int rc = hom_mul_run(c, t, ct1, ct2, batch, shared);
  if (rc) return rc;
  return cp_hom_mul_relin_run(c, key, (cupoly_gi *const *)out.table, (cupoly_gi *)out.base, T,
                              batch, shared);
}

// The whole batch with scratch ws of hom_mul_cover_ws(e, batch, key):
// allocates nothing (a graph may capture it: the second stream joins the
// capture through the fork event and is joined back before the return).  The
// halves go to two streams only when the context already has its second
// stream (the allocating entry points make it); otherwise one pipeline runs.
static int hom_mul_pipeline(cupoly_ctx *c, const cupoly_relin_key *key, cp_cts out, cp_cts ct1,
                            cp_cts ct2, int batch, void *ws) {
  if (!c->aux || !hom_mul_splits(c->e, batch)) return hom_mul_one(c, key, out, ct1, ct2, batch, ws);
  const int a = (batch + 1) / 2, b = batch - a;
  cupoly_ctx *x = c->aux;
  x->wide_slots = c->wide_slots;
  x->scale_chain = c->scale_chain;
  CP_HIP(hipEventRecord(c->ev_fork, c->stream));
  CP_HIP(hipStreamWaitEvent(x->stream, c->ev_fork, 0));
  int rc = hom_mul_one(c, key, out, ct1, ct2, a, ws);
  if (!rc)
    rc = hom_mul_one(x, key, cts_from(out, a), cts_from(ct1, a), cts_from(ct2, a), b,
                     (char *)ws + pipeline_ws(c->e, a, key != nullptr));
  // joined back also after a failed launch, so the stream state stays whole
  const hipError_t e1 = hipEventRecord(c->ev_join, x->stream);
  const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(c->stream, c->ev_join, 0) : e1;
  if (rc) return rc;
  return e2 == hipSuccess ? CUPOLY_OK : cp_check_hip(e2, "hom_mul: join");
}

// The pipeline for gathered operands given as device pointer tables, with
// scratch ws of cp_hom_mul_ws_bytes(e, batch, key != nullptr) bytes; key =
// nullptr: tensor product + rescale into out[i] = [3][k][n/2], else also
// relinearized into out[i] = [2][k][n/2].  Allocates nothing (a graph may
// capture it).
int cp_hom_mul_tables_run(cupoly_ctx *c, const cupoly_relin_key *key, cupoly_gi *const *d_out,
                          cupoly_gi *const *d_ct1, cupoly_gi *const *d_ct2, int batch, void *ws) {
  cp_cts o{nullptr, 0, (cp_gi *const *)d_out}, a{nullptr, 0, (cp_gi *const *)d_ct1},
      b{nullptr, 0, (cp_gi *const *)d_ct2};
  return hom_mul_pipeline(c, key, o, a, b, batch, ws);
}

// A batch with a (pair, residue) unit for every CU keeps the device busy on
// its own: running it beside another such batch only shares the CUs.
bool cp_hom_mul_fills(const cupoly_engine *e, int batch) {
  return (long)batch * (e->k + e->kl) >= e->ncu;
}

// Everything a captured cp_hom_mul_tables_run of `batch` pairs needs besides
// its scratch: the second stream, when the batch splits.
int cp_hom_mul_prepare(cupoly_ctx *c, int batch) {
  return hom_mul_splits(c->e, batch) ? cp_ctx_aux(c) : CUPOLY_OK;
}

size_t cp_hom_mul_ws_bytes(const cupoly_engine *e, int batch, bool relin) {
  return relin ? hom_mul_relin_ws_bytes(e, batch) : hom_mul_ws_bytes(e, batch);
}

static int hom_mul_relin_alloc_run(cupoly_ctx *c, const cupoly_relin_key *key,
                                   cupoly_gi *const *d_out_table, cupoly_gi *out_base, cp_cts ct1,
                                   cp_cts ct2, int batch) {
  if (batch > 65535 / 4) return cp_fail(CUPOLY_ERR_ARG, "hom_mul: batch > 16383");
  void *ws;
  int rc = cp_workspace(c, hom_mul_relin_ws_bytes(c->e, batch), &ws);
  if (!rc && hom_mul_splits(c->e, batch)) rc = cp_ctx_aux(c);
  if (rc) return rc;
  const size_t k = c->e->k, ng = c->e->ng;
  cp_cts o{(cp_gi *)out_base, 2 * k * ng, (cp_gi *const *)d_out_table};
  return hom_mul_pipeline(c, key, o, ct1, ct2, batch, ws);
}

static int hom_mul_alloc_run(cupoly_ctx *c, cp_cts out, cp_cts ct1, cp_cts ct2, int batch) {
  if (batch > 65535 / 4) return cp_fail(CUPOLY_ERR_ARG, "hom_mul: batch > 16383");
  void *ws;
  int rc = cp_workspace(c, hom_mul_ws_bytes(c->e, batch), &ws);
  if (!rc && hom_mul_splits(c->e, batch)) rc = cp_ctx_aux(c);
  if (rc) return rc;
  return hom_mul_pipeline(c, nullptr, out, ct1, ct2, batch, ws);
}

extern "C" int cupoly_hom_mul_batch(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *ct1,
                                    const cupoly_gi *ct2, int batch) {
  if (c && batch == 0) return CUPOLY_OK;  // empty batch: pointers may be null
  if (!c || !out || !ct1 || !ct2 || batch < 0)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_batch: bad argument");
  const size_t k = c->e->k, ng = c->e->ng;
  cp_cts o{(cp_gi *)out, 3 * k * ng, nullptr}, a{(cp_gi *)ct1, 2 * k * ng, nullptr},
      b{(cp_gi *)ct2, 2 * k * ng, nullptr};
  return hom_mul_alloc_run(c, o, a, b, batch);
}

// Stages the gathered form's pointer tables (out, ct1, ct2) on the context's
// stream, from a pinned slot that is not rewritten until its previous copy
// has finished; *d receives the device table.
static int stage_ptr_tables(cupoly_ctx *c, cupoly_gi *const *out, const cupoly_gi *const *ct1,
                            const cupoly_gi *const *ct2, int batch, cp_gi ***d) {
  // the limit hom_mul_alloc_run enforces, checked before the tables grow
  if (batch > 65535 / 4) return cp_fail(CUPOLY_ERR_ARG, "hom_mul: batch > 16383");
  for (int i = 0; i < batch; i++)
    if (!out[i] || !ct1[i] || !ct2[i]) return cp_fail(CUPOLY_ERR_ARG, "hom_mul_gather: null entry");
  const size_t need = (size_t)batch * 3 * sizeof(void *);
  int rc = grow_ptr_tables(c, need);
  if (rc) return rc;
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
  *d = (cp_gi **)c->d_ptrs;
  return CUPOLY_OK;
}

extern "C" int cupoly_hom_mul_gather(cupoly_ctx *c, cupoly_gi *const *out,
                                     const cupoly_gi *const *ct1, const cupoly_gi *const *ct2,
                                     int batch) {
  if (c && batch == 0) return CUPOLY_OK;  // empty batch: tables may be null
  if (!c || !out || !ct1 || !ct2 || batch < 0)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_gather: bad argument");
  cp_gi **d;
  int rc = stage_ptr_tables(c, out, ct1, ct2, batch, &d);
  if (rc) return rc;
  cp_cts o{nullptr, 0, d}, a{nullptr, 0, d + batch}, b{nullptr, 0, d + 2 * batch};
  return hom_mul_alloc_run(c, o, a, b, batch);
}

extern "C" int cupoly_hom_mul_relin_batch(cupoly_ctx *c, const cupoly_relin_key *key,
                                          cupoly_gi *out, const cupoly_gi *ct1,
                                          const cupoly_gi *ct2, int batch) {
  if (c && key && batch == 0) return CUPOLY_OK;
  if (!c || !key || !out || !ct1 || !ct2 || batch < 0)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_relin_batch: bad argument");
  if (cp_relin_key_engine(key) != c->e)
    return cp_fail(CUPOLY_ERR_STATE, "hom_mul_relin_batch: key of another engine");
  const size_t k = c->e->k, ng = c->e->ng;
  cp_cts a{(cp_gi *)ct1, 2 * k * ng, nullptr}, b{(cp_gi *)ct2, 2 * k * ng, nullptr};
  return hom_mul_relin_alloc_run(c, key, nullptr, out, a, b, batch);
}

extern "C" int cupoly_hom_mul_relin_gather(cupoly_ctx *c, const cupoly_relin_key *key,
                                           cupoly_gi *const *out, const cupoly_gi *const *ct1,
                                           const cupoly_gi *const *ct2, int batch) {
  if (c && key && batch == 0) return CUPOLY_OK;
  if (!c || !key || !out || !ct1 || !ct2 || batch < 0)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_relin_gather: bad argument");
  if (cp_relin_key_engine(key) != c->e)
    return cp_fail(CUPOLY_ERR_STATE, "hom_mul_relin_gather: key of another engine");
  cp_gi **d;
  int rc = stage_ptr_tables(c, out, ct1, ct2, batch, &d);
  if (rc) return rc;
  cp_cts a{nullptr, 0, d + batch}, b{nullptr, 0, d + 2 * batch};
  return hom_mul_relin_alloc_run(c, key, (cupoly_gi *const *)d, nullptr, a, b, batch);
}

extern "C" int cupoly_hom_mul_reserve(cupoly_ctx *c, int batch) {
  if (!c || batch < 0 || batch > 65535 / 4)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_reserve: bad argument");
  if (batch == 0) return CUPOLY_OK;
  void *ws;  // room for the relinearizing form, the larger of the two
  int rc = cp_workspace(c, hom_mul_relin_ws_bytes(c->e, batch), &ws);
  if (!rc && hom_mul_splits(c->e, batch)) rc = cp_ctx_aux(c);
  if (rc) return rc;
  return grow_ptr_tables(c, (size_t)batch * 3 * sizeof(void *));
}

extern "C" int cupoly_hom_mul_relin_workspace_bytes(const cupoly_engine *e, int batch,
                                                    uint64_t *bytes) {
  if (!e || !bytes || batch < 0 || batch > 65535 / 4)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_relin_workspace_bytes: bad argument");
  *bytes = hom_mul_relin_ws_bytes(e, batch);
  return CUPOLY_OK;
}

extern "C" int cupoly_hom_mul_set_tuning(cupoly_ctx *c, long wide_slots, int scale_chain) {
  if (!c || wide_slots < 0 || scale_chain < 0 || scale_chain > 1)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_set_tuning: bad argument");
  c->wide_slots = wide_slots ? wide_slots : 1L << 16;
  c->scale_chain = scale_chain != 0;
  return CUPOLY_OK;
}

extern "C" int cupoly_hom_mul_workspace_bytes(const cupoly_engine *e, int batch, uint64_t *bytes) {
  if (!e || !bytes || batch < 0 || batch > 65535 / 4)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_workspace_bytes: bad argument");
  *bytes = hom_mul_ws_bytes(e, batch);
  return CUPOLY_OK;
}

extern "C" int cupoly_hom_mul_batch_noalloc(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *ct1,
                                            const cupoly_gi *ct2, int batch) {
  if (c && batch == 0) return CUPOLY_OK;  // empty batch: pointers may be null
  if (!c || !out || !ct1 || !ct2 || batch < 0)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_batch_noalloc: bad argument");
  if (batch > 65535 / 4) return cp_fail(CUPOLY_ERR_ARG, "hom_mul_batch_noalloc: batch > 16383");
  const size_t need = hom_mul_ws_bytes(c->e, batch);
  if (need > (size_t)CUPOLY_NOALLOC_SCRATCH_BYTES)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_batch_noalloc: " + std::to_string(batch) +
                                       " pairs need " + std::to_string(need) +
                                       " bytes of scratch; the maximum is " +
                                       std::to_string((size_t)CUPOLY_NOALLOC_SCRATCH_BYTES));
  void *ws = scratch_of(c->e->device);
  if (!ws) return cp_fail(CUPOLY_ERR_STATE, "hom_mul_batch_noalloc: no scratch on this device");
  const size_t k = c->e->k, ng = c->e->ng;
  cp_cts o{(cp_gi *)out, 3 * k * ng, nullptr}, a{(cp_gi *)ct1, 2 * k * ng, nullptr},
      b{(cp_gi *)ct2, 2 * k * ng, nullptr};
  return hom_mul_pipeline(c, nullptr, o, a, b, batch, ws);
}

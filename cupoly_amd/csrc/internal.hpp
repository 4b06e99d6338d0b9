// Internal declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/cupoly_amd.h"
#include "modarith.hpp"

// Pointers typed in the global address space: code reached through an
// out-of-line call (or pointers read from memory) would otherwise address
// tables with flat instructions.
#define CP_GLOBAL __attribute__((address_space(1)))
typedef unsigned long long cp_u64x2 __attribute__((ext_vector_type(2)));

// Operands read once and results written once by a kernel (as opposed to
// its tables and its own parked data, which are re-read): nontemporal
// accesses (the nt cache policy), which keep them from displacing the tables
// from the caches.
__device__ __forceinline__ cp_gi ld_gi_nt(const CP_GLOBAL cp_gi *ptr) {
  const cp_u64x2 v = __builtin_nontemporal_load((const CP_GLOBAL cp_u64x2 *)ptr);
  return cp_gi{v.x, v.y};
}
__device__ __forceinline__ void st_gi_nt(CP_GLOBAL cp_gi *ptr, const cp_gi &v) {
  __builtin_nontemporal_store(cp_u64x2{v.re, v.imag}, (CP_GLOBAL cp_u64x2 *)ptr);
}
#define CP_CPOL_NT 2  // cache-policy operand of the LDS-DMA builtin: nontemporal

#define CP_MAX_RES 30   // primes in the BFV bucket (settings.h:108-110)
#define CP_MAX_BASE 14  // residues per base: MAX_COPRIMES_IN_A_BASE - 1 (cudaengine.cu:1270-1276)

// HPS constants, one device copy per engine.  Shoup pairs are {w, w'}.
struct cp_hps {
  int k, kl;
  uint64_t t;
  cp_prime q[CP_MAX_BASE], b[CP_MAX_BASE];
  uint64_t invQqi[CP_MAX_BASE][2];               // (Q/q_j)^-1 mod q_j
  uint64_t Qqi[CP_MAX_BASE][2];                  // (Q/q_j) mod q_j
  uint64_t Qqibi[CP_MAX_BASE][CP_MAX_BASE][2];   // [i][j] (Q/q_j) mod b_i
  uint64_t Qbi[CP_MAX_BASE][2];                  // Q mod b_i
  double rcp_q[CP_MAX_BASE];                     // RD(1 / RN(q_j))
  uint64_t invBbi[CP_MAX_BASE][2];               // (B/b_i)^-1 mod b_i
  uint64_t Bbiqi[CP_MAX_BASE][CP_MAX_BASE][2];   // [j][i] (B/b_i) mod q_j
  uint64_t Bqi[CP_MAX_BASE][2];                  // B mod q_j
  uint64_t wint[CP_MAX_BASE][CP_MAX_BASE][2];    // [i][j] omega_int mod b_i
  uint64_t wfrac[CP_MAX_BASE][2];                // [j] {lo, hi} of omega_frac
  uint64_t lam[CP_MAX_BASE][2];                  // lambda_i
  uint64_t twint[CP_MAX_BASE];                   // simple scaling, integer part
  uint64_t twfrac[CP_MAX_BASE][2];               // simple scaling, fraction * 2^128
};

// CRT constants of base Q for reconstruction, and 2^64 mod p for every
// residue (multi-word reduction).  Multi-word values are little-endian
// 64-bit limbs.
struct cp_crt {
  int k, nres;
  uint64_t Q[CP_MAX_BASE + 1];                // Q (k limbs used, the rest 0)
  uint64_t M[CP_MAX_BASE][CP_MAX_BASE];       // [i] Q / q_i
  double rcp_q[CP_MAX_BASE];                  // RN(1 / q_i)
  uint64_t two64[2 * CP_MAX_BASE][2];         // [r] 2^64 mod p_r, Shoup pair
};

struct cupoly_engine {
  int k = 0, kl = 0, n = 0, ng = 0, logng = 0, nres = 0;
  uint64_t t = 0;
  int device = 0;
  int ncu = 256;  // compute units of `device`
  std::vector<uint64_t> primes;      // Q then B
  cp_prime *d_primes = nullptr;      // [nres]
  // DGT tables, per residue r (all of Q then B):
  uint64_t *d_tw_fwd = nullptr;      // [nres][ng/2][2]   omega^e, Shoup
  uint64_t *d_tw_inv = nullptr;      // [nres][ng/2][2]   omega^-e, Shoup
  uint64_t *d_twist_fwd = nullptr;   // [nres][ng][6]     xi^n as three Shoup pairs
  uint64_t *d_twist_inv = nullptr;   // [nres][ng][6]     xi^-n / ng
  // split-basis product tables (dgt_dev.hpp), per residue r:
  uint64_t *d_split_tw = nullptr;    // [nres][ng][4]     xi+^g, xi-^g (Shoup pairs)
  uint64_t *d_split_tw_h = nullptr;  // [nres][ng][4]     the same times 1/(2 ng)
  uint64_t *d_split_root = nullptr;  // [nres][ng][4]     r_u^-1, r_v^-1 at node 2^s + group
  uint64_t *d_split_k = nullptr;     // [nres][8]         J, 1/2N, 1/2JN (Shoup pairs)
  cp_hps h_hps;                      // host copy
  cp_hps *d_hps = nullptr;
  cp_crt h_crt;                      // host copy
  cp_crt *d_crt = nullptr;
};

struct cupoly_ctx {
  cupoly_engine *e = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  uint64_t *d_aux = nullptr;   // uint64 [kl][n]  (ctx->d_coefs_B in the reference)
  cp_gi *d_tmp = nullptr;      // [nres][ng]      (ctx->d_tmp_data / d_aux_coefs)
  cp_gi *d_tmp2 = nullptr;     // [nres][ng]
  void *d_ws = nullptr;        // grow-only workspace for batched pipelines
  size_t ws_bytes = 0;
  void *d_ptrs = nullptr;      // gathered-request pointer tables
  size_t ptr_bytes = 0;
  // pinned staging ring for the pointer tables: slot i is reused only after
  // the copy that read it has completed (ring_ev[i])
  static constexpr int kRing = 8;
  void *h_ring = nullptr;
  size_t ring_slot_bytes = 0;
  hipEvent_t ring_ev[kRing] = {};
  bool ring_used[kRing] = {};
  int ring_next = 0;
  // batched-product launch choices (cupoly_polymul_set_tuning / _autotune);
  // 0 = the built-in choice
  int tune_rows = 0;        // residue polynomials per workgroup, n = 2^11 .. 2^13
  int tune_persistent = 0;  // n = 2^15: 1 one workgroup per product, -1 two-kernel split
  // hom-mul: below this many output slots (batch * 3 * n/2) the HPS steps run
  // one thread per (slot, residue) instead of one per slot
  long wide_slots = 1L << 16;
  // hom-mul t/Q scaling: the reference's addmod chain on every slot instead of
  // the exact sums (which give the same words)
  bool scale_chain = false;
  // a second stream (default priority, bufferless context) and the events
  // that fork it from `stream` and join it back: a hom-mul batch that fills
  // the CUs twice over runs its two halves on the two streams, so one half's
  // HPS steps and the other's transforms share the CUs.  Made on first need
  // (cp_ctx_aux), by the calls that may allocate.
  cupoly_ctx *aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};
int cp_ctx_aux(cupoly_ctx *c);

// error plumbing
int cp_fail(int code, const std::string &msg);
int cp_check_hip(hipError_t err, const char *what);
#define CP_HIP(call)                                   \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return cp_check_hip(_e, #call); \
  } while (0)
#define CP_LAUNCH_CHECK() CP_HIP(hipGetLastError())

// base -> (first residue index, residue count); returns false on bad base
bool cp_base_range(const cupoly_engine *e, int base, int *off, int *cnt);

// kernel launchers implemented in the .hip files
int cp_dgt_launch(cupoly_ctx *c, cp_gi *data, int roff, int nres, int npoly, int direction);
int cp_polymul_fused_launch(cupoly_ctx *c, cp_gi *out, const cp_gi *a, const cp_gi *b, int batch);
int cp_dgt_setup(cupoly_engine *e);
int cp_homomul_setup(cupoly_engine *e);
int cp_encrypt_setup(cupoly_engine *e);
int cp_workspace(cupoly_ctx *c, size_t bytes, void **out);
int cp_relin_setup(cupoly_engine *e);
// relinearization (relin.hip): scratch beyond its input, and the run over
// T = [batch][3][k][n/2] coefficient-domain words into out (a device table of
// `batch` pointers, or out_base + i * 2 k n/2 when the table is null)
size_t cp_relin_ws_bytes(const cupoly_engine *e, int batch);
const cupoly_engine *cp_relin_key_engine(const cupoly_relin_key *key);
// hom-mul pipeline over device pointer tables (homomul.hip); allocates nothing
size_t cp_hom_mul_ws_bytes(const cupoly_engine *e, int batch, bool relin);
bool cp_hom_mul_fills(const cupoly_engine *e, int batch);
int cp_hom_mul_prepare(cupoly_ctx *c, int batch);
int cp_hom_mul_tables_run(cupoly_ctx *c, const cupoly_relin_key *key, cupoly_gi *const *d_out,
                          cupoly_gi *const *d_ct1, cupoly_gi *const *d_ct2, int batch, void *ws);
int cp_hom_mul_relin_run(cupoly_ctx *c, const cupoly_relin_key *key, cupoly_gi *const *d_out_table,
                         cupoly_gi *out_base, const cp_gi *T, int batch, void *ws);
int cp_pointwise_mul_launch(cupoly_ctx *c, cp_gi *out, const cp_gi *a, const cp_gi *b, int nres);

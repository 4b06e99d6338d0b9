/*
 * cupoly_amd.h - C ABI of the MI355X (gfx950) polynomial hot path.
 *
 * Every entry point replaces one interface of the reference library
 * (pdroalves/cupoly); the comment on each cites the reference declaration it
 * stands in for.  Plain pointers and sizes only: device buffers are HIP device
 * pointers, streams are hipStream_t passed as void*.  All functions return
 * CUPOLY_OK (0) or a negative error code; cupoly_last_error() gives the text.
 *
 * Data layout (same as the reference, include/cuPoly/settings.h:70-80 and
 * src/arithmetic/polynomial.cu:206-217): a polynomial of ring degree n over a
 * base with nres residues is cupoly_gi[nres][n/2]; coefficient j of residue r
 * is data[r*n/2 + j].re for j < n/2 and data[r*n/2 + j - n/2].imag otherwise.
 * The DGT domain stores X_k at position bitrev(k), exactly as the reference.
 */
#ifndef CUPOLY_AMD_H
#define CUPOLY_AMD_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint64_t re, imag; } cupoly_gi; /* GaussianInteger, settings.h:70-80 */

typedef struct cupoly_engine cupoly_engine;     /* CUDAEngine + DGTEngine state */
typedef struct cupoly_ctx cupoly_ctx;           /* Context (stream + scratch)   */
typedef struct cupoly_uneven cupoly_uneven;     /* plan for an uneven batch      */
typedef struct cupoly_relin_key cupoly_relin_key; /* resident relinearization key */
typedef struct cupoly_hom_mul_plan cupoly_hom_mul_plan; /* mixed-shape multiply batch  */

/* Error codes (the reference asserts / throws std::runtime_error / exit(1)). */
enum {
  CUPOLY_OK = 0,
  CUPOLY_ERR_ARG = -1,        /* null pointer, bad size or enum value          */
  CUPOLY_ERR_UNSUPPORTED = -2,/* unsupported n, base or t (runtime_error)      */
  CUPOLY_ERR_HIP = -3,        /* a HIP runtime call failed                     */
  CUPOLY_ERR_STATE = -4       /* engine not initialised / wrong object         */
};

/* Bases, include/cuPoly/settings.h:37 (enum bases). */
enum { CUPOLY_QBASE = 0, CUPOLY_BBASE = 1, CUPOLY_QBBASE = 2, CUPOLY_TBASE = 3 };
/* DGT directions, settings.h:41 (enum dgt_direction). */
enum { CUPOLY_FORWARD = 0, CUPOLY_INVERSE = 1 };
/* Scalar ops, settings.h:33 (enum add_mode_t). */
enum { CUPOLY_ADD = 0, CUPOLY_SUB = 1, CUPOLY_MUL = 2, CUPOLY_MULMUL = 3, CUPOLY_ADDADD = 4 };
/* Sampler kinds: the reference's enum pdistributions (sampler.h:32-41) values,
 * re-specified on Philox4x32-10, plus TERNARY (uniform {-1, 0, 1}).
 * HAMMINGWEIGHT draws exactly CUPOLY_HAMMINGWEIGHT_H nonzero +-1 coefficients
 * (settings.h:43), ZO exactly n/2. */
enum {
  CUPOLY_DISCRETE_GAUSSIAN = 0,
  CUPOLY_HAMMINGWEIGHT = 1,
  CUPOLY_NARROW = 2,
  CUPOLY_ZO = 3,
  CUPOLY_BINARY = 4,
  CUPOLY_UNIFORM = 5,
  CUPOLY_TERNARY = 6
};
#define CUPOLY_HAMMINGWEIGHT_H 64

const char *cupoly_last_error(void);
const char *cupoly_version(void);   /* tool/version.h */

/* ---- engine: CUDAEngine::init(k, kl, M, t) / destroy ---------------------
 * cudaengine.h:109-129.  n = ring degree (the reference's M / nphi); the
 * engine's slot count is n/2 (CUDAEngine::N).  Q = first k primes of the 63-bit
 * bucket, B = next kl (cudaengine.cu:1249-1289).  Unlike the reference, any
 * t >= 2 is accepted (t = 65537 included) and several engines may coexist. */
int cupoly_engine_create(int k, int kl, int n, uint64_t t, cupoly_engine **out);
int cupoly_engine_destroy(cupoly_engine *e);
/* CUDAEngine::get_n_residues(base), cudaengine.h:406. */
int cupoly_get_n_residues(const cupoly_engine *e, int base);
/* CUDAEngine::N (= n/2), CUDAEngine::RNSPrimes / RNSBPrimes / t. */
int cupoly_engine_info(const cupoly_engine *e, int *n, int *ng, int *k, int *kl, uint64_t *t);
int cupoly_engine_primes(const cupoly_engine *e, uint64_t *out /* k+kl */);
/* HPS constants as computed by the engine (CUDAEngine::precompute,
 * cudaengine.cu:1337-1511); for tests.  name: "invQqi" "Qqi" "Qqibi" "Qbi" "invBbi"
 * "Bbiqi" "Bqi" "lambda" "wint" "wfrac" "twint" "twfrac"; returns count. */
int cupoly_engine_constant(const cupoly_engine *e, const char *name, uint64_t *out, int cap);
/* The same constants for Q = first k bucket primes, B = next kl, computed on
 * the host only (no device, no engine); "rcp_q" gives the FP64 bits of
 * RD(1/RN(q_j)). */
int cupoly_hps_constant(int k, int kl, uint64_t t, const char *name, uint64_t *out, int cap);

/* Device self-test of the arithmetic primitives whose gfx950 forms are inline
 * asm (no reference counterpart: a diagnostic for tests).  in = n records of
 * 4 words (x, y, w, z); out = n records of 12 words: high products, Shoup
 * products, products of variables, 160-bit multiply-accumulate sums mod p and
 * round-down FMAs (csrc/selftest.hip lists them).  p must lie in
 * (2^63 - 2^30, 2^63).  Asynchronous on the context's stream. */
int cupoly_arith_selftest(cupoly_ctx *c, const uint64_t *in, uint64_t *out, int n, uint64_t p);

/* ---- context: Context() / ~Context(), context.cpp:19-67 ------------------
 * stream = 0 creates a private stream (as the reference); otherwise the
 * context uses the caller's hipStream_t. */
int cupoly_ctx_create(cupoly_engine *e, void *stream, cupoly_ctx **out);
int cupoly_ctx_destroy(cupoly_ctx *c);
void *cupoly_ctx_stream(const cupoly_ctx *c);   /* Context::get_stream() */
int cupoly_ctx_sync(const cupoly_ctx *c);

/* ---- DGT: DGTEngine::execute_dgt(data, base, direction, ctx), dgt.h:141 --
 * In place over all residues of `base`. */
int cupoly_dgt(cupoly_ctx *c, cupoly_gi *data, int base, int direction);

/* ---- pointwise (DGT domain), dgt.h:158-281 --------------------------------
 * base selects the residue range (Q, B or QB); all on the context's stream. */
int cupoly_add_dgt(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, const cupoly_gi *b, int base);
int cupoly_sub_dgt(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, const cupoly_gi *b, int base);
int cupoly_mul_dgt_gi(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, const cupoly_gi *b, int base);
/* execute_mul_dgt_u64: b holds one uint64 per residue slot ([nres][n/2]). */
int cupoly_mul_dgt_u64(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, const uint64_t *b, int base);
int cupoly_muladd_dgt(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, const cupoly_gi *b,
                      const cupoly_gi *add, int base);
int cupoly_double_add_dgt(cupoly_ctx *c, cupoly_gi *c1, const cupoly_gi *a1, const cupoly_gi *b1,
                          cupoly_gi *c2, const cupoly_gi *a2, const cupoly_gi *b2, int base);
/* execute_dr2_dgt: ct21 = ct01*ct11, ct22 = ct01*ct12 + ct02*ct11, ct23 = ct02*ct12. */
int cupoly_dr2_dgt(cupoly_ctx *c, cupoly_gi *ct21, cupoly_gi *ct22, cupoly_gi *ct23,
                   const cupoly_gi *ct01, const cupoly_gi *ct02, const cupoly_gi *ct11,
                   const cupoly_gi *ct12, int base);

/* ---- scalar ops, cudaengine.h:159-211 ------------------------------------
 * execute_polynomial_negate; op_by_int (ADD/SUB/MUL by x, x reduced per
 * residue); double_op_by_int (MULMUL: b1 = a1*x1, b2 = a2*x2; ADDADD). */
int cupoly_negate(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, int base);
int cupoly_op_by_int(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, uint64_t x, int op, int base);
int cupoly_double_op_by_int(cupoly_ctx *c, cupoly_gi *b1, const cupoly_gi *a1, uint64_t x1,
                            cupoly_gi *b2, const cupoly_gi *a2, uint64_t x2, int op, int base);

/* ---- HPS, cudaengine.h:213-296 (coefficient-domain inputs) --------------- */
/* execute_polynomial_basis_ext_Q_to_B(a, ctx): in place, Q residues in, B out
 * (buffer sized for kl residues). */
int cupoly_basis_ext_q_to_b(cupoly_ctx *c, cupoly_gi *a);
/* execute_polynomial_complex_scaling(b, a_Q, a_B, ctx): b = uint64[kl][n]. */
int cupoly_complex_scaling(cupoly_ctx *c, uint64_t *b, const cupoly_gi *a_q, const cupoly_gi *a_b);
/* execute_polynomial_basis_ext_B_to_Q(b, a, ctx): a = uint64[kl][n] -> b Q. */
int cupoly_basis_ext_b_to_q(cupoly_ctx *c, cupoly_gi *b, const uint64_t *a);
/* execute_polynomial_simple_scaling(a, ctx): Q -> T in place, round(t x / Q) mod t.
 * t a power of two (the reference's masked sums) or any t < 2^32 (exact
 * reductions mod t; the reference rejects such t, known-issues.md item 1). */
int cupoly_simple_scaling(cupoly_ctx *c, cupoly_gi *a);

/* ---- poly_t-level compositions (polynomial.cu) ----------------------------
 * DGT-domain inputs and outputs, like the reference's poly_* wrappers.       */
/* poly_basis_extension_Q_to_B (polynomial.cu:407-457): out_b = B part. */
int cupoly_poly_basis_ext_q_to_b(cupoly_ctx *c, cupoly_gi *out_b, const cupoly_gi *a_q);
/* poly_complex_scaling_tDivQ (polynomial.cu:350-404); inputs are not
 * destroyed (the reference transforms them in place). */
int cupoly_poly_complex_scaling(cupoly_ctx *c, cupoly_gi *out_q, const cupoly_gi *a_q,
                                const cupoly_gi *a_b);

/* ---- Sampler::sample(ctx, p, kind), sampler.h:77-87 ------------------------
 * Counter-based Philox4x32-10: the value of coefficient j of `stream` under
 * `seed` is fixed, independent of launch shape.  Output in the DGT domain of
 * `base` (as the reference: sample, then forward DGT). */
int cupoly_sample(cupoly_ctx *c, cupoly_gi *out, int kind, uint64_t seed, uint64_t stream, int base);
/* Same draws in the coefficient domain (no DGT). */
int cupoly_sample_coeff(cupoly_ctx *c, cupoly_gi *out, int kind, uint64_t seed, uint64_t stream,
                        int base);

/* ---- coefficient edges: poly_crt / poly_icrt on the device ------------------
 * poly_crt (polynomial.cu:190-219): unsigned multi-word coefficients,
 * little-endian 64-bit limbs, coefs[ncoefs][limbs] (ncoefs <= n; missing high
 * coefficients are 0) -> residues of `base` (Q, B or QB) in the coefficient-
 * domain layout [nres][n/2] GIs (coefficient j in re of slot j, j + n/2 in
 * imag).  poly_icrt (polynomial.cu:221-291) for base Q: residues [k][n/2] GIs
 * -> x mod Q in [0, Q) as out[n][k] limbs. */
int cupoly_crt(cupoly_ctx *c, cupoly_gi *out, const uint64_t *coefs, int ncoefs, int limbs,
               int base);
int cupoly_icrt(cupoly_ctx *c, uint64_t *out, const cupoly_gi *in);

/* ---- BFV word decomposition and dot product ------------------------------
 * poly_xi_bfv -> CUDAEngine::execute_xi_rns (cudaengine.cu:992-1035):
 * residue i of out[i] = a_i (Q/q_i)^-1 mod q_i for i < k (base Q input).
 * poly_rho_bfv -> execute_rho_bfv_rns (cudaengine.cu:1037-1062): residue i of
 * out[i] = a_i (Q/q_i) mod q_i for i < k; a in base Q or QB (B residues are
 * not used).  Only residue i of each out[i] is written, as in the reference.
 * out is a host array of k device pointers to polynomials of the input's base. */
int cupoly_xi_bfv(cupoly_ctx *c, cupoly_gi *const *out, const cupoly_gi *a);
int cupoly_rho_bfv(cupoly_ctx *c, cupoly_gi *const *out, const cupoly_gi *a, int base);
/* poly_dot (polynomial.cu:1439-1451): out = sum_{i < count} a[i] * b[i],
 * slot-wise in the DGT domain of `base`; a, b host arrays of device pointers. */
int cupoly_dot(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *const *a, const cupoly_gi *const *b,
               int count, int base);

/* ---- batched products (no reference counterpart: the reference issues one
 * poly_mul per product, 7 launches each, polynomial.cu:892-914) ------------
 * out[i] = a[i] * b[i] mod (X^n + 1, q_r) for `batch` products of the engine's
 * base-Q shape, coefficient domain in and out, [batch][k][n/2] GIs; out may be
 * a or b.  One fused launch: forward DGT of a and b, pointwise product,
 * inverse DGT.  batch * k must stay below 2^31 (CUPOLY_ERR_ARG otherwise). */
int cupoly_polymul_batch(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, const cupoly_gi *b,
                         int batch);
/* Launch choices of cupoly_polymul_batch on this context (no reference
 * counterpart): `rows` residue polynomials per workgroup for n = 2^11 .. 2^13
 * (1, 2, 4 or 8 with rows * n/2 <= 8192; 0 = built-in), and `persistent`:
 *  - n = 2^11 .. 2^14 with one residue polynomial per workgroup: 1 = the
 *    persistent streaming form (a grid the GPU holds at once; operands copied
 *    into LDS while the previous transform ends), -1 = one workgroup per
 *    residue product; 0 = streaming at n = 2^14 when the batch has more
 *    residue products than the GPU has CUs, one-shot otherwise;
 *  - n = 2^15: 1 = one workgroup per residue product, persistent, -1 =
 *    two-kernel column/block split; 0 = persistent when the batch has at
 *    least one residue product per CU.
 * Other ring degrees take persistent = 0 only (CUPOLY_ERR_ARG otherwise). */
int cupoly_polymul_set_tuning(cupoly_ctx *c, int rows, int persistent);
/* Autotuner for the batched product: times every available (rows,
 * persistent) choice for this ring degree over `reps` launches on the
 * caller's operands (out must not be a or b), keeps the fastest on the
 * context, reports it in *rows / *persistent (either may be NULL) and leaves
 * the product in out. */
int cupoly_polymul_autotune(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, const cupoly_gi *b,
                            int batch, int reps, int *rows, int *persistent);
/* The same products one launch per op, in the reference's order
 * (fwd a, fwd b, mul, inv) through cupoly_dgt / cupoly_mul_dgt_gi; a and b are
 * left in the DGT domain afterwards. */
int cupoly_polymul_plain(cupoly_ctx *c, cupoly_gi *out, cupoly_gi *a, cupoly_gi *b, int batch);

/* Batched BFV multiplication core (no single reference entry point: it is the
 * composition poly_basis_extension_Q_to_B x4 -> poly_dr2 -> inverse DGT ->
 * poly_complex_scaling_tDivQ x3, polynomial.cu:350-457, 951-999).  For each
 * pair i: ct1[i] = (c0, c1), ct2[i] = (d0, d1), [batch][2][k][n/2] GIs, base Q,
 * coefficient domain; out[i] = [3][k][n/2]: round(t/Q (c0 d0, c0 d1 + c1 d0,
 * c1 d1)) mod Q, coefficient domain.  Uses a context-owned workspace of
 * cupoly_hom_mul_workspace_bytes(e, batch) bytes (allocated on first use,
 * grown as needed).  batch <= 16383 (CUPOLY_ERR_ARG otherwise). */
int cupoly_hom_mul_batch(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *ct1, const cupoly_gi *ct2,
                         int batch);
/* Same computation for gathered operands (a serving batch of requests whose
 * buffers are not contiguous): host arrays of `batch` device pointers, out[i]
 * -> [3][k][n/2], ct1[i], ct2[i] -> [2][k][n/2].  The pointer tables are copied
 * to the device on the context's stream before the call returns. */
int cupoly_hom_mul_gather(cupoly_ctx *c, cupoly_gi *const *out, const cupoly_gi *const *ct1,
                          const cupoly_gi *const *ct2, int batch);
/* Grows the context's hom-mul workspace and gathered-form pointer tables to
 * hold `batch` pairs, so later hom_mul_batch / hom_mul_gather calls (and
 * their relinearizing forms) of up to that many pairs on this context
 * allocate nothing (no reference counterpart:
 * the reference allocates its Context scratch once, context.cpp:19-67; a
 * serving loop calls this before it starts timing requests). */
int cupoly_hom_mul_reserve(cupoly_ctx *c, int batch);
/* cupoly_hom_mul_batch with no allocation of any kind (no reference
 * counterpart: it serves callers that cannot allocate device memory, such as
 * code under HIP graph capture or a tight serving loop).  The pipeline's
 * scratch is a fixed device-resident area of CUPOLY_NOALLOC_SCRATCH_BYTES
 * (128 MiB) per GPU, held from the first engine created on that GPU until the
 * process ends; batches whose scratch (cupoly_hom_mul_workspace_bytes) exceeds
 * it fail with CUPOLY_ERR_ARG (at n = 2^14, k = 4: 16 pairs).  Every call on
 * one GPU shares that area: calls that overlap on one GPU (on different
 * streams, or from different threads) race and must be serialised by the
 * caller.  Same results as cupoly_hom_mul_batch. */
#define CUPOLY_NOALLOC_SCRATCH_BYTES (128u << 20)
int cupoly_hom_mul_batch_noalloc(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *ct1,
                                 const cupoly_gi *ct2, int batch);
/* Launch choices of the hom-mul pipeline on this context (no reference
 * counterpart): below `wide_slots` output slots (batch * 3 * n/2; 0 = the
 * built-in 65536) the HPS steps run one thread per (slot, residue);
 * scale_chain = 1 computes every t/Q residue (here and in
 * cupoly_complex_scaling on this context) with the reference's addmod chain
 * (cudaengine.cu:401-433) instead of the exact sums, which give the same
 * words (the default, 0, takes the chain only where the reference's operands
 * arrive unreduced). */
int cupoly_hom_mul_set_tuning(cupoly_ctx *c, long wide_slots, int scale_chain);
/* Device bytes of the hom-mul workspace for `batch` pairs of e's shape (no
 * reference counterpart; the reference sizes its Context scratch from the
 * ring shape, context.cpp:19-67), enough for every batch of up to `batch`
 * pairs: per pair 16 * n/2 * (4 kl + 3 (k + kl)) bytes in the fused form
 * (n = 2^11 .. 2^14, batches of at least ceil(CUs / (k + kl)) pairs) and
 * 16 * n/2 * 7 (k + kl) bytes otherwise. */
int cupoly_hom_mul_workspace_bytes(const cupoly_engine *e, int batch, uint64_t *bytes);

/* ---- relinearization (no single reference entry point: the reference
 * supplies its pieces, poly_xi_bfv / poly_rho_bfv, polynomial.cu:610-693 ->
 * cudaengine.cu:992-1062, and poly_dot, polynomial.cu:1439-1451, and leaves
 * key switching to the scheme built on it) ----------------------------------
 * Key for secret s (base Q, coefficient domain, [k][n/2] GIs holding the
 * residues of small integers): rlk_i = (b_i, a_i), i < k, a_i uniform mod Q
 * (Philox stream stream0 + 2i under seed), e_i discrete Gaussian (stream
 * stream0 + 2i + 1), b_i = -(a_i s + e_i) + s^2 (Q/q_i) mod Q; kept on the
 * device in the DGT domain, [k][2][k][n/2] GIs ([i][0] = b_i, [i][1] = a_i),
 * until cupoly_relin_key_destroy.  Creation synchronises the stream. */
int cupoly_relin_key_create(cupoly_ctx *c, const cupoly_gi *s, uint64_t seed, uint64_t stream0,
                            cupoly_relin_key **out);
int cupoly_relin_key_destroy(cupoly_relin_key *key);
/* Device address of the key's [k][2][k][n/2] GIs (read-only). */
int cupoly_relin_key_data(const cupoly_relin_key *key, const cupoly_gi **data);
/* Copies the key's words to dst (device, [k][2][k][n/2] GIs) on c's stream. */
int cupoly_relin_key_copy(cupoly_ctx *c, const cupoly_relin_key *key, cupoly_gi *dst);
/* in[i] = (y0, y1, y2), [batch][3][k][n/2] GIs, base Q, coefficient domain ->
 * out[i] = (y0 + sum_j D_j b_j, y1 + sum_j D_j a_j) mod Q, [batch][2][k][n/2],
 * with the RNS digits D_j = [y2_j (Q/q_j)^-1]_{q_j} (the reference's xi)
 * written to every residue; out must not be in.  Then
 * out0 + out1 s = y0 + y1 s + y2 s^2 - sum_j D_j e_j (mod Q). */
int cupoly_relinearize(cupoly_ctx *c, const cupoly_relin_key *key, cupoly_gi *out,
                       const cupoly_gi *in, int batch);
/* The full BFV multiplication: cupoly_hom_mul_batch followed by
 * cupoly_relinearize; out[i] = [2][k][n/2].  Workspace:
 * cupoly_hom_mul_relin_workspace_bytes (cupoly_hom_mul_reserve reserves it). */
int cupoly_hom_mul_relin_batch(cupoly_ctx *c, const cupoly_relin_key *key, cupoly_gi *out,
                               const cupoly_gi *ct1, const cupoly_gi *ct2, int batch);
/* Gathered form (serving): out[i] -> [2][k][n/2], ct1[i], ct2[i] -> [2][k][n/2]. */
int cupoly_hom_mul_relin_gather(cupoly_ctx *c, const cupoly_relin_key *key,
                                cupoly_gi *const *out, const cupoly_gi *const *ct1,
                                const cupoly_gi *const *ct2, int batch);
int cupoly_hom_mul_relin_workspace_bytes(const cupoly_engine *e, int batch, uint64_t *bytes);

/* Mixed-shape multiplication batches (no reference counterpart: the
 * reference holds one ring shape per process, cudaengine.h:71-93).  count
 * requests; request i has ring shape shape[i] in [0, n_shapes), the shape of
 * engines[shape[i]], operands ct1[i], ct2[i] ([2][k][n/2] GIs of that shape)
 * and output out[i]: [2][k][n/2] when keys[shape[i]] is a relinearization
 * key of that engine (tensor product, HPS rescale, relinearization), [3][k][n/2]
 * when keys is NULL or keys[shape[i]] is NULL (tensor product and rescale).
 * Every shape's pipeline is captured as one branch of a single hipGraph, with
 * its own context, stream, workspace and device pointer tables allocated
 * here; the buffers must stay valid for the plan's lifetime, and each run
 * reads the operands they hold at that time.  Creation synchronises c's
 * stream. */
int cupoly_hom_mul_plan_create(cupoly_ctx *c, cupoly_engine *const *engines,
                               const cupoly_relin_key *const *keys, int n_shapes,
                               const int *shape, cupoly_gi *const *out,
                               const cupoly_gi *const *ct1, const cupoly_gi *const *ct2,
                               int count, cupoly_hom_mul_plan **plan);
/* One hipGraphLaunch of the plan on c's stream. */
int cupoly_hom_mul_plan_run(cupoly_hom_mul_plan *plan);
int cupoly_hom_mul_plan_destroy(cupoly_hom_mul_plan *plan);

/* Encryption-shaped step for a batch: out[i] = a[i] * s_i + e_i in the DGT
 * domain of base Q, s_i ternary (stream stream0 + 2i), e_i discrete Gaussian
 * (stream stream0 + 2i + 1), both drawn on the device (Sampler::sample x2 +
 * poly_mul_add, sampler.cu:576-635, polynomial.cu:916-948); a, out
 * [batch][k][n/2].  For n <= 16384 two launches: s_i and e_i drawn once per
 * instance into the context's workspace (4 bytes per slot), then one fused
 * transform + multiply-add launch per residue.  batch * k must stay below
 * 2^31. */
int cupoly_sample_muladd_batch(cupoly_ctx *c, cupoly_gi *out, const cupoly_gi *a, int batch,
                               uint64_t seed, uint64_t stream0);

/* ---- uneven batches (no reference counterpart: the reference holds one
 * (n, k) per process, cudaengine.h:71-93) -----------------------------------
 * `count` independent products out[i] = a[i] * b[i] mod (X^n[i] + 1), residues
 * 0..k[i]-1, coefficient domain, [k[i]][n[i]/2] GIs each, n[i] in 2^11..2^15.
 * engines[] holds one engine per ring degree used (k of each >= the largest
 * k[i] of that degree); the plan keeps device copies of the work list and runs
 * on c's stream.  Pointers must stay valid for the plan's lifetime. */
int cupoly_uneven_create(cupoly_ctx *c, cupoly_engine *const *engines, int n_engines, const int *n,
                         const int *k, cupoly_gi *const *out, const cupoly_gi *const *a,
                         const cupoly_gi *const *b, int count, cupoly_uneven **plan);
/* One persistent launch (plus a counter-reset kernel), or the captured graph. */
int cupoly_uneven_run(cupoly_uneven *plan);
/* Capture reset + persistent kernel into a hipGraph; later runs replay it. */
int cupoly_uneven_capture(cupoly_uneven *plan);
/* The same products one launch per op (forward a, forward b, mul, inverse),
 * product by product; a and b are left in the DGT domain. */
int cupoly_uneven_run_plain(cupoly_uneven *plan);
/* Tunables: persistent grid size and tile order (0: most expensive first,
 * 1: input order).  autotune times every candidate `reps` times on the plan's
 * own batch and keeps the fastest. */
int cupoly_uneven_set_tuning(cupoly_uneven *plan, int grid, int order);
int cupoly_uneven_get_tuning(const cupoly_uneven *plan, int *grid, int *order, int *ntiles);
int cupoly_uneven_autotune(cupoly_uneven *plan, int reps, int *best_grid, int *best_order,
                           float *best_ms);
int cupoly_uneven_destroy(cupoly_uneven *plan);

#ifdef __cplusplus
}
#endif
#endif /* CUPOLY_AMD_H */

// cupoly_amd.hpp - C++ facade with the reference's host API, over the C ABI.
//
// A caller written against cuPoly's device-side API (include/cuPoly/cuda/
// cudaengine.h, dgt.h, sampler.h, arithmetic/context.h, arithmetic/
// polynomial.h) keeps its names and argument order:
//
//   CUDAEngine::init(k, kl, M, t);  Context ctx;  poly_t a, b, c;
//   Sampler::sample(&ctx, &a, UNIFORM);  poly_mul(&ctx, &c, &a, &b);
//   DGTEngine::execute_dgt(c.d_coefs, QBase, INVERSE, &ctx);
//
// Differences a caller sees:
//   * errors throw std::runtime_error with the library's message instead of
//     asserting or calling exit(1) (settings.h:144-152);
//   * streams are hipStream_t; the DGTEngine pointwise calls that take a
//     stream run on a context bound to that stream;
//   * poly_t keeps residues, not NTL ZZ coefficients: host data enters and
//     leaves through poly_import_residues / poly_export_residues /
//     poly_get_residues (the reference's own residue wire format,
//     polynomial.cu:1291-1393);
//   * Sampler draws are counter-based (Philox4x32-10): Sampler::sample takes
//     an optional (seed, stream) and otherwise uses a per-process counter.
//
// Header-only; link with libcupoly_amd.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cupoly_amd.h"

typedef cupoly_gi GaussianInteger;
enum bases { QBase = CUPOLY_QBASE, BBase = CUPOLY_BBASE, QBBase = CUPOLY_QBBASE, TBase = CUPOLY_TBASE };
enum dgt_direction { FORWARD = CUPOLY_FORWARD, INVERSE = CUPOLY_INVERSE };
enum add_mode_t { ADD = CUPOLY_ADD, SUB = CUPOLY_SUB, MUL = CUPOLY_MUL, MULMUL = CUPOLY_MULMUL,
                  ADDADD = CUPOLY_ADDADD };
enum pdistributions {
  DISCRETE_GAUSSIAN = CUPOLY_DISCRETE_GAUSSIAN,
  HAMMINGWEIGHT = CUPOLY_HAMMINGWEIGHT,
  NARROW = CUPOLY_NARROW,
  ZO = CUPOLY_ZO,
  BINARY = CUPOLY_BINARY,
  UNIFORM = CUPOLY_UNIFORM,
  TERNARY = CUPOLY_TERNARY
};
enum poly_states { HOSTSTATE, RNSSTATE, BOTHSTATE };

namespace cupoly_detail {
inline void check(int rc) {
  if (rc < 0) throw std::runtime_error(cupoly_last_error());
}
inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
// A temporary device buffer, freed on every exit path.  hipFree waits for the
// device work still using it, so an exception between its uses and the
// caller's stream synchronisation leaks nothing and frees nothing early.
struct DeviceTemp {
  void *p = nullptr;
  DeviceTemp(size_t bytes, const char *what) { hip_check(hipMalloc(&p, bytes), what); }
  ~DeviceTemp() {
    if (p) (void)hipFree(p);
  }
  DeviceTemp(const DeviceTemp &) = delete;
  DeviceTemp &operator=(const DeviceTemp &) = delete;
  template <class T>
  T *as() const {
    return (T *)p;
  }
};
// Waits for a stream when it goes out of scope, so that host buffers declared
// before it outlive every asynchronous copy into or out of them, whichever
// way the function is left.
struct SyncOnExit {
  hipStream_t s;
  ~SyncOnExit() { (void)hipStreamSynchronize(s); }
};
}  // namespace cupoly_detail

// CUDAEngine: the reference's process-wide engine (cudaengine.h:65-407),
// backed by one cupoly_engine.
class CUDAEngine {
 public:
  inline static cupoly_engine *engine = nullptr;
  inline static int N = 0;  // n / 2 Gaussian slots
  inline static uint64_t t = 0;
  inline static bool is_init = false;
  inline static std::vector<uint64_t> RNSPrimes, RNSBPrimes;

  static void init(int k, int kl, int M, uint64_t plaintext) {
    if (is_init) destroy();
    cupoly_detail::check(cupoly_engine_create(k, kl, M, plaintext, &engine));
    N = M / 2;
    t = plaintext;
    std::vector<uint64_t> all(k + kl);
    cupoly_detail::check(cupoly_engine_primes(engine, all.data()));
    RNSPrimes.assign(all.begin(), all.begin() + k);
    RNSBPrimes.assign(all.begin() + k, all.end());
    is_init = true;
  }
  static void init(int k, int M, uint64_t plaintext) { init(k, k + 1, M, plaintext); }
  static void destroy();
  static int get_n_residues(int base) {
    int r = cupoly_get_n_residues(engine, base);
    cupoly_detail::check(r);
    return r;
  }
};

// Context: stream + scratch (context.h:36-109).
class Context {
 public:
  cupoly_ctx *ctx = nullptr;
  explicit Context(hipStream_t stream = nullptr) {
    if (!CUDAEngine::is_init) throw std::runtime_error("Context: CUDAEngine not initialised");
    cupoly_detail::check(cupoly_ctx_create(CUDAEngine::engine, (void *)stream, &ctx));
  }
  ~Context() {
    if (ctx) cupoly_ctx_destroy(ctx);
  }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  hipStream_t get_stream() { return (hipStream_t)cupoly_ctx_stream(ctx); }
  void sync() { cupoly_detail::check(cupoly_ctx_sync(ctx)); }
};

namespace cupoly_detail {
inline std::map<hipStream_t, Context *> &stream_contexts() {
  static std::map<hipStream_t, Context *> m;
  return m;
}
inline Context *for_stream(hipStream_t s) {
  auto &m = stream_contexts();
  auto it = m.find(s);
  if (it != m.end()) return it->second;
  Context *c = new Context(s);
  m[s] = c;
  return c;
}
}  // namespace cupoly_detail

inline void CUDAEngine::destroy() {
  for (auto &kv : cupoly_detail::stream_contexts()) delete kv.second;
  cupoly_detail::stream_contexts().clear();
  if (engine) cupoly_engine_destroy(engine);
  engine = nullptr;
  is_init = false;
  N = 0;
  RNSPrimes.clear();
  RNSBPrimes.clear();
}

// DGTEngine (dgt.h:90-282).
class DGTEngine {
 public:
  static void execute_dgt(GaussianInteger *data, int base, int direction, Context *ctx) {
    cupoly_detail::check(cupoly_dgt(ctx->ctx, data, base, direction));
  }
  static void execute_add_dgt(GaussianInteger *c, const GaussianInteger *a,
                              const GaussianInteger *b, int base, hipStream_t s) {
    cupoly_detail::check(cupoly_add_dgt(cupoly_detail::for_stream(s)->ctx, c, a, b, base));
  }
  static void execute_sub_dgt(GaussianInteger *c, const GaussianInteger *a,
                              const GaussianInteger *b, int base, hipStream_t s) {
    cupoly_detail::check(cupoly_sub_dgt(cupoly_detail::for_stream(s)->ctx, c, a, b, base));
  }
  static void execute_double_add_dgt(GaussianInteger *c1, const GaussianInteger *a1,
                                     const GaussianInteger *b1, GaussianInteger *c2,
                                     const GaussianInteger *a2, const GaussianInteger *b2,
                                     int base, hipStream_t s) {
    cupoly_detail::check(
        cupoly_double_add_dgt(cupoly_detail::for_stream(s)->ctx, c1, a1, b1, c2, a2, b2, base));
  }
  static void execute_dr2_dgt(GaussianInteger *ct21, GaussianInteger *ct22, GaussianInteger *ct23,
                              const GaussianInteger *ct01, const GaussianInteger *ct02,
                              const GaussianInteger *ct11, const GaussianInteger *ct12, int base,
                              hipStream_t s) {
    cupoly_detail::check(cupoly_dr2_dgt(cupoly_detail::for_stream(s)->ctx, ct21, ct22, ct23, ct01,
                                        ct02, ct11, ct12, base));
  }
  static void execute_mul_dgt_gi(GaussianInteger *c, const GaussianInteger *a,
                                 const GaussianInteger *b, int base, hipStream_t s) {
    cupoly_detail::check(cupoly_mul_dgt_gi(cupoly_detail::for_stream(s)->ctx, c, a, b, base));
  }
  static void execute_mul_dgt_u64(GaussianInteger *c, const GaussianInteger *a, const uint64_t *b,
                                  int base, hipStream_t s) {
    cupoly_detail::check(cupoly_mul_dgt_u64(cupoly_detail::for_stream(s)->ctx, c, a, b, base));
  }
  static void execute_muladd_dgt(GaussianInteger *d, const GaussianInteger *a,
                                 const GaussianInteger *b, const GaussianInteger *c, int base,
                                 hipStream_t s) {
    cupoly_detail::check(cupoly_muladd_dgt(cupoly_detail::for_stream(s)->ctx, d, a, b, c, base));
  }
};

// poly_t (polynomial.h:55-66), residues only.
struct poly_t {
  bool init = false;
  GaussianInteger *d_coefs = nullptr;
  int status = BOTHSTATE;
  int base = QBase;
};

inline size_t poly_get_residues_size(int base) {
  return (size_t)CUDAEngine::N * CUDAEngine::get_n_residues(base) * sizeof(GaussianInteger);
}

inline void poly_init(Context *ctx, poly_t *a, int base = QBase) {
  if (a->init) return;  // idempotent (polynomial.cu:51-52)
  a->base = base;
  cupoly_detail::hip_check(hipMalloc((void **)&a->d_coefs, poly_get_residues_size(base)),
                           "poly_init");
  cupoly_detail::hip_check(
      hipMemsetAsync(a->d_coefs, 0, poly_get_residues_size(base), ctx->get_stream()), "poly_init");
  a->init = true;
  a->status = BOTHSTATE;
}

inline void poly_free(Context *ctx, poly_t *a) {
  if (!a || !a->init) return;
  ctx->sync();
  (void)hipFree(a->d_coefs);
  a->d_coefs = nullptr;
  a->init = false;
}

inline void poly_clear(Context *ctx, poly_t *a) {
  poly_init(ctx, a, a->base);
  cupoly_detail::hip_check(
      hipMemsetAsync(a->d_coefs, 0, poly_get_residues_size(a->base), ctx->get_stream()),
      "poly_clear");
}

inline void poly_copy(Context *ctx, poly_t *b, poly_t *a) {
  poly_init(ctx, b, a->base);
  cupoly_detail::hip_check(hipMemcpyAsync(b->d_coefs, a->d_coefs, poly_get_residues_size(a->base),
                                          hipMemcpyDeviceToDevice, ctx->get_stream()),
                           "poly_copy");
  b->status = a->status;
}

// poly_copy_to_device / poly_copy_to_host (polynomial.cu:105-188) bring the
// reference's host ZZ coefficients and its device residues into step.  This
// poly_t holds only the device residues, which are always current; host
// values are read with poly_get_residues / poly_get_integers.  Both calls
// only check that the polynomial exists.
inline void poly_copy_to_device(Context *, poly_t *a) {
  if (!a || !a->init) throw std::runtime_error("poly_copy_to_device: polynomial not initialised");
}
inline void poly_copy_to_host(Context *, poly_t *a) {
  if (!a || !a->init) throw std::runtime_error("poly_copy_to_host: polynomial not initialised");
}

inline void poly_add(Context *ctx, poly_t *c, poly_t *a, poly_t *b) {
  poly_init(ctx, c, a->base);
  cupoly_detail::check(cupoly_add_dgt(ctx->ctx, c->d_coefs, a->d_coefs, b->d_coefs, a->base));
  c->status = RNSSTATE;
}
inline void poly_sub(Context *ctx, poly_t *c, poly_t *a, poly_t *b) {
  poly_init(ctx, c, a->base);
  cupoly_detail::check(cupoly_sub_dgt(ctx->ctx, c->d_coefs, a->d_coefs, b->d_coefs, a->base));
  c->status = RNSSTATE;
}
inline void poly_double_add(Context *ctx, poly_t *c1, poly_t *a1, poly_t *b1, poly_t *c2,
                            poly_t *a2, poly_t *b2) {
  poly_add(ctx, c1, a1, b1);
  poly_add(ctx, c2, a2, b2);
}
inline void poly_mul(Context *ctx, poly_t *c, poly_t *a, poly_t *b) {
  poly_init(ctx, c, a->base);
  cupoly_detail::check(cupoly_mul_dgt_gi(ctx->ctx, c->d_coefs, a->d_coefs, b->d_coefs, a->base));
  c->status = RNSSTATE;
}
inline void poly_mul_add(Context *ctx, poly_t *d, poly_t *a, poly_t *b, poly_t *c) {
  poly_init(ctx, d, a->base);
  cupoly_detail::check(
      cupoly_muladd_dgt(ctx->ctx, d->d_coefs, a->d_coefs, b->d_coefs, c->d_coefs, a->base));
  d->status = RNSSTATE;
}
inline void poly_dr2(Context *ctx, poly_t *ct21, poly_t *ct22, poly_t *ct23, poly_t *ct01,
                     poly_t *ct02, poly_t *ct11, poly_t *ct12) {
  for (poly_t *o : {ct21, ct22, ct23}) poly_init(ctx, o, ct01->base);
  cupoly_detail::check(cupoly_dr2_dgt(ctx->ctx, ct21->d_coefs, ct22->d_coefs, ct23->d_coefs,
                                      ct01->d_coefs, ct02->d_coefs, ct11->d_coefs, ct12->d_coefs,
                                      ct01->base));
  ct21->status = ct22->status = ct23->status = RNSSTATE;
}
// poly_xi_bfv / poly_rho_bfv (polynomial.cu:610-693): c is an array of k
// polynomials; residue i of c[i] receives the scaled residue i of a.
inline void poly_xi_bfv(Context *ctx, poly_t *c, poly_t *a) {
  const int k = CUDAEngine::get_n_residues(QBase);
  std::vector<cupoly_gi *> out(k);
  for (int i = 0; i < k; i++) {
    poly_init(ctx, &c[i], QBase);
    out[i] = c[i].d_coefs;
  }
  cupoly_detail::check(cupoly_xi_bfv(ctx->ctx, out.data(), a->d_coefs));
  for (int i = 0; i < k; i++) c[i].status = RNSSTATE;
}
inline void poly_rho_bfv(Context *ctx, poly_t *c, poly_t *a) {
  const int k = CUDAEngine::get_n_residues(QBase);
  std::vector<cupoly_gi *> out(k);
  for (int i = 0; i < k; i++) {
    poly_init(ctx, &c[i], a->base);
    out[i] = c[i].d_coefs;
  }
  cupoly_detail::check(cupoly_rho_bfv(ctx->ctx, out.data(), a->d_coefs, a->base));
  for (int i = 0; i < k; i++) c[i].status = RNSSTATE;
}
// poly_dot (polynomial.cu:1439-1451): c = sum_{i < k} a[i] * b[i].
inline void poly_dot(Context *ctx, poly_t *c, poly_t *a, poly_t *b, const int k) {
  poly_init(ctx, c, a[0].base);
  std::vector<const cupoly_gi *> pa(k), pb(k);
  for (int i = 0; i < k; i++) {
    pa[i] = a[i].d_coefs;
    pb[i] = b[i].d_coefs;
  }
  cupoly_detail::check(cupoly_dot(ctx->ctx, c->d_coefs, pa.data(), pb.data(), k, a[0].base));
  c->status = RNSSTATE;
}
inline void poly_mul_int(Context *ctx, poly_t *c, poly_t *a, uint64_t x) {
  poly_init(ctx, c, a->base);
  cupoly_detail::check(cupoly_op_by_int(ctx->ctx, c->d_coefs, a->d_coefs, x, MUL, a->base));
  c->status = RNSSTATE;
}
inline void poly_add_int(Context *ctx, poly_t *c, poly_t *a, uint64_t x) {
  poly_init(ctx, c, a->base);
  cupoly_detail::check(cupoly_op_by_int(ctx->ctx, c->d_coefs, a->d_coefs, x, ADD, a->base));
  c->status = RNSSTATE;
}
inline void poly_sub_int(Context *ctx, poly_t *c, poly_t *a, uint64_t x) {
  poly_init(ctx, c, a->base);
  cupoly_detail::check(cupoly_op_by_int(ctx->ctx, c->d_coefs, a->d_coefs, x, SUB, a->base));
  c->status = RNSSTATE;
}
inline void poly_double_mul_int(Context *ctx, poly_t *c1, poly_t *a1, poly_t *c2, poly_t *a2,
                                uint64_t x1, uint64_t x2) {
  poly_init(ctx, c1, a1->base);
  poly_init(ctx, c2, a2->base);
  cupoly_detail::check(cupoly_double_op_by_int(ctx->ctx, c1->d_coefs, a1->d_coefs, x1,
                                               c2->d_coefs, a2->d_coefs, x2, MULMUL, a1->base));
  c1->status = c2->status = RNSSTATE;
}
inline void poly_double_add_int(Context *ctx, poly_t *c1, poly_t *a1, poly_t *c2, poly_t *a2,
                                uint64_t x1, uint64_t x2) {
  poly_init(ctx, c1, a1->base);
  poly_init(ctx, c2, a2->base);
  cupoly_detail::check(cupoly_double_op_by_int(ctx->ctx, c1->d_coefs, a1->d_coefs, x1,
                                               c2->d_coefs, a2->d_coefs, x2, ADDADD, a1->base));
  c1->status = c2->status = RNSSTATE;
}
inline void poly_negate(Context *ctx, poly_t *a) {
  cupoly_detail::check(cupoly_negate(ctx->ctx, a->d_coefs, a->d_coefs, a->base));
}
inline void poly_basis_extension_Q_to_B(Context *ctx, poly_t *a_B, poly_t *a_Q) {
  poly_init(ctx, a_B, BBase);
  cupoly_detail::check(cupoly_poly_basis_ext_q_to_b(ctx->ctx, a_B->d_coefs, a_Q->d_coefs));
  a_B->status = RNSSTATE;
}
inline void poly_complex_scaling_tDivQ(Context *ctx, poly_t *b, poly_t *a_Q, poly_t *a_B) {
  poly_init(ctx, b, QBase);
  cupoly_detail::check(
      cupoly_poly_complex_scaling(ctx->ctx, b->d_coefs, a_Q->d_coefs, a_B->d_coefs));
  b->status = RNSSTATE;
}

// poly_basis_extension_Q_to_QB (polynomial.cu:459-527): b = a's Q residues
// followed by its extension to B, both in the DGT domain.
inline void poly_basis_extension_Q_to_QB(Context *ctx, poly_t *b, poly_t *a) {
  if (a->base != QBase) throw std::runtime_error("poly_basis_extension_Q_to_QB: a must be in base Q");
  poly_init(ctx, b, QBBase);
  if (b->base != QBBase) throw std::runtime_error("poly_basis_extension_Q_to_QB: b must be in base QB");
  const size_t qoff = (size_t)CUDAEngine::get_n_residues(QBase) * CUDAEngine::N;
  cupoly_detail::check(cupoly_poly_basis_ext_q_to_b(ctx->ctx, b->d_coefs + qoff, a->d_coefs));
  cupoly_detail::hip_check(hipMemcpyAsync(b->d_coefs, a->d_coefs, poly_get_residues_size(QBase),
                                          hipMemcpyDeviceToDevice, ctx->get_stream()),
                           "poly_basis_extension_Q_to_QB");
  b->status = RNSSTATE;
}

// poly_simple_scaling_tDivQ (polynomial.cu:310-346): b = round(t a / Q) mod t
// in base T, coefficient domain (a transform over base T is the identity, so
// poly_get_residues reads the 2N values back as they are).  a is unchanged.
inline void poly_simple_scaling_tDivQ(Context *ctx, poly_t *b, poly_t *a) {
  if (a->base != QBase) throw std::runtime_error("poly_simple_scaling_tDivQ: a must be in base Q");
  poly_init(ctx, b, QBase);
  if (b->d_coefs != a->d_coefs)
    cupoly_detail::hip_check(hipMemcpyAsync(b->d_coefs, a->d_coefs, poly_get_residues_size(QBase),
                                            hipMemcpyDeviceToDevice, ctx->get_stream()),
                             "poly_simple_scaling_tDivQ");
  cupoly_detail::check(cupoly_dgt(ctx->ctx, b->d_coefs, QBase, INVERSE));
  cupoly_detail::check(cupoly_simple_scaling(ctx->ctx, b->d_coefs));
  b->status = RNSSTATE;
  b->base = TBase;
}

// Coefficient-domain residues of a, [nres][2N] (re half then imag half per
// residue), as poly_export_residues lays them out (polynomial.cu:1321-1330).
inline std::vector<uint64_t> poly_get_residues(Context *ctx, poly_t *a) {
  const int nres = CUDAEngine::get_n_residues(a->base), N = CUDAEngine::N;
  const size_t bytes = poly_get_residues_size(a->base);
  cupoly_detail::DeviceTemp buf(bytes, "poly_get_residues");
  GaussianInteger *tmp = buf.as<GaussianInteger>();
  cupoly_detail::hip_check(
      hipMemcpyAsync(tmp, a->d_coefs, bytes, hipMemcpyDeviceToDevice, ctx->get_stream()),
      "poly_get_residues");
  cupoly_detail::check(cupoly_dgt(ctx->ctx, tmp, a->base, INVERSE));
  std::vector<GaussianInteger> h((size_t)nres * N);
  cupoly_detail::SyncOnExit guard{ctx->get_stream()};
  cupoly_detail::hip_check(
      hipMemcpyAsync(h.data(), tmp, bytes, hipMemcpyDeviceToHost, ctx->get_stream()),
      "poly_get_residues");
  ctx->sync();
  std::vector<uint64_t> out((size_t)nres * 2 * N);
  for (int r = 0; r < nres; r++)
    for (int c = 0; c < N; c++) {
      out[(size_t)r * 2 * N + c] = h[(size_t)r * N + c].re;
      out[(size_t)r * 2 * N + N + c] = h[(size_t)r * N + c].imag;
    }
  return out;
}

// poly_get_residue (polynomial.cu:1538-1592): the 2N coefficient-domain
// values of residue id (re half then imag half).  Written to b when b is
// given, else to a new malloc'd array the caller frees.
inline uint64_t *poly_get_residue(Context *ctx, uint64_t *b, poly_t *a, int id) {
  if (id < 0 || id >= CUDAEngine::get_n_residues(a->base))
    throw std::runtime_error("poly_get_residue: residue id out of range");
  const std::vector<uint64_t> v = poly_get_residues(ctx, a);
  const size_t w = (size_t)2 * CUDAEngine::N;
  uint64_t *r = b ? b : (uint64_t *)malloc(w * sizeof(uint64_t));
  if (!r) throw std::runtime_error("poly_get_residue: out of host memory");
  std::copy(v.begin() + (size_t)id * w, v.begin() + (size_t)(id + 1) * w, r);
  return r;
}

// poly_residue_to_string (polynomial.cu:1489-1536): residue id as "v0,v1,...,"
// (every value followed by a comma).
inline std::string poly_residue_to_string(Context *ctx, poly_t *a, int id) {
  std::vector<uint64_t> v((size_t)2 * CUDAEngine::N);
  poly_get_residue(ctx, v.data(), a, id);
  std::ostringstream os;
  for (uint64_t x : v) os << x << ",";
  return os.str();
}

inline std::string poly_export_residues(Context *ctx, poly_t *a) {
  std::vector<uint64_t> v = poly_get_residues(ctx, a);
  std::ostringstream os;
  for (size_t i = 0; i < v.size(); i++) os << (i ? " " : "") << (int64_t)v[i];
  return os.str();
}

// Residues (coefficient domain, [nres][2N]) -> a new poly_t in the DGT domain.
inline poly_t *poly_import_residues(Context *ctx, const std::vector<uint64_t> &v, int base = QBase) {
  const int nres = CUDAEngine::get_n_residues(base), N = CUDAEngine::N;
  std::vector<GaussianInteger> h((size_t)nres * N, GaussianInteger{0, 0});
  for (size_t i = 0; i < v.size() && i < (size_t)nres * 2 * N; i++) {
    const int cid = (int)(i % (2 * N)), rid = (int)(i / (2 * N));
    if (cid < N)
      h[(size_t)rid * N + cid].re = v[i];
    else
      h[(size_t)rid * N + cid - N].imag = v[i];
  }
  poly_t *a = new poly_t;
  try {
    cupoly_detail::SyncOnExit guard{ctx->get_stream()};
    poly_init(ctx, a, base);
    cupoly_detail::hip_check(hipMemcpyAsync(a->d_coefs, h.data(), poly_get_residues_size(base),
                                            hipMemcpyHostToDevice, ctx->get_stream()),
                             "poly_import_residues");
    cupoly_detail::check(cupoly_dgt(ctx->ctx, a->d_coefs, base, FORWARD));
    ctx->sync();
  } catch (...) {
    poly_free(ctx, a);
    delete a;
    throw;
  }
  a->status = RNSSTATE;
  return a;
}

inline poly_t *poly_import_residues(Context *ctx, const std::string &s, int base = QBase) {
  std::istringstream is(s);
  std::vector<uint64_t> v;
  int64_t x;
  while (is >> x) v.push_back((uint64_t)x);
  return poly_import_residues(ctx, v, base);
}

// Host coefficient integers in and out (the reference keeps NTL ZZ
// coefficients in poly_t::coefs and converts with poly_crt / poly_icrt,
// polynomial.cu:190-291): little-endian 64-bit limbs, `limbs` per
// coefficient in, k per coefficient out (values in [0, Q)); the conversion
// runs on the device.
inline poly_t *poly_import_integers(Context *ctx, const std::vector<uint64_t> &coefs, int limbs,
                                    int base = QBase) {
  poly_t *p = new poly_t;
  try {
    poly_init(ctx, p, base);
    const int ncoefs = (int)(coefs.size() / limbs);
    cupoly_detail::DeviceTemp buf(coefs.size() * 8 + 8, "poly_import_integers");
    uint64_t *d = buf.as<uint64_t>();
    cupoly_detail::hip_check(hipMemcpyAsync(d, coefs.data(), coefs.size() * 8,
                                            hipMemcpyHostToDevice, ctx->get_stream()),
                             "poly_import_integers");
    cupoly_detail::check(cupoly_crt(ctx->ctx, p->d_coefs, d, ncoefs, limbs, base));
    cupoly_detail::check(cupoly_dgt(ctx->ctx, p->d_coefs, base, FORWARD));
    ctx->sync();
  } catch (...) {
    poly_free(ctx, p);
    delete p;
    throw;
  }
  p->status = RNSSTATE;
  return p;
}
inline std::vector<uint64_t> poly_get_integers(Context *ctx, poly_t *a) {
  const int k = CUDAEngine::get_n_residues(QBase), N = CUDAEngine::N;
  const size_t bytes = poly_get_residues_size(QBase), obytes = (size_t)2 * N * k * 8;
  cupoly_detail::DeviceTemp tbuf(bytes, "poly_get_integers"), dbuf(obytes, "poly_get_integers");
  GaussianInteger *tmp = tbuf.as<GaussianInteger>();
  uint64_t *d = dbuf.as<uint64_t>();
  cupoly_detail::hip_check(
      hipMemcpyAsync(tmp, a->d_coefs, bytes, hipMemcpyDeviceToDevice, ctx->get_stream()),
      "poly_get_integers");
  cupoly_detail::check(cupoly_dgt(ctx->ctx, tmp, QBase, INVERSE));
  cupoly_detail::check(cupoly_icrt(ctx->ctx, d, tmp));
  std::vector<uint64_t> out((size_t)2 * N * k);
  cupoly_detail::SyncOnExit guard{ctx->get_stream()};
  cupoly_detail::hip_check(
      hipMemcpyAsync(out.data(), d, obytes, hipMemcpyDeviceToHost, ctx->get_stream()),
      "poly_get_integers");
  ctx->sync();
  return out;
}

// poly_set_coeff / poly_get_coeff (polynomial.h:461-470).  The reference
// passes NTL ZZ values; here a coefficient is an unsigned integer as
// little-endian 64-bit limbs.  set: coefficient `index` (0 <= index < n) of a
// becomes the integer, reduced mod every prime of a's base (the polynomial is
// returned to the coefficient domain, the residues written, and transformed
// back, as the reference's host path does).  get: coefficient `index` of a
// (base Q) in [0, Q), k limbs.
inline void poly_set_coeff(Context *ctx, poly_t *a, int index, const std::vector<uint64_t> &value) {
  const int nres = CUDAEngine::get_n_residues(a->base), N = CUDAEngine::N;
  if (index < 0 || index >= 2 * N) throw std::runtime_error("poly_set_coeff: index out of range");
  poly_init(ctx, a, a->base);
  std::vector<uint64_t> primes;  // the primes of a's base, in residue order
  if (a->base != BBase) primes = CUDAEngine::RNSPrimes;
  if (a->base != QBase)
    primes.insert(primes.end(), CUDAEngine::RNSBPrimes.begin(), CUDAEngine::RNSBPrimes.end());
  std::vector<uint64_t> res(nres);
  for (int r = 0; r < nres; r++) {
    const uint64_t p = primes[r];
    unsigned __int128 acc = 0;  // Horner from the top limb: exact, acc < p
    for (size_t l = value.size(); l-- > 0;) acc = ((acc << 64) + value[l]) % p;
    res[r] = (uint64_t)acc;
  }
  cupoly_detail::check(cupoly_dgt(ctx->ctx, a->d_coefs, a->base, INVERSE));
  // coefficient j lives in re of slot j (j < N) or imag of slot j - N
  uint64_t *dst = index < N ? &a->d_coefs[index].re : &a->d_coefs[index - N].imag;
  cupoly_detail::SyncOnExit guard{ctx->get_stream()};
  cupoly_detail::hip_check(hipMemcpy2DAsync(dst, (size_t)N * sizeof(GaussianInteger), res.data(),
                                            sizeof(uint64_t), sizeof(uint64_t), nres,
                                            hipMemcpyHostToDevice, ctx->get_stream()),
                           "poly_set_coeff");
  cupoly_detail::check(cupoly_dgt(ctx->ctx, a->d_coefs, a->base, FORWARD));
  a->status = RNSSTATE;
}
inline std::vector<uint64_t> poly_get_coeff(Context *ctx, poly_t *a, int index) {
  const int k = CUDAEngine::get_n_residues(QBase), N = CUDAEngine::N;
  if (a->base != QBase) throw std::runtime_error("poly_get_coeff: a must be in base Q");
  if (index < 0 || index >= 2 * N) throw std::runtime_error("poly_get_coeff: index out of range");
  const std::vector<uint64_t> all = poly_get_integers(ctx, a);
  return std::vector<uint64_t>(all.begin() + (size_t)index * k, all.begin() + (size_t)(index + 1) * k);
}

inline bool poly_are_equal(Context *ctx, poly_t *a, poly_t *b) {
  if (a->base != b->base) return false;
  return poly_get_residues(ctx, a) == poly_get_residues(ctx, b);
}

// poly_select_residue (polynomial.cu:1611-1643): every residue of p becomes
// residue id read as an integer polynomial.  The values of residue id lie in
// [0, p_id) and are reduced mod each prime of p's base on the way in, so every
// residue stays in [0, p_i) when p_i < p_id (the reference copies the raw
// values over).
inline void poly_select_residue(Context *ctx, poly_t *p, int id) {
  const int nres = CUDAEngine::get_n_residues(p->base), N = CUDAEngine::N;
  if (!p->init || id < 0 || id >= nres)
    throw std::runtime_error("poly_select_residue: bad argument");
  if (nres == 1) return;
  cupoly_detail::DeviceTemp buf((size_t)2 * N * sizeof(uint64_t), "poly_select_residue");
  uint64_t *d = buf.as<uint64_t>();
  cupoly_detail::check(cupoly_dgt(ctx->ctx, p->d_coefs, p->base, INVERSE));
  // residue id as 2N one-limb integers: coefficient j from re[j], j + N from imag[j]
  const GaussianInteger *r = p->d_coefs + (size_t)id * N;
  cupoly_detail::hip_check(hipMemcpy2DAsync(d, sizeof(uint64_t), &r->re, sizeof(GaussianInteger),
                                            sizeof(uint64_t), N, hipMemcpyDeviceToDevice,
                                            ctx->get_stream()),
                           "poly_select_residue");
  cupoly_detail::hip_check(hipMemcpy2DAsync(d + N, sizeof(uint64_t), &r->imag,
                                            sizeof(GaussianInteger), sizeof(uint64_t), N,
                                            hipMemcpyDeviceToDevice, ctx->get_stream()),
                           "poly_select_residue");
  cupoly_detail::check(cupoly_crt(ctx->ctx, p->d_coefs, d, 2 * N, 1, p->base));
  cupoly_detail::check(cupoly_dgt(ctx->ctx, p->d_coefs, p->base, FORWARD));
  ctx->sync();
}

// poly_discard_qbase (polynomial.cu:1645-1660): a QB polynomial keeps only its
// B residues (moved to the front) and becomes a B polynomial; B is left as is.
// The move goes one residue at a time in ascending order on the context's
// stream: residue i lands on [i N, (i + 1) N), which holds no residue that is
// still to be moved (the single overlapping copy the reference issues has no
// defined result).
inline void poly_discard_qbase(Context *ctx, poly_t *p) {
  if (p->base == BBase) return;
  if (p->base != QBBase) throw std::runtime_error("poly_discard_qbase: p must be in base QB");
  const int k = CUDAEngine::get_n_residues(QBase), kl = CUDAEngine::get_n_residues(BBase);
  const size_t N = CUDAEngine::N;
  for (int i = 0; i < kl; i++)
    cupoly_detail::hip_check(hipMemcpyAsync(p->d_coefs + i * N, p->d_coefs + (k + i) * N,
                                            N * sizeof(GaussianInteger), hipMemcpyDeviceToDevice,
                                            ctx->get_stream()),
                             "poly_discard_qbase");
  p->base = BBase;
}

// Sampler (sampler.h:49-99), counter-based.
class Sampler {
 public:
  inline static uint64_t seed = 42;  // SEED, sampler.h:44
  inline static uint64_t next_stream = 0;
  static void init(Context *) {}
  static void reset(Context *) { next_stream = 0; }
  static void destroy() {}
  static void sample(Context *ctx, poly_t *p, int kind) { sample(ctx, p, kind, seed, next_stream++); }
  static void sample(Context *ctx, poly_t *p, int kind, uint64_t s, uint64_t stream) {
    poly_init(ctx, p, p->base);
    cupoly_detail::check(cupoly_sample(ctx->ctx, p->d_coefs, kind, s, stream, p->base));
    p->status = RNSSTATE;
    p->init = true;
  }
  static poly_t *sample(Context *ctx, int kind) {
    poly_t *p = new poly_t;
    sample(ctx, p, kind);
    return p;
  }
};

// Engine and context lifecycle, parameter tables and RNS constants.
//
// Mirrors CUDAEngine::init / gen_rns_primes / precompute
// (src/cuda/cudaengine.cu:1249-1511, 1681-1731) and DGTEngine::init
// (src/cuda/dgt.cu:1694-1982), with three differences of design:
//   * no process-wide singleton: each cupoly_engine owns its tables, so
//     several ring shapes can be resident on the device at once;
//   * every RNS constant is computed with modular arithmetic on 64-bit words
//     (no NTL): e.g. floor(t I_j B / q_j) mod b_i = -(t I_j B mod q_j) q_j^-1
//     mod b_i because B = 0 mod b_i, and the 128-bit fractions come from two
//     128/64 divisions.  The results are the exact values the reference's NTL
//     code approximates with 512-bit RR (equal whenever its precision
//     suffices, i.e. for kl <= 5);
//   * the DGT tables are the powers the reference stores (omega^e, xi^n and
//     their inverses, with 1/N folded into the inverse twist), each stored with
//     its Shoup quotient so kernels need no Barrett step for them.
#include <math.h>
#include <string.h>

#include <mutex>

#include "internal.hpp"
#include "tables63.inc"

static thread_local std::string g_last_error;

int cp_fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int cp_check_hip(hipError_t err, const char *what) {
  return cp_fail(CUPOLY_ERR_HIP, std::string(what) + ": " + hipGetErrorString(err));
}

extern "C" const char *cupoly_last_error(void) { return g_last_error.c_str(); }
extern "C" const char *cupoly_version(void) { return "cupoly_amd 0.1.0 (gfx950)"; }

bool cp_base_range(const cupoly_engine *e, int base, int *off, int *cnt) {
  switch (base) {
    case CUPOLY_QBASE: *off = 0; *cnt = e->k; return true;
    case CUPOLY_BBASE: *off = e->k; *cnt = e->kl; return true;
    case CUPOLY_QBBASE: *off = 0; *cnt = e->k + e->kl; return true;
    default: return false;
  }
}

static uint64_t mulm(uint64_t a, uint64_t b, uint64_t p) {
  return (uint64_t)(((cp_u128)a * b) % p);
}
static uint64_t invm(uint64_t a, uint64_t p) { return cp_powmod_host(a, p - 2, p); }

// RD(1 / RN(q)) computed exactly: take the correctly rounded quotient and step
// down one ulp when it exceeds 1/qd (decided with a 128-bit integer product).
static double rcp_rd_exact(uint64_t q) {
  double qd = (double)q;  // round to nearest
  double r = 1.0 / qd;    // round to nearest
  int er, eq;
  double fr = frexp(r, &er), fq = frexp(qd, &eq);
  uint64_t mr = (uint64_t)ldexp(fr, 53), mq = (uint64_t)ldexp(fq, 53);
  // r * qd = mr * mq * 2^(er + eq - 106); compare with 1
  cp_u128 prod = (cp_u128)mr * mq;
  int sh = 106 - (er + eq);  // r*qd > 1  <=>  prod > 2^sh
  bool above = (sh >= 128) ? false : (prod > ((cp_u128)1 << sh));
  if (above) r = nextafter(r, 0.0);
  return r;
}

static void shoup_pair(uint64_t w, uint64_t p, uint64_t *out) {
  out[0] = w;
  out[1] = cp_shoup_factor(w, p);
}

static int size_index(int ng) {
  for (int i = 0; i < CUPOLY_TABLE_NSIZES; i++)
    if (cupoly_table_sizes[i] == ng) return i;
  return -1;
}

// Multi-word (little-endian 64-bit limbs) constants of the CRT of base Q.
static void mp_mul_u64(uint64_t *x, int len, uint64_t m) {
  cp_u128 carry = 0;
  for (int l = 0; l < len; l++) {
    const cp_u128 v = (cp_u128)x[l] * m + carry;
    x[l] = (uint64_t)v;
    carry = v >> 64;
  }
}

static void compute_crt(cp_crt &c, const uint64_t *primes, int k, int nres) {
  memset(&c, 0, sizeof(c));
  c.k = k;
  c.nres = nres;
  c.Q[0] = 1;
  for (int j = 0; j < k; j++) mp_mul_u64(c.Q, CP_MAX_BASE + 1, primes[j]);
  for (int i = 0; i < k; i++) {
    c.M[i][0] = 1;
    for (int j = 0; j < k; j++)
      if (j != i) mp_mul_u64(c.M[i], CP_MAX_BASE, primes[j]);
    c.rcp_q[i] = 1.0 / (double)primes[i];
  }
  for (int r = 0; r < nres; r++) {
    const uint64_t p = primes[r];
    shoup_pair((uint64_t)(((cp_u128)1 << 64) % p), p, c.two64[r]);
  }
}

static void compute_hps(cp_hps &h, const uint64_t *primes, int k, int kl, uint64_t t) {
  memset(&h, 0, sizeof(h));
  h.k = k;
  h.kl = kl;
  h.t = t;
  const uint64_t *q = primes, *b = primes + k;
  for (int j = 0; j < k; j++) h.q[j] = cp_make_prime(q[j]);
  for (int i = 0; i < kl; i++) h.b[i] = cp_make_prime(b[i]);
  // products of all primes of a base except one, reduced mod m
  auto prodQ_except = [&](int skip, uint64_t m) {
    uint64_t r = 1 % m;
    for (int l = 0; l < k; l++)
      if (l != skip) r = mulm(r, q[l] % m, m);
    return r;
  };
  auto prodB_except = [&](int skip, uint64_t m) {
    uint64_t r = 1 % m;
    for (int l = 0; l < kl; l++)
      if (l != skip) r = mulm(r, b[l] % m, m);
    return r;
  };
  for (int j = 0; j < k; j++) {
    uint64_t Qj = prodQ_except(j, q[j]);  // (Q/q_j) mod q_j
    shoup_pair(invm(Qj, q[j]), q[j], h.invQqi[j]);
    shoup_pair(Qj, q[j], h.Qqi[j]);
    shoup_pair(prodB_except(-1, q[j]), q[j], h.Bqi[j]);  // B mod q_j
    h.rcp_q[j] = rcp_rd_exact(q[j]);
    for (int i = 0; i < kl; i++) {
      shoup_pair(prodQ_except(j, b[i]), b[i], h.Qqibi[i][j]);
      shoup_pair(prodB_except(i, q[j]), q[j], h.Bbiqi[j][i]);
    }
  }
  for (int i = 0; i < kl; i++) {
    shoup_pair(prodQ_except(-1, b[i]), b[i], h.Qbi[i]);  // Q mod b_i
    uint64_t Bi = prodB_except(i, b[i]);                 // (B/b_i) mod b_i
    shoup_pair(invm(Bi, b[i]), b[i], h.invBbi[i]);
    // lambda_i = t * ((B Q / b_i) mod b_i)^-1 * (B / b_i) mod b_i   (cudaengine.cu:1466-1470)
    uint64_t QBi = mulm(Bi, prodQ_except(-1, b[i]), b[i]);
    uint64_t lam = mulm(mulm(t % b[i], invm(QBi, b[i]), b[i]), Bi, b[i]);
    shoup_pair(lam, b[i], h.lam[i]);
  }
  // omega_j = t * ((Q B / q_j) mod q_j)^-1 * B / q_j  (cudaengine.cu:1472-1488)
  for (int j = 0; j < k; j++) {
    uint64_t QBj = mulm(prodQ_except(j, q[j]), prodB_except(-1, q[j]), q[j]);
    uint64_t Ij = invm(QBj, q[j]);
    // r = (t * Ij * B) mod q_j
    uint64_t r = mulm(mulm(t % q[j], Ij, q[j]), prodB_except(-1, q[j]), q[j]);
    // frac * 2^128 = floor(r * 2^128 / q_j)
    cp_u128 hi = ((cp_u128)r << 64) / q[j];
    uint64_t rem = (uint64_t)(((cp_u128)r << 64) % q[j]);
    cp_u128 lo = ((cp_u128)rem << 64) / q[j];
    h.wfrac[j][0] = (uint64_t)lo;
    h.wfrac[j][1] = (uint64_t)hi;
    for (int i = 0; i < kl; i++) {
      // floor(t Ij B / q_j) mod b_i = (0 - r) * q_j^-1 mod b_i, as t Ij B = 0 mod b_i
      uint64_t rb = r % b[i];
      uint64_t v = mulm((b[i] - rb) % b[i], invm(q[j] % b[i], b[i]), b[i]);
      shoup_pair(v, b[i], h.wint[i][j]);
    }
    // simple scaling: t * ((Q/q_j) mod q_j)^-1 / q_j  (cudaengine.cu:1491-1507)
    cp_u128 a = (cp_u128)t * invm(prodQ_except(j, q[j]), q[j]);
    h.twint[j] = (uint64_t)(a / q[j]);
    uint64_t ra = (uint64_t)(a % q[j]);
    cp_u128 fh = ((cp_u128)ra << 64) / q[j];
    uint64_t frem = (uint64_t)(((cp_u128)ra << 64) % q[j]);
    cp_u128 fl = ((cp_u128)frem << 64) / q[j];
    h.twfrac[j][0] = (uint64_t)fl;
    h.twfrac[j][1] = (uint64_t)fh;
  }
}

extern "C" int cupoly_engine_create(int k, int kl, int n, uint64_t t, cupoly_engine **out) {
  if (!out) return cp_fail(CUPOLY_ERR_ARG, "engine_create: null out");
  *out = nullptr;
  if (k < 1 || kl < 1 || k >= 15 || kl >= 15 || k + kl + 1 >= CP_MAX_RES)
    return cp_fail(CUPOLY_ERR_ARG, "engine_create: need 1 <= k, kl <= 14 and k + kl <= 28 "
                                   "(cudaengine.cu:1270-1293)");
  if (n < 32 || (n & (n - 1)) || size_index(n / 2) < 0)
    return cp_fail(CUPOLY_ERR_UNSUPPORTED, "DGT is not working for such N (n must be a power "
                                           "of two in [32, 65536])");
  if (t < 2) return cp_fail(CUPOLY_ERR_UNSUPPORTED, "engine_create: plaintext modulus t < 2");
  cupoly_engine *e = new cupoly_engine();
  e->k = k;
  e->kl = kl;
  e->n = n;
  e->ng = n / 2;
  e->logng = 0;
  while ((1 << e->logng) < e->ng) e->logng++;
  e->nres = k + kl;
  e->t = t;
  for (int r = 0; r < e->nres; r++) {
    uint64_t p = cupoly_table_primes[r];
    if (!(p > (1ULL << 62) && p < (1ULL << 63) && (1ULL << 63) - p < (1ULL << 30))) {
      delete e;
      return cp_fail(CUPOLY_ERR_UNSUPPORTED, "prime outside (2^63 - 2^30, 2^63)");
    }
    e->primes.push_back(p);
  }
  compute_hps(e->h_hps, e->primes.data(), k, kl, t);
  compute_crt(e->h_crt, e->primes.data(), k, e->nres);
  hipError_t err = hipGetDevice(&e->device);
  if (err == hipSuccess)
    err = hipDeviceGetAttribute(&e->ncu, hipDeviceAttributeMultiprocessorCount, e->device);
  if (err != hipSuccess) {
    delete e;
    return cp_check_hip(err, "hipGetDevice");
  }
  // --- DGT tables (host build, one upload) ---
  const int ng = e->ng, nres = e->nres, si = size_index(ng);
  std::vector<uint64_t> twf((size_t)nres * ng), twi((size_t)nres * ng);
  std::vector<uint64_t> tsf((size_t)nres * ng * 6), tsi((size_t)nres * ng * 6);
  std::vector<cp_prime> prs(nres);
  for (int r = 0; r < nres; r++) {
    const uint64_t p = e->primes[r];
    prs[r] = cp_make_prime(p);
    const uint64_t om = cp_powmod_host(cupoly_table_proot[r], (p - 1) / ng, p);
    const uint64_t omi = invm(om, p);
    uint64_t w = 1, wi = 1;
    for (int x = 0; x < ng / 2; x++) {
      shoup_pair(w, p, &twf[((size_t)r * ng / 2 + x) * 2]);
      shoup_pair(wi, p, &twi[((size_t)r * ng / 2 + x) * 2]);
      w = mulm(w, om, p);
      wi = mulm(wi, omi, p);
    }
    const uint64_t xr = cupoly_table_nthroot[r][si][0], xim = cupoly_table_nthroot[r][si][1];
    const uint64_t yr = cupoly_table_invnthroot[r][si][0], yim = cupoly_table_invnthroot[r][si][1];
    uint64_t ar = 1, ai = 0;                   // xi^n
    uint64_t br = invm((uint64_t)ng, p), bi = 0;  // xi^-n / ng
    for (int x = 0; x < ng; x++) {
      uint64_t *o = &tsf[((size_t)r * ng + x) * 6];
      shoup_pair(ar, p, o);
      shoup_pair(ai, p, o + 2);
      shoup_pair((ar + ai) % p, p, o + 4);
      uint64_t *v = &tsi[((size_t)r * ng + x) * 6];
      shoup_pair(br, p, v);
      shoup_pair(bi, p, v + 2);
      shoup_pair((br + bi) % p, p, v + 4);
      uint64_t nr = (mulm(ar, xr, p) + p - mulm(ai, xim, p)) % p;
      uint64_t ni = (mulm(ar, xim, p) + mulm(ai, xr, p)) % p;
      ar = nr;
      ai = ni;
      nr = (mulm(br, yr, p) + p - mulm(bi, yim, p)) % p;
      ni = (mulm(br, yim, p) + mulm(bi, yr, p)) % p;
      br = nr;
      bi = ni;
    }
  }
  // split-basis product tables (see dgt_dev.hpp): J = omega^(ng/4) squares to
  // -1; xi+- = xi.re +- J xi.im; the inverse roots r^-1 of the split tree of
  // X^ng -+ J come from its leaves xi+-^-1 omega^-bitrev(l) by squaring:
  // r^-1 at (s, g) is the leaf-side value at (s + 1, 2g).
  std::vector<uint64_t> spt((size_t)nres * ng * 4), spth((size_t)nres * ng * 4),
      spr((size_t)nres * ng * 4), spk((size_t)nres * 8, 0);
  {
    int lg = 0;
    while ((1 << lg) < ng) lg++;
    std::vector<uint64_t> pw(ng), cur(ng), nxt(ng);
    for (int r = 0; r < nres; r++) {
      const uint64_t p = e->primes[r];
      const uint64_t om = cp_powmod_host(cupoly_table_proot[r], (p - 1) / ng, p);
      const uint64_t omi = invm(om, p);
      const uint64_t J = cp_powmod_host(om, (uint64_t)ng / 4, p);
      const uint64_t xr = cupoly_table_nthroot[r][si][0], xim = cupoly_table_nthroot[r][si][1];
      const uint64_t xs[2] = {(xr + mulm(J, xim, p)) % p, (xr + p - mulm(J, xim, p)) % p};
      if (mulm(J, J, p) != p - 1 || cp_powmod_host(xs[0], (uint64_t)ng, p) != J ||
          cp_powmod_host(xs[1], (uint64_t)ng, p) != p - J) {
        delete e;
        return cp_fail(CUPOLY_ERR_UNSUPPORTED, "split basis: root tables inconsistent for this prime");
      }
      // the second operand of a product is twisted by h xi+-^g (h = 1/2ng): the
      // product's spectrum then carries the inverse's 1/2ng already
      const uint64_t h = invm((uint64_t)(2 * ng) % p, p);
      uint64_t f[2] = {1, 1};
      for (int g = 0; g < ng; g++) {
        for (int c = 0; c < 2; c++) {
          shoup_pair(f[c], p, &spt[((size_t)r * ng + g) * 4 + 2 * c]);
          shoup_pair(mulm(f[c], h, p), p, &spth[((size_t)r * ng + g) * 4 + 2 * c]);
          f[c] = mulm(f[c], xs[c], p);
        }
      }
      pw[0] = 1;
      for (int x = 1; x < ng; x++) pw[x] = mulm(pw[x - 1], omi, p);
      for (int c = 0; c < 2; c++) {
        const uint64_t xinv = invm(xs[c], p);
        for (int l = 0; l < ng; l++) {
          int b = 0;
          for (int i = 0; i < lg; i++) b |= ((l >> i) & 1) << (lg - 1 - i);
          cur[l] = mulm(xinv, pw[b], p);
        }
        for (int s = lg - 1; s >= 0; s--) {
          for (int g = 0; g < (1 << s); g++) {
            const uint64_t ri = cur[2 * g];
            shoup_pair(ri, p, &spr[((size_t)r * ng + (1 << s) + g) * 4 + 2 * c]);
            nxt[g] = mulm(ri, ri, p);
          }
          std::swap(cur, nxt);
        }
      }
      shoup_pair(J, p, &spk[(size_t)r * 8]);
      shoup_pair(h, p, &spk[(size_t)r * 8 + 2]);
      shoup_pair(mulm(h, invm(J, p), p), p, &spk[(size_t)r * 8 + 4]);
    }
  }
  auto up = [&](void **dst, const void *src, size_t bytes) -> int {
    CP_HIP(hipMalloc(dst, bytes));
    CP_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return CUPOLY_OK;
  };
  int rc;
  if ((rc = up((void **)&e->d_primes, prs.data(), prs.size() * sizeof(cp_prime))) ||
      (rc = up((void **)&e->d_tw_fwd, twf.data(), twf.size() * 8)) ||
      (rc = up((void **)&e->d_tw_inv, twi.data(), twi.size() * 8)) ||
      (rc = up((void **)&e->d_twist_fwd, tsf.data(), tsf.size() * 8)) ||
      (rc = up((void **)&e->d_twist_inv, tsi.data(), tsi.size() * 8)) ||
      (rc = up((void **)&e->d_split_tw, spt.data(), spt.size() * 8)) ||
      (rc = up((void **)&e->d_split_tw_h, spth.data(), spth.size() * 8)) ||
      (rc = up((void **)&e->d_split_root, spr.data(), spr.size() * 8)) ||
      (rc = up((void **)&e->d_split_k, spk.data(), spk.size() * 8)) ||
      (rc = up((void **)&e->d_hps, &e->h_hps, sizeof(cp_hps))) ||
      (rc = up((void **)&e->d_crt, &e->h_crt, sizeof(cp_crt))) || (rc = cp_dgt_setup(e)) ||
      (rc = cp_homomul_setup(e)) || (rc = cp_encrypt_setup(e)) ||
      (rc = cp_relin_setup(e))) {
    std::string msg = g_last_error;
    cupoly_engine_destroy(e);
    return cp_fail(rc, msg);
  }
  *out = e;
  return CUPOLY_OK;
}

extern "C" int cupoly_engine_destroy(cupoly_engine *e) {
  if (!e) return cp_fail(CUPOLY_ERR_ARG, "engine_destroy: null");
  (void)hipDeviceSynchronize();
  void *ptrs[] = {e->d_primes,    e->d_tw_fwd,   e->d_tw_inv,       e->d_twist_fwd,
                  e->d_twist_inv, e->d_split_tw, e->d_split_root, e->d_split_k,
                  e->d_hps,       e->d_crt,      e->d_split_tw_h};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  delete e;
  return CUPOLY_OK;
}

extern "C" int cupoly_get_n_residues(const cupoly_engine *e, int base) {
  if (!e) return cp_fail(CUPOLY_ERR_STATE, "get_n_residues: engine not initialised");
  switch (base) {
    case CUPOLY_QBASE: return e->k;
    case CUPOLY_BBASE: return e->kl;
    case CUPOLY_QBBASE: return e->k + e->kl;
    case CUPOLY_TBASE: return 1;
    default: return cp_fail(CUPOLY_ERR_UNSUPPORTED, "Unknown base");
  }
}

extern "C" int cupoly_engine_info(const cupoly_engine *e, int *n, int *ng, int *k, int *kl,
                                  uint64_t *t) {
  if (!e) return cp_fail(CUPOLY_ERR_STATE, "engine_info: null engine");
  if (n) *n = e->n;
  if (ng) *ng = e->ng;
  if (k) *k = e->k;
  if (kl) *kl = e->kl;
  if (t) *t = e->t;
  return CUPOLY_OK;
}

extern "C" int cupoly_engine_primes(const cupoly_engine *e, uint64_t *out) {
  if (!e || !out) return cp_fail(CUPOLY_ERR_ARG, "engine_primes: null");
  for (int r = 0; r < e->nres; r++) out[r] = e->primes[r];
  return e->nres;
}

static int hps_list(const cp_hps &h, const char *name, uint64_t *out, int cap) {
  std::vector<uint64_t> v;
  std::string s(name);
  const int k = h.k, kl = h.kl;
  if (s == "invQqi") for (int j = 0; j < k; j++) v.push_back(h.invQqi[j][0]);
  else if (s == "Qqi") for (int j = 0; j < k; j++) v.push_back(h.Qqi[j][0]);
  else if (s == "Qqibi") for (int i = 0; i < kl; i++) for (int j = 0; j < k; j++) v.push_back(h.Qqibi[i][j][0]);
  else if (s == "Qbi") for (int i = 0; i < kl; i++) v.push_back(h.Qbi[i][0]);
  else if (s == "invBbi") for (int i = 0; i < kl; i++) v.push_back(h.invBbi[i][0]);
  else if (s == "Bbiqi") for (int j = 0; j < k; j++) for (int i = 0; i < kl; i++) v.push_back(h.Bbiqi[j][i][0]);
  else if (s == "Bqi") for (int j = 0; j < k; j++) v.push_back(h.Bqi[j][0]);
  else if (s == "lambda") for (int i = 0; i < kl; i++) v.push_back(h.lam[i][0]);
  else if (s == "wint") for (int i = 0; i < kl; i++) for (int j = 0; j < k; j++) v.push_back(h.wint[i][j][0]);
  else if (s == "wfrac") for (int i = 0; i < kl; i++) for (int j = 0; j < k; j++) { v.push_back(h.wfrac[j][0]); v.push_back(h.wfrac[j][1]); }
  else if (s == "twint") for (int j = 0; j < k; j++) v.push_back(h.twint[j]);
  else if (s == "twfrac") for (int j = 0; j < k; j++) { v.push_back(h.twfrac[j][0]); v.push_back(h.twfrac[j][1]); }
  else if (s == "rcp_q") for (int j = 0; j < k; j++) { uint64_t bits; memcpy(&bits, &h.rcp_q[j], 8); v.push_back(bits); }
  else return cp_fail(CUPOLY_ERR_ARG, "unknown constant name " + s);
  if ((int)v.size() > cap) return cp_fail(CUPOLY_ERR_ARG, "constant: buffer too small");
  for (size_t x = 0; x < v.size(); x++) out[x] = v[x];
  return (int)v.size();
}

extern "C" int cupoly_engine_constant(const cupoly_engine *e, const char *name, uint64_t *out,
                                      int cap) {
  if (!e || !name || !out) return cp_fail(CUPOLY_ERR_ARG, "engine_constant: null");
  return hps_list(e->h_hps, name, out, cap);
}

extern "C" int cupoly_hps_constant(int k, int kl, uint64_t t, const char *name, uint64_t *out,
                                   int cap) {
  if (!name || !out) return cp_fail(CUPOLY_ERR_ARG, "hps_constant: null");
  if (k < 1 || kl < 1 || k >= 15 || kl >= 15 || k + kl + 1 >= CP_MAX_RES || t < 2)
    return cp_fail(CUPOLY_ERR_ARG, "hps_constant: bad (k, kl, t)");
  cp_hps h;
  compute_hps(h, cupoly_table_primes, k, kl, t);
  return hps_list(h, name, out, cap);
}

// ---------------------------------------------------------------------------
// Context: stream + scratch (context.cpp:19-67)
// ---------------------------------------------------------------------------
extern "C" int cupoly_ctx_create(cupoly_engine *e, void *stream, cupoly_ctx **out) {
  if (!e || !out) return cp_fail(CUPOLY_ERR_ARG, "ctx_create: null");
  *out = nullptr;
  cupoly_ctx *c = new cupoly_ctx();
  c->e = e;
  hipError_t err = hipSuccess;
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    err = hipStreamCreate(&c->stream);  // default priority
    c->own_stream = true;
  }
  if (err == hipSuccess)
    err = hipMalloc((void **)&c->d_aux, (size_t)e->kl * e->n * sizeof(uint64_t));
  if (err == hipSuccess) err = hipMalloc((void **)&c->d_tmp, (size_t)e->nres * e->ng * sizeof(cp_gi));
  if (err == hipSuccess) err = hipMalloc((void **)&c->d_tmp2, (size_t)e->nres * e->ng * sizeof(cp_gi));
  if (err != hipSuccess) {
    cupoly_ctx_destroy(c);
    return cp_check_hip(err, "ctx_create");
  }
  *out = c;
  return CUPOLY_OK;
}

// The second stream is made on first need only: every stream a process
// creates takes a share of the device's few hardware queues, and contexts
// that never split a batch (a serving loop's lanes, say) measured slower when
// each also held an idle second stream (p50 latency 0.82 -> 1.45 ms).
int cp_ctx_aux(cupoly_ctx *c) {
  if (c->aux) return CUPOLY_OK;
  cupoly_ctx *x = new cupoly_ctx();
  x->e = c->e;
  x->own_stream = true;
  hipError_t err = hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking);  // default priority
  if (err == hipSuccess && !c->ev_fork)
    err = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (err == hipSuccess && !c->ev_join)
    err = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  if (err != hipSuccess) {
    if (x->stream) (void)hipStreamDestroy(x->stream);
    delete x;
    return cp_check_hip(err, "ctx: second stream");
  }
  c->aux = x;
  return CUPOLY_OK;
}

extern "C" int cupoly_ctx_destroy(cupoly_ctx *c) {
  if (!c) return cp_fail(CUPOLY_ERR_ARG, "ctx_destroy: null");
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->d_aux) (void)hipFree(c->d_aux);
  if (c->d_tmp) (void)hipFree(c->d_tmp);
  if (c->d_tmp2) (void)hipFree(c->d_tmp2);
  if (c->d_ws) (void)hipFree(c->d_ws);
  if (c->d_ptrs) (void)hipFree(c->d_ptrs);
  if (c->h_ring) (void)hipHostFree(c->h_ring);
  for (int i = 0; i < cupoly_ctx::kRing; i++)
    if (c->ring_ev[i]) (void)hipEventDestroy(c->ring_ev[i]);
  if (c->aux) (void)cupoly_ctx_destroy(c->aux);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CUPOLY_OK;
}

extern "C" void *cupoly_ctx_stream(const cupoly_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int cupoly_ctx_sync(const cupoly_ctx *c) {
  if (!c) return cp_fail(CUPOLY_ERR_ARG, "ctx_sync: null");
  CP_HIP(hipStreamSynchronize(c->stream));
  return CUPOLY_OK;
}

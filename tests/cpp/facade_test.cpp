// Drop-in check of the C++ facade (include/cupoly_amd.hpp): the sequence a
// cuPoly caller writes (tests/unit_tests/test_bfv_mode.cu:1180-1244, Mul and
// Add) runs unchanged against libcupoly_amd.so.
//
// usage: facade_test <case file>
// The case file (written by tests/test_gpu_facade.py from the CPU oracle) is
// raw little-endian uint64: k, kl, n, t, then a, b, a*b, a+b as
// coefficient-domain residues [k][n] (re half then imag half per residue),
// then a extended to base B [kl][n], round(t a / Q) mod t [n], a residue id
// and residue id of a reduced mod every prime of Q [k][n].
#include <cstdio>
#include <algorithm>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/cupoly_amd.hpp"

typedef unsigned __int128 u128;

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <case file>\n", argv[0]);
    return 2;
  }
  std::ifstream f(argv[1], std::ios::binary);
  uint64_t hdr[4];
  f.read((char *)hdr, sizeof(hdr));
  const int k = (int)hdr[0], kl = (int)hdr[1], n = (int)hdr[2];
  const uint64_t t = hdr[3];
  const size_t cnt = (size_t)k * n;
  const size_t bcnt = (size_t)kl * n;
  std::vector<uint64_t> a(cnt), b(cnt), prod(cnt), sum(cnt), a_b(bcnt), scaled(n), sel(cnt);
  uint64_t sid = 0;
  f.read((char *)a.data(), cnt * 8);
  f.read((char *)b.data(), cnt * 8);
  f.read((char *)prod.data(), cnt * 8);
  f.read((char *)sum.data(), cnt * 8);
  f.read((char *)a_b.data(), bcnt * 8);
  f.read((char *)scaled.data(), (size_t)n * 8);
  f.read((char *)&sid, 8);
  f.read((char *)sel.data(), cnt * 8);
  if (!f) {
    std::fprintf(stderr, "short case file\n");
    return 2;
  }
  try {
    CUDAEngine::init(k, kl, n, t);
    {
      Context ctx;
      poly_t *pa = poly_import_residues(&ctx, a);
      poly_t *pb = poly_import_residues(&ctx, b);
      poly_t c, d, e;
      poly_mul(&ctx, &c, pa, pb);
      poly_add(&ctx, &d, pa, pb);
      if (poly_get_residues(&ctx, &c) != prod) {
        std::fprintf(stderr, "poly_mul mismatch\n");
        return 1;
      }
      if (poly_get_residues(&ctx, &d) != sum) {
        std::fprintf(stderr, "poly_add mismatch\n");
        return 1;
      }
      // the wire format round-trips
      poly_t *pc = poly_import_residues(&ctx, poly_export_residues(&ctx, &c));
      if (!poly_are_equal(&ctx, pc, &c)) {
        std::fprintf(stderr, "export/import mismatch\n");
        return 1;
      }
      Sampler::sample(&ctx, &e, UNIFORM);
      // integer coefficients through the device CRT: x mod Q round-trips
      {
        const std::vector<uint64_t> xs = poly_get_integers(&ctx, &c);
        poly_t *pi = poly_import_integers(&ctx, xs, k);
        const bool ok = poly_are_equal(&ctx, pi, &c);
        poly_free(&ctx, pi);
        delete pi;
        if (!ok) {
          std::fprintf(stderr, "integer round trip mismatch\n");
          return 1;
        }
      }
      // poly_set_coeff / poly_get_coeff (the reference's SetCoeff test,
      // test_bfv_mode.cu:1108-1117): coefficient 3 becomes 42, a coefficient
      // in the second half becomes a three-limb integer, every other stays
      {
        poly_t cc;  // a copy: c stays as it is for the checks below
        poly_init(&ctx, &cc);
        poly_copy(&ctx, &cc, &c);
        const std::vector<uint64_t> before = poly_get_integers(&ctx, &cc);
        poly_set_coeff(&ctx, &cc, 3, {42});
        const int hi = CUDAEngine::N + 5;
        const std::vector<uint64_t> big = {0x0123456789abcdefULL, 0xfedcba9876543210ULL, 7};
        poly_set_coeff(&ctx, &cc, hi, big);
        const std::vector<uint64_t> after = poly_get_integers(&ctx, &cc);
        std::vector<uint64_t> want42(k, 0);
        want42[0] = 42;
        bool ok = poly_get_coeff(&ctx, &cc, 3) == want42;
        // the three-limb value reduced mod Q (k limbs): compare residues instead
        const std::vector<uint64_t> got_hi = poly_get_coeff(&ctx, &cc, hi);
        for (int r = 0; r < k && ok; r++) {
          const uint64_t q = CUDAEngine::RNSPrimes[r];
          u128 x = 0, y = 0;
          for (int l = 2; l >= 0; l--) x = ((x << 64) + big[l]) % q;
          for (int l = k - 1; l >= 0; l--) y = ((y << 64) + got_hi[l]) % q;
          ok = x == y;
        }
        for (int j = 0; j < 2 * CUDAEngine::N && ok; j++)
          if (j != 3 && j != hi)
            for (int l = 0; l < k; l++) ok = ok && after[(size_t)j * k + l] == before[(size_t)j * k + l];
        poly_free(&ctx, &cc);
        if (!ok) {
          std::fprintf(stderr, "poly_set_coeff / poly_get_coeff mismatch\n");
          return 1;
        }
      }
      // slot-wise and scalar wrappers against residue arithmetic done here
      // (the DGT is linear, and an integer x is the constant polynomial x)
      {
        const std::vector<uint64_t> &Q = CUDAEngine::RNSPrimes;
        const uint64_t x1 = 123456789123456789ULL, x2 = 987654321ULL;
        auto modq = [&](int r, u128 v) { return (uint64_t)(v % Q[r]); };
        std::vector<uint64_t> w_sub(cnt), w_neg(cnt), w_mul1(cnt), w_mul2(cnt), w_add1(a),
            w_sub1(a), w_add2(b), w_bb(cnt), w_muladd(cnt);
        for (int r = 0; r < k; r++) {
          const uint64_t q = Q[r], y1 = x1 % q, y2 = x2 % q;
          for (int j = 0; j < n; j++) {
            const size_t i = (size_t)r * n + j;
            w_sub[i] = modq(r, (u128)a[i] + q - b[i]);
            w_neg[i] = modq(r, (u128)q - a[i]);
            w_mul1[i] = modq(r, (u128)a[i] * y1);
            w_mul2[i] = modq(r, (u128)b[i] * y2);
            w_bb[i] = modq(r, (u128)b[i] + b[i]);
            w_muladd[i] = modq(r, (u128)prod[i] + a[i]);
          }
          const size_t i0 = (size_t)r * n;  // coefficient 0
          w_add1[i0] = modq(r, (u128)a[i0] + y1);
          w_sub1[i0] = modq(r, (u128)a[i0] + q - y1);
          w_add2[i0] = modq(r, (u128)b[i0] + y2);
        }
        poly_t s, ng, m1, ad1, sb1, da1, da2, dm1, dm2, di1, di2, ma, cl, cc;
        poly_sub(&ctx, &s, pa, pb);
        poly_copy(&ctx, &ng, pa);
        poly_negate(&ctx, &ng);
        poly_mul_int(&ctx, &m1, pa, x1);
        poly_add_int(&ctx, &ad1, pa, x1);
        poly_sub_int(&ctx, &sb1, pa, x1);
        poly_double_add(&ctx, &da1, pa, pb, &da2, pb, pb);
        poly_double_mul_int(&ctx, &dm1, pa, &dm2, pb, x1, x2);
        poly_double_add_int(&ctx, &di1, pa, &di2, pb, x1, x2);
        poly_mul_add(&ctx, &ma, pa, pb, pa);
        poly_copy(&ctx, &cl, pa);
        poly_clear(&ctx, &cl);
        poly_copy(&ctx, &cc, pb);
        struct {
          poly_t *p;
          const std::vector<uint64_t> *want;
          const char *name;
        } checks[] = {{&s, &w_sub, "poly_sub"},
                      {&ng, &w_neg, "poly_negate"},
                      {&m1, &w_mul1, "poly_mul_int"},
                      {&ad1, &w_add1, "poly_add_int"},
                      {&sb1, &w_sub1, "poly_sub_int"},
                      {&da1, &sum, "poly_double_add (first)"},
                      {&da2, &w_bb, "poly_double_add (second)"},
                      {&dm1, &w_mul1, "poly_double_mul_int (first)"},
                      {&dm2, &w_mul2, "poly_double_mul_int (second)"},
                      {&di1, &w_add1, "poly_double_add_int (first)"},
                      {&di2, &w_add2, "poly_double_add_int (second)"},
                      {&ma, &w_muladd, "poly_mul_add"},
                      {&cc, &b, "poly_copy"}};
        for (auto &ch : checks)
          if (poly_get_residues(&ctx, ch.p) != *ch.want) {
            std::fprintf(stderr, "%s mismatch\n", ch.name);
            return 1;
          }
        for (uint64_t v : poly_get_residues(&ctx, &cl))
          if (v != 0) {
            std::fprintf(stderr, "poly_clear left a nonzero residue\n");
            return 1;
          }
        for (poly_t *p : {&s, &ng, &m1, &ad1, &sb1, &da1, &da2, &dm1, &dm2, &di1, &di2, &ma, &cl,
                          &cc})
          poly_free(&ctx, p);
      }
      // HPS wrappers: Q -> B, Q -> QB then dropping Q, and t/Q simple scaling
      {
        poly_t xb, xqb, ts;
        poly_basis_extension_Q_to_B(&ctx, &xb, pa);
        poly_basis_extension_Q_to_QB(&ctx, &xqb, pa);
        std::vector<uint64_t> qb = poly_get_residues(&ctx, &xqb);
        std::vector<uint64_t> want_qb(a);
        want_qb.insert(want_qb.end(), a_b.begin(), a_b.end());
        bool ok = poly_get_residues(&ctx, &xb) == a_b && qb == want_qb;
        poly_discard_qbase(&ctx, &xqb);
        ok = ok && xqb.base == BBase && poly_get_residues(&ctx, &xqb) == a_b;
        poly_simple_scaling_tDivQ(&ctx, &ts, pa);
        ok = ok && ts.base == TBase && poly_get_residues(&ctx, &ts) == scaled;
        ok = ok && poly_get_residues(&ctx, pa) == a;  // the input is left as it was
        poly_free(&ctx, &xb);
        poly_free(&ctx, &xqb);
        poly_free(&ctx, &ts);
        if (!ok) {
          std::fprintf(stderr, "basis extension / discard / simple scaling mismatch\n");
          return 1;
        }
      }
      // one residue read back, and one residue copied over all the others
      {
        std::vector<uint64_t> r((size_t)n);  // 2N values
        poly_get_residue(&ctx, r.data(), pa, (int)sid);
        bool ok = std::equal(r.begin(), r.end(), a.begin() + sid * n);
        const std::string rs = poly_residue_to_string(&ctx, pa, (int)sid);
        ok = ok && rs.size() > 0 && rs.back() == ',' &&
             rs.substr(0, rs.find(',')) == std::to_string(a[sid * n]);
        poly_t x;
        poly_copy(&ctx, &x, pa);
        poly_select_residue(&ctx, &x, (int)sid);
        ok = ok && poly_get_residues(&ctx, &x) == sel;
        poly_free(&ctx, &x);
        if (!ok) {
          std::fprintf(stderr, "get_residue / select_residue mismatch\n");
          return 1;
        }
      }
      // word decomposition (test_bfv_mode.cu:492-546): sum_i rho(a)[i] xi(b)[i] = a b,
      // through poly_dot and through the reference's poly_mul_add loop
      {
        const int kk = CUDAEngine::get_n_residues(QBase);
        poly_t *A = new poly_t[kk];
        poly_t *B = new poly_t[kk];
        poly_rho_bfv(&ctx, A, pa);
        poly_xi_bfv(&ctx, B, pb);
        poly_t acc, acc2;
        poly_dot(&ctx, &acc, A, B, kk);
        poly_init(&ctx, &acc2);
        for (int i = 0; i < kk; i++) poly_mul_add(&ctx, &acc2, &A[i], &B[i], &acc2);
        const bool ok = poly_get_residues(&ctx, &acc) == prod && poly_get_residues(&ctx, &acc2) == prod;
        for (int i = 0; i < kk; i++) {
          poly_free(&ctx, &A[i]);
          poly_free(&ctx, &B[i]);
        }
        delete[] A;
        delete[] B;
        poly_free(&ctx, &acc);
        poly_free(&ctx, &acc2);
        if (!ok) {
          std::fprintf(stderr, "rho/xi/dot mismatch\n");
          return 1;
        }
      }
      poly_free(&ctx, pa);
      poly_free(&ctx, pb);
      poly_free(&ctx, pc);
      poly_free(&ctx, &c);
      poly_free(&ctx, &d);
      poly_free(&ctx, &e);
      delete pa;
      delete pb;
      delete pc;
    }
    CUDAEngine::destroy();
  } catch (const std::exception &ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  std::printf("facade ok\n");
  return 0;
}

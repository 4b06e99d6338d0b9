// CPU check of the modular primitives in cupoly_amd/csrc/modarith.hpp (the
// same inline functions the gfx950 kernels use, in their host forms: the
// carry-chained asm of the paired device forms is checked on the GPU by the
// parity tests) against unsigned __int128 %.
// usage: modarith_test <primes...>; prints "modarith ok" or the first failure.
#define CP_HD static inline
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../cupoly_amd/csrc/modarith.hpp"

static uint64_t rng_state = 0x9E3779B97F4A7C15ULL;
static uint64_t next64() {  // splitmix64
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int main(int argc, char **argv) {
  long checks = 0;
  for (int i = 1; i < argc; i++) {
    const uint64_t p = strtoull(argv[i], nullptr, 10);
    const cp_prime pr = cp_make_prime(p);
    std::vector<uint64_t> edge = {0, 1, 2, p - 1, p - 2, p, p + 1, 2 * p - 1, (1ULL << 63) - 1,
                                  1ULL << 63, ~0ULL, ~0ULL - 1};
    for (int it = 0; it < 400000; it++) {
      uint64_t a = it < 144 ? edge[it % 12] : next64();
      uint64_t w = it < 144 ? edge[(it / 12) % 12] % p : next64() % p;
      const uint64_t want = (uint64_t)(((cp_u128)a * w) % p);
      const uint64_t wp = cp_shoup_factor(w, p);
      const uint64_t lazy = cp_shoup_lazy(a, w, wp, p);
      if (lazy >= 2 * p || lazy % p != want || cp_shoup(a, w, wp, p) != want) {
        std::printf("shoup FAIL p=%llu a=%llu w=%llu\n", (unsigned long long)p,
                    (unsigned long long)a, (unsigned long long)w);
        return 1;
      }
      const uint64_t b = it < 144 ? w : next64() >> 1;  // b < 2^63
      const uint64_t want2 = (uint64_t)(((cp_u128)a * b) % p);
      if (cp_mulmod(a, b, pr) != want2) {
        std::printf("mulmod FAIL p=%llu a=%llu b=%llu\n", (unsigned long long)p,
                    (unsigned long long)a, (unsigned long long)b);
        return 1;
      }
      // canonical a times any 64-bit b (execute_mul_dgt_u64)
      const uint64_t a3 = it < 144 ? edge[it % 12] % p : next64() % p;
      const uint64_t b3 = it < 144 ? edge[(it / 12) % 12] : next64();
      if (cp_mulmod(a3, b3, pr) != (uint64_t)(((cp_u128)a3 * b3) % p)) {
        std::printf("mulmod (b < 2^64) FAIL p=%llu a=%llu b=%llu\n", (unsigned long long)p,
                    (unsigned long long)a3, (unsigned long long)b3);
        return 1;
      }
      // the paired forms the kernels use: two products of variables
      // (cp_fold127 on the 128-bit words), the column form of the 64-bit high
      // product, and two Shoup products
      uint64_t m1, m2;
      cp_mulmod2(a, b, a3, b3, pr, m1, m2);
      if (m1 != want2 || m2 != (uint64_t)(((cp_u128)a3 * b3) % p)) {
        std::printf("mulmod2 FAIL p=%llu\n", (unsigned long long)p);
        return 1;
      }
      if (cp_mulhi_col(a, wp) != cp_mulhi(a, wp) || cp_mulhi_col(b3, a) != cp_mulhi(b3, a)) {
        std::printf("mulhi_col FAIL a=%llu\n", (unsigned long long)a);
        return 1;
      }
      uint64_t s1, s2;
      cp_shoup2(a, w, wp, b3, w, wp, p, s1, s2);
      if (s1 != want || s2 != (uint64_t)(((cp_u128)b3 * w) % p)) {
        std::printf("shoup2 FAIL p=%llu\n", (unsigned long long)p);
        return 1;
      }
      const uint64_t r = next64() % (2 * p);
      if (cp_red2(r, p) != r % p) {
        std::printf("red2 FAIL\n");
        return 1;
      }
      // exact sums of 1..40 products of any 64-bit words (all-ones words in
      // the edge iterations, so that the top word carries)
      if (it % 16 == 0) {
        cp_acc acc;
        cp_acc_init(acc);
        uint64_t want4 = 0;
        const int terms = 1 + (int)(next64() % 40);
        for (int j = 0; j < terms; j++) {
          const uint64_t x = it < 144 ? ~0ULL : next64(), y = it < 144 ? ~0ULL : next64();
          cp_acc_mac(acc, x, y);
          want4 = (uint64_t)(((cp_u128)want4 + (uint64_t)(((cp_u128)x * y) % p)) % p);
        }
        if (cp_acc_reduce(acc, p) != want4) {
          std::printf("acc FAIL p=%llu terms=%d\n", (unsigned long long)p, terms);
          return 1;
        }
        checks++;
      }
      checks += 4;
    }
  }
  std::printf("modarith ok (%ld checks)\n", checks);
  return 0;
}

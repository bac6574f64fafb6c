// Host check of the merge's claimed-position bitmaps (mlp_claims.h, the code
// the kernels include) against a byte-per-position brute force: random sets
// of disjoint spans, then random interval and point queries.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#define __host__
#define __device__
#define MLP_CLAIM_OR(p, m) (*(p) |= (m))
#include "mlp_claims.h"

int main() {
  std::mt19937 rng(1);
  long checks = 0, bad = 0;
  for (int trial = 0; trial < 3000; ++trial) {
    const int n = 5 + (int)(rng() % 5000);
    std::vector<uint32_t> bits(claim_words(n), 0);
    std::vector<char> ref(n, 0);
    const int nspans = (int)(rng() % 40);
    for (int k = 0; k < nspans; ++k) {
      const int a = (int)(rng() % n);
      const int len = (rng() % 4 == 0) ? (int)(rng() % n) : (int)(rng() % 70);
      const int b = std::min(n - 1, a + len);
      bool free_span = true;
      for (int p = a; p <= b; ++p) free_span = free_span && !ref[p];
      if (!free_span) continue;  // kept spans are disjoint
      claims_set(bits.data(), n, a, b);
      for (int p = a; p <= b; ++p) ref[p] = 1;
    }
    for (int q = 0; q < 300; ++q) {
      const int a = (int)(rng() % n);
      const int len = (rng() % 3 == 0) ? (int)(rng() % n) : (int)(rng() % 80);
      const int b = std::min(n - 1, a + len);
      bool r = false;
      for (int p = a; p <= b; ++p) r = r || ref[p];
      ++checks;
      if (claims_any(bits.data(), n, a, b) != r) ++bad;
      const int p = (int)(rng() % n);
      ++checks;
      if (claimed(bits.data(), n, p) != (ref[p] != 0)) ++bad;
    }
  }
  std::printf("checks %ld mismatches %ld\n", checks, bad);
  return bad != 0;
}

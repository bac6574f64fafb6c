// mlp_claims.h -- positions claimed by the moves a merge keeps.
//
// Included by the HIP kernels (inside their anonymous namespace, with
// MLP_CLAIM_OR bound to a device atomicOr) and by a host harness of the CPU
// tests (bound to a plain |=), so the same code is checked against brute
// force on the host.
#ifndef MLP_CLAIMS_H
#define MLP_CLAIMS_H

// Positions claimed by kept moves (kept spans are disjoint intervals):
//   level 0  one bit per position (claim_words0(n) words), set for the
//            positions of a span's first and last word;
//   some     one bit per level-0 word: the word holds a claimed position;
//   full     one bit per level-0 word: the word lies inside one span, between
//            its first and last word (its level-0 bits are not set).
// Setting or testing a span of m positions touches O(1 + m / 1024) words.
__host__ __device__ constexpr int claim_words0(int n) { return (n + 32) / 32 + 1; }
__host__ __device__ constexpr int claim_words1(int n) { return claim_words0(n) / 32 + 1; }
__host__ __device__ constexpr int claim_words(int n) {
  return claim_words0(n) + 2 * claim_words1(n);
}

__host__ __device__ inline bool range_any(const uint32_t* bits, int a, int b) {
  const int wa = a >> 5, wb = b >> 5;
  for (int w = wa; w <= wb; ++w) {
    uint32_t m = 0xffffffffu;
    if (w == wa) m &= 0xffffffffu << (a & 31);
    if (w == wb) m &= 0xffffffffu >> (31 - (b & 31));
    if (bits[w] & m) return true;
  }
  return false;
}

__host__ __device__ inline void range_set(uint32_t* bits, int a, int b) {
  const int wa = a >> 5, wb = b >> 5;
  for (int w = wa; w <= wb; ++w) {
    uint32_t m = 0xffffffffu;
    if (w == wa) m &= 0xffffffffu << (a & 31);
    if (w == wb) m &= 0xffffffffu >> (31 - (b & 31));
    MLP_CLAIM_OR(&bits[w], m);
  }
}

__host__ __device__ inline bool full_word(const uint32_t* bits, int n, int w) {
  const uint32_t* full = bits + claim_words0(n) + claim_words1(n);
  return (full[w >> 5] >> (w & 31)) & 1u;
}

// Is position p claimed?  (both words read before either is tested: no
// short-circuit, so the two loads are in flight together)
__host__ __device__ inline bool claimed(const uint32_t* bits, int n, int p) {
  const uint32_t w = bits[p >> 5];
  const bool f = full_word(bits, n, p >> 5);
  return (((w >> (p & 31)) & 1u) != 0) | f;
}

// Is any position of [a, b] claimed?  (n: the tour's positions; the four
// end words are read before any is tested)
__host__ __device__ inline bool claims_any(const uint32_t* bits, int n, int a, int b) {
  const int wa = a >> 5, wb = b >> 5;
  const uint32_t ma = 0xffffffffu << (a & 31), mb = 0xffffffffu >> (31 - (b & 31));
  const uint32_t xa = bits[wa], xb = bits[wb];
  const bool fa = full_word(bits, n, wa), fb = full_word(bits, n, wb);
  if (wa == wb) return ((xa & ma & mb) != 0) | fa;
  if (((xa & ma) != 0) | ((xb & mb) != 0) | fa | fb) return true;
  // the words between: any claimed position at all
  return wb - wa > 1 && range_any(bits + claim_words0(n), wa + 1, wb - 1);
}

__host__ __device__ inline void claims_set(uint32_t* bits, int n, int a, int b) {
  const int wa = a >> 5, wb = b >> 5;
  if (wb - wa <= 1) {
    range_set(bits, a, b);
  } else {
    MLP_CLAIM_OR(&bits[wa], 0xffffffffu << (a & 31));
    MLP_CLAIM_OR(&bits[wb], 0xffffffffu >> (31 - (b & 31)));
    range_set(bits + claim_words0(n) + claim_words1(n), wa + 1, wb - 1);
  }
  range_set(bits + claim_words0(n), wa, wb);
}

#endif  // MLP_CLAIMS_H

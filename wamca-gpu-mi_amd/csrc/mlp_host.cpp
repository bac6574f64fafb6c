// mlp_host.cpp -- host-side inputs of the local search: MT19937, synthetic
// instances, TSPLIB EUC_2D files and start tours.  Plain C++ (no device code);
// exported through include/mlp_host.h.
//
//   MTRandom                mtrand.cpp:51-121, mtrand.h:168-195, 294-312
//   TSPLIB EUC_2D + inear   tsplib.cpp:236-240, 458-497, 604-687;
//                           mlproblem.cpp:77-151 (coordShift)
//   MLSolution::random      mlsolution.cpp:753-874 (alpha RCL over sorted CL)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "mlp_host.h"

namespace {

struct MT {
  uint32_t s[624];
  int p = 624;
  explicit MT(uint32_t seed) {
    s[0] = seed;
    for (int i = 1; i < 624; ++i) s[i] = 1812433253U * (s[i - 1] ^ (s[i - 1] >> 30)) + (uint32_t)i;
  }
  void regen() {
    static const uint32_t mag[2] = {0u, 0x9908b0dfU};
    for (int i = 0; i < 624; ++i) {
      const uint32_t y = (s[i] & 0x80000000U) | (s[(i + 1) % 624] & 0x7fffffffU);
      s[i] = s[(i + 397) % 624] ^ (y >> 1) ^ mag[y & 1u];
    }
    p = 0;
  }
  uint32_t next() {
    if (p >= 624) regen();
    uint32_t x = s[p++];
    x ^= x >> 11;
    x ^= (x << 7) & 0x9d2c5680U;
    x ^= (x << 15) & 0xefc60000U;
    return x ^ (x >> 18);
  }
  uint32_t upto(uint32_t max) { return (uint32_t)(((max + 1.0) * next()) / 4294967296.0); }
  double open01() { return ((double)next() + 0.5) * (1.0 / 4294967296.0); }
  double co01() { return (double)next() * (1.0 / 4294967296.0); }
};

int rnd_half_away(double v) { return v >= 0 ? (int)(v + 0.5) : (int)(v - 0.5); }

int host_dist(int x1, int y1, int x2, int y2, int round) {
  const float s = (float)((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
  return (int)(sqrtf(s) + (round ? 0.5F : 0.0F));
}

}  // namespace

extern "C" {

uint32_t mlp_mt_first(uint32_t seed, int32_t skip) {
  MT m(seed);
  for (int q = 0; q < skip; ++q) m.next();
  return m.next();
}

int mlp_gen_uniform(int32_t n, uint32_t seed, int32_t maxc, int32_t* xs, int32_t* ys) {
  if (n < 1 || maxc < 0 || !xs || !ys) return -1;
  MT m(seed);
  for (int p = 0; p < n; ++p) {
    xs[p] = (int32_t)m.upto((uint32_t)maxc);
    ys[p] = (int32_t)m.upto((uint32_t)maxc);
  }
  return 0;
}

int mlp_gen_clustered(int32_t n, uint32_t seed, int32_t ncent, int32_t span, double sigma,
                      int32_t* xs, int32_t* ys) {
  if (n < 1 || ncent < 1 || span < 0 || !xs || !ys) return -1;
  MT m(seed);
  std::vector<int> cx((size_t)ncent), cy((size_t)ncent);
  for (int c = 0; c < ncent; ++c) {
    cx[(size_t)c] = (int)m.upto((uint32_t)span);
    cy[(size_t)c] = (int)m.upto((uint32_t)span);
  }
  int xmin = INT32_MAX, ymin = INT32_MAX;
  for (int p = 0; p < n; ++p) {
    const int c = (int)m.upto((uint32_t)(ncent - 1));
    const double u1 = m.open01(), u2 = m.open01();
    const double r = std::sqrt(-2.0 * std::log(u1)), a = 6.283185307179586 * u2;
    xs[p] = cx[(size_t)c] + rnd_half_away(sigma * r * std::cos(a));
    ys[p] = cy[(size_t)c] + rnd_half_away(sigma * r * std::sin(a));
    xmin = std::min(xmin, (int)xs[p]);
    ymin = std::min(ymin, (int)ys[p]);
  }
  const int sx = xmin < 0 ? -xmin : 0, sy = ymin < 0 ? -ymin : 0;
  for (int p = 0; p < n; ++p) {
    xs[p] += sx;
    ys[p] += sy;
  }
  return 0;
}

// n = round(exp(U[ln lo, ln hi])) for a job stream; one draw per call index.
int32_t mlp_gen_loguniform_n(uint32_t seed, int32_t index, int32_t lo, int32_t hi) {
  MT m(seed);
  for (int q = 0; q < index; ++q) m.next();
  const double u = m.co01();
  const double v = std::exp(std::log((double)lo) + u * (std::log((double)hi) - std::log((double)lo)));
  int n = (int)std::lround(v);
  return std::min(hi, std::max(lo, n));
}

int mlp_tsplib_load(const char* path, int32_t* xs, int32_t* ys, int32_t cap, int32_t* n_out) {
  if (!path || !n_out) return -1;
  std::ifstream in(path);
  if (!in.is_open()) return -2;
  std::string line;
  int dim = -1;
  bool euc = false, coords = false;
  int got = 0;
  while (std::getline(in, line)) {
    std::string key = line;
    const size_t colon = key.find(':');
    std::string val;
    if (colon != std::string::npos) {
      val = key.substr(colon + 1);
      key = key.substr(0, colon);
    }
    auto trim = [](std::string& s) {
      s.erase(0, s.find_first_not_of(" \t\r"));
      s.erase(s.find_last_not_of(" \t\r") + 1);
    };
    trim(key);
    trim(val);
    if (!coords) {
      if (key == "DIMENSION") dim = std::atoi(val.c_str());
      else if (key == "EDGE_WEIGHT_TYPE") euc = (val == "EUC_2D");
      else if (key.rfind("NODE_COORD_SECTION", 0) == 0) {
        if (!euc || dim < 1) return -3;  // only EUC_2D (mlproblem.cpp:94-95)
        if (dim > cap) {
          *n_out = dim;
          return -4;
        }
        coords = true;
      }
      continue;
    }
    if (key.rfind("EOF", 0) == 0) break;
    std::istringstream ls(line);
    int id;
    float x, y;  // tsplib.h:56 Float = float
    if (!(ls >> id >> x >> y)) continue;
    if (got >= dim) break;
    // inear in float (tsplib.cpp:236-240)
    xs[got] = x >= 0 ? (int)(x + 0.5F) : (int)(x - 0.5F);
    ys[got] = y >= 0 ? (int)(y + 0.5F) : (int)(y - 0.5F);
    ++got;
  }
  if (!coords || got != dim) return -5;
  *n_out = dim;
  return 0;
}

int mlp_shift_nonnegative(int32_t n, int32_t* xs, int32_t* ys) {
  int xmin = INT32_MAX, ymin = INT32_MAX;
  for (int p = 0; p < n; ++p) {
    xmin = std::min(xmin, (int)xs[p]);
    ymin = std::min(ymin, (int)ys[p]);
  }
  const int sx = xmin < 0 ? -xmin : 0, sy = ymin < 0 ? -ymin : 0;
  for (int p = 0; p < n; ++p) {
    xs[p] += sx;
    ys[p] += sy;
  }
  return 0;
}

int mlp_host_dist(int32_t x1, int32_t y1, int32_t x2, int32_t y2, int32_t round) {
  return host_dist(x1, y1, x2, y2, round);
}

// Latency of a tour of N positions over client coordinates, int64
// (mlsolution.cpp:215-228 without the uint32 wrap).
int64_t mlp_host_latency(const int32_t* xs, const int32_t* ys, int32_t round, const int32_t* tour,
                         int32_t N) {
  int64_t L = 0;
  for (int p = 1; p < N; ++p) {
    const int a = tour[p - 1], b = tour[p];
    L += (int64_t)(N - p) * host_dist(xs[a], ys[a], xs[b], ys[b], round);
  }
  return L;
}

int mlp_tour_perm(uint32_t seed, int32_t n, int32_t* tour) {
  if (n < 2 || !tour) return -1;
  MT m(seed);
  tour[0] = 0;
  for (int p = 1; p < n; ++p) tour[p] = p;
  // MTRandom::shuffle over positions 1..n-1: j = rand(len-i-1) (mtrand.h:301-312)
  int32_t* d = tour + 1;
  const uint32_t len = (uint32_t)(n - 1);
  for (uint32_t i = 0; i + 1 < len; ++i) {
    const uint32_t j = m.upto(len - i - 1);
    std::swap(d[i], d[j]);
  }
  tour[n] = 0;
  return 0;
}

// Greedy randomized construction with an alpha-RCL over the candidates sorted
// by distance to the last client; the state of an MT seeded with `seed` is
// advanced by `skip` constructions first (the reference draws consecutive
// start tours from one generator).
int mlp_tour_random(const int32_t* xs, const int32_t* ys, int32_t n, int32_t round, uint32_t seed,
                    int32_t skip, float alpha, int32_t* tour) {
  if (n < 2 || !xs || !ys || !tour) return -1;
  MT m(seed);
  std::vector<int> cand((size_t)n - 1);
  std::vector<int32_t> dl((size_t)n);
  for (int rep = 0; rep <= skip; ++rep) {
    for (int c = 1; c < n; ++c) cand[(size_t)c - 1] = c;
    unsigned count = (unsigned)(n - 1);
    int last = 0, pos = 0;
    tour[pos++] = 0;
    while (count > 0) {
      unsigned pick = 0;
      if (count > 1) {
        const unsigned rcl = (unsigned)std::ceil(alpha * count);
        pick = rcl > 0 ? m.upto(rcl - 1) : 0;
        for (unsigned q = 0; q < count; ++q) {
          const int c = cand[q];
          dl[(size_t)c] = host_dist(xs[last], ys[last], xs[c], ys[c], round);
        }
        int32_t* dd = dl.data();
        if (alpha < 0.5F) {
          std::sort(cand.begin(), cand.begin() + count, [dd](int a, int b) { return dd[a] > dd[b]; });
          pick = count - pick - 1;
        } else {
          std::sort(cand.begin(), cand.begin() + count, [dd](int a, int b) { return dd[a] < dd[b]; });
        }
      }
      const int cur = cand[pick];
      tour[pos++] = cur;
      last = cur;
      for (unsigned q = pick; q + 1 < count; ++q) cand[q] = cand[q + 1];
      --count;
    }
    tour[pos] = 0;
  }
  return 0;
}

// Same randomized construction with the candidate of rank `pick` found by
// selection instead of a full sort, ties in distance ordered by client id:
// O(n^2) instead of O(n^2 log n), for synthetic tours of large instances.
// (Tie order differs from std::sort's, so tours differ from
// mlp_tour_random's wherever distances tie.)
int mlp_tour_random_fast(const int32_t* xs, const int32_t* ys, int32_t n, int32_t round,
                         uint32_t seed, float alpha, int32_t* tour) {
  if (n < 2 || !xs || !ys || !tour) return -1;
  MT m(seed);
  std::vector<int> cand((size_t)n - 1);
  std::vector<int64_t> key((size_t)n - 1);
  for (int c = 1; c < n; ++c) cand[(size_t)c - 1] = c;
  unsigned count = (unsigned)(n - 1);
  int last = 0, pos = 0;
  tour[pos++] = 0;
  while (count > 0) {
    unsigned pick = 0;
    if (count > 1) {
      const unsigned rcl = (unsigned)std::ceil(alpha * count);
      pick = rcl > 0 ? m.upto(rcl - 1) : 0;
      if (alpha < 0.5F) pick = count - pick - 1;
    }
    for (unsigned q = 0; q < count; ++q) {
      const int c = cand[q];
      key[q] = ((int64_t)host_dist(xs[last], ys[last], xs[c], ys[c], round) << 32) | (uint32_t)c;
    }
    std::nth_element(key.begin(), key.begin() + pick, key.begin() + count);
    const int cur = (int)(uint32_t)key[pick];
    tour[pos++] = cur;
    last = cur;
    for (unsigned q = 0; q < count; ++q)
      if (cand[q] == cur) {
        cand[q] = cand[count - 1];
        break;
      }
    --count;
  }
  tour[pos] = 0;
  return 0;
}

}  // extern "C"

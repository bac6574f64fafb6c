// =============================================================================
//  mlp_oracle.cpp -- CPU ORACLE (test infrastructure, NOT the product)
//
//  A plain, sequential restatement of the reference's Minimum Latency Problem
//  local-search evaluation (LabIC-UFF/wamca-gpu-mi).  Only tests/, the smoke()
//  entry point and bench.py's cpu_baseline leg may load this library, and then
//  only as the checker of the HIP path (or as the timed CPU baseline).  The
//  product library never links, loads or calls it.
//
//  Every function cites the reference file:line it restates.  All latency and
//  delta arithmetic is int64 (the reference uses uint32/int32 with wrap-around;
//  the two agree modulo 2^32, and exactly while |delta| < 2^31).
//
//  Pins (each checked in tests/test_oracle_*.py):
//   * MT19937          : tests/mtrand.Test.cpp:16-47 raw outputs and %10 runs
//   * distance         : glibc sqrtf, exhaustive s < 2^24 vs a double-precision
//                        restatement; 21 logged start-tour latencies
//   * constructor      : berlin52 seed 0 alpha 0.5 -> 494309
//   * ADS              : closed form from prefix sums == ldsUpdate recurrences
//   * move deltas      : ADS concatenation == positional (ODC) form ==
//                        brute force L(apply(tour)) - L(tour), every move
//   * per-row / merge  : results/*.txt best move, improving-row count and
//                        merged (sum, count); berlin52 ODC swap pass 367403
//   * TOTAL entries    : results/*.txt TOTAL sections: best entry, improving
//                        entry count and GPU-GPU merged (sum, count) for the
//                        logged block size
//   * DVND loop        : no reference implementation; pinned by brute force
//                        local optimality of the final tour (parity unpinned
//                        against reference outputs -- none exist)
// =============================================================================
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

// ---- MT19937 (mtrand.cpp:51-61 seed, :99-121 genState; mtrand.h:168-182) ----
struct Mt {
  uint32_t s[624];
  int p;
};

void mt_seed(Mt* m, uint32_t seed) {
  m->s[0] = seed;
  for (int i = 1; i < 624; ++i)
    m->s[i] = 1812433253U * (m->s[i - 1] ^ (m->s[i - 1] >> 30)) + (uint32_t)i;
  m->p = 624;
}

void mt_gen(Mt* m) {
  static const uint32_t mag[2] = {0u, 0x9908b0dfU};
  int i;
  uint32_t y;
  for (i = 0; i < 624 - 397; ++i) {
    y = (m->s[i] & 0x80000000U) | (m->s[i + 1] & 0x7fffffffU);
    m->s[i] = m->s[i + 397] ^ (y >> 1) ^ mag[y & 1u];
  }
  for (; i < 623; ++i) {
    y = (m->s[i] & 0x80000000U) | (m->s[i + 1] & 0x7fffffffU);
    m->s[i] = m->s[i + (397 - 624)] ^ (y >> 1) ^ mag[y & 1u];
  }
  y = (m->s[623] & 0x80000000U) | (m->s[0] & 0x7fffffffU);
  m->s[623] = m->s[396] ^ (y >> 1) ^ mag[y & 1u];
  m->p = 0;
}

uint32_t mt_rand(Mt* m) {
  if (m->p >= 624) mt_gen(m);
  uint32_t x = m->s[m->p++];
  x ^= (x >> 11);
  x ^= (x << 7) & 0x9d2c5680U;
  x ^= (x << 15) & 0xefc60000U;
  return x ^ (x >> 18);
}

// mtrand.h:192-195: uniform integer in [0,max], computed in double.
uint32_t mt_randmax(Mt* m, uint32_t max) {
  return (uint32_t)(((max + 1.0) * mt_rand(m)) / (4294967295.0 + 1.0));
}

// mtrand.h:294-297
double mt_real_oo(Mt* m) { return (((double)mt_rand(m)) + 0.5) * (1.0 / 4294967296.0); }
double mt_real_co(Mt* m) { return ((double)mt_rand(m)) * (1.0 / 4294967296.0); }

// ---- distance (mlproblem.cpp:88, 159-162) --------------------------------------
// s is formed in int, stored to a float (round to nearest), sqrtf is the
// correctly rounded IEEE square root, dr in {0, 0.5} is added in float and the
// sum truncated.
inline int dist_xy(int x1, int y1, int x2, int y2, int round) {
  float s = (float)((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
  float dr = round ? 0.5F : 0.0F;
  return (int)(sqrtf(s) + dr);
}

inline int64_t D(const int32_t* W, int n, const int32_t* t, int a, int b) {
  return W[(int64_t)t[a] * n + t[b]];
}

// Prefix time T[p] = sum_{q<=p} w[q] (mlsolution.cpp:451-458), int64.
void prefix_time(const int32_t* W, int n, const int32_t* t, int N, int64_t* T) {
  T[0] = 0;
  for (int p = 1; p < N; ++p) T[p] = T[p - 1] + D(W, n, t, p - 1, p);
}

// ADS cost matrix C (mlsolution.cpp:460-478), int64, row-major N x N.
void ads_cost(const int64_t* T, int N, int64_t* C) {
  for (int i = 0; i < N; ++i) {
    int64_t* r = C + (int64_t)i * N;
    r[i] = 0;
    for (int j = i + 1; j < N; ++j) r[j] = r[j - 1] + (T[j] - T[i]);
  }
  for (int i = N - 1; i >= 0; --i) {
    int64_t* r = C + (int64_t)i * N;
    for (int j = i - 1; j >= 0; --j) r[j] = r[j + 1] + (T[i] - T[j]);
  }
}

// Latency L = sum_{p=1}^{N-1} (N-p) w[p] (mlsolution.cpp:215-228).
int64_t latency(const int32_t* W, int n, const int32_t* t, int N) {
  int64_t c = 0;
  for (int p = 1; p < N; ++p) c += (int64_t)(N - p) * D(W, n, t, p - 1, p);
  return c;
}

// ---- move ids (mlads.h:144-176) -------------------------------------------------
// nb 0 = swap, 1 = 2-opt, 2+k-1 = or-opt-k.
inline int nb_k(int nb) { return nb >= 2 ? nb - 1 : 0; }

// ---- row maps (reference launch grids and thread->move index maps) -------------
// rowmap 0 = ADS kernels (mlkswap.cu:205-210, mlk2opt.cu:168-174,
// mlkoropt.cu:183-247), rowmap 1 = ODC kernels (mlkswap.cu:764-768,
// mlk2opt.cu:559-563, mlkoropt.cu:696-723).  tour = 1 (costTour) throughout.
int rows_of(int nb, int rowmap, int N) {
  if (nb == 0) return rowmap == 0 ? N - 4 : ((N + 1) / 2) - 1;  // mlkswap.cu:991,1003
  if (nb == 1) return (N + 1) / 2;                              // mlk2opt.cu:661,673
  return N - nb_k(nb) - 2;                                      // mlkoropt.cu:819,831
}

int ctx_of(int nb, int rowmap, int N) {
  if (nb == 0) return rowmap == 0 ? N : N - 4;
  if (nb == 1) return N - 3;
  int k = nb_k(nb);
  return rowmap == 0 ? N - k - 1 : N - k - 2;
}

// Returns true and sets (i,j) when (row,ctx) enumerates a move.
bool move_of(int nb, int rowmap, int N, int by, int ctx, int* pi, int* pj) {
  int i, j;
  if (nb == 0) {
    if (rowmap == 0) {  // mlkswap.cu:205-210
      i = by + 1;
      j = ctx;
      if (!((j > i + 1) && (j < N - 1))) return false;
    } else {  // mlkswap.cu:764-768
      if (!(ctx < N - 4)) return false;
      bool k = ctx >= by;
      i = k ? by + 1 : N - 1 - by - 2;
      j = k ? ctx + 3 : N - 1 - ctx - 1;
    }
  } else if (nb == 1) {  // mlk2opt.cu:168-174 (== ODC :559-563 for tour=1)
    if (!(ctx < N - 3)) return false;
    bool k = ctx >= by;
    i = k ? ctx - by + 1 : by - ctx;
    j = k ? ctx + 2 : N - ctx - 2;
  } else {
    int k = nb_k(nb);
    if (rowmap == 0) {  // mlkoropt.cu:183-247
      if (!(ctx < N - k - 1)) return false;
      if (ctx > by) {
        i = by + 1;
        j = ctx + 1;
      } else {
        i = by + 2;
        j = ctx + 1;
      }
    } else {  // mlkoropt.cu:696-723
      if (!(ctx < N - 1 - k - 1)) return false;
      if (ctx >= by) {
        i = ctx - by + 1;
        j = ctx + 2;
      } else {
        i = N - ctx - k - 1;
        j = by - ctx;
      }
    }
  }
  *pi = i;
  *pj = j;
  return true;
}

// ---- move cost by O(1) subsequence concatenation over the ADS -------------------
// Restates the per-thread cost of the ADS kernels: swap mlkswap.cu:210-278,
// 2-opt mlk2opt.cu:168-219, or-opt mlkoropt.cu:187-300.
// Cm(a, b) is the ADS cost entry C[a][b], Dd(a, b) the distance between the
// clients at tour positions a and b; T is the prefix time and L the latency.
template <class CM, class DD>
int64_t delta_concat(int nb, int N, const int64_t* T, int64_t L, int i, int j, CM Cm, DD Dd) {
  const int last = N - 1;
  int64_t dist, cost, time, wait;
  if (nb == 0) {
    dist = Dd(i - 1, j);
    cost = Cm(0, i - 1) + 1 * (T[i - 1] + dist);
    time = T[i - 1] + dist;
    dist = Dd(j, i + 1);
    wait = j - i - 1;
    cost += Cm(i + 1, j - 1) + wait * (time + dist);
    time += (T[j - 1] - T[i + 1]) + dist;
    dist = Dd(j - 1, i);
    cost += time + dist;
    time += dist;
    if (j + 1 <= last) {
      dist = Dd(i, j + 1);
      wait = last - j;
      cost += Cm(j + 1, last) + wait * (time + dist);
    }
    return cost - L;
  }
  if (nb == 1) {
    dist = Dd(i - 1, j);
    wait = j - i + 1;
    cost = Cm(0, i - 1) + Cm(j, i) + wait * (T[i - 1] + dist);
    time = T[i - 1] + (T[j] - T[i]) + dist;
    if (j + 1 <= last) {
      dist = Dd(i, j + 1);
      wait = last - j;
      cost += Cm(j + 1, last) + wait * (time + dist);
    }
    return cost - L;
  }
  const int k = nb_k(nb);
  if (i < j) {
    dist = Dd(i - 1, i + k);
    wait = j - i;
    cost = Cm(0, i - 1) + Cm(i + k, j + k - 1) + wait * (T[i - 1] + dist);
    time = T[i - 1] + (T[j + k - 1] - T[i + k]) + dist;
    dist = Dd(j + k - 1, i);
    cost += Cm(i, i + k - 1) + (int64_t)k * (time + dist);
    time += (T[i + k - 1] - T[i]) + dist;
    if (j + k <= last) {
      dist = Dd(i + k - 1, j + k);
      wait = last - j - k + 1;
      cost += Cm(j + k, last) + wait * (time + dist);
    }
  } else {
    dist = Dd(j - 1, i);
    cost = Cm(0, j - 1) + Cm(i, i + k - 1) + (int64_t)k * (T[j - 1] + dist);
    time = T[j - 1] + (T[i + k - 1] - T[i]) + dist;
    dist = Dd(i + k - 1, j);
    wait = i - j;
    cost += Cm(j, i - 1) + wait * (time + dist);
    time += (T[i - 1] - T[j]) + dist;
    if (i + k <= last) {
      dist = Dd(i - 1, i + k);
      wait = last - i - k + 1;
      cost += Cm(i + k, last) + wait * (time + dist);
    }
  }
  return cost - L;
}

int64_t delta_ads(int nb, const int32_t* W, int n, const int32_t* t, int N, const int64_t* T,
                  const int64_t* C, int64_t L, int i, int j) {
  return delta_concat(
      nb, N, T, L, i, j, [&](int a, int b) { return C[(int64_t)a * N + b]; },
      [&](int a, int b) { return D(W, n, t, a, b); });
}

// ---- the ADS in O(N) prefix form (closed form of ldsUpdate) ---------------------
// With S[p] = sum_{q<=p} T[q] (S[-1] = 0) the recurrences of
// mlsolution.cpp:460-478 sum to
//   C[a][b] = (S[b] - S[a]) - (b - a) T[a]              for a < b,
//   C[a][b] = (a - b) T[a] - (S[a-1] - S[b-1])          for a > b,   C[a][a] = 0.
// Distances are computed from tour-ordered coordinates (no n x n matrix), so a
// tour of N positions needs O(N) memory: the path for instances whose C or W
// matrices would not fit (n = 18000: 2.6 GB of C per evaluation).
struct PrefixAds {
  int N = 0, round = 0;
  std::vector<int32_t> px, py;
  std::vector<int64_t> T, S;
  int64_t L = 0;
  void build(const int32_t* xs, const int32_t* ys, int rnd, const int32_t* t, int n_pos) {
    N = n_pos;
    round = rnd;
    px.resize(N);
    py.resize(N);
    T.assign(N, 0);
    S.assign(N, 0);
    for (int p = 0; p < N; ++p) {
      px[p] = xs[t[p]];
      py[p] = ys[t[p]];
    }
    L = 0;
    for (int p = 1; p < N; ++p) {
      const int64_t w = d(p - 1, p);
      T[p] = T[p - 1] + w;
      L += (int64_t)(N - p) * w;
    }
    S[0] = T[0];
    for (int p = 1; p < N; ++p) S[p] = S[p - 1] + T[p];
  }
  int64_t d(int a, int b) const { return dist_xy(px[a], py[a], px[b], py[b], round); }
  int64_t C(int a, int b) const {
    if (a < b) return (S[b] - S[a]) - (int64_t)(b - a) * T[a];
    if (a > b) return (int64_t)(a - b) * T[a] - ((a > 0 ? S[a - 1] : 0) - (b > 0 ? S[b - 1] : 0));
    return 0;
  }
  int64_t delta(int nb, int i, int j) const {
    return delta_concat(
        nb, N, T.data(), L, i, j, [&](int a, int b) { return C(a, b); },
        [&](int a, int b) { return d(a, b); });
  }
};

// ---- move cost by positional edge weights (ODC kernels) ------------------------
// swap mlkswap.cu:770-785, 2-opt mlk2opt.cu:565-571, or-opt mlkoropt.cu:704-736.
int64_t delta_odc(int nb, const int32_t* W, int n, const int32_t* t, int N, int i, int j) {
  auto w = [&](int p) -> int64_t { return D(W, n, t, p - 1, p); };
  const int64_t sz = N;
  int64_t cost;
  if (nb == 0) {
    cost = (sz - i) * (D(W, n, t, i - 1, j) - w(i));
    cost += (sz - i - 1) * (D(W, n, t, j, i + 1) - w(i + 1));
    cost += (sz - j) * (D(W, n, t, j - 1, i) - w(j));
    if (j + 1 < N) cost += (sz - j - 1) * (D(W, n, t, i, j + 1) - w(j + 1));
    return cost;
  }
  if (nb == 1) {
    cost = (sz - i) * (D(W, n, t, i - 1, j) - w(i));
    if (j + 1 < N) cost += (sz - j - 1) * (D(W, n, t, i, j + 1) - w(j + 1));
    for (int k = 1; k <= j - i; ++k) cost += (sz - i - k) * (w(j - k + 1) - w(i + k));
    return cost;
  }
  const int64_t k = nb_k(nb);
  int64_t s;
  if (i < j) {
    cost = (sz - i) * (D(W, n, t, i - 1, i + k) - w(i));
    cost += (sz - j - k) * (D(W, n, t, i + k - 1, j + k) - w(j + k));
    cost += (sz - j) * D(W, n, t, j + k - 1, i) - (sz - i - k) * w(i + k);
    s = 0;
    for (int l = i + 1; l < i + k; ++l) s += w(l);
    cost += (int64_t)(i - j) * s;
    s = 0;
    for (int l = i + k + 1; l < j + k; ++l) s += w(l);
    cost += k * s;
  } else {
    cost = (sz - j) * (D(W, n, t, j - 1, i) - w(j));
    cost += (sz - i - k) * (D(W, n, t, i - 1, i + k) - w(i + k));
    cost += (sz - j - k) * D(W, n, t, i + k - 1, j) - (sz - i) * w(i);
    s = 0;
    for (int l = i + 1; l < i + k; ++l) s += w(l);
    cost += (int64_t)(i - j) * s;
    s = 0;
    for (int l = j + 1; l < i; ++l) s += w(l);
    cost -= k * s;
  }
  return cost;
}

// ---- apply (mlkswap.cu:1079-1081, mlk2opt.cu:740-748, mlkoropt.cu:899-931) -----
void apply_move(int nb, int32_t* t, int i, int j) {
  if (nb == 0) {
    std::swap(t[i], t[j]);
    return;
  }
  if (nb == 1) {
    while (i < j) std::swap(t[i++], t[j--]);
    return;
  }
  const int k = nb_k(nb);
  int32_t tmp[64];
  for (int l = 0; l < k; ++l) tmp[l] = t[i + l];
  if (i < j) {
    int d = i;
    for (int l = i + k; l < j + k; ++l) t[d++] = t[l];
    for (int l = 0; l < k; ++l) t[d++] = tmp[l];
  } else if (i > j) {
    for (int l = i - 1; l >= j; --l) t[l + k] = t[l];
    for (int l = 0; l < k; ++l) t[j + l] = tmp[l];
  }
}

// ---- same-neighborhood merge predicates (mlkernel.cu:51-71) --------------------
bool can_merge(int nb, int i1, int j1, int i2, int j2) {
  if (nb == 0)
    return (std::abs(i1 - i2) > 1) && (std::abs(i1 - j2) > 1) && (std::abs(j1 - i2) > 1) &&
           (std::abs(j1 - j2) > 1);
  if (nb == 1) return (j1 < i2 - 1) || (i1 > j2 + 1) || (j2 < i1 - 1) || (i2 > j1 + 1);
  const int k = nb_k(nb);
  int mn1 = std::min(i1, j1), mx1 = std::max(i1, j1);
  int mn2 = std::min(i2, j2), mx2 = std::max(i2, j2);
  return (mx1 + k < mn2) || (mn1 > mx2 + k);
}

struct RowBest {
  int64_t cost;
  int32_t i, j, ctx;
};

const int64_t kInf = INT64_MAX;

// Per-row best (strict '<' in ctx order: lowest ctx wins ties).  Cost formula:
// ADS concatenation.  Rows without a move keep (kInf, 0, 0, -1).
void eval_rows(int nb, int rowmap, const int32_t* W, int n, const int32_t* t, int N,
               std::vector<RowBest>& out) {
  const int rows = rows_of(nb, rowmap, N);
  out.assign(rows > 0 ? rows : 0, RowBest{kInf, 0, 0, -1});
  if (rows <= 0) return;
  std::vector<int64_t> T(N), C((int64_t)N * N);
  prefix_time(W, n, t, N, T.data());
  ads_cost(T.data(), N, C.data());
  const int64_t L = latency(W, n, t, N);
  const int nctx = ctx_of(nb, rowmap, N);
  for (int by = 0; by < rows; ++by) {
    RowBest b{kInf, 0, 0, -1};
    for (int ctx = 0; ctx < nctx; ++ctx) {
      int i, j;
      if (!move_of(nb, rowmap, N, by, ctx, &i, &j)) continue;
      int64_t c = delta_ads(nb, W, n, t, N, T.data(), C.data(), L, i, j);
      if (c < b.cost) b = RowBest{c, i, j, ctx};
    }
    out[by] = b;
  }
}

// eval_rows over the prefix form (same enumeration and tie rule), rows split
// over OpenMP threads (threads <= 0: the OpenMP default).
void eval_rows_prefix(int nb, int rowmap, const PrefixAds& P, std::vector<RowBest>& out,
                      int threads) {
  const int N = P.N;
  const int rows = rows_of(nb, rowmap, N);
  out.assign(rows > 0 ? rows : 0, RowBest{kInf, 0, 0, -1});
  if (rows <= 0) return;
  const int nctx = ctx_of(nb, rowmap, N);
  const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 8) num_threads(nt)
  for (int by = 0; by < rows; ++by) {
    RowBest b{kInf, 0, 0, -1};
    for (int ctx = 0; ctx < nctx; ++ctx) {
      int i, j;
      if (!move_of(nb, rowmap, N, by, ctx, &i, &j)) continue;
      const int64_t c = P.delta(nb, i, j);
      if (c < b.cost) b = RowBest{c, i, j, ctx};
    }
    out[by] = b;
  }
}

// Stable sort by cost (thrust::sort with OBICmp, mlkernel.cu:254-256) followed
// by the greedy filter of testSwap/test2Opt/testOrOpt (mlkernel.cu:73-170):
// an improving move survives iff no earlier surviving improving move conflicts.
// Writes the surviving row indices in sorted order; returns their count.
int merge_rows(int nb, const std::vector<RowBest>& rows, std::vector<int>& kept, int64_t* sum) {
  std::vector<int> ord(rows.size());
  for (size_t r = 0; r < rows.size(); ++r) ord[r] = (int)r;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int a, int b) { return rows[a].cost < rows[b].cost; });
  kept.clear();
  int64_t s = 0;
  for (int r : ord) {
    const RowBest& m = rows[r];
    if (m.cost >= 0) break;
    bool ok = true;
    for (int q : kept) {
      const RowBest& e = rows[q];
      if (!can_merge(nb, e.i, e.j, m.i, m.j)) {
        ok = false;
        break;
      }
    }
    if (ok) {
      kept.push_back(r);
      s += m.cost;
    }
  }
  if (sum) *sum = s;
  return (int)kept.size();
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------------
// RNG
// ---------------------------------------------------------------------------------
int or_mt_state_bytes(void) { return (int)sizeof(Mt); }
void or_mt_seed(void* m, uint32_t seed) { mt_seed((Mt*)m, seed); }
uint32_t or_mt_rand(void* m) { return mt_rand((Mt*)m); }
uint32_t or_mt_randmax(void* m, uint32_t max) { return mt_randmax((Mt*)m, max); }
double or_mt_real_oo(void* m) { return mt_real_oo((Mt*)m); }
double or_mt_real_co(void* m) { return mt_real_co((Mt*)m); }

// MTRandom::shuffle (mtrand.h:301-312): j = rand(n-i-1), not i + rand(...).
void or_mt_shuffle_i32(void* m, int32_t* data, uint32_t n) {
  for (uint32_t i = 0; i + 1 < n; ++i) {
    uint32_t j = mt_randmax((Mt*)m, n - i - 1);
    std::swap(data[i], data[j]);
  }
}

// ---------------------------------------------------------------------------------
// Distances
// ---------------------------------------------------------------------------------
int or_dist(int x1, int y1, int x2, int y2, int round) { return dist_xy(x1, y1, x2, y2, round); }

// d(s) for a given squared length s (the part of the formula after the int s).
int or_dist_s(int32_t s, int round) {
  float f = (float)s;
  return (int)(sqrtf(f) + (round ? 0.5F : 0.0F));
}

// Client-indexed n x n weight matrix (mlproblem.cpp:153-166).
void or_weights(const int32_t* xs, const int32_t* ys, int n, int round, int32_t* W) {
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) W[(int64_t)a * n + b] = dist_xy(xs[a], ys[a], xs[b], ys[b], round);
}

// ---------------------------------------------------------------------------------
// Synthetic instances (generator definition shared with the product host code;
// the reference has none, so this is a specification, not a restatement).
// ---------------------------------------------------------------------------------
void or_instance_uniform(int n, uint32_t seed, int maxc, int32_t* xs, int32_t* ys) {
  Mt m;
  mt_seed(&m, seed);
  for (int p = 0; p < n; ++p) {
    xs[p] = (int32_t)mt_randmax(&m, (uint32_t)maxc);
    ys[p] = (int32_t)mt_randmax(&m, (uint32_t)maxc);
  }
}

// tsplib.cpp:236-240 inear, applied to a double.
static int inear_d(double v) { return v >= 0 ? (int)(v + 0.5) : (int)(v - 0.5); }

void or_instance_clustered(int n, uint32_t seed, int ncent, int span, double sigma, int32_t* xs,
                           int32_t* ys) {
  Mt m;
  mt_seed(&m, seed);
  std::vector<int> cx(ncent), cy(ncent);
  for (int c = 0; c < ncent; ++c) {
    cx[c] = (int)mt_randmax(&m, (uint32_t)span);
    cy[c] = (int)mt_randmax(&m, (uint32_t)span);
  }
  for (int p = 0; p < n; ++p) {
    int c = (int)mt_randmax(&m, (uint32_t)(ncent - 1));
    double u1 = mt_real_oo(&m), u2 = mt_real_oo(&m);
    double r = std::sqrt(-2.0 * std::log(u1));
    double a = 6.283185307179586 * u2;
    xs[p] = cx[c] + inear_d(sigma * r * std::cos(a));
    ys[p] = cy[c] + inear_d(sigma * r * std::sin(a));
  }
  // coordinate shift to non-negative values (mlproblem.cpp:137-151)
  int xmin = INT32_MAX, ymin = INT32_MAX;
  for (int p = 0; p < n; ++p) {
    xmin = std::min(xmin, (int)xs[p]);
    ymin = std::min(ymin, (int)ys[p]);
  }
  int sx = xmin < 0 ? -xmin : 0, sy = ymin < 0 ? -ymin : 0;
  for (int p = 0; p < n; ++p) {
    xs[p] += sx;
    ys[p] += sy;
  }
}

// ---------------------------------------------------------------------------------
// Tours
// ---------------------------------------------------------------------------------
// MLSolution::random (mlsolution.cpp:753-874) with costTour = true: n clients,
// tour of N = n + 1 positions ending in the depot copy.  Uses the caller's RNG
// state so consecutive calls follow the reference's solution sequence.
void or_random_tour(void* rngp, const int32_t* W, int n, float alpha, int32_t* tour) {
  Mt* rng = (Mt*)rngp;
  std::vector<int> cand;
  for (int c = 1; c < n; ++c) cand.push_back(c);
  unsigned clCount = (unsigned)(n - 1);
  int last = 0, pos = 0;
  tour[pos++] = 0;
  int* cl = cand.data();
  while (clCount > 0) {
    unsigned rclIndex;
    if (clCount > 1) {
      unsigned rclCount = (unsigned)std::ceil(alpha * clCount);
      rclIndex = (rclCount > 0) ? mt_randmax(rng, rclCount - 1) : 0;
      const int32_t* wl = W + (int64_t)last * n;
      if (alpha < 0.5) {
        std::sort(cl, cl + clCount, [wl](int a, int b) { return wl[a] > wl[b]; });
        rclIndex = clCount - rclIndex - 1;
      } else {
        std::sort(cl, cl + clCount, [wl](int a, int b) { return wl[a] < wl[b]; });
      }
    } else {
      rclIndex = 0;
    }
    int cur = cl[rclIndex];
    tour[pos++] = cur;
    last = cur;
    for (unsigned i = rclIndex; i + 1 < clCount; ++i) cl[i] = cl[i + 1];
    clCount--;
  }
  tour[pos++] = 0;  // depot copy
}

// Random permutation tour: [0, shuffle(1..n-1), 0] with MTRandom::shuffle.
void or_perm_tour(uint32_t seed, int n, int32_t* tour) {
  Mt m;
  mt_seed(&m, seed);
  tour[0] = 0;
  for (int p = 1; p < n; ++p) tour[p] = p;
  or_mt_shuffle_i32(&m, tour + 1, (uint32_t)(n - 1));
  tour[n] = 0;
}

int64_t or_latency(const int32_t* W, int n, const int32_t* tour, int N) {
  return latency(W, n, tour, N);
}

// Reference uint32 latency (wraps modulo 2^32, mlsolution.cpp:215-228).
uint32_t or_latency_u32(const int32_t* W, int n, const int32_t* tour, int N) {
  uint32_t c = 0, k = (uint32_t)N - 1;
  for (int p = 1; p < N; ++p, --k) c += k * (uint32_t)W[(int64_t)tour[p - 1] * n + tour[p]];
  return c;
}

// ---------------------------------------------------------------------------------
// ADS (mlsolution.cpp:419-486): T[N] and C[N][N] with the reference's uint32
// wrap-around, exactly as the host ADS buffer stores them.
// ---------------------------------------------------------------------------------
void or_ads_u32(const int32_t* W, int n, const int32_t* tour, int N, uint32_t* T, uint32_t* C) {
  T[0] = 0;
  for (int j = 1; j < N; ++j) T[j] = T[j - 1] + (uint32_t)W[(int64_t)tour[j - 1] * n + tour[j]];
  for (int i = 0; i < N; ++i) {
    uint32_t* r = C + (int64_t)i * N;
    r[i] = 0;
    for (int j = i + 1; j < N; ++j) r[j] = r[j - 1] + (T[j] - T[i]);
  }
  for (int i = N - 1; i >= 0; --i) {
    uint32_t* r = C + (int64_t)i * N;
    for (int j = i - 1; j >= 0; --j) r[j] = r[j + 1] + (T[i] - T[j]);
  }
}

// ---------------------------------------------------------------------------------
// Moves
// ---------------------------------------------------------------------------------
int or_rows(int nb, int rowmap, int N) { return rows_of(nb, rowmap, N); }
int or_ctxs(int nb, int rowmap, int N) { return ctx_of(nb, rowmap, N); }
int or_move_of(int nb, int rowmap, int N, int row, int ctx, int* i, int* j) {
  return move_of(nb, rowmap, N, row, ctx, i, j) ? 1 : 0;
}

int64_t or_delta_ads(int nb, const int32_t* W, int n, const int32_t* tour, int N, int i, int j) {
  std::vector<int64_t> T(N), C((int64_t)N * N);
  prefix_time(W, n, tour, N, T.data());
  ads_cost(T.data(), N, C.data());
  return delta_ads(nb, W, n, tour, N, T.data(), C.data(), latency(W, n, tour, N), i, j);
}

int64_t or_delta_odc(int nb, const int32_t* W, int n, const int32_t* tour, int N, int i, int j) {
  return delta_odc(nb, W, n, tour, N, i, j);
}

// Brute force: latency after applying the move minus latency before.
int64_t or_delta_brute(int nb, const int32_t* W, int n, const int32_t* tour, int N, int i, int j) {
  std::vector<int32_t> t(tour, tour + N);
  apply_move(nb, t.data(), i, j);
  return latency(W, n, t.data(), N) - latency(W, n, tour, N);
}

// All moves of a row map in (row, ctx) order with the three cost restatements;
// returns the number of moves written (arrays sized by the caller, or NULL to count).
int64_t or_enumerate(int nb, int rowmap, const int32_t* W, int n, const int32_t* tour, int N,
                     int32_t* rows, int32_t* ctxs, int32_t* is, int32_t* js, int64_t* dads,
                     int64_t* dodc, int64_t* dbrute) {
  const int R = rows_of(nb, rowmap, N), X = ctx_of(nb, rowmap, N);
  std::vector<int64_t> T(N), C;
  int64_t L = 0;
  if (dads) {
    C.resize((int64_t)N * N);
    prefix_time(W, n, tour, N, T.data());
    ads_cost(T.data(), N, C.data());
    L = latency(W, n, tour, N);
  }
  int64_t m = 0;
  for (int by = 0; by < R; ++by)
    for (int ctx = 0; ctx < X; ++ctx) {
      int i, j;
      if (!move_of(nb, rowmap, N, by, ctx, &i, &j)) continue;
      if (rows) {
        rows[m] = by;
        ctxs[m] = ctx;
        is[m] = i;
        js[m] = j;
      }
      if (dads) dads[m] = delta_ads(nb, W, n, tour, N, T.data(), C.data(), L, i, j);
      if (dodc) dodc[m] = delta_odc(nb, W, n, tour, N, i, j);
      if (dbrute) dbrute[m] = or_delta_brute(nb, W, n, tour, N, i, j);
      ++m;
    }
  return m;
}

// Per-row best records; returns the number of rows.
int or_eval_rows(int nb, int rowmap, const int32_t* W, int n, const int32_t* tour, int N,
                 int64_t* cost, int32_t* is, int32_t* js, int32_t* ctxs) {
  std::vector<RowBest> rb;
  eval_rows(nb, rowmap, W, n, tour, N, rb);
  for (size_t r = 0; r < rb.size(); ++r) {
    cost[r] = rb[r].cost;
    is[r] = rb[r].i;
    js[r] = rb[r].j;
    ctxs[r] = rb[r].ctx;
  }
  return (int)rb.size();
}

// TOTAL mode (kernelSwapTotal mlkswap.cu:334-526, kernel2OptTotal
// mlk2opt.cu:270-447, kernelOrOptTotal mlkoropt.cu:368-612): one launch row per
// ADS row, B threads; thread t of row `row` scans the row-local indices
// ctx = t, t + B, ... and keeps the first strictly smallest cost; its record
// lands at entry row * B + t (mlkswap.cu:526).  Entries whose indices hold no
// move are (INT64_MAX, 0, 0).  Returns rows * B (entries written when the
// arrays are non-NULL).
int64_t or_eval_total(int nb, const int32_t* W, int n, const int32_t* tour, int N, int B,
                      int64_t* cost, int32_t* is, int32_t* js) {
  const int R = rows_of(nb, 0, N), X = ctx_of(nb, 0, N);
  if (R <= 0 || B <= 0) return 0;
  const int64_t M = (int64_t)R * B;
  if (!cost) return M;
  std::vector<int64_t> T(N), C((int64_t)N * N);
  prefix_time(W, n, tour, N, T.data());
  ads_cost(T.data(), N, C.data());
  const int64_t L = latency(W, n, tour, N);
  for (int64_t e = 0; e < M; ++e) {
    cost[e] = kInf;
    is[e] = js[e] = 0;
  }
  for (int by = 0; by < R; ++by)
    for (int ctx = 0; ctx < X; ++ctx) {
      int i, j;
      if (!move_of(nb, 0, N, by, ctx, &i, &j)) continue;
      const int64_t e = (int64_t)by * B + ctx % B;
      const int64_t c = delta_ads(nb, W, n, tour, N, T.data(), C.data(), L, i, j);
      if (c < cost[e]) {
        cost[e] = c;
        is[e] = i;
        js[e] = j;
      }
    }
  return M;
}

// Merge of a per-row list (any source): surviving row ids in sorted order.
int or_merge(int nb, int nrows, const int64_t* cost, const int32_t* is, const int32_t* js,
             int32_t* kept_rows, int64_t* sum) {
  std::vector<RowBest> rb(nrows);
  for (int r = 0; r < nrows; ++r) rb[r] = RowBest{cost[r], is[r], js[r], 0};
  std::vector<int> kept;
  int c = merge_rows(nb, rb, kept, sum);
  for (int q = 0; q < c; ++q) kept_rows[q] = kept[q];
  return c;
}

void or_apply(int nb, int32_t* tour, int i, int j) { apply_move(nb, tour, i, j); }

int or_can_merge(int nb, int i1, int j1, int i2, int j2) { return can_merge(nb, i1, j1, i2, j2); }

// One neighborhood step on a tour: evaluate, merge, apply.  Returns the number
// of applied moves; *sum receives their summed delta.
int or_step(int nb, int rowmap, const int32_t* W, int n, int32_t* tour, int N, int64_t* sum) {
  std::vector<RowBest> rb;
  eval_rows(nb, rowmap, W, n, tour, N, rb);
  std::vector<int> kept;
  int c = merge_rows(nb, rb, kept, sum);
  for (int q : kept) apply_move(nb, tour, rb[q].i, rb[q].j);
  return c;
}

// DVND fixed point: rounds over neighborhood ids 0 .. nbs-1 ({swap, 2-opt,
// or-opt-1, 2, 3, ...}), each chained on the tour left by the previous one
// (WamcaExperiment.hpp:414-416, 360-362; kMax neighborhoods,
// WamcaExperiment.hpp:126-129), until a round applies no move.  Returns the
// final latency.
int64_t or_dvnd_nbs(int rowmap, const int32_t* W, int n, int32_t* tour, int N, int max_rounds,
                    int nbs, int32_t* rounds_out, int64_t* moves_out) {
  int rounds = 0;
  int64_t moves = 0;
  while (max_rounds <= 0 || rounds < max_rounds) {
    int applied = 0;
    for (int nb = 0; nb < nbs; ++nb) {
      if (rows_of(nb, rowmap, N) <= 0) continue;
      int64_t s;
      applied += or_step(nb, rowmap, W, n, tour, N, &s);
    }
    ++rounds;
    moves += applied;
    if (applied == 0) break;
  }
  if (rounds_out) *rounds_out = rounds;
  if (moves_out) *moves_out = moves;
  return latency(W, n, tour, N);
}

int64_t or_dvnd(int rowmap, const int32_t* W, int n, int32_t* tour, int N, int max_rounds,
                int32_t* rounds_out, int64_t* moves_out) {
  return or_dvnd_nbs(rowmap, W, n, tour, N, max_rounds, 5, rounds_out, moves_out);
}

// ---------------------------------------------------------------------------------
// O(N)-memory forms over client coordinates (xs, ys, distance rounding rnd):
// the same evaluation, merge and DVND as above with the ADS in prefix form,
// rows split over `threads` OpenMP threads (<= 0: the OpenMP default).
// ---------------------------------------------------------------------------------
int64_t or_delta_prefix(int nb, const int32_t* xs, const int32_t* ys, int rnd,
                        const int32_t* tour, int N, int i, int j) {
  PrefixAds P;
  P.build(xs, ys, rnd, tour, N);
  return P.delta(nb, i, j);
}

int64_t or_latency_xy(const int32_t* xs, const int32_t* ys, int rnd, const int32_t* tour, int N) {
  PrefixAds P;
  P.build(xs, ys, rnd, tour, N);
  return P.L;
}

int or_eval_rows_xy(int nb, int rowmap, const int32_t* xs, const int32_t* ys, int rnd,
                    const int32_t* tour, int N, int64_t* cost, int32_t* is, int32_t* js,
                    int32_t* ctxs, int threads) {
  PrefixAds P;
  P.build(xs, ys, rnd, tour, N);
  std::vector<RowBest> rb;
  eval_rows_prefix(nb, rowmap, P, rb, threads);
  for (size_t r = 0; r < rb.size(); ++r) {
    cost[r] = rb[r].cost;
    is[r] = rb[r].i;
    js[r] = rb[r].j;
    if (ctxs) ctxs[r] = rb[r].ctx;
  }
  return (int)rb.size();
}

// DVND fixed point as or_dvnd_nbs, over coordinates.
int64_t or_dvnd_xy(int rowmap, const int32_t* xs, const int32_t* ys, int rnd, int32_t* tour,
                   int N, int max_rounds, int nbs, int32_t* rounds_out, int64_t* moves_out,
                   int threads) {
  int rounds = 0;
  int64_t moves = 0;
  PrefixAds P;
  std::vector<RowBest> rb;
  std::vector<int> kept;
  while (max_rounds <= 0 || rounds < max_rounds) {
    int applied = 0;
    for (int nb = 0; nb < nbs; ++nb) {
      if (rows_of(nb, rowmap, N) <= 0) continue;
      P.build(xs, ys, rnd, tour, N);
      eval_rows_prefix(nb, rowmap, P, rb, threads);
      int64_t s;
      const int c = merge_rows(nb, rb, kept, &s);
      for (int q : kept) apply_move(nb, tour, rb[q].i, rb[q].j);
      applied += c;
    }
    ++rounds;
    moves += applied;
    if (applied == 0) break;
  }
  if (rounds_out) *rounds_out = rounds;
  if (moves_out) *moves_out = moves;
  PrefixAds F;
  F.build(xs, ys, rnd, tour, N);
  return F.L;
}

// Per-row bests of every neighborhood in nb_mask (ids 0..4) for `count` tours
// of N positions each (concatenated), OpenMP over (tour x row).  Output for
// tour q and neighborhood nb starts at q * stride + off[nb], where off[] is the
// prefix sum of rows_of(nb, rowmap, N) over the ids in the mask and stride
// their total.  Returns the moves evaluated.
int64_t or_eval_batch_xy(uint32_t nb_mask, int rowmap, const int32_t* xs, const int32_t* ys,
                         int rnd, const int32_t* tours, int count, int N, int64_t* cost,
                         int32_t* is, int32_t* js, int threads) {
  int off[5], stride = 0;
  for (int nb = 0; nb < 5; ++nb) {
    off[nb] = stride;
    if ((nb_mask >> nb) & 1u) stride += std::max(0, rows_of(nb, rowmap, N));
  }
  std::vector<PrefixAds> P((size_t)count);
  const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int q = 0; q < count; ++q) P[(size_t)q].build(xs, ys, rnd, tours + (int64_t)q * N, N);
  const int64_t items = (int64_t)count * stride;
  int64_t moves = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt) reduction(+ : moves)
  for (int64_t it = 0; it < items; ++it) {
    const int q = (int)(it / stride);
    int r = (int)(it % stride), nb = 0;
    while (nb < 4 && (!((nb_mask >> nb) & 1u) || r >= off[nb] + rows_of(nb, rowmap, N))) ++nb;
    const int by = r - off[nb];
    const int nctx = ctx_of(nb, rowmap, N);
    RowBest b{kInf, 0, 0, -1};
    for (int ctx = 0; ctx < nctx; ++ctx) {
      int i, j;
      if (!move_of(nb, rowmap, N, by, ctx, &i, &j)) continue;
      const int64_t c = P[(size_t)q].delta(nb, i, j);
      ++moves;
      if (c < b.cost) b = RowBest{c, i, j, ctx};
    }
    cost[it] = b.cost;
    is[it] = b.i;
    js[it] = b.j;
  }
  return moves;
}

int or_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"

// dvnd_abi.cpp -- the reference's extern "C" DVND interface (include/dvnd.h)
// over the MI355X C ABI.  Host-only parts (noConflict, getNoConflictMoves,
// applyMoves, bestNeighbor with justCalc) never touch the GPU.
#include <algorithm>
#include <climits>
#include <csignal>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "dvnd.h"
#include "mlp_gpu.h"
#include "mlp_host.h"

namespace {

std::mutex g_mu;

struct Problem {
  std::string file;
  std::vector<int32_t> xs, ys;
  int n = 0;  // clients; tours have N = n + 1 positions
};

// Loads a TSPLIB EUC_2D file the way MLProblem::load does with costTour,
// !distRound and !coordShift (dvnd.cpp:163-175).
std::shared_ptr<Problem> load_problem(const char* file) {
  if (!file) return nullptr;
  auto p = std::make_shared<Problem>();
  p->file = file;
  int32_t n = 0;
  std::vector<int32_t> xs(1 << 15), ys(1 << 15);
  int st = mlp_tsplib_load(file, xs.data(), ys.data(), (int32_t)xs.size(), &n);
  if (st == -4) {
    xs.resize((size_t)n);
    ys.resize((size_t)n);
    st = mlp_tsplib_load(file, xs.data(), ys.data(), n, &n);
  }
  if (st != 0) {
    fprintf(stderr, "dvnd: cannot load TSPLIB EUC_2D instance '%s' (status %d)\n", file, st);
    return nullptr;
  }
  xs.resize((size_t)n);
  ys.resize((size_t)n);
  p->xs = std::move(xs);
  p->ys = std::move(ys);
  p->n = n;
  return p;
}

bool strict() {
  const char* e = getenv("MLP_DVND_STRICT");
  return !(e && e[0] == '0');
}

int rowmap_env() {
  const char* e = getenv("MLP_DVND_ROWMAP");
  return (e && (strcmp(e, "ads") == 0 || strcmp(e, "ADS") == 0)) ? MLP_ROWMAP_ADS : MLP_ROWMAP_ODC;
}

// Problem of bestNeighbor: the first one loaded serves every call
// (dvnd.cu:23), unless MLP_DVND_STRICT=0 keys them by hostCode.
std::shared_ptr<Problem> g_first;
std::map<unsigned, std::shared_ptr<Problem>> g_by_host;
// Problems of applyMoves: getProblem(file) with hostCode 0 (dvnd.cpp:18).
std::shared_ptr<Problem> g_apply;

struct DeviceState {
  mlp_ctx* ctx = nullptr;
  std::map<const Problem*, std::pair<int32_t, int32_t>> inst_slot;  // problem -> (inst, slot)
};
std::map<int, DeviceState> g_dev;

bool valid_tour(const Problem& p, const int* t, unsigned N) {
  if (N != (unsigned)p.n + 1 || t[0] != 0 || t[N - 1] != 0) return false;
  std::vector<char> seen((size_t)p.n, 0);
  for (unsigned q = 1; q + 1 < N; ++q) {
    if (t[q] <= 0 || t[q] >= p.n || seen[(size_t)t[q]]) return false;
    seen[(size_t)t[q]] = 1;
  }
  return true;
}

uint32_t latency_u32(const Problem& p, const int32_t* t, int N) {
  return (uint32_t)(uint64_t)mlp_host_latency(p.xs.data(), p.ys.data(), 0, t, N);
}

// MLMove64 bit-fields (mlads.h:190-195): id:4, i:14, j:14, cost int32.
struct Move64 {
  unsigned id, i, j;
  int cost;
};

Move64 pack64(unsigned short id, unsigned i, unsigned j, int cost) {
  return Move64{id & 0xfu, i & 0x3fffu, j & 0x3fffu, cost};
}

int oropt_k(unsigned id) { return id == 2 ? 1 : (id == 3 ? 2 : 3); }

// Cross-neighborhood compatibility, literal to dvnd.cu:132-175 (including its
// asymmetries).  Bit-field operands promote to int.
bool no_conflict(const Move64& m1, const Move64& m2) {
  const int i1 = (int)m1.i, j1 = (int)m1.j, i2 = (int)m2.i, j2 = (int)m2.j;
  if (m1.id == 0) {
    if (m2.id == 0)
      return (std::abs(i1 - i2) > 1) && (std::abs(i1 - j2) > 1) && (std::abs(j1 - i2) > 1) &&
             (std::abs(j1 - j2) > 1);
    if (m2.id == 1)
      return ((i1 < i2 - 1) || (i1 > j2 - 1)) && ((j1 < i2 - 1) || (j1 > j2 + 1));
    const int k2 = oropt_k(m2.id);
    return (j1 < std::min(i2, j2) - 1) || (i1 > std::max(i2, j2) + k2) ||
           ((i1 < std::min(i2, j2) - 1) && (j1 > std::max(i2, j2) + k2));
  }
  if (m1.id == 1) {
    if (m2.id == 0) return ((i2 < i1 - 1) || (i2 > j1 + 1)) && ((j2 < i1 - 1) || (j2 > j1 + 1));
    if (m2.id == 1) return (j1 < i2 - 1) || (i1 > j2 + 1) || (j2 > i1 - 1) || (i2 > j1 + 1);
    const int k2 = oropt_k(m2.id);
    return (i1 > std::max(i2, j2) + k2) || (j1 < std::min(i2, j2) - 1);
  }
  const int k1 = oropt_k(m1.id);
  if (m2.id == 0)
    return (j2 < std::min(i1, i2) - 1) || (i2 > std::max(i1, i2) + k1) ||
           ((i2 < std::min(i1, i2) - 1) && (j2 > std::max(i1, i2) + k1));
  if (m2.id == 1) return (j2 < std::min(i1, j1) - 1) || (i2 > std::max(i1, j1) + k1);
  const int k2 = oropt_k(m2.id);
  return (std::max(i1, j1) + k1 < std::min(i2, j2)) || (std::min(i1, j1) > std::max(i2, j2) + k2);
}

// Host application of one move as applyMoves does (dvnd.cpp:23-130): ids 0
// and 1 swap positions i and j, ids 2.. move the segment of k = id - 1
// clients starting at i so that it starts at j.
void host_apply(int32_t* t, unsigned N, unsigned id, int i, int j) {
  if (id == 0 || id == 1) {
    if (i >= 0 && j >= 0 && (unsigned)i < N && (unsigned)j < N) std::swap(t[i], t[j]);
    return;
  }
  const int k = (int)id - 1;
  if (i < 1 || j < 1 || i + k > (int)N || j + k > (int)N) return;  // outside the tour
  std::vector<int32_t> tmp(t + i, t + i + k);
  if (i < j) {
    int d = i;
    for (int l = i + k; l < j + k; ++l) t[d++] = t[l];
    for (int l = 0; l < k; ++l) t[d++] = tmp[(size_t)l];
  } else if (i > j) {
    for (int l = i - 1; l >= j; --l) t[l + k] = t[l];
    for (int l = 0; l < k; ++l) t[j + l] = tmp[(size_t)l];
  }
}

}  // namespace

extern "C" {

bool noConflict(unsigned short id1, unsigned int i1, unsigned int j1, unsigned short id2,
                unsigned int i2, unsigned int j2) {
  return no_conflict(pack64(id1, i1, j1, 0), pack64(id2, i2, j2, 0));
}

// betterNoConflict (dvnd.cu:85-130): sort (std::sort on {move, index}),
// descending when (melhorParaPior ? maximize : !maximize); for each move the
// first later conflicting move removes the later one (melhorParaPior) or the
// earlier one, then the scan moves to the next i.  Removed moves still take
// part in later conflict tests, as in the reference.
int getNoConflictMoves(unsigned int numberOfMoves, unsigned short* ids, unsigned int* is,
                       unsigned int* js, int* costs, int* selectedMoves, int* impValue,
                       bool maximize, bool melhorParaPior) {
  if (numberOfMoves == 0) {
    if (impValue) *impValue = 0;
    return 0;
  }
  if (!ids || !is || !js || !costs || !selectedMoves || !impValue) return -1;
  std::vector<Move64> mv(numberOfMoves);
  for (unsigned q = 0; q < numberOfMoves; ++q) mv[q] = pack64(ids[q], is[q], js[q], costs[q]);
  struct MoveIndex {
    Move64* move;
    unsigned int index;
  };
  std::vector<MoveIndex> ix(numberOfMoves);
  for (unsigned q = 0; q < numberOfMoves; ++q) ix[q] = MoveIndex{&mv[q], q};
  if (melhorParaPior ? maximize : !maximize)
    std::sort(ix.begin(), ix.end(),
              [](const MoveIndex& x, const MoveIndex& y) { return x.move->cost > y.move->cost; });
  else
    std::sort(ix.begin(), ix.end(),
              [](const MoveIndex& x, const MoveIndex& y) { return x.move->cost < y.move->cost; });
  const int n = (int)numberOfMoves;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      if (!no_conflict(*ix[(size_t)i].move, *ix[(size_t)j].move)) {
        if (melhorParaPior) ix[(size_t)j].index = (unsigned)-1;
        else ix[(size_t)i].index = (unsigned)-1;
        break;
      }
  int len = 0, imp = 0;
  for (int i = 0; i < n; ++i)
    if (ix[(size_t)i].index != (unsigned)-1) {
      selectedMoves[len++] = (int)ix[(size_t)i].index;
      imp += ix[(size_t)i].move->cost;
    }
  *impValue = imp;
  return len;
}

unsigned int applyMoves(char* file, int* solution, unsigned int solutionSize,
                        unsigned int numberOfMoves, unsigned short* ids, unsigned int* is,
                        unsigned int* js, int* costs) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_apply) g_apply = load_problem(file);
  if (!g_apply || !solution || solutionSize != (unsigned)g_apply->n + 1) return UINT_MAX;
  if (numberOfMoves && (!ids || !is || !js)) return UINT_MAX;
  std::vector<int32_t> t(solution, solution + solutionSize);
  for (unsigned q = 0; q < numberOfMoves; ++q)
    if (ids[q] < 30)  // MLP_MAX_NEIGHBOR (consts.h:19)
      host_apply(t.data(), solutionSize, ids[q], (int)is[q], (int)js[q]);
  for (unsigned q = 0; q < solutionSize; ++q) solution[q] = t[q];
  return latency_u32(*g_apply, t.data(), (int)solutionSize);
}

unsigned int bestNeighbor(char* file, int* solution, unsigned int solutionSize, int neighborhood,
                          bool justCalc, unsigned int hostCode, unsigned int* useMoves,
                          unsigned short* ids, unsigned int* is, unsigned int* js, int* costs,
                          bool useMultipleGpu, unsigned int deviceCount, int* solutionResp) {
  std::lock_guard<std::mutex> lk(g_mu);
  std::shared_ptr<Problem> prob;
  if (strict()) {
    if (!g_first) g_first = load_problem(file);
    prob = g_first;
  } else {
    auto& slot = g_by_host[hostCode];
    if (!slot) slot = load_problem(file);
    prob = slot;
  }
  if (!prob || !solution || !valid_tour(*prob, solution, solutionSize)) {
    fprintf(stderr, "dvnd: bestNeighbor needs a tour of N = n + 1 positions of the loaded problem\n");
    return UINT_MAX;
  }
  if (justCalc) return latency_u32(*prob, solution, (int)solutionSize);
  if (neighborhood < 0 || neighborhood > 4) {
    fprintf(stderr, "dvnd: neighborhood %d is not one of 0..4\n", neighborhood);
    return UINT_MAX;
  }
  const bool record = useMoves ? (*useMoves != 0) : false;
  if (!useMoves && strict()) {
    fprintf(stderr, "dvnd: useMoves is NULL (the reference dereferences it, dvnd.cu:49)\n");
    return UINT_MAX;
  }
  const int device = useMultipleGpu && deviceCount ? (int)(hostCode % deviceCount) : 0;
  DeviceState& ds = g_dev[device];
  if (!ds.ctx && mlp_ctx_create(device, &ds.ctx) != MLP_OK) {
    fprintf(stderr, "dvnd: no HIP device %d\n", device);
    return UINT_MAX;
  }
  auto it = ds.inst_slot.find(prob.get());
  if (it == ds.inst_slot.end()) {
    int32_t inst = -1, slot = -1;
    if (mlp_instance_create(ds.ctx, prob->xs.data(), prob->ys.data(), prob->n, 0, &inst) ||
        mlp_slot_create(ds.ctx, inst, &slot)) {
      fprintf(stderr, "dvnd: %s\n", mlp_ctx_error(ds.ctx));
      return UINT_MAX;
    }
    it = ds.inst_slot.emplace(prob.get(), std::make_pair(inst, slot)).first;
  }
  const int32_t slot = it->second.second;
  const int nb = neighborhood;
  std::vector<mlp_move> mv((size_t)solutionSize);
  int32_t count = 0;
  std::vector<int32_t> t(solution, solution + solutionSize);
  int st = mlp_set_tours(ds.ctx, &slot, 1, t.data());
  if (!st) st = mlp_eval(ds.ctx, &slot, 1, 1u << nb, rowmap_env(), 0);
  if (!st) st = mlp_merge_apply(ds.ctx, &slot, 1, nb);
  if (!st) st = mlp_get_applied(ds.ctx, slot, nb, mv.data(), (int32_t)mv.size(), &count);
  if (!st) st = mlp_get_tours(ds.ctx, &slot, 1, t.data());
  if (st) {
    fprintf(stderr, "dvnd: %s\n", mlp_ctx_error(ds.ctx));
    return UINT_MAX;
  }
  if (!record) {
    for (unsigned q = 0; q < solutionSize; ++q) solution[q] = t[q];
  } else {
    if (solutionResp)
      for (unsigned q = 0; q < solutionSize; ++q) solutionResp[q] = t[q];
    const unsigned take = std::min<unsigned>((unsigned)count, *useMoves);
    for (unsigned q = 0; q < take; ++q) {
      if (ids) ids[q] = (unsigned short)mv[q].id;
      if (is) is[q] = (unsigned)mv[q].i;
      if (js) js[q] = (unsigned)mv[q].j;
      if (costs) costs[q] = (int)mv[q].cost;
    }
    *useMoves = take;
  }
  return latency_u32(*prob, t.data(), (int)solutionSize);
}

unsigned int bestNeighborSimple(char* file, int* solution, unsigned int solutionSize,
                                int neighborhood) {
  return bestNeighbor(file, solution, solutionSize, neighborhood);
}

// Fault-injection hook of the reference: the external driver uses it to test
// its own crash handling (dvnd.cpp:230-233).
void segFault(void) { raise(SIGSEGV); }

}  // extern "C"

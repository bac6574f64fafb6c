// mlkernel.cpp -- the reference's C++ surface (include/wamca/mlkernel.hpp)
// over the MI355X C ABI.  Host bookkeeping follows the reference's classes;
// every device step goes through mlp_gpu.h.
#include "wamca/mlkernel.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <istream>
#include <ostream>
#include <sstream>

#include "mlp_host.h"

namespace wamca {

// mlads.cpp:41-72 (ids this build supports)
const char* nameMove[15] = {"SWAP",   "2OPT",   "OROPT1", "OROPT2",  "OROPT3",
                            "OROPT4", "OROPT5", "OROPT6", "OROPT7",  "OROPT8",
                            "OROPT9", "OROPT10", "OROPT11", "OROPT12", "OROPT13"};

// Default TOTAL block: the largest multiple of 64 (one wavefront) not above
// min(size, 1024).
static uint default_total_block(uint size) {
  const uint b = std::min<uint>(size, 1024u) / 64u * 64u;
  return b ? b : std::max<uint>(size, 1u);
}

namespace {

void check(mlp_ctx* ctx, int st, const char* what) {
  if (st != MLP_OK) {
    std::string m = std::string(what) + ": " + mlp_strerror(st);
    if (ctx) m += std::string(" (") + mlp_ctx_error(ctx) + ")";
    throw Error(m);
  }
}

int rows_of(int nb, int rowmap, int N) {
  if (nb == 0) return rowmap == MLP_ROWMAP_ADS ? N - 4 : ((N + 1) / 2) - 1;
  if (nb == 1) return (N + 1) / 2;
  return N - (nb - 1) - 2;
}

}  // namespace

// ---- MTRandom (mtrand.cpp:51-121, mtrand.h:168-182) --------------------------
void MTRandom::seed(uint s) {
  rseed_ = s;
  state_[0] = s;
  for (int i = 1; i < 624; ++i)
    state_[i] = 1812433253U * (state_[i - 1] ^ (state_[i - 1] >> 30)) + (uint)i;
  p_ = 624;
}

void MTRandom::regen() {
  static const uint mag[2] = {0u, 0x9908b0dfU};
  for (int i = 0; i < 624; ++i) {
    const uint y = (state_[i] & 0x80000000U) | (state_[(i + 1) % 624] & 0x7fffffffU);
    state_[i] = state_[(i + 397) % 624] ^ (y >> 1) ^ mag[y & 1u];
  }
  p_ = 0;
}

uint MTRandom::rand() {
  if (p_ >= 624) regen();
  uint x = state_[p_++];
  x ^= (x >> 11);
  x ^= (x << 7) & 0x9d2c5680U;
  x ^= (x << 15) & 0xefc60000U;
  return x ^ (x >> 18);
}

// ---- MLProblem (mlproblem.cpp:58-207) ---------------------------------------
MLProblem::MLProblem(bool ct, bool dr, bool cs) : costTour(ct), distRound(dr), coordShift(cs) {}

MLProblem::~MLProblem() {
  if (ctx_) mlp_ctx_destroy(ctx_);
}

void MLProblem::load(const char* fname) {
  std::vector<int32_t> xs(1 << 15), ys(1 << 15);
  int32_t n = 0;
  int st = mlp_tsplib_load(fname, xs.data(), ys.data(), (int32_t)xs.size(), &n);
  if (st == -4) {
    xs.resize((size_t)n);
    ys.resize((size_t)n);
    st = mlp_tsplib_load(fname, xs.data(), ys.data(), n, &n);
  }
  if (st != 0) throw Error(std::string("Error opening instance file: '") + fname + "'");
  if (coordShift) mlp_shift_nonnegative(n, xs.data(), ys.data());
  const char* base = strrchr(fname, '/');
  base = base ? base + 1 : fname;
  snprintf(filename, sizeof filename, "%s", fname);
  snprintf(name, sizeof name, "%s", base);
  char* dot = strrchr(name, '.');
  if (dot) *dot = 0;
  xs.resize((size_t)n);
  ys.resize((size_t)n);
  create(xs.data(), ys.data(), n, name);
}

void MLProblem::create(const int32_t* xs, const int32_t* ys, int n, const char* nm) {
  if (n < 2) throw Error("an instance needs at least 2 clients");
  n_ = n;
  xs_.assign(xs, xs + n);
  ys_.assign(ys, ys + n);
  if (nm != name) snprintf(name, sizeof name, "%s", nm);
  size = (uint)n + (costTour ? 1u : 0u);
  clients.resize(size);
  for (uint i = 0; i < size; ++i) {
    const int ii = (int)(i % (uint)n);
    clients[i] = MLClientData{(ushort)ii, xs_[(size_t)ii], ys_[(size_t)ii]};
  }
  build_weights();
  if (inst_ >= 0 && ctx_) {
    mlp_instance_destroy(ctx_, inst_);
    inst_ = -1;
  }
}

void MLProblem::build_weights() {
  w_.assign((size_t)n_ * n_, 0);
  maxWeight = 0;
  for (int a = 0; a < n_; ++a)
    for (int b = 0; b < n_; ++b) {
      const int d = mlp_host_dist(xs_[(size_t)a], ys_[(size_t)a], xs_[(size_t)b], ys_[(size_t)b],
                                  distRound ? 1 : 0);
      w_[(size_t)a * n_ + b] = d;
      maxWeight = std::max(maxWeight, d);
    }
}

mlp_ctx* MLProblem::ctx() {
  if (!ctx_) check(nullptr, mlp_ctx_create(0, &ctx_), "mlp_ctx_create");
  return ctx_;
}

int32_t MLProblem::instance() {
  if (!costTour) throw Error("only costTour problems (tours ending at the depot) are supported");
  if (inst_ < 0)
    check(ctx(), mlp_instance_create(ctx(), xs_.data(), ys_.data(), n_, distRound ? 1 : 0, &inst_),
          "mlp_instance_create");
  return inst_;
}

// ---- MLSolution (mlsolution.cpp) ---------------------------------------------
MLSolution::MLSolution(MLProblem& prob, int) : problem(prob) {
  clients.assign(problem.size + 1, 0);
  weights.assign(problem.size + 1, 0);
}

void MLSolution::assign(const MLSolution* sol, bool) {
  if (sol == this) return;
  move = sol->move;
  cost = sol->cost;
  cost64 = sol->cost64;
  clientCount = sol->clientCount;
  clients = sol->clients;
  weights = sol->weights;
}

// mlsolution.cpp:215-228 (uint arithmetic wraps like the reference's)
void MLSolution::update() {
  cost = 0;
  cost64 = 0;
  uint k = clientCount - 1;
  weights[0] = 0;
  for (uint i = 1; i < clientCount; ++i, --k) {
    weights[i] = (ushort)dist(i - 1, i);
    cost += k * weights[i];
    cost64 += (int64_t)k * weights[i];
  }
}

uint MLSolution::costCalc() const {
  uint c = 0, k = clientCount - 1;
  for (uint i = 1; i < clientCount; ++i, --k) c += k * weights[i];
  return c;
}

// mlsolution.cpp:753-874
void MLSolution::random(MTRandom& rng, float alpha) {
  const uint n = problem.size - (problem.costTour ? 1u : 0u);
  std::vector<ushort> cand;
  for (uint c = 1; c < n; ++c) cand.push_back((ushort)c);
  uint clCount = n - 1;
  clear();
  ushort last = 0;
  add(0);
  while (clCount > 0) {
    uint rclIndex = 0;
    if (clCount > 1) {
      const uint rclCount = (uint)std::ceil(alpha * clCount);
      rclIndex = rclCount > 0 ? rng.rand(rclCount - 1) : 0;
      const MLProblem& P = problem;
      const ushort l = last;
      if (alpha < 0.5F) {
        std::sort(cand.begin(), cand.begin() + clCount,
                  [&P, l](ushort a, ushort b) { return P.weight(l, a) > P.weight(l, b); });
        rclIndex = clCount - rclIndex - 1;
      } else {
        std::sort(cand.begin(), cand.begin() + clCount,
                  [&P, l](ushort a, ushort b) { return P.weight(l, a) < P.weight(l, b); });
      }
    }
    const ushort cur = cand[rclIndex];
    add(cur);
    last = cur;
    for (uint i = rclIndex; i + 1 < clCount; ++i) cand[i] = cand[i + 1];
    --clCount;
  }
  if (problem.costTour) add(0);
  update();
}

void MLSolution::sequence() {
  clear();
  const uint n = problem.size - (problem.costTour ? 1u : 0u);
  for (uint i = 0; i < n; ++i) add((ushort)i);
  if (problem.costTour) add(0);
  update();
}

// mlsolution.cpp:684-751
void MLSolution::save(std::ostream& os) const {
  os << clientCount << '\n' << cost << '\n';
  for (uint i = 0; i < clientCount; ++i) os << clients[i] << (i + 1 < clientCount ? " " : "");
  os << '\n';
}

void MLSolution::load(std::istream& is) {
  uint v;
  clear();
  is >> v;
  if (v != problem.size) {
    const bool error = problem.costTour ? (problem.size - 1 != v) : (problem.size + 1 != v);
    if (error) throw Error("Error loading solution: invalid problem size");
  }
  is >> v;  // discard cost
  for (uint i = 0; i < problem.size; ++i) {
    is >> v;
    add((ushort)v);
  }
  update();
}

std::string MLSolution::str(bool showCost) const {
  std::ostringstream ss;
  ss << '<';
  for (uint i = 0; i < clientCount; ++i) ss << clients[i] << (i + 1 < clientCount ? "," : "");
  ss << '>';
  if (showCost) ss << " = " << cost << (problem.costTour ? " (TOUR)" : " (PATH)");
  return ss.str();
}

// ---- MLKernel -----------------------------------------------------------------
MLKernel::MLKernel(MLProblem& p, bool total, int kid, uint ktag)
    : problem(p), id(kid), tag(ktag), name(kid >= 0 && kid < 15 ? nameMove[kid] : "OROPT"),
      isTotal(total) {
  if (kid < 0 || kid > (int)MLMI_OROPT13) throw Error("neighborhood id must be 0..14");
}

MLKernel::~MLKernel() {
  try {
    term();
  } catch (...) {
  }
}

void MLKernel::init(bool solCreate) {
  solDestroy = solCreate;
  if (solCreate) solution = new MLSolution(problem);
  // or-opt-k for k > 3 needs device buffers for ids 5..14 (slots created from now on)
  if (id >= 5) check(problem.ctx(), mlp_set_neighborhoods(problem.ctx(), 15), "mlp_set_neighborhoods");
  check(problem.ctx(), mlp_slot_create(problem.ctx(), problem.instance(), &slot_), "mlp_slot_create");
  // sizes of mlkernel.cu:442-556
  solSize = problem.size;
  adsRowElems = solSize;
  adsDataSize = 128u + (solSize + 3u) * adsRowElems * 4u;
  reserveMoves(isTotal ? (size_t)solSize * solSize : (size_t)solSize);
  check(problem.ctx(), mlp_timer_create(problem.ctx(), &timer_), "mlp_timer_create");
  defineKernelGrid();
}

// moveData and transBuffer hold at least `entries` records.
void MLKernel::reserveMoves(size_t entries) {
  entries = std::max<size_t>(entries, 1);
  if (moveData && entries * 8 <= moveDataSize) return;
  if (entries * 8 > 0xffffffffull) throw Error("move buffer above 4 GiB");
  void* dev = nullptr;
  check(problem.ctx(), mlp_dev_alloc(problem.ctx(), (int64_t)entries * 8, &dev), "moveData");
  if (moveData) mlp_dev_free(problem.ctx(), moveData);
  moveData = (MLMovePack*)dev;
  moveDataSize = (uint)(entries * 8);
  host_.assign(entries, MLMove64{0, 0, 0, 0});
  transBuffer.p_move64 = host_.data();
  transBufferSize = moveDataSize;
}

void MLKernel::term() {
  if (slot_ >= 0) mlp_slot_destroy(problem.ctx(), slot_);
  slot_ = -1;
  if (moveData) mlp_dev_free(problem.ctx(), moveData);
  moveData = nullptr;
  moveDataSize = transBufferSize = 0;
  if (timer_ >= 0) mlp_timer_destroy(problem.ctx(), timer_);
  timer_ = -1;
  if (solDestroy && solution) delete solution;
  solution = nullptr;
}

void MLKernel::timerStart() { check(problem.ctx(), mlp_timer_start(problem.ctx(), timer_), "timerStart"); }

void MLKernel::timerStop() { check(problem.ctx(), mlp_timer_stop(problem.ctx(), timer_), "timerStop"); }

unsigned long long MLKernel::timerElapsed() {
  float ms = 0.0F;
  check(problem.ctx(), mlp_timer_elapsed(problem.ctx(), timer_, &ms), "timerElapsed");
  time = (unsigned long long)(ms * 1000.0F);
  return time;
}

void MLKernel::sendSolution() {
  if (solution->clientCount != problem.size) throw Error("solution is not a full tour");
  std::vector<int32_t> t(solution->clients.begin(), solution->clients.begin() + problem.size);
  check(problem.ctx(), mlp_set_tours(problem.ctx(), &slot_, 1, t.data()), "sendSolution");
}

void MLKernel::defineKernelGrid() {
  if (isTotal) {  // rows * block entries (mlkswap.cu:1022-1024)
    rowmap = MLP_ROWMAP_ADS;
    if (!totalBlock) totalBlock = default_total_block(problem.size);
    moveElems = (uint)std::max(0, rows_of(id, rowmap, (int)problem.size)) * totalBlock;
  } else {
    moveElems = (uint)std::max(0, rows_of(id, rowmap, (int)problem.size));
  }
  reserveMoves(moveElems);
}

// The evaluation, then its records written into moveData by the device.
void MLKernel::launchKernel() {
  ++callCount;
  mlp_ctx* c = problem.ctx();
  if (isTotal) {
    check(c, mlp_eval_total(c, &slot_, 1, id, (int32_t)totalBlock), "launchKernel");
    int64_t M = 0;
    reserveMoves((size_t)std::max(0, rows_of(id, MLP_ROWMAP_ADS, (int)problem.size)) * totalBlock);
    check(c, mlp_total_moves64_device(c, slot_, (uint64_t*)moveData, moveDataSize / 8, &M),
          "launchKernel");
    moveElems = (uint)M;
  } else {
    check(c, mlp_eval(c, &slot_, 1, 1u << id, rowmap, 0), "launchKernel");
    int32_t R = 0;
    check(c, mlp_rows_moves64_device(c, slot_, id, (uint64_t*)moveData, moveDataSize / 8, &R),
          "launchKernel");
    moveElems = (uint)R;
  }
}

void MLKernel::sync() { check(problem.ctx(), mlp_sync(problem.ctx()), "sync"); }

// mlkernel.h:240-244: moveElems records from moveData to transBuffer.
void MLKernel::recvResult() {
  check(problem.ctx(),
        mlp_copy_to_host(problem.ctx(), transBuffer.p_void, moveData, (int64_t)moveElems * 8),
        "recvResult");
}

// Device merge and apply; moveData then holds the merged list in the
// reference's post-merge form (mlkernel.cu:254-384 sorts moveData by cost and
// zeroes the cost of every move the greedy rejects): the kept moves first, in
// merge order, with their costs, every other entry with cost 0, so a caller
// that applies the entries with cost < 0 in order applies exactly the kept
// moves (WamcaExperiment.hpp:229-290).
void MLKernel::mergeGPU() {
  mlp_ctx* c = problem.ctx();
  if (isTotal)
    check(c, mlp_merge_total(c, &slot_, 1), "mergeGPU");
  else
    check(c, mlp_merge_apply(c, &slot_, 1, id), "mergeGPU");
  std::vector<mlp_move> mv(std::max<uint>(moveElems, 1));
  int32_t count = 0;
  check(c, mlp_get_applied(c, slot_, id, mv.data(), (int32_t)mv.size(), &count), "mergeGPU");
  std::vector<MLMove64> list(std::max<uint>(moveElems, 1), MLMove64{(unsigned)id, 0, 0, 0});
  for (int q = 0; q < count; ++q)
    list[(size_t)q] = MLMove64{(unsigned)id, (unsigned)mv[(size_t)q].i & 0x3fffu,
                               (unsigned)mv[(size_t)q].j & 0x3fffu, (int)mv[(size_t)q].cost};
  check(c, mlp_copy_to_device(c, moveData, list.data(), (int64_t)moveElems * 8), "mergeGPU");
  ++mergeCount;
}

// mlkernel.cu:617-647: first minimum of the record list.
int MLKernel::bestMove(MLMove& move) {
  uint b = 0;
  const MLMove64* t = transBuffer.p_move64;
  for (uint i = 1; i < moveElems; ++i)
    if (t[i].cost < t[b].cost) b = i;
  move.id = (MLMoveId)t[b].id;
  move.i = t[b].i;
  move.j = t[b].j;
  move.cost = t[b].cost;
  return move.cost;
}

namespace {
// mlkernel.cu:649-728
bool can_merge(const MLMove64& m1, const MLMove64& m2) {
  const int i1 = m1.i, j1 = m1.j, i2 = m2.i, j2 = m2.j;
  auto swapm = [](int a1, int b1, int a2, int b2) {
    return std::abs(a1 - a2) > 1 && std::abs(a1 - b2) > 1 && std::abs(b1 - a2) > 1 &&
           std::abs(b1 - b2) > 1;
  };
  if (m1.id == MLMI_SWAP && m2.id == MLMI_SWAP) return swapm(i1, j1, i2, j2);
  if (m1.id == MLMI_2OPT && m2.id == MLMI_2OPT)
    return (j1 < i2 - 1) || (i1 > j2 + 1) || (j2 < i1 - 1) || (i2 > j1 + 1);
  const int k1 = (int)m1.id - 1, k2 = (int)m2.id - 1;
  return (std::max(i1, j1) + k1 < std::min(i2, j2)) || (std::min(i1, j1) > std::max(i2, j2) + k2);
}
}  // namespace

// Host greedy merge (mlkernel.cu:732-864): std::sort by cost, keep a move iff
// no earlier kept move conflicts.  Sorts transBuffer in place like the
// reference.
// In TOTAL mode the reference's conflict graph takes at most the first
// 120000 improving moves (mlkernel.cu:778-787); the same limit applies here,
// so this merge can differ from mergeGPU exactly where the reference's does.
int MLKernel::mergeGreedy(MLMove64* merge, int& count) {
  MLMove64* mv = transBuffer.p_move64;
  int n = (int)moveElems;
  std::sort(mv, mv + n, [](const MLMove64& a, const MLMove64& b) { return a.cost < b.cost; });
  if (isTotal) n = std::min(n, 120000);
  count = 0;
  int cost = 0;
  std::vector<const MLMove64*> kept;
  for (int i = 0; i < n; ++i) {
    if (mv[i].cost >= 0) break;
    bool ok = true;
    for (const MLMove64* k : kept)
      if (!can_merge(*k, mv[i])) {
        ok = false;
        break;
      }
    if (!ok) continue;
    kept.push_back(&mv[i]);
    merge[count++] = mv[i];
    cost += mv[i].cost;
  }
  ++mergeCount;
  return cost;
}

// mlkswap.cu:1066-1106
void MLKernelSwap::applyMove(MLMove& move) {
  const int i = move.i, j = move.j;
  MLSolution& s = *solution;
  s.cost += (uint)move.cost;
  s.move = move;
  std::swap(s.clients[(size_t)i], s.clients[(size_t)j]);
  s.weights[(size_t)i] = (ushort)s.dist(i - 1, i);
  s.weights[(size_t)i + 1] = (ushort)s.dist(i, i + 1);
  s.weights[(size_t)j] = (ushort)s.dist(j - 1, j);
  if ((uint)j + 1 < s.clientCount) s.weights[(size_t)j + 1] = (ushort)s.dist(j, j + 1);
}

// mlk2opt.cu:725-768
void MLKernel2Opt::applyMove(MLMove& move) {
  MLSolution& s = *solution;
  s.cost += (uint)move.cost;
  s.move = move;
  int i = move.i, j = move.j;
  if (i < j) {
    while (i < j) std::swap(s.clients[(size_t)i++], s.clients[(size_t)j--]);
    j = move.j;
    if ((uint)j + 1 < s.clientCount) ++j;
    for (i = move.i; i <= j; ++i) s.weights[(size_t)i] = (ushort)s.dist((uint)i - 1, (uint)i);
  }
}

// mlkoropt.cu:886-950
void MLKernelOrOpt::applyMove(MLMove& move) {
  MLSolution& s = *solution;
  const int k = (int)tag;
  s.cost += (uint)move.cost;
  s.move = move;
  int i = move.i, j = move.j;
  std::vector<ushort> tmp(s.clients.begin() + i, s.clients.begin() + i + k);
  if (i < j) {
    for (int l = i + k; l < j + k; ++l) s.clients[(size_t)i++] = s.clients[(size_t)l];
    for (int l = 0; l < k; ++l) s.clients[(size_t)i++] = tmp[(size_t)l];
    i = move.i;
    j = move.j + k;
    if ((uint)j >= s.clientCount) --j;
  } else if (i > j) {
    for (int l = i - 1; l >= j; --l) s.clients[(size_t)l + k] = s.clients[(size_t)l];
    for (int l = 0; l < k; ++l) s.clients[(size_t)l + j] = tmp[(size_t)l];
    i = move.j;
    j = move.i + k;
    if ((uint)j >= s.clientCount) --j;
  }
  for (; i <= j; ++i) s.weights[(size_t)i] = (ushort)s.dist((uint)i - 1, (uint)i);
}

// ---- WAMCAExperiment (WamcaExperiment.hpp) ------------------------------------
WAMCAExperiment::WAMCAExperiment(MLProblem& p, int seed) : problem(p) {
  rng.seed(seed != 0 ? (uint)seed : (uint)time(nullptr) + 1u);
  kernels[0] = new MLKernelSwap(problem);
  kernels[1] = new MLKernel2Opt(problem);
  for (uint k = 1; k <= 3; ++k) kernels[k + 1] = new MLKernelOrOpt(problem, k);
  for (MLKernel* kn : kernels) kn->init(true);
  solDevice = new MLSolution(problem);
}

WAMCAExperiment::~WAMCAExperiment() {
  for (MLKernel*& kn : kernels) {
    delete kn;
    kn = nullptr;
  }
  delete solDevice;
}

// WamcaExperiment.hpp:292-365 with runGpuGpu (:229-290)
void WAMCAExperiment::runKernel(int k, std::vector<MLMove>* moves) {
  MLKernel* kernel = kernels[k];
  kernel->setSolution(solDevice);
  kernel->sendSolution();
  kernel->defineKernelGrid();
  kernel->launchKernel();
  kernel->sync();
  kernel->mergeGPU();
  kernel->recvResult();
  kernel->sync();
  for (uint it = 0; it < kernel->moveElems; ++it) {
    const MLMove64& m = kernel->transBuffer.p_move64[it];
    if (m.cost < 0) {
      MLMove mv{(MLMoveId)m.id, (int)m.i, (int)m.j, m.cost};
      kernel->applyMove(mv);
      if (moves) moves->push_back(mv);
    }
  }
  kernel->getSolution(solDevice);
  solDevice->update();
}

// WamcaExperiment.hpp:368-581 (mMax forced to 1 as in the reference)
unsigned int WAMCAExperiment::runWAMCA2016(int, int kMin, int kMax, int* solution,
                                           unsigned int solutionSize, std::vector<MLMove>* moves,
                                           int* solutionResp) {
  if (!solution) {
    solDevice->random(rng, 0.50F);
  } else {
    solDevice->clientCount = solutionSize;
    for (unsigned si = 0; si < solutionSize; ++si) solDevice->clients[si] = (ushort)solution[si];
    solDevice->update();
  }
  for (int k = kMin; k < kMax; ++k) runKernel(k, moves);
  if (solution && !moves)
    for (unsigned si = 0; si < solutionSize; ++si) solution[si] = solDevice->clients[si];
  if (solutionResp && moves)
    for (unsigned si = 0; si < solutionSize; ++si) solutionResp[si] = solDevice->clients[si];
  solDevice->update();
  return solDevice->costCalc();
}

}  // namespace wamca

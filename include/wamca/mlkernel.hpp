// mlkernel.hpp -- the reference's C++ surface for the MLP local search
// (MLProblem / MLSolution / MTRandom / MLKernel{Swap,2Opt,OrOpt} /
// WAMCAExperiment), implemented over the MI355X C ABI (mlp_gpu.h) in
// libwamca_mi.so.  Plain C++17, no HIP headers: link with -lwamca_mi.
//
// Reference interfaces mirrored (LabIC-UFF/wamca-gpu-mi):
//   MTRandom                     mtrand.h:168-312, mtrand.cpp:51-121
//   MLProblem(costTour, distRound, coordShift), load()   mlproblem.h:61-151, mlproblem.cpp:58-207
//   MLSolution: clients/weights/cost, update, costCalc, random, sequence, save/load
//                                mlsolution.h, mlsolution.cpp:215-302, 684-906
//   MLMove64 / MLMovePack / MLMove / MLMoveId            mlads.h:144-216
//   MLKernel: init, term, setSolution, getSolution, sendSolution, defineKernelGrid,
//             launchKernel, sync, recvResult, mergeGPU, bestMove, mergeGreedy, applyMove
//                                mlkernel.h:104-318, mlkernel.cu:254-864
//   MLKernelSwap / MLKernel2Opt / MLKernelOrOpt           mlkswap.h:29-48, mlk2opt.h:29-48, mlkoropt.h:29-48
//   WAMCAExperiment::runKernel / runWAMCA2016             WamcaExperiment.hpp:292-581
//
// Differences: device work goes through the C ABI (the ADS is built on the
// device from the tour; sendSolution uploads the tour only; moveData is
// device memory of the library's context, cleared with mlp_dev_memset or
// hipMemset); latencies are
// kept as uint32 like the reference (cost64 holds the exact value); errors
// throw wamca::Error instead of exiting.  TOTAL kernels (isTotal = true)
// use `totalBlock` threads per row: the reference takes the block size from
// the CUDA occupancy calculator (mlkswap.cu:1009-1014), which has no MI355X
// counterpart, so it is a field here (0: the largest multiple of 64 not above
// min(size, 1024)); set it to a logged blck(...) value to reproduce that run.
// Move ids go up to MLMI_OROPT13 (the 4-bit id of MLMove64 holds 0..15 and
// the device keeps buffers for ids 0..14).
#ifndef WAMCA_MLKERNEL_HPP_
#define WAMCA_MLKERNEL_HPP_

#include <cstdint>
#include <iosfwd>
#include <stdexcept>
#include <string>
#include <vector>

#include "../mlp_gpu.h"

namespace wamca {

using uint = unsigned int;
using ushort = unsigned short;

struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

// mlads.h:144-176
enum MLMoveId {
  MLMI_SWAP,
  MLMI_2OPT,
  MLMI_OROPT1,
  MLMI_OROPT2,
  MLMI_OROPT3,
  MLMI_OROPT4,
  MLMI_OROPT5,
  MLMI_OROPT6,
  MLMI_OROPT7,
  MLMI_OROPT8,
  MLMI_OROPT9,
  MLMI_OROPT10,
  MLMI_OROPT11,
  MLMI_OROPT12,
  MLMI_OROPT13,
};

// mlads.h:190-216: 64-bit packed move (id:4 | i:14 | j:14, int32 cost)
struct MLMove64 {
  unsigned int id : 4;
  unsigned int i : 14;
  unsigned int j : 14;
  int cost;
};
union MLMovePack {
  MLMove64 s;
  uint64_t w;
};
struct MLMove {
  MLMoveId id;
  int i;
  int j;
  int cost;
};
static_assert(sizeof(MLMove64) == 8, "MLMove64 is 8 bytes");

// mlads.h:221-237: typed views of one buffer address
union MLPointer {
  const void* p_cvoid;
  void* p_void;
  char* p_char;
  unsigned char* p_byte;
  short* p_short;
  ushort* p_ushort;
  int* p_int;
  uint* p_uint;
  long* p_long;
  unsigned long* p_ulong;
  long long* p_llong;
  unsigned long long* p_ullong;
  MLMove* p_move;
  MLMove64* p_move64;
  MLMovePack* p_mpack;
};

extern const char* nameMove[15];  // mlads.cpp:41-72, ids 0..14

class MTRandom {
 public:
  MTRandom() { seed(5489U); }
  explicit MTRandom(uint s) { seed(s); }
  void seed(uint s);
  uint getSeed() const { return rseed_; }
  uint rand();
  uint rand(uint max) { return (uint)(((max + 1.0) * rand()) / 4294967296.0); }
  uint rand(uint min, uint max) { return min + rand(max - min); }
  double realCO() { return ((double)rand()) * (1.0 / 4294967296.0); }
  double realOO() { return (((double)rand()) + 0.5) * (1.0 / 4294967296.0); }
  template <typename T>
  void shuffle(T* data, uint n) {
    for (uint i = 0; i + 1 < n; ++i) {
      const uint j = rand(n - i - 1);
      T t = data[i];
      data[i] = data[j];
      data[j] = t;
    }
  }

 private:
  void regen();
  uint state_[624];
  int p_ = 624;
  uint rseed_ = 0;
};

struct MLClientData {
  ushort id;
  int x, y;
};

class MLProblem {
 public:
  MLProblem(bool costTour = true, bool distRound = false, bool coordShift = false);
  ~MLProblem();
  MLProblem(const MLProblem&) = delete;
  MLProblem& operator=(const MLProblem&) = delete;
  void load(const char* fname);
  void create(const int32_t* xs, const int32_t* ys, int n, const char* name = "synthetic");
  int weight(int a, int b) const { return w_[(size_t)a * n_ + (size_t)b]; }

  bool costTour, distRound, coordShift;
  uint size = 0;  // positions of a solution (clients + costTour)
  char name[64] = {0};
  char filename[256] = {0};
  std::vector<MLClientData> clients;  // size entries; the depot copy repeats client 0
  int maxWeight = 0;

  // device side, created lazily by the first kernel init on a context
  mlp_ctx* ctx();
  int32_t instance();
  int clientCount() const { return n_; }
  const int32_t* xs() const { return xs_.data(); }
  const int32_t* ys() const { return ys_.data(); }

 private:
  void build_weights();
  int n_ = 0;
  std::vector<int32_t> xs_, ys_, w_;
  mlp_ctx* ctx_ = nullptr;
  int32_t inst_ = -1;
};

class MLSolution {
 public:
  explicit MLSolution(MLProblem& prob, int flags = 0);
  void clear() {
    clientCount = 0;
    cost = 0x7fffffff;
  }
  void add(ushort c) { clients[clientCount++] = c; }
  void assign(const MLSolution* sol, bool ads = true);
  void update();
  uint costCalc() const;
  void random(MTRandom& rng, float alpha);
  void sequence();
  int dist(uint i, uint j) const { return problem.weight(clients[i], clients[j]); }
  void save(std::ostream& os) const;
  void load(std::istream& is);
  std::string str(bool showCost = false) const;

  MLProblem& problem;
  std::vector<ushort> clients;
  std::vector<ushort> weights;
  uint clientCount = 0;
  uint cost = 0x7fffffff;
  int64_t cost64 = 0;  // exact latency from the last update()
  MLMove move{MLMI_SWAP, 0, 0, 0};
};

class MLKernel {
 public:
  MLKernel(MLProblem& problem, bool isTotal, int kid, uint ktag = 0);
  virtual ~MLKernel();
  MLKernel(const MLKernel&) = delete;
  MLKernel& operator=(const MLKernel&) = delete;

  void init(bool solCreate);
  void term();
  void resetStats() {
    callCount = mergeCount = imprvCount = 0;
    timeMove = 0;
  }
  void setSolution(MLSolution* sol, bool ads = true) { solution->assign(sol, ads); }
  void getSolution(MLSolution* sol, bool ads = true) { sol->assign(solution, ads); }
  void sendSolution();
  void defineKernelGrid();
  void launchKernel();
  void sync();
  void recvResult();
  void mergeGPU();
  int bestMove(MLMove& move);
  int mergeGreedy(MLMove64* merge, int& count);
  virtual void applyMove(MLMove& move) = 0;
  // HIP events on the problem's stream around the calls in between
  // (mlkernel.h:263-276); timerElapsed waits for the stop event and returns
  // microseconds
  void timerStart();
  void timerStop();
  unsigned long long timerElapsed();

  MLProblem& problem;
  int id;
  uint tag;
  const char* name;
  bool isTotal;
  int rowmap = MLP_ROWMAP_ODC;  // the shipped build's kernels; MLP_ROWMAP_ADS for the logged ones
                                // (TOTAL kernels always use the ADS row map)
  uint totalBlock = 0;          // TOTAL: threads (entries) per row; 0 = default, set before init
  unsigned long long time = 0;  // last timerElapsed(), microseconds

  // Sizes and buffers of mlkernel.h:117-136 (set by init, mlkernel.cu:442-556).
  uint solSize = 0;      // positions of a tour (problem.size)
  uint adsDataSize = 0;  // bytes of the reference's ADS buffer for solSize: 128 + 4 N (N + 3);
                         // the device keeps the ADS in O(N) prefix form instead
                         // (mlp_ads_materialize builds this layout on request)
  uint adsRowElems = 0;  // = solSize
  // Device buffer of MLMove64 records (N entries; TOTAL: N * N): launchKernel
  // leaves the per-row (per-entry) records in it, mergeGPU the merged list
  // (kept moves first in merge order, every other entry with cost 0);
  // recvResult copies moveElems entries to transBuffer.  Callers may clear it
  // (mlp_dev_memset) before a launch, as the reference's callers do.
  MLMovePack* moveData = nullptr;
  uint moveDataSize = 0;  // bytes
  uint moveElems = 0;     // records of the last launch (rows, or rows * totalBlock)
  MLPointer transBuffer{};  // host copy of moveData (transBufferSize bytes)
  uint transBufferSize = 0;
  MLSolution* solution = nullptr;
  bool solDestroy = false;
  uint callCount = 0, mergeCount = 0, imprvCount = 0;
  unsigned long long timeMove = 0;

 protected:
  void reserveMoves(size_t entries);
  std::vector<MLMove64> host_;  // storage behind transBuffer
  int32_t slot_ = -1;
  int32_t timer_ = -1;
};

class MLKernelSwap : public MLKernel {
 public:
  explicit MLKernelSwap(MLProblem& p, bool isTotal = false) : MLKernel(p, isTotal, MLMI_SWAP) {}
  void applyMove(MLMove& move) override;
};

class MLKernel2Opt : public MLKernel {
 public:
  explicit MLKernel2Opt(MLProblem& p, bool isTotal = false) : MLKernel(p, isTotal, MLMI_2OPT) {}
  void applyMove(MLMove& move) override;
};

class MLKernelOrOpt : public MLKernel {
 public:
  MLKernelOrOpt(MLProblem& p, uint tag, bool isTotal = false)
      : MLKernel(p, isTotal, (int)MLMI_OROPT1 + (int)tag - 1, tag) {}
  void applyMove(MLMove& move) override;
};

// In-repo driver: one pass over neighborhoods [kMin, kMax) chained on the
// solution (WamcaExperiment.hpp:292-581).
class WAMCAExperiment {
 public:
  WAMCAExperiment(MLProblem& problem, int seed);
  ~WAMCAExperiment();
  void runKernel(int k, std::vector<MLMove>* moves = nullptr);
  unsigned int runWAMCA2016(int mMax = 3, int kMin = 0, int kMax = 3, int* solution = nullptr,
                            unsigned int solutionSize = 0, std::vector<MLMove>* moves = nullptr,
                            int* solutionResp = nullptr);
  MLProblem& problem;
  MTRandom rng;
  MLKernel* kernels[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  MLSolution* solDevice;
};

}  // namespace wamca

#endif  // WAMCA_MLKERNEL_HPP_

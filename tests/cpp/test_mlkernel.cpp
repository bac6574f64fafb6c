// test_mlkernel.cpp -- the reference's gtest cases (tests/mtrand.Test.cpp,
// tests/mlproblem.Test.cpp, tests/mlsolution.Test.cpp,
// tests/WamcaExperiment.Test.cpp) restated against include/wamca/mlkernel.hpp,
// with a minimal check harness (googletest is not available).
//
//   test_mlkernel cpu <instances-dir>         CPU cases
//   test_mlkernel gpu <instances-dir>         GPU cases (need a device)
//   test_mlkernel chain <tsp> <seed> <rowmap> one runWAMCA2016 pass over the
//                                             five kernels; prints the result
//   test_mlkernel total <tsp> <seed> <sols> <kid>:<block> ...
//                                             the logged TOTAL section of each of
//                                             the first <sols> random solutions
//                                             (WamcaExperiment.hpp:368-581)
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "wamca/mlkernel.hpp"

using namespace wamca;

// The reference's callers reach the device through gpu* macros (gpu.h:51-134);
// here they map to the HIP runtime the library runs on.
#define gpuMemset hipMemset
#define gpuDeviceSynchronize hipDeviceSynchronize

static int g_fail = 0, g_checks = 0;
#define EXPECT_EQ(a, b)                                                                     \
  do {                                                                                      \
    ++g_checks;                                                                             \
    const long long va_ = (long long)(a), vb_ = (long long)(b);                             \
    if (va_ != vb_) {                                                                       \
      ++g_fail;                                                                             \
      std::printf("FAIL %s:%d: %s == %lld, expected %lld\n", __FILE__, __LINE__, #a, va_, vb_); \
    }                                                                                       \
  } while (0)

static std::string g_dir;

static MLProblem& berlin52() {
  static MLProblem* p = nullptr;
  if (!p) {
    p = new MLProblem(true, false, false);
    p->load((g_dir + "/01_berlin52.tsp").c_str());
  }
  return *p;
}

// tests/mtrand.Test.cpp:11-48
static void test_mtrand() {
  MTRandom rng;
  rng.seed(0);
  EXPECT_EQ(rng.rand(), 2357136044u);
  EXPECT_EQ(rng.rand(), 2546248239u);
  EXPECT_EQ(rng.rand(), 3071714933u);
  EXPECT_EQ(rng.rand(), 3626093760u);
  EXPECT_EQ(rng.rand(), 2588848963u);
  rng.seed(0);
  const int s0[5] = {4, 9, 3, 0, 3};
  for (int q = 0; q < 5; ++q) EXPECT_EQ(rng.rand() % 10, s0[q]);
  rng.seed(0);
  EXPECT_EQ(rng.rand() % 10, 4);
  rng.seed(1);
  const int s1[5] = {5, 9, 4, 8, 3};
  for (int q = 0; q < 5; ++q) EXPECT_EQ(rng.rand() % 10, s1[q]);
  EXPECT_EQ(rng.getSeed(), 1);
}

// tests/mlproblem.Test.cpp:11-28
static void test_mlproblem() {
  MLProblem& p = berlin52();
  EXPECT_EQ(p.size, 53);
  EXPECT_EQ(p.clientCount(), 52);
  EXPECT_EQ(p.clients[52].id, 0);
  EXPECT_EQ(std::string(p.name) == "01_berlin52", 1);
  bool threw = false;
  try {
    MLProblem q;
    q.load((g_dir + "/does-not-exist.tsp").c_str());
  } catch (const Error&) {
    threw = true;
  }
  EXPECT_EQ(threw, 1);
}

// tests/mlsolution.Test.cpp:40-80
static void test_mlsolution() {
  MLProblem& p = berlin52();
  MLSolution s(p);
  EXPECT_EQ(s.clientCount, 0);
  MTRandom rng;
  rng.seed(0);
  s.random(rng, 0.50F);
  EXPECT_EQ(s.clientCount, 53);
  s.update();
  EXPECT_EQ(s.costCalc(), 494309);
  EXPECT_EQ(s.cost, 494309);
  EXPECT_EQ(s.cost64, 494309);
  EXPECT_EQ(s.clients[0], 0);
  EXPECT_EQ(s.clients[52], 0);
  // save / load round trip (mlsolution.cpp:684-751)
  std::stringstream ss;
  s.save(ss);
  MLSolution t(p);
  t.load(ss);
  EXPECT_EQ(t.clientCount, 53);
  EXPECT_EQ(t.cost, 494309);
  bool same = true;
  for (uint i = 0; i < 53; ++i) same = same && t.clients[i] == s.clients[i];
  EXPECT_EQ(same, 1);
  // a size that is neither size nor size-1 is rejected
  std::stringstream bad("7\n0\n0 1 2 3 4 5 0\n");
  bool threw = false;
  try {
    t.load(bad);
  } catch (const Error&) {
    threw = true;
  }
  EXPECT_EQ(threw, 1);
  MLSolution q(p);
  q.sequence();
  EXPECT_EQ(q.clientCount, 53);
  EXPECT_EQ(q.clients[51], 51);
}

// tests/WamcaExperiment.Test.cpp:70-132: one ODC swap pass with the device
// merge and the host apply loop of runGpuGpu (:6-58).
static void test_wamca_init_kernel() {
  MLProblem& p = berlin52();
  EXPECT_EQ(p.size, 53);
  MTRandom rng;
  rng.seed(0);
  MLSolution* solDevice = new MLSolution(p);
  solDevice->random(rng, 0.50F);
  solDevice->update();
  EXPECT_EQ(solDevice->costCalc(), 494309);
  MLKernel* kernel = new MLKernelSwap(p);
  kernel->init(true);
  kernel->setSolution(solDevice);
  kernel->sendSolution();
  EXPECT_EQ(solDevice->costCalc(), 494309);
  kernel->defineKernelGrid();
  EXPECT_EQ(kernel->moveElems, 26);  // ceil(53/2) - 1 ODC swap rows
  // the members the reference's callers use (WamcaExperiment.hpp:238,339,
  // tests/WamcaExperiment.Test.cpp:18,115): sizes, the device move buffer
  // cleared before the launch, the pinned-style host view, the event timers
  EXPECT_EQ(kernel->solSize, 53);
  EXPECT_EQ(kernel->adsRowElems, 53);
  EXPECT_EQ(kernel->adsDataSize, 128 + 4 * 53 * 56);
  EXPECT_EQ(kernel->moveDataSize, 53 * 8);
  EXPECT_EQ(kernel->transBufferSize, kernel->moveDataSize);
  EXPECT_EQ(gpuMemset(kernel->moveData, 0, kernel->moveDataSize), hipSuccess);
  EXPECT_EQ(gpuDeviceSynchronize(), hipSuccess);
  MLMove64* h_moves = kernel->transBuffer.p_move64;
  kernel->timerStart();
  kernel->launchKernel();
  kernel->timerStop();
  kernel->sync();
  EXPECT_EQ(kernel->timerElapsed() == kernel->time, 1);
  // the per-row records the launch left in moveData
  std::vector<MLMove64> dev_rows(kernel->moveElems);
  EXPECT_EQ(hipMemcpy(dev_rows.data(), kernel->moveData, kernel->moveElems * 8,
                      hipMemcpyDeviceToHost), hipSuccess);
  kernel->recvResult();
  int best_row = 0;
  for (uint it = 0; it < kernel->moveElems; ++it) {
    EXPECT_EQ(h_moves[it].cost, dev_rows[it].cost);
    EXPECT_EQ(h_moves[it].i * 100000 + h_moves[it].j, dev_rows[it].i * 100000 + dev_rows[it].j);
    if (h_moves[it].cost < h_moves[best_row].cost) best_row = (int)it;
  }
  MLMove bm;
  EXPECT_EQ(kernel->bestMove(bm), h_moves[best_row].cost);
  kernel->mergeGPU();
  kernel->recvResult();
  kernel->sync();
  EXPECT_EQ(kernel->transBuffer.p_move64, h_moves);
  int impr = 0, count = 0;
  for (uint it = 0; it < kernel->moveElems; ++it) {
    const MLMove64& m = kernel->transBuffer.p_move64[it];
    if (m.cost < 0) {
      MLMove mv{(MLMoveId)m.id, (int)m.i, (int)m.j, m.cost};
      impr += m.cost;
      ++count;
      kernel->applyMove(mv);
    }
  }
  kernel->getSolution(solDevice);
  solDevice->update();
  EXPECT_EQ(solDevice->costCalc(), 367403);
  EXPECT_EQ((long long)494309 + impr, 367403);
  // the per-row records: bestMove returns the first minimum; mergeGreedy keeps
  // improving moves whose costs add up to its return value
  kernel->setSolution(solDevice);
  kernel->sendSolution();
  kernel->launchKernel();
  kernel->sync();
  kernel->recvResult();
  EXPECT_EQ(kernel->moveElems, 26);
  int mn = kernel->transBuffer.p_move64[0].cost;
  for (uint it = 1; it < kernel->moveElems; ++it) mn = std::min(mn, kernel->transBuffer.p_move64[it].cost);
  MLMove best;
  EXPECT_EQ(kernel->bestMove(best), mn);
  EXPECT_EQ(best.cost, mn);
  std::vector<MLMove64> merged(kernel->moveElems);
  int mcount = 0;
  const int mcost = kernel->mergeGreedy(merged.data(), mcount);
  int msum = 0;
  for (int q = 0; q < mcount; ++q) msum += merged[(size_t)q].cost;
  EXPECT_EQ(mcost, msum);
  EXPECT_EQ(mcount > 0 || mn >= 0, 1);
  delete kernel;
  delete solDevice;
}

static int run_chain(const char* tsp, int seed, int rowmap) {
  MLProblem p(true, false, false);
  p.load(tsp);
  WAMCAExperiment exp(p, seed);
  for (MLKernel* k : exp.kernels) {
    k->rowmap = rowmap;
    k->defineKernelGrid();
  }
  exp.solDevice->random(exp.rng, 0.50F);
  std::vector<int> sol(exp.solDevice->clients.begin(), exp.solDevice->clients.begin() + p.size);
  std::printf("start");
  for (int c : sol) std::printf(" %d", c);
  std::printf("\n");
  std::vector<MLMove> moves;
  std::vector<int> resp(p.size);
  const unsigned v = exp.runWAMCA2016(1, 0, 5, sol.data(), p.size, &moves, resp.data());
  std::printf("cost %u\nfinal", v);
  for (int c : resp) std::printf(" %d", c);
  std::printf("\nmoves %zu\n", moves.size());
  for (const MLMove& m : moves) std::printf("%d %d %d %d\n", (int)m.id, m.i, m.j, m.cost);
  return 0;
}

// For each solution and kernel: best entry, improving entries, GPU merge and
// CPU merge (sum, count), as the reference's TOTAL section logs them.
static int run_total(const char* tsp, int seed, int sols, int argc, char** argv) {
  MLProblem p(true, false, false);
  p.load(tsp);
  MTRandom rng;
  rng.seed((uint)seed);
  MLSolution sol(p);
  for (int q = 1; q <= sols; ++q) {
    sol.random(rng, 0.50F);
    for (int a = 0; a < argc; ++a) {
      int kid = 0, block = 0;
      if (std::sscanf(argv[a], "%d:%d", &kid, &block) != 2) return 2;
      MLKernel* k = kid == 0   ? (MLKernel*)new MLKernelSwap(p, true)
                    : kid == 1 ? (MLKernel*)new MLKernel2Opt(p, true)
                               : (MLKernel*)new MLKernelOrOpt(p, (uint)kid - 1, true);
      k->totalBlock = (uint)block;
      k->init(true);
      k->setSolution(&sol);
      k->sendSolution();
      k->launchKernel();
      k->recvResult();
      MLMove best;
      k->bestMove(best);
      int imp = 0;
      for (uint e = 0; e < k->moveElems; ++e) imp += k->transBuffer.p_move64[e].cost < 0;
      const uint elems = k->moveElems;
      std::vector<MLMove64> merged(elems);
      int ccount = 0;
      const int csum = k->mergeGreedy(merged.data(), ccount);
      k->mergeGPU();
      k->recvResult();  // the merged list from moveData (WamcaExperiment.hpp:229-290)
      long long gsum = 0;
      int gcount = 0;
      for (uint e = 0; e < k->moveElems && k->transBuffer.p_move64[e].cost < 0; ++e) {
        gsum += k->transBuffer.p_move64[e].cost;
        ++gcount;
      }
      std::printf("%d %s elems %u best %d %d %d improving %d gpu %lld %d cpu %d %d\n", q, k->name,
                  elems, (int)best.i, (int)best.j, best.cost, imp, gsum, gcount, csum, ccount);
      delete k;
    }
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s cpu|gpu <instances-dir> | chain <tsp> <seed> <rowmap>\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  try {
    if (mode == "total" && argc > 5)
      return run_total(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), argc - 5, argv + 5);
    if (mode == "chain") return run_chain(argv[2], argc > 3 ? std::atoi(argv[3]) : 1,
                                          argc > 4 ? std::atoi(argv[4]) : MLP_ROWMAP_ODC);
    g_dir = argv[2];
    if (mode == "cpu") {
      test_mtrand();
      test_mlproblem();
      test_mlsolution();
    } else if (mode == "gpu") {
      test_wamca_init_kernel();
    } else {
      return 2;
    }
  } catch (const std::exception& e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 1;
  }
  std::printf("%s: %d checks, %d failed\n", mode.c_str(), g_checks, g_fail);
  return g_fail ? 1 : 0;
}

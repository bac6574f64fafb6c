/*
 * dvnd.h -- the reference's extern "C" DVND interface, served by the MI355X
 * library (libwamca_mi.so).  Same symbol names, C types and argument meaning
 * as the reference, for an external DVND driver that loads it by name.
 *
 *   bestNeighborSimple   dvnd.cuh:33, dvnd.cu:8-10
 *   bestNeighbor         dvnd.cuh:34-36, dvnd.cu:12-71
 *   noConflict           dvnd.cuh:40, dvnd.cu:73-82, 132-175
 *   getNoConflictMoves   dvnd.h:22-23, dvnd.cpp:155-161 (+ betterNoConflict dvnd.cu:85-130)
 *   applyMoves           dvnd.h:19-20, dvnd.cpp:16-145
 *   segFault             dvnd.cpp:230-233
 *
 * Behaviour kept from the reference:
 *   - instances are TSPLIB EUC_2D files, tours of N = n + 1 positions ending
 *     with the depot, truncated distances (costTour, !distRound, !coordShift,
 *     dvnd.cpp:168-170);
 *   - the first problem bestNeighbor loads serves every later call (a static
 *     initialised once, dvnd.cu:23); applyMoves caches problems per file path
 *     of hostCode 0 (dvnd.cpp:163-175);
 *   - bestNeighbor runs one neighborhood (0 swap, 1 2-opt, 2..4 or-opt-1..3)
 *     with the GPU merge and returns the uint32 latency of the new tour; the
 *     new tour goes to solutionResp when moves are requested (*useMoves != 0),
 *     otherwise back into solution;
 *   - noConflict keeps the reference's asymmetric cross-neighborhood tests
 *     (swap x 2-opt uses move2.j - 1, 2-opt x 2-opt tests move2.j >
 *     move1.i - 1, or-opt x swap mixes move1.i / move2.i);
 *   - applyMoves applies id 1 (2-opt) as a swap of the end points.
 * Deliberate differences: no exit() or exception crosses the ABI (errors
 * return UINT_MAX / -1 and are reported on stderr); calls are serialised by a
 * mutex; the kernels use the shipped build's ODC row partition unless the
 * environment sets MLP_DVND_ROWMAP=ads.  MLP_DVND_STRICT=0 (default 1 =
 * reference behaviour) makes a NULL useMoves mean "no moves" instead of the
 * reference's NULL dereference, and keys problems by hostCode.
 */
#ifndef MLP_DVND_H_
#define MLP_DVND_H_

#ifdef __cplusplus
extern "C" {
#define MLP_DEFAULT(v) = v
#else
#include <stdbool.h>
#define MLP_DEFAULT(v)
#endif

unsigned int bestNeighborSimple(char* file, int* solution, unsigned int solutionSize,
                                int neighborhood);

unsigned int bestNeighbor(char* file, int* solution, unsigned int solutionSize, int neighborhood,
                          bool justCalc MLP_DEFAULT(false), unsigned int hostCode MLP_DEFAULT(0),
                          unsigned int* useMoves MLP_DEFAULT(0), unsigned short* ids MLP_DEFAULT(0),
                          unsigned int* is MLP_DEFAULT(0), unsigned int* js MLP_DEFAULT(0),
                          int* costs MLP_DEFAULT(0), bool useMultipleGpu MLP_DEFAULT(false),
                          unsigned int deviceCount MLP_DEFAULT(1),
                          int* solutionResp MLP_DEFAULT(0));

bool noConflict(unsigned short id1, unsigned int i1, unsigned int j1, unsigned short id2,
                unsigned int i2, unsigned int j2);

int getNoConflictMoves(unsigned int numberOfMoves, unsigned short* ids, unsigned int* is,
                       unsigned int* js, int* costs, int* selectedMoves, int* impValue,
                       bool maximize MLP_DEFAULT(false), bool melhorParaPior MLP_DEFAULT(false));

unsigned int applyMoves(char* file, int* solution, unsigned int solutionSize,
                        unsigned int numberOfMoves, unsigned short* ids, unsigned int* is,
                        unsigned int* js, int* costs);

void segFault(void);

#undef MLP_DEFAULT

#ifdef __cplusplus
}
#endif

#endif /* MLP_DVND_H_ */

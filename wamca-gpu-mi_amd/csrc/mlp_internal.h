// mlp_internal.h -- data layout shared by the HIP kernels and the host runtime.
//
// A slot is one device-resident solution.  Its ADS is kept in O(N) prefix form
// (the closed form of the reference's O(N^2) ADS, mlsolution.cpp:419-486):
//   xy[p]  packed tour-ordered coordinates (x << 16 | y) of client tour[p]
//   w[p]   edge weight d(tour[p-1], tour[p]), w[0] = 0       (mlsolution.cpp:224-226)
//   T[p]   prefix duration  sum_{q<=p} w[q]                  (ADS "T" row)
//   Q[p]   prefix moment    sum_{q<=p} q * w[q]
// C[i][j] of the reference equals (S[j]-S[i]) - (j-i) T[i] with S the prefix
// of T; the move costs below only need T and Q.  Both are kept as int64 and as
// their low 32 bits (int32 path, exact while every |delta| < 2^31).
#ifndef MLP_INTERNAL_H_
#define MLP_INTERNAL_H_

#include <stdint.h>

#define MLP_NB_COUNT 5   // swap, 2-opt, or-opt-1/2/3: the tiled sweep's neighborhoods
#define MLP_NB_MAX 15    // ids up to 14: or-opt-k for k <= 13 (MLMove64 id is 4 bits)
#define MLP_KEY_SHIFT 20                       // row-local index bits in a key
#define MLP_KEY_MASK ((1LL << MLP_KEY_SHIFT) - 1)
#define MLP_KEY_NONE INT64_MAX                 // row without a move

// info[] words of a slot
enum {
  MLP_INFO_LATENCY = 0,   // int64 latency of the current tour
  MLP_INFO_TTOT = 1,      // T[N-1]
  MLP_INFO_PATH32 = 2,    // MLP_P32_* flags: neighborhoods whose int32 path is exact
  MLP_INFO_ROUND_MOVES = 3,  // moves applied since the last round reset
  MLP_INFO_TOTAL_MOVES = 4,  // moves applied since the tour was set
  MLP_INFO_ACTIVE = 5,    // 1 while a DVND run on this slot continues
  MLP_INFO_ROUNDS = 6,    // DVND rounds run since the tour was set
  MLP_INFO_MCOUNT = 7,    // improving rows gathered by a many-workgroup merge (zero between merges)
  MLP_INFO_KEPT0 = 8,     // kept-move count of nb (8 + nb)
  MLP_INFO_MG_APPLY = 23, // moves the last many-workgroup merge applied (its rebuild runs if > 0)
  MLP_INFO_PARITY0 = 24,  // row buffer (0/1) holding the last eval of nb (24 + nb)
  MLP_INFO_WORDS = 40
};

// info[PATH32] bits: 2-opt / swap and or-opt deltas fit int32; swap and
// or-opt deltas below 2^26 (keys packed with a 5-bit column iteration);
// T[N-1] below 2^24 (products of a position and a prefix time are exact in
// 24-bit multiplies)
enum { MLP_P32_2OPT = 1, MLP_P32_SWAP_OROPT = 2, MLP_PK_SWAP_OROPT = 4, MLP_T24 = 8 };

struct SlotDev {
  int32_t N;      // positions (n + 1)
  int32_t n;      // clients
  int32_t round;  // 0 trunc, 1 nint
  int32_t maxd;   // upper bound of any distance in the instance
  int32_t rowcap; // entries per row buffer (rows[nb] holds two buffers)
  const int32_t* xs;  // instance coordinates (client indexed)
  const int32_t* ys;
  int32_t* tour;  // [N]
  int32_t* xy;    // [N]
  int32_t* w;     // [N]
  int32_t* t32;   // [N]
  int32_t* q32;   // [N]
  int32_t* g32;   // [N] 2-opt i-side constant -(N-p) w[p] - Z[p]      (mod 2^32)
  int32_t* h32;   // [N] 2-opt j-side constant Z[p] - (N-p-1) w[p+1]  (mod 2^32)
  int64_t* t64;   // [N]
  int64_t* q64;   // [N]
  int64_t* info;  // [MLP_INFO_WORDS]
  // per-row best keys (delta << 20 | ctx), two buffers of rowcap entries.
  // info[PARITY0 + nb] = p names the buffer holding the last result; the
  // other one is all KEY_NONE.  An eval writes buffer p ^ 1 (column tiles
  // fold with atomicMin) and resets buffer p.  A standalone eval gets p ^ 1
  // from the host and its first tile records it; inside a DVND round the
  // eval reads p from info and the merge that follows flips it.
  int64_t* rows[MLP_NB_MAX];
  int64_t* kept_cost[MLP_NB_MAX]; // merged moves, sorted
  int32_t* kept_ij[MLP_NB_MAX];   // (i, j, row) triplets
  int32_t nbs;                    // neighborhoods with buffers (ids 0 .. nbs-1)
  void* mscratch;  // the merge's working arrays when they exceed LDS (layout of its LDS)
};

// Z[p] = 2 Q[p] - (2p + 1) T[p] turns the 2-opt internal sum into
// Z[j] - Z[i] + (j - i)(T[i] + T[j]).

// (slot, neighborhood) segment of an eval launch: rows of the row map are cut
// into blocks of MLP_EVAL_ROWS_PER_WG rows; blk0 is the first block index.
#define MLP_EVAL_ROWS_PER_WG 4
struct EvalSeg {
  int32_t slot;
  int32_t nb;
  int32_t rows;
  int32_t blk0;
  int32_t parity;  // buffer written by a standalone eval
};
#define MLP_EVALSEG_INTS 5

// Sweep segments: kind 0 = fixed-i row blocks (swap + or-opt-1/2/3 rows of
// the ADS row map sharing distance vectors; with mask bit 1 also the 2-opt
// moves (i, j > i) of the block's rows, folded per diagonal row), kind 1 =
// 2-opt diagonal rows.
#define MLP_2OPT_ROWS 4
struct SweepSeg {
  int32_t slot;
  int32_t kind;
  int32_t mask;    // neighborhoods of kind 0 to evaluate (bits 0, 2, 3, 4; bit 1 with bit 0)
  int32_t blk0;
  int32_t ncols;   // column tiles per row block
  int32_t parity;  // bit nb: buffer written by a standalone eval
};
#define MLP_SWEEPSEG_INTS 6
// Tile shapes (rows x columns) of the sweep: from 4 x 256 (most workgroups,
// for launches of one or a few tours) to 16 x 1024 (fewest distance
// evaluations per move, for large batches).
#define MLP_TILE_4x256 0
#define MLP_TILE_8x512 1
#define MLP_TILE_16x1024 2
#define MLP_TILE_8x256 3
#define MLP_TILE_4x512 4
#define MLP_TILE_COUNT 5
// Launch flag beside the shape: the fixed-row tiles also evaluate 2-opt
// (shapes of eight or more rows, 16 x 1024 then run as 12 x 1024; the launch
// has no 2-opt diagonal tiles).
#define MLP_TILE_TWO 0x100
// ... or the ODC row map for swap and or-opt (16 x 1024 then runs as
// 8 x 512, and 8 x 256 as 16 x 128: an or-opt diagonal gets one key from
// each tile it crosses, and all of a tour's tiles fold at once, so the row
// count of the tile sets the contention on each ODC row).
#define MLP_TILE_ODC 0x200

#define MLP_SEGS_INLINE 32
struct EvalSegsInline {
  EvalSeg s[MLP_SEGS_INLINE];
};
struct SweepSegsInline {
  SweepSeg s[MLP_SEGS_INLINE];
};

// A fused DVND launch: `rounds` rounds of {eval nb, merge nb} for nb = 0..4
// over `count` slots as one persistent grid draining a work queue.  Items are
// numbered round-major, then by phase (eval swap, merge swap, eval 2-opt, ...,
// merge or-opt-3), then by slot (and tile); an item only waits for items of
// its own slot with smaller numbers (the merge of the previous phase, the
// tiles of its own eval phase), so the queue always drains.
struct FusedPlan {
  int32_t count, per_round, rounds, tail, pcap, ncap;
  int32_t phase[11];          // first item of each phase within a round; [10] = per_round
  const int32_t* slots;       // [count]
  const int32_t* prefix;      // [5][count + 1]: eval items of nb before slot q
  const int32_t* tiles;       // [5][count]: eval tiles of nb for slot q
  const int32_t* ncols;       // [count][2]: column tiles of the fixed-row / 2-opt tiles
  int32_t group;              // tiles per eval item (one wait, one signal)
  int32_t* sync;              // [4 + 2 count]: work, exits, timeout flag, -, done[count], merged[count]
};
// positions per workgroup of the many-workgroup merge stages (kBlock threads
// x 8 positions)
#define MLP_MG_CHUNK 2048
#define MLP_FUSED_R 8
#define MLP_FUSED_C 512
// Streams (slot groups) a DVND call may use (64 instances, n in [50, 2000]:
// 42.3 ms with one, 37.7 with two, 51.6 with three, 48.6 with four).
#define MLP_DVND_GROUPS 2

#ifndef MLP_FUSED_G
#define MLP_FUSED_G 8
#endif

#define MLP_SLOTS_INLINE 128
// 64-bit word ranges of a slot block holding row buffers (set to KEY_NONE
// when the slot is created)
struct SlotInitRows {
  int64_t start[MLP_NB_MAX];
  int64_t count[MLP_NB_MAX];
  int32_t nbs;
};

struct SlotListInline {
  int32_t id[MLP_SLOTS_INLINE];
};

#endif  // MLP_INTERNAL_H_

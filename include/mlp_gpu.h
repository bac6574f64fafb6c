/*
 * mlp_gpu.h -- C ABI of the MI355X Minimum Latency Problem local-search engine.
 *
 * Plain C types only (no HIP, no torch).  Every call returns an int status
 * (MLP_OK == 0, negative on error; mlp_ctx_error() holds the message) and never
 * throws or exits.  A context is externally synchronised: one host thread at a
 * time per context.  Device work is queued on the context's HIP stream; calls
 * that hand results back to host memory synchronise that stream first.
 *
 * Model.  An *instance* holds n client coordinates (client 0 = depot) and the
 * distance rounding mode.  A *slot* is a device-resident solution of one
 * instance: a tour of N = n + 1 positions [0, perm(1..n-1), 0], its ADS (the
 * prefix arrays of the reference's auxiliary data structure), the per-row best
 * move of each neighborhood, and the last merged move set.  Neighborhoods:
 * 0 swap, 1 2-opt, 2/3/4 or-opt-1/2/3 (ids as mlads.h:144-176 MLMoveId).
 * Row maps: 0 = the reference's ADS kernel grids, 1 = its ODC kernel grids.
 *
 * Reference interfaces these entry points replace (LabIC-UFF/wamca-gpu-mi):
 *   MLProblem::load distances           mlproblem.cpp:153-166  -> mlp_instance_create
 *   MLSolution::update / ldsUpdate      mlsolution.cpp:215-228, 419-486
 *                                                              -> mlp_set_tours (device ADS)
 *   MLKernel::sendSolution              mlkernel.h:229-235     -> mlp_set_tours
 *   MLKernel{Swap,2Opt,OrOpt}::launchKernel
 *                                       mlkswap.cu:1038-1064, mlk2opt.cu:705-723,
 *                                       mlkoropt.cu:863-884    -> mlp_eval
 *   MLKernel::recvResult / bestMove     mlkernel.h:240-244, mlkernel.cu:617-647
 *                                                              -> mlp_get_rows / mlp_get_moves64
 *   MLKernel::mergeGPU + applyMove      mlkernel.cu:254-384, mlkswap.cu:1066-1106,
 *                                       mlk2opt.cu:725-768, mlkoropt.cu:886-950
 *                                                              -> mlp_merge_apply
 *   WAMCAExperiment::runKernel chain    WamcaExperiment.hpp:292-365, 368-581
 *                                                              -> mlp_dvnd
 *   MLSolution::ldsUpdate ADS buffer    mlads.h:38-74 layout   -> mlp_ads_materialize
 */
#ifndef MLP_GPU_H_
#define MLP_GPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MLP_OK = 0,
  MLP_EINVAL = -1,   /* bad argument (size, id, tour not a permutation ...) */
  MLP_ENOMEM = -2,   /* device or host allocation failed / admission refused */
  MLP_EDEVICE = -3,  /* HIP runtime error */
  MLP_ERANGE = -4,   /* value outside the representable range (N, delta) */
  MLP_ESTATE = -5    /* call out of order (e.g. rows read before an eval) */
};

enum { MLP_NB_SWAP = 0, MLP_NB_2OPT = 1, MLP_NB_OROPT1 = 2, MLP_NB_OROPT2 = 3, MLP_NB_OROPT3 = 4 };
enum { MLP_ROWMAP_ADS = 0, MLP_ROWMAP_ODC = 1 };

/* Largest supported tour size N (positions).  Per-row keys pack the row-local
 * index in 20 bits and the delta in the remaining 43. */
#define MLP_MAX_N 32768

typedef struct mlp_ctx mlp_ctx;

/* A merged move as returned to the host (the reference's MLMove with an int64
 * cost, mlads.h:211-216). */
typedef struct {
  int64_t cost;
  int32_t id;
  int32_t i;
  int32_t j;
  int32_t row;
} mlp_move;

/* Context: device selection, stream, kernels. */
int mlp_ctx_create(int device, mlp_ctx** out);
int mlp_ctx_destroy(mlp_ctx* ctx);
const char* mlp_ctx_error(const mlp_ctx* ctx);
const char* mlp_strerror(int status);
/* The HIP device of the context. */
int mlp_ctx_device(mlp_ctx* ctx, int32_t* device);
/* Without waiting: *idle = 1 when every call queued on the context stream has
 * finished on the device. */
int mlp_query(mlp_ctx* ctx, int32_t* idle);
int mlp_sync(mlp_ctx* ctx);
/* Device memory held by instances and slots (live) and reserved by the
 * context's pool (released blocks are reused, not returned to HIP). */
int mlp_pool_stats(mlp_ctx* ctx, int64_t* live_bytes, int64_t* reserved_bytes);
int mlp_device_info(mlp_ctx* ctx, int32_t* cu_count, int64_t* hbm_bytes, char* name, int32_t name_len);

/* Device memory free and total on the context's GPU (hipMemGetInfo). */
int mlp_mem_info(mlp_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);

/* Instances.  round = 0 truncates, 1 rounds to nearest (dr = 0.5).
 * Coordinates may be any int32 values whose span (max - min) is at most 32767
 * in x and in y; an axis whose values leave [0, 32767] is stored translated
 * to [0, span], which leaves every distance unchanged (the reference reads TSPLIB coordinates as they are,
 * mlproblem.cpp:110-117, 153-166).  MLP_ERANGE: a larger span, n + 1 above
 * MLP_MAX_N, or N (N + 32) maxd >= 2^43 (maxd bounds a distance; the row keys
 * hold delta * 2^20 + index in 64 bits). */
int mlp_instance_create(mlp_ctx* ctx, const int32_t* xs, const int32_t* ys, int32_t n,
                        int32_t round, int32_t* inst);
/* As mlp_instance_create, but the coordinates are staged in pinned host
 * memory and their copy is only enqueued on the context stream (the call does
 * not wait for queued kernels); xs and ys may be reused as soon as it
 * returns. */
int mlp_instance_create_async(mlp_ctx* ctx, const int32_t* xs, const int32_t* ys, int32_t n,
                              int32_t round, int32_t* inst);
int mlp_instance_destroy(mlp_ctx* ctx, int32_t inst);
/* The checks mlp_instance_create and mlp_set_tours apply to one job (n
 * clients, a tour of n + 1 positions), without device work: MLP_OK, or the
 * status and message those calls would return. */
int mlp_check_job(mlp_ctx* ctx, const int32_t* xs, const int32_t* ys, int32_t n, int32_t round,
                  const int32_t* tour);
/* Device-computed client-indexed distance matrix (n*n int32) -- used to check
 * the device distance function against the host one. */
int mlp_instance_weights(mlp_ctx* ctx, int32_t inst, int32_t* w_out);

/* HBM bytes a slot of an n-client instance holds (admission control). */
int64_t mlp_slot_footprint(int32_t n);

/* Slots. */
int mlp_slot_create(mlp_ctx* ctx, int32_t inst, int32_t* slot);
int mlp_slot_destroy(mlp_ctx* ctx, int32_t slot);

/* Upload tours (count slots, tours concatenated, each N ints) and build the
 * ADS on the device.  Each tour must start and end with 0 and visit every
 * client once. */
int mlp_set_tours(mlp_ctx* ctx, const int32_t* slots, int32_t count, const int32_t* tours);
int mlp_get_tours(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t* tours);
int mlp_get_latency(mlp_ctx* ctx, const int32_t* slots, int32_t count, int64_t* latency);

/* Evaluate the neighborhoods in nb_mask (bit nb) for every slot: one per-row
 * best move per row of the row map (ties -> lowest row-local index).
 * persistent != 0 drains a device work queue of (slot, neighborhood, row
 * block) items with a persistent grid instead of one workgroup per item. */
int mlp_eval(mlp_ctx* ctx, const int32_t* slots, int32_t count, uint32_t nb_mask,
             int32_t rowmap, int32_t persistent);

/* Per-row results of the last eval of (slot, nb): cost (INT64_MAX for a row
 * without moves), i, j.  *rows receives the row count; arrays hold MLP_MAX_N. */
int mlp_get_rows(mlp_ctx* ctx, int32_t slot, int32_t nb, int64_t* cost, int32_t* i, int32_t* j,
                 int32_t* rows);

/* Same rows in the reference's 64-bit wire format MLMove64 (mlads.h:190-206):
 * low word id:4 | i:14 | j:14, high word int32 cost.  Returns MLP_ERANGE if a
 * cost does not fit in int32 or N exceeds 16384. */
int mlp_get_moves64(mlp_ctx* ctx, int32_t slot, int32_t nb, uint64_t* moves, int32_t* rows);

/* Device greedy merge of the last eval of nb (stable sort by cost, keep an
 * improving move iff no earlier kept move of the same neighborhood conflicts)
 * and application of the kept moves; the ADS is rebuilt for the new tour. */
int mlp_merge_apply(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t nb);

/* The moves kept by the last mlp_merge_apply of (slot, nb), in sorted order. */
int mlp_get_applied(mlp_ctx* ctx, int32_t slot, int32_t nb, mlp_move* moves, int32_t cap,
                    int32_t* count);

/* TOTAL mode (the reference's isTotal kernels: kernelSwapTotal
 * mlkswap.cu:334-526, kernel2OptTotal mlk2opt.cu:270-447, kernelOrOptTotal
 * mlkoropt.cu:368-612, launched with `block` threads per ADS row): entry
 * row * block + t holds the best move of nb among the row-local indices t,
 * t + block, t + 2 block, ... of that row (ties: lowest index), so a row
 * yields `block` entries instead of one.  block is 1..MLP_TOTAL_MAX_BLOCK;
 * a block at least the row length gives every move an entry of its own.
 * The entries stay with the slot until its next mlp_eval_total or tour
 * upload.  ADS row map only (the reference has no ODC TOTAL kernels). */
#define MLP_TOTAL_MAX_BLOCK 16384
int mlp_eval_total(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t nb,
                   int32_t block);

/* The entries of the slot's last mlp_eval_total: cost (INT64_MAX for an
 * entry without a move), i, j for the first min(cap, rows * block) entries;
 * *entries receives rows * block.  Any array may be NULL. */
int mlp_get_total(mlp_ctx* ctx, int32_t slot, int64_t* cost, int32_t* i, int32_t* j,
                  int64_t cap, int64_t* entries);

/* Device greedy merge over all entries of each slot's last mlp_eval_total
 * (stable sort by cost, keep an improving entry iff no earlier kept move
 * conflicts: mergeGPU mlkernel.cu:254-384 with testSwapTotal /
 * test2OptTotal / testOrOptTotal :183-250) and application of the kept
 * moves; the ADS is rebuilt.  mlp_get_applied(slot, nb) then lists the kept
 * moves, with the entry index in place of the row. */
int mlp_merge_total(mlp_ctx* ctx, const int32_t* slots, int32_t count);

/* DVND to a local optimum: rounds of {eval, merge, apply} over swap, 2-opt,
 * or-opt-1/2/3 until a round applies no move (or max_rounds, 0 = unbounded).
 * rounds_out (per slot, may be NULL) receives the rounds each slot ran. */
int mlp_dvnd(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t rowmap,
             int32_t max_rounds, int32_t* rounds_out);

/* DVND in steps, for callers that interleave other work between rounds (a
 * serving loop admitting new jobs).  mlp_dvnd_start marks the slots active
 * with no rounds run; mlp_dvnd_rounds enqueues `rounds` rounds over the
 * slots (asynchronous; stopped slots are skipped on the device: a slot stops
 * after a round that applied no move, or after max_rounds > 0 rounds);
 * mlp_dvnd_status synchronises and reports, per slot, whether it is still
 * active and how many rounds it ran.  mlp_dvnd runs these to completion. */
int mlp_dvnd_start(mlp_ctx* ctx, const int32_t* slots, int32_t count);
int mlp_dvnd_rounds(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t rowmap,
                    int32_t rounds, int32_t max_rounds);
int mlp_dvnd_status(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t* active,
                    int32_t* rounds);
/* Rounds per convergence check (1..64, default 4), capture of a chunk of
 * rounds into a HIP graph replayed while the active set is unchanged
 * (default on), persistent evaluation grids (default off). */
int mlp_dvnd_config(mlp_ctx* ctx, int32_t chunk, int32_t use_graph, int32_t persistent);
/* on != 0: each DVND round is one persistent launch (default off): a grid of
 * resident workgroups drains a queue of every slot's evaluation tiles and
 * merges in round order, each item waiting only for its own slot's earlier
 * items, so slots of different sizes advance independently within the round.
 * Applies to the ADS row map with the five neighborhoods when every slot's
 * merge fits in LDS (tours of 8 .. about 4000 positions); other launches run
 * the per-neighborhood form.  Results are the same in both forms.  Each
 * dependency costs an agent-scope release and acquire (a write-back and an
 * invalidation of the XCD's L2), so on gfx950 the per-neighborhood launches
 * replayed as a graph, which get coherence at kernel boundaries, are faster
 * (64 instances, n in [50, 2000]: 0.036 s against 0.072 s). */
int mlp_dvnd_fused(mlp_ctx* ctx, int32_t on);
/* groups = 2 (default): mlp_dvnd and mlp_dvnd_rounds split the slots into
 * two groups (the largest slots holding 40% of the work of a round, and the
 * rest) and run the second group's rounds on a second stream, so the merges
 * of one group (one workgroup per slot) overlap the evaluation of the other;
 * in mlp_dvnd each group is checked for convergence as soon as its own chunk
 * of rounds ends; work queued afterwards on the context stream follows both
 * groups (64 instances, n in [50, 2000]: 41.5 ms with one group, 35.5 ms
 * with two).
 * groups = 1: one stream.  Results are the same either way.  Not used with
 * the fused or persistent forms or inside a caller's graph capture. */
int mlp_dvnd_groups(mlp_ctx* ctx, int32_t groups);

/* Device buffers and copies for callers without HIP headers (the C++
 * surface's MLKernel::moveData, mlkernel.h:128-130, allocated in
 * MLKernel::init, mlkernel.cu:551-556).  mlp_dev_alloc returns device memory
 * of the context's GPU, released by mlp_dev_free or with the context;
 * mlp_dev_memset is enqueued on the context stream; the copies are ordered
 * on it and return when done. */
int mlp_dev_alloc(mlp_ctx* ctx, int64_t bytes, void** dev);
int mlp_dev_free(mlp_ctx* ctx, void* dev);
int mlp_dev_memset(mlp_ctx* ctx, void* dev, int32_t value, int64_t bytes);
int mlp_copy_to_host(mlp_ctx* ctx, void* host, const void* dev, int64_t bytes);
int mlp_copy_to_device(mlp_ctx* ctx, void* dev, const void* host, int64_t bytes);

/* The rows of the slot's last evaluation of nb, as MLMove64 words written by
 * the device into the device array dev (cap entries; *rows receives the row
 * count): the per-row records the reference's kernels leave in moveData
 * (mlkswap.cu:298-331, mlk2opt.cu:230-266, mlkoropt.cu:313-346), the cost as
 * the reference's int32 and COST_INFTY with move word 0 for a row without a
 * move.  Enqueued on the context stream.  MLP_ERANGE when N > 16384. */
int mlp_rows_moves64_device(mlp_ctx* ctx, int32_t slot, int32_t nb, uint64_t* dev, int64_t cap,
                            int32_t* rows);
/* The same for the entries of the slot's last mlp_eval_total. */
int mlp_total_moves64_device(mlp_ctx* ctx, int32_t slot, uint64_t* dev, int64_t cap,
                             int64_t* entries);

/* Timers of HIP event pairs on the context stream (MLKernel::timerStart /
 * timerStop / timerElapsed, mlkernel.h:263-276).  elapsed waits for the stop
 * event. */
int mlp_timer_create(mlp_ctx* ctx, int32_t* timer);
int mlp_timer_start(mlp_ctx* ctx, int32_t timer);
int mlp_timer_stop(mlp_ctx* ctx, int32_t timer);
int mlp_timer_elapsed(mlp_ctx* ctx, int32_t timer, float* ms);
int mlp_timer_destroy(mlp_ctx* ctx, int32_t timer);

/* mlp_eval for one tour held in device memory, with no allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  inst is an instance made beforehand; tour (N = n + 1
 * ints), cost (int64), is and js (int32) are device arrays of the caller.
 * The rows of every neighborhood in nb_mask (ids 0..4, in id order,
 * rows_of(nb, rowmap, N) each) land in cost / is / js as mlp_get_rows returns
 * them (INT64_MAX for a row without a move).  The slot's working arrays live
 * in a file-scope device array of 4 MiB that each GPU holds until the process
 * ends (tours up to MLP_STATIC_MAX_N positions; MLP_ERANGE beyond); the call
 * only enqueues kernels on the context stream (no allocation, copy, memset or
 * synchronisation), so it can be captured into a graph.  Every call on one
 * GPU shares that GPU's array and keeps the last call's contents: calls that
 * overlap on one GPU (two contexts, or two streams) race.  The tour is not
 * checked (it is in device memory): a tour that is not [0, perm, 0] gives
 * unspecified rows, but every access stays inside the instance. */
#define MLP_STATIC_MAX_N 8192
int mlp_eval_static(mlp_ctx* ctx, int32_t inst, const int32_t* tour, uint32_t nb_mask,
                    int32_t rowmap, int64_t* cost, int32_t* is, int32_t* js);

/* The reference's ADS buffer (mlads.h:38-74) for the slot's current tour, built
 * on the device: 128-byte header, coords, [client|weight], T, C[N][N] (uint32
 * with the reference's wrap-around).  bytes = 128 + 4*N*(N+3). */
int mlp_ads_materialize(mlp_ctx* ctx, int32_t slot, uint32_t* host_out, int64_t bytes);

/* Device-side distance function over a range of squared lengths s (for
 * exhaustive checks against the host): out[q] = d(s0 + q). */
int mlp_dist_table(mlp_ctx* ctx, int32_t s0, int32_t count, int32_t round, int32_t* out);
/* Mismatch count of the kernels' square root against the library's correctly
 * rounded sqrtf for every float(s), s in [0, 2^31) (0 expected). */
int mlp_sqrt_selfcheck(mlp_ctx* ctx, int64_t* mismatches);
/* Same for an explicit list of squared lengths. */
int mlp_dist_list(mlp_ctx* ctx, const int32_t* s, int32_t count, int32_t round, int32_t* out);

/* Sweep tile shape for mlp_eval with the ADS row map: -1 automatic (chosen
 * from the launch's total move count, or by the table mlp_autotune fills),
 * 0 = 4 rows x 256 columns, 1 = 8 x 512, 2 = 16 x 1024, 3 = 8 x 256,
 * 4 = 4 x 512.  Results do not depend on the shape. */
/* Or-opt-k beyond k = 3: neighborhood id k + 1, up to id 14 (k = 13, the
 * largest id MLMove64 holds).  Slots created after mlp_set_neighborhoods(ctx,
 * count) (5 <= count <= 15) have row buffers for ids 0 .. count-1; their
 * evaluation of ids >= 5 runs one wave per row.  mlp_dvnd_neighborhoods sets
 * the ids a DVND round runs (0 .. count-1, default 5).  mlp_slot_bytes gives
 * the device bytes of a slot created now (mlp_slot_footprint assumes 5). */
int mlp_set_neighborhoods(mlp_ctx* ctx, int32_t count);
int mlp_dvnd_neighborhoods(mlp_ctx* ctx, int32_t count);
int mlp_slot_bytes(mlp_ctx* ctx, int32_t n, int64_t* bytes);
/* Capture of the context stream into a replayable HIP graph: between
 * mlp_graph_begin and mlp_graph_end, mlp_eval, mlp_eval_each, mlp_merge_apply
 * and mlp_dvnd_rounds are recorded instead of run (an eval also records the
 * clearing of the row buffers it folds into, so every replay starts from
 * empty rows).  Calls that synchronise or copy to the host fail while a
 * capture is open.  A graph keeps the slots it was recorded on; replays run
 * in stream order with other work. */
int mlp_graph_begin(mlp_ctx* ctx);
int mlp_graph_end(mlp_ctx* ctx, int32_t* graph);
int mlp_graph_launch(mlp_ctx* ctx, int32_t graph);
int mlp_graph_destroy(mlp_ctx* ctx, int32_t graph);
/* The merge keeps its working arrays in LDS when they fit (about N <= 8190;
 * TOTAL merge: about N <= 13100) and in the slot's device memory otherwise;
 * on = 1 uses device memory for every size (same results; for checking the
 * large-N path on small tours). */
int mlp_merge_force_global(mlp_ctx* ctx, int32_t on);
int mlp_set_tile(mlp_ctx* ctx, int32_t tile);
/* Autotuner of the sweep tile shape: times every shape on an evaluation of
 * nb_mask over the slots (reps launches each, after a warm-up; the slots' row
 * buffers are overwritten), records the fastest for (nb_mask, log2 of the
 * launch's move count) in the context's table, which the automatic shape
 * choice consults (nearest tuned size of the same mask).  us_per_launch, if
 * given, receives 5 timings (one per shape). */
int mlp_autotune(mlp_ctx* ctx, const int32_t* slots, int32_t count, uint32_t nb_mask,
                 int32_t reps, int32_t* best_tile, float* us_per_launch);
/* The table: 32 masks x 48 log2-size buckets of int8 (-1 = untuned). */
#define MLP_TUNE_TABLE_BYTES (32 * 48)
int mlp_tune_get(mlp_ctx* ctx, int8_t* table);
int mlp_tune_set(mlp_ctx* ctx, const int8_t* table);

/* A timed region on the context stream: a HIP event at begin and one at
 * end; end synchronises and returns the elapsed milliseconds. */
int mlp_region_begin(mlp_ctx* ctx);
int mlp_region_end(mlp_ctx* ctx, float* ms);
/* One eval launch per slot (the one-solution-per-launch form), enqueued from
 * a host loop with nothing between the launches. */
int mlp_eval_each(mlp_ctx* ctx, const int32_t* slots, int32_t count, uint32_t nb_mask,
                  int32_t rowmap);
/* Device-to-device copies of `bytes` (reps times): a kernel of known traffic
 * for calibrating memory counters. */
int mlp_copy_probe(mlp_ctx* ctx, int64_t bytes, int32_t reps);
/* Fixed-cost probes with a sweep's launch shape (grid workgroups of 256
 * threads, the sweep's kernel-argument table), reps launches back to back:
 * kind 0 empty, 1 + LDS and a barrier, 2 + the dependent loads a sweep tile
 * starts with (the slot's table entry, info and xy words), 3 + four 64-bit
 * atomic minima per workgroup.  For kernel traces; no result. */
int mlp_launch_probe(mlp_ctx* ctx, int32_t slot, int32_t kind, int32_t grid, int32_t reps);
/* Per-launch eval timing: between begin and end every mlp_eval launch is
 * bracketed by its own pair of HIP events (up to capacity launches); end
 * synchronises and returns the summed kernel time and the launch count. */
int mlp_timing_begin(mlp_ctx* ctx, int32_t capacity);
int mlp_timing_end(mlp_ctx* ctx, double* total_ms, int32_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* MLP_GPU_H_ */

// mlp_launch.h -- host launch wrappers of the kernels in mlp_kernels.hip.
#ifndef MLP_LAUNCH_H_
#define MLP_LAUNCH_H_

#include <hip/hip_runtime.h>

#include <algorithm>

#include "mlp_internal.h"

hipError_t mlp_launch_prep(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                           const int32_t* list_dev, int count);
hipError_t mlp_launch_dvnd_start(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                                 const int32_t* list_dev, int count);
hipError_t mlp_launch_round_begin(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                                  const int32_t* list_dev, int count, int* counters,
                                  int ncounters);
hipError_t mlp_launch_round_end(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                                const int32_t* list_dev, int count, int max_rounds);
hipError_t mlp_launch_eval(hipStream_t st, const SlotDev* table, const EvalSegsInline& inl,
                           const EvalSeg* segs_dev, int nsegs, int total, int rowmap,
                           int* counter, int grid, int skip_inactive);
// flags bit 0: DVND launch (inactive slots skipped); bit 1: filtered folds
// of the diagonal rows (launches with many more tiles than resident
// workgroups); bits 8..31 (with bit 2: two segments per slot): the items per
// slot when every slot of a device segment list has the same segments
hipError_t mlp_launch_sweep(hipStream_t st, int tile, const SlotDev* table,
                            const SweepSegsInline& inl, const SweepSeg* segs_dev, int nsegs,
                            int total, int* counter, int grid, int flags);
void mlp_tile_shape(int tile, int* rows, int* cols);
size_t mlp_merge_lds_bytes(int pcap, int ncap);
hipError_t mlp_merge_setup(size_t max_dyn_lds);
// tail > 0: the last merge of a DVND round also ends the round (max_rounds = tail - 1).
hipError_t mlp_launch_merge_apply(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                                  const int32_t* list_dev, int count, int nb, int rowmap,
                                  int pcap, int ncap, int skip_inactive, int tail);
// The merge of slots whose working arrays exceed LDS over many workgroups
// per slot (four launches; max_n: the largest N among the slots).
hipError_t mlp_launch_merge_gmem(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                                 const int32_t* list_dev, int count, int nb, int rowmap,
                                 int skip_inactive, int tail, int max_n);
// Fused DVND rounds (FusedPlan): dynamic LDS of a launch, its limit, the launch.
size_t mlp_fused_lds_bytes(int pcap, int ncap);
hipError_t mlp_fused_setup(size_t max_dyn_lds);
hipError_t mlp_launch_fused(hipStream_t st, const SlotDev* table, const FusedPlan& plan, int grid);
// TOTAL mode: entries rows * B of nb, their sort keys, the merge + apply.
size_t mlp_total_merge_lds_bytes(int N);
size_t mlp_total_static_lds();
hipError_t mlp_total_setup(size_t max_dyn_lds);
hipError_t mlp_launch_total_eval(hipStream_t st, const SlotDev* table, int id, int nb, int rows,
                                 int B, int64_t* out);
hipError_t mlp_launch_total_keys(hipStream_t st, const int64_t* tk, int64_t M, int B,
                                 uint64_t bias, uint64_t* ku, int32_t* ve);
hipError_t mlp_launch_total_merge(hipStream_t st, const SlotDev* table, int id, int nb, int N,
                                  int B, int64_t M, uint64_t bias, const uint64_t* ku,
                                  const int32_t* ve, int gmem);
// Stable LSD radix sort of (key, value) pairs on bits [0, end_bit) of the
// keys (mlp_sort.hip).  tmp == nullptr: *tmp_bytes receives the scratch size.
hipError_t mlp_sort_pairs_u64(void* tmp, size_t* tmp_bytes, const uint64_t* kin, uint64_t* kout,
                              const int32_t* vin, int32_t* vout, int n, int end_bit,
                              hipStream_t st);
hipError_t mlp_launch_ads(hipStream_t st, const SlotDev* table, int id, int N, uint32_t* S,
                          uint32_t* out);
// MLMove64 words of the rows of the last evaluation of nb / of TOTAL entries.
hipError_t mlp_launch_rows_moves64(hipStream_t st, const SlotDev* table, int id, int nb,
                                  int rowmap, int R, uint64_t* out);
hipError_t mlp_launch_rows_out(hipStream_t st, const SlotDev* table, int id, int mask, int rowmap,
                              int total, int64_t* cost, int32_t* i, int32_t* j);
hipError_t mlp_launch_total_moves64(hipStream_t st, const int64_t* tk, int64_t M, int B, int nb,
                                   int N, uint64_t* out);
hipError_t mlp_launch_dist_table(hipStream_t st, int s0, int count, int round, int32_t* out);
hipError_t mlp_launch_dist_list(hipStream_t st, const int32_t* s, int count, int round,
                                int32_t* out);
hipError_t mlp_launch_sqrt_selfcheck(hipStream_t st, uint32_t s0, uint32_t count,
                                     unsigned long long* bad);
hipError_t mlp_launch_weights(hipStream_t st, const int32_t* xs, const int32_t* ys, int n,
                              int round, int32_t* out);
hipError_t mlp_launch_clear_rows(hipStream_t st, const SlotDev* table, const SweepSegsInline& inl,
                                 const SweepSeg* segs_dev, int nsegs);
hipError_t mlp_launch_probe(hipStream_t st, const SlotDev* table, int slot, int kind, int grid,
                            int64_t* out);
hipError_t mlp_launch_fill_i64(hipStream_t st, int64_t* p, int64_t count, int64_t v);
hipError_t mlp_launch_set_slot(hipStream_t st, SlotDev* table, int id, const SlotDev& d);
hipError_t mlp_launch_slot_init(hipStream_t st, SlotDev* table, int id, const SlotDev& d,
                                int64_t* mem, int64_t words, const SlotInitRows& rows);
hipError_t mlp_launch_gather_tours(hipStream_t st, const SlotDev* table, const int32_t* ids,
                                   int count, int32_t* out);
hipError_t mlp_launch_scatter_tours(hipStream_t st, const SlotDev* table, const int32_t* ids,
                                    int count, const int32_t* in);
hipError_t mlp_launch_gather_info(hipStream_t st, const SlotDev* table, const int32_t* ids,
                                  int count, int word, int64_t* out);

#endif  // MLP_LAUNCH_H_

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
hipError_t mlp_launch_sweep(hipStream_t st, int tile, const SlotDev* table,
                            const SweepSegsInline& inl, const SweepSeg* segs_dev, int nsegs,
                            int total, int* counter, int grid, int skip_inactive);
void mlp_tile_shape(int tile, int* rows, int* cols);
size_t mlp_merge_lds_bytes(int pcap, int ncap);
hipError_t mlp_merge_setup(size_t max_dyn_lds);
hipError_t mlp_launch_merge_apply(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                                  const int32_t* list_dev, int count, int nb, int rowmap,
                                  int pcap, int ncap, int skip_inactive, int gmem);
hipError_t mlp_launch_ads(hipStream_t st, const SlotDev* table, int id, int N, uint32_t* S,
                          uint32_t* out);
hipError_t mlp_launch_dist_table(hipStream_t st, int s0, int count, int round, int32_t* out);
hipError_t mlp_launch_dist_list(hipStream_t st, const int32_t* s, int count, int round,
                                int32_t* out);
hipError_t mlp_launch_sqrt_selfcheck(hipStream_t st, uint32_t s0, uint32_t count,
                                     unsigned long long* bad);
hipError_t mlp_launch_weights(hipStream_t st, const int32_t* xs, const int32_t* ys, int n,
                              int round, int32_t* out);
hipError_t mlp_launch_clear_rows(hipStream_t st, const SlotDev* table, const SweepSegsInline& inl,
                                 const SweepSeg* segs_dev, int nsegs);
hipError_t mlp_launch_fill_i64(hipStream_t st, int64_t* p, int64_t count, int64_t v);
hipError_t mlp_launch_set_slot(hipStream_t st, SlotDev* table, int id, const SlotDev& d);
hipError_t mlp_launch_gather_info(hipStream_t st, const SlotDev* table, const int32_t* ids,
                                  int count, int word, int64_t* out);

#endif  // MLP_LAUNCH_H_

// mlp_kernels.hip -- gfx950 kernels of the MLP local search.
//
//   k_prep          tour -> device ADS (xy, w, T, Q, latency)  [MLSolution::update
//                   + ldsUpdate, mlsolution.cpp:215-228, 419-486]
//   k_eval          per-row best move of swap / 2-opt / or-opt-k over a
//                   (slot, neighborhood, row block) work list, one wave per row,
//                   O(1) integer cost per move, wave argmin on (delta, index)
//                   [kernelSwap / kernel2Opt / kernelOrOpt, mlkswap.cu:54-332,
//                   mlk2opt.cu:41-267, mlkoropt.cu:48-366]
//   k_sweep         the same rows over LDS-tiled distance blocks: fixed-row
//                   tiles (swap, or-opt-1/2/3, optionally 2-opt) folding into
//                   either row map, 2-opt diagonal tiles
//   k_merge_apply   stable sort of the improving rows, greedy non-overlapping
//                   subset, parallel application, ADS rebuild, one workgroup
//                   per slot with its working arrays in LDS
//                   [MLKernel::mergeGPU mlkernel.cu:254-384 + testSwap/test2Opt/
//                   testOrOpt :73-170 + applyMove mlkswap.cu:1066, mlk2opt.cu:725,
//                   mlkoropt.cu:886]
//   k_mg_*          the same merge for tours beyond LDS, its O(N) steps over
//                   many workgroups per slot
//   k_dvnd_fused    a DVND round as one persistent launch
//   k_total_*       TOTAL mode (per-thread-slot entries, sort, merge)
//   k_ads_*         the reference's materialised ADS buffer (mlads.h:38-74)
//
// Integer arithmetic throughout; the only floating-point step is the distance
// d = int(sqrtf(float(dx^2 + dy^2)) + dr) with an IEEE correctly rounded
// square root (mlproblem.cpp:159-162).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "mlp_gpu.h"
#include "mlp_internal.h"

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// distance
// ---------------------------------------------------------------------------
// IEEE correctly rounded square root of f = float(s), s an integer in
// [0, 2^31): the hardware estimate followed by the two-residual correction to
// the nearest float (the same correction LLVM emits for sqrtf); the denormal
// scaling and special-value handling of the general routine are not needed
// for these inputs.  mlp_sqrt_selfcheck compares it with __builtin_sqrtf for
// every s in [0, 2^31).
__device__ __forceinline__ float sqrt_rn_int(float f) {
  float r = __builtin_amdgcn_sqrtf(f);
  const float rm = __int_as_float(__float_as_int(r) - 1);
  const float rp = __int_as_float(__float_as_int(r) + 1);
  const float em = __builtin_fmaf(-rm, r, f);
  const float ep = __builtin_fmaf(-rp, r, f);
  r = em <= 0.0F ? rm : r;
  r = ep > 0.0F ? rp : r;
  return r;
}

__device__ __forceinline__ int dist_s(int s, float dr) {
  const float f = (float)(uint32_t)s;  // round to nearest
  return (int)(sqrt_rn_int(f) + dr);  // RN add, truncating convert
}

// Global-address-space load of a slot array element (the slot table stores
// generic pointers; this keeps the access a global_load).
template <typename T>
__device__ __forceinline__ T ldg(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

// Packed coordinates (x << 16 | y, x and y in [0, 32767]) as two signed
// 16-bit lanes: one packed subtraction gives dx and dy exactly (|d| <= 32767)
// and one 16-bit dot product dx^2 + dy^2 < 2^31 (v_pk_sub_i16, v_dot2).
typedef short short2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int dist_xy(int a, int b, float dr) {
  const short2v d = __builtin_bit_cast(short2v, a) - __builtin_bit_cast(short2v, b);
  return dist_s(__builtin_amdgcn_sdot2(d, d, 0, false), dr);
}

// ---------------------------------------------------------------------------
// row maps (the reference's launch grids and thread -> move maps, tour = 1)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int nb_k(int nb) { return nb >= 2 ? nb - 1 : 0; }

__device__ __forceinline__ int rows_of(int nb, int rowmap, int N) {
  if (nb == 0) return rowmap == 0 ? N - 4 : ((N + 1) >> 1) - 1;  // mlkswap.cu:991,1003
  if (nb == 1) return (N + 1) >> 1;                              // mlk2opt.cu:661
  return N - nb_k(nb) - 2;                                       // mlkoropt.cu:819
}

// Valid row-local index range [lo, hi) of a row.
__device__ __forceinline__ void ctx_range(int nb, int rowmap, int N, int row, int& lo, int& hi) {
  if (nb == 0) {
    if (rowmap == 0) { lo = row + 3; hi = N - 1; }  // j = ctx in [i + 2, N - 2]
    else { lo = 0; hi = N - 4; }
  } else if (nb == 1) {
    lo = 0; hi = N - 3;
  } else {
    lo = 0;
    hi = rowmap == 0 ? N - nb_k(nb) - 1 : N - nb_k(nb) - 2;
  }
}

// (row, ctx) -> (i, j); ctx must lie in ctx_range.
__device__ __forceinline__ void move_ij(int nb, int rowmap, int N, int row, int ctx, int& i, int& j) {
  if (nb == 0) {
    if (rowmap == 0) { i = row + 1; j = ctx; }  // mlkswap.cu:205-206
    else if (ctx >= row) { i = row + 1; j = ctx + 3; }  // mlkswap.cu:766-768
    else { i = N - row - 3; j = N - ctx - 2; }
  } else if (nb == 1) {  // mlk2opt.cu:170-174
    if (ctx >= row) { i = ctx - row + 1; j = ctx + 2; }
    else { i = row - ctx; j = N - ctx - 2; }
  } else {
    const int k = nb_k(nb);
    if (rowmap == 0) {  // mlkoropt.cu:187-247
      i = ctx > row ? row + 1 : row + 2;
      j = ctx + 1;
    } else if (ctx >= row) {  // mlkoropt.cu:697-723
      i = ctx - row + 1; j = ctx + 2;
    } else {
      i = N - ctx - k - 1; j = row - ctx;
    }
  }
}

// ---------------------------------------------------------------------------
// O(1) move deltas from the prefix ADS.  Each equals L(apply(tour)) - L(tour)
// (the reference's ADS concatenation value, mlkswap.cu:210-278 ...).  P32
// computes modulo 2^32, exact when the tour's bound puts every |delta| < 2^31.
// ---------------------------------------------------------------------------
template <bool P32>
struct Arith;
template <>
struct Arith<true> {
  using U = uint32_t;
  __device__ static int64_t fin(U v) { return (int64_t)(int32_t)v; }
  __device__ static U T(const SlotDev& s, int p) { return (U)s.t32[p]; }
  __device__ static U Q(const SlotDev& s, int p) { return (U)s.q32[p]; }
};
template <>
struct Arith<false> {
  using U = int64_t;
  __device__ static int64_t fin(U v) { return v; }
  __device__ static U T(const SlotDev& s, int p) { return s.t64[p]; }
  __device__ static U Q(const SlotDev& s, int p) { return s.q64[p]; }
};

// swap(i, j), j >= i + 2 (ODC positional form, mlkswap.cu:770-785)
template <bool P32>
__device__ __forceinline__ int64_t delta_swap(const SlotDev& s, int i, int j, float dr) {
  using A = Arith<P32>;
  using U = typename A::U;
  const int N = s.N;
  const int a = s.xy[i - 1], xi = s.xy[i], b = s.xy[i + 1];
  const int jm = s.xy[j - 1], xj = s.xy[j], jp = s.xy[j + 1];
  const int t1 = __mul24(N - i, dist_xy(a, xj, dr) - s.w[i]);
  const int t2 = __mul24(N - i - 1, dist_xy(xj, b, dr) - s.w[i + 1]);
  const int t3 = __mul24(N - j, dist_xy(jm, xi, dr) - s.w[j]);
  const int t4 = __mul24(N - j - 1, dist_xy(xi, jp, dr) - s.w[j + 1]);
  return A::fin((U)t1 + (U)t2 + (U)t3 + (U)t4);
}

// 2-opt(i, j): reverse [i..j] (mlk2opt.cu:565-571 with the internal sum in
// closed form: sum_{p=i+1}^{j} (2p - i - j - 1) w[p] = 2(Q[j]-Q[i]) - (i+j+1)(T[j]-T[i]))
template <bool P32>
__device__ __forceinline__ int64_t delta_2opt(const SlotDev& s, int i, int j, float dr) {
  using A = Arith<P32>;
  using U = typename A::U;
  const int N = s.N;
  const int t1 = __mul24(N - i, dist_xy(s.xy[i - 1], s.xy[j], dr) - s.w[i]);
  const int t2 = __mul24(N - j - 1, dist_xy(s.xy[i], s.xy[j + 1], dr) - s.w[j + 1]);
  const U inner = (U)2 * (A::Q(s, j) - A::Q(s, i)) - (U)(i + j + 1) * (A::T(s, j) - A::T(s, i));
  return A::fin((U)t1 + (U)t2 + inner);
}

// or-opt-k(i, j): segment [i..i+k-1] moved to start at j (mlkoropt.cu:704-736)
template <bool P32>
__device__ __forceinline__ int64_t delta_oropt(const SlotDev& s, int k, int i, int j, float dr) {
  using A = Arith<P32>;
  using U = typename A::U;
  const int N = s.N;
  if (i < j) {
    const int t1 = __mul24(N - i, dist_xy(s.xy[i - 1], s.xy[i + k], dr) - s.w[i]);
    const int t2 = __mul24(N - j - k, dist_xy(s.xy[i + k - 1], s.xy[j + k], dr) - s.w[j + k]);
    const int t3 = __mul24(N - j, dist_xy(s.xy[j + k - 1], s.xy[i], dr)) -
                   __mul24(N - i - k, s.w[i + k]);
    const U seg = A::T(s, i + k - 1) - A::T(s, i);
    const U blk = A::T(s, j + k - 1) - A::T(s, i + k);
    return A::fin((U)t1 + (U)t2 + (U)t3 + (U)(i - j) * seg + (U)k * blk);
  } else {
    const int t1 = __mul24(N - j, dist_xy(s.xy[j - 1], s.xy[i], dr) - s.w[j]);
    const int t2 = __mul24(N - i - k, dist_xy(s.xy[i - 1], s.xy[i + k], dr) - s.w[i + k]);
    const int t3 = __mul24(N - j - k, dist_xy(s.xy[i + k - 1], s.xy[j], dr)) -
                   __mul24(N - i, s.w[i]);
    const U seg = A::T(s, i + k - 1) - A::T(s, i);
    const U blk = A::T(s, i - 1) - A::T(s, j);
    return A::fin((U)t1 + (U)t2 + (U)t3 + (U)(i - j) * seg - (U)k * blk);
  }
}

// Wave-wide minimum of an int32 with DPP (row_shr 1/2/4/8 inside each
// 16-lane row, then row_bcast 15 and 31 across rows); the result is read from
// lane 63 and is uniform.  Lanes whose DPP source lies outside the row see the
// identity INT32_MAX.
__device__ __forceinline__ int32_t wave_min_i32(int32_t v) {
  v = min(v, __builtin_amdgcn_update_dpp(INT32_MAX, v, 0x111, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT32_MAX, v, 0x112, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT32_MAX, v, 0x114, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT32_MAX, v, 0x118, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT32_MAX, v, 0x142, 0xa, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(INT32_MAX, v, 0x143, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(v, 63);
}

// Lexicographic (delta, index) minimum over the wave from per-lane bests
// (lanes without a move hold index -1): the minimum delta first, then the
// smallest index among the lanes holding it.  Returns the packed row key.
__device__ __forceinline__ int64_t wave_argmin(int32_t bd, int32_t bc) {
  const int32_t m = wave_min_i32(bc >= 0 ? bd : INT32_MAX);
  const int32_t c = wave_min_i32((bc >= 0 && bd == m) ? bc : INT32_MAX);
  return c == INT32_MAX ? MLP_KEY_NONE
                        : (int64_t)m * (int64_t)(1 << MLP_KEY_SHIFT) + (int64_t)c;
}

// Wave-wide maximum (DPP, as wave_min_i32).
__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
  v = max(v, __builtin_amdgcn_update_dpp(INT32_MIN, v, 0x111, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT32_MIN, v, 0x112, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT32_MIN, v, 0x114, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT32_MIN, v, 0x118, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT32_MIN, v, 0x142, 0xa, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT32_MIN, v, 0x143, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(v, 63);
}

// (delta, index) with the minimum delta and, among equal ones, the highest
// index (lanes without a move hold index -1); (INT32_MAX, -1) when none.
__device__ __forceinline__ void wave_argmin_hi(int32_t bd, int32_t bc, int32_t& m, int32_t& c) {
  m = wave_min_i32(bc >= 0 ? bd : INT32_MAX);
  c = wave_max_i32((bc >= 0 && bd == m) ? bc : -1);
}

__device__ __forceinline__ int64_t wave_min64(int64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int64_t o = __shfl_xor((long long)v, off, kWave);
    v = o < v ? o : v;
  }
  return v;
}

__device__ __forceinline__ float round_dr(int round) { return round ? 0.5F : 0.0F; }

// One row: lanes stride over the row-local index, keep the smallest key
// (delta << 20 | ctx) -> lowest delta, then lowest index; wave reduction.
template <int KIND, bool P32>
__device__ __forceinline__ void eval_row(const SlotDev& s, int nb, int rowmap, int row, int lane, int64_t* out) {
  const int N = s.N;
  const float dr = round_dr(s.round);
  int lo, hi;
  ctx_range(nb, rowmap, N, row, lo, hi);
  int64_t best = MLP_KEY_NONE;
  const int k = nb_k(nb);
  for (int ctx = lo + lane; ctx < hi; ctx += kWave) {
    int i, j;
    move_ij(nb, rowmap, N, row, ctx, i, j);
    int64_t d;
    if (KIND == 0) d = delta_swap<P32>(s, i, j, dr);
    else if (KIND == 1) d = delta_2opt<P32>(s, i, j, dr);
    else d = delta_oropt<P32>(s, k, i, j, dr);
    const int64_t key = d * (int64_t)(1 << MLP_KEY_SHIFT) + ctx;
    best = key < best ? key : best;
  }
  best = wave_min64(best);
  if (lane == 0) *out = best;
}

template <bool P32>
__device__ __forceinline__ void eval_row_dispatch(const SlotDev& s, int nb, int rowmap, int row, int lane,
                                  int64_t* out) {
  if (nb == 0) eval_row<0, P32>(s, nb, rowmap, row, lane, out);
  else if (nb == 1) eval_row<1, P32>(s, nb, rowmap, row, lane, out);
  else eval_row<2, P32>(s, nb, rowmap, row, lane, out);
}

__device__ __forceinline__ EvalSeg find_seg(const EvalSeg* segs, int nsegs, int item) {
  int lo = 0, hi = nsegs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].blk0 <= item) lo = mid;
    else hi = mid - 1;
  }
  return segs[lo];
}

// Inline (kernel-argument) segment tables.  The argument stays in the
// kernarg segment and a dynamic index is one load from it.  One or two
// segments (the launches of a single slot) are read together and selected;
// otherwise the blk0 column is read in one batch and counted, then the
// segment is loaded.
template <typename Seg, typename Inl>
__device__ __forceinline__ Seg find_inline(const Inl& inl, int nsegs, int item) {
  if (nsegs <= 2) {
    const Seg a = inl.s[0], b = inl.s[1];
    return (nsegs == 2 && item >= b.blk0) ? b : a;
  }
  int q = 0;
#pragma unroll
  for (int t = 1; t < MLP_SEGS_INLINE; ++t) q += (t < nsegs && inl.s[t].blk0 <= item) ? 1 : 0;
  return inl.s[q];
}

__device__ __forceinline__ EvalSeg find_seg_inline(const EvalSegsInline& inl, int nsegs, int item) {
  return find_inline<EvalSeg>(inl, nsegs, item);
}

__device__ __forceinline__ void eval_item(const SlotDev* __restrict__ table, const EvalSeg& sg, int item,
                          int rowmap, bool dvnd) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int row = (item - sg.blk0) * MLP_EVAL_ROWS_PER_WG + wave;
  const SlotDev& s = table[sg.slot];
  if (dvnd && s.info[MLP_INFO_ACTIVE] == 0) return;
  const int par = dvnd ? 1 - (int)(s.info[MLP_INFO_PARITY0 + sg.nb] & 1) : sg.parity;
  int64_t* cur = s.rows[sg.nb] + (par ? s.rowcap : 0);
  int64_t* oth = s.rows[sg.nb] + (par ? 0 : s.rowcap);
  if (item == sg.blk0) {  // first block: clear the tails of both buffers, record parity
    for (int r = sg.rows + (int)threadIdx.x; r < s.rowcap; r += kBlock) {
      cur[r] = MLP_KEY_NONE;
      oth[r] = MLP_KEY_NONE;
    }
    if (!dvnd && threadIdx.x == 0) s.info[MLP_INFO_PARITY0 + sg.nb] = par;
  }
  if (row >= sg.rows) return;
  if (lane == 0) oth[row] = MLP_KEY_NONE;
  bool p32;
  if (sg.nb <= 4) {
    p32 = (s.info[MLP_INFO_PATH32] & (sg.nb == 1 ? MLP_P32_2OPT : MLP_P32_SWAP_OROPT)) != 0;
  } else {  // or-opt-k, k > 3: |delta| <= k Ttot + (k + 4) N maxd (one more N maxd margin)
    const int64_t k = nb_k(sg.nb);
    p32 = k * s.info[MLP_INFO_TTOT] + (k + 5) * (int64_t)s.N * s.maxd < 2147483647LL;
  }
  if (p32)
    eval_row_dispatch<true>(s, sg.nb, rowmap, row, lane, cur + row);
  else eval_row_dispatch<false>(s, sg.nb, rowmap, row, lane, cur + row);
}

__device__ __forceinline__ int list_at(const SlotListInline& inl, const int32_t* dev, int q) {
  return dev ? dev[q] : inl.id[q];
}

// ---------------------------------------------------------------------------
// ADS prefix build for one slot by one workgroup (tour may live in LDS).
// Segments of kBlock * kPrepPer positions: thread t owns kPrepPer
// consecutive positions, gathers their coordinates (all loads issued before
// the first wait), forms w, T and Q prefixes in registers, and one
// block-wide exclusive scan of the per-thread totals places them.
// ---------------------------------------------------------------------------
constexpr int kPrepPer = 8;

// 64-bit value of lane (lane - shift within a row of 16, or a row
// broadcast) through DPP on both halves; lanes without a source get 0.
template <int CTRL, int RMASK>
__device__ __forceinline__ int64_t dpp64(int64_t v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(uint64_t)v, CTRL, RMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)((uint64_t)v >> 32), CTRL, RMASK,
                                             0xf, false);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Inclusive wave scan (Hillis-Steele over rows of 16 lanes: shifts 1, 2, 4,
// 8; then row 0's total into row 1 and row 2's into row 3, then rows 0-1's
// total into rows 2 and 3), in registers.
__device__ __forceinline__ int64_t wave_scan_incl(int64_t v, int lane) {
  (void)lane;
  v += dpp64<0x111, 0xf>(v);
  v += dpp64<0x112, 0xf>(v);
  v += dpp64<0x114, 0xf>(v);
  v += dpp64<0x118, 0xf>(v);
  v += dpp64<0x142, 0xa>(v);
  v += dpp64<0x143, 0xc>(v);
  return v;
}

__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
  v = wave_scan_incl(v, 0);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)v, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), 63);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

struct PrepLds {
  int64_t wt[kBlock / kWave][2];
  int64_t carry[2];  // T, Q before this segment
  int64_t lsum[kBlock / kWave];
};

// One segment of kBlock * kPrepPer positions from `base` (thread t owns
// kPrepPer consecutive positions): w, T and Q (offset by carryT, carryQ,
// the prefixes before the segment) and the 2-opt constants G, H; WRITE
// false only sums.  segT / segQ receive the segment's sums of w and p w, lacc
// gains the thread's share of sum_p T[p].  Ends with every thread past its
// reads of P.wt.
// xyc: packed coordinates indexed by client (the merge's copy of the old
// tour's xy, so the new tour needs no gather from the instance arrays), or
// nullptr to read the instance arrays.
template <bool WRITE>
__device__ __forceinline__ void prep_segment(const SlotDev& s, const int32_t* tour, const int32_t* xyc,
                                             int base, int64_t carryT, int64_t carryQ, PrepLds& P,
                                             int64_t& segT, int64_t& segQ, int64_t& lacc) {
  const int N = s.N, tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
  const float dr = round_dr(s.round);
  const int p0 = base + tid * kPrepPer;
  // xy of positions p0 - 1 .. p0 + kPrepPer (clamped into the tour): every
  // client id is loaded before the first coordinate gather
  int cl[kPrepPer + 2], xy[kPrepPer + 2];
#pragma unroll
  for (int e = 0; e < kPrepPer + 2; ++e) cl[e] = tour[min(max(p0 - 1 + e, 0), N - 1)];
#pragma unroll
  for (int e = 0; e < kPrepPer + 2; ++e) {
    // client ids clamped into the instance: a tour read from caller memory
    // never indexes outside the coordinate arrays
    const int c = min(max(cl[e], 0), s.n - 1);
    xy[e] = xyc ? xyc[c] : (ldg(s.xs + c) << 16) | ldg(s.ys + c);
  }
  int w[kPrepPer + 1];  // w[e]: edge into position p0 + e
#pragma unroll
  for (int e = 0; e <= kPrepPer; ++e) {
    const int p = p0 + e;
    w[e] = (p >= 1 && p < N) ? dist_xy(xy[e], xy[e + 1], dr) : 0;
  }
  int64_t lt[kPrepPer], lq[kPrepPer], at = 0, aq = 0;
#pragma unroll
  for (int e = 0; e < kPrepPer; ++e) {
    at += w[e];
    aq += (int64_t)(p0 + e) * w[e];
    lt[e] = at;
    lq[e] = aq;
  }
  const int64_t it = wave_scan_incl(at, lane), iq = wave_scan_incl(aq, lane);
  if (lane == kWave - 1) {
    P.wt[wave][0] = it;
    P.wt[wave][1] = iq;
  }
  __syncthreads();
  int64_t offT = carryT + it - at, offQ = carryQ + iq - aq;
  segT = 0;
  segQ = 0;
#pragma unroll
  for (int q = 0; q < kBlock / kWave; ++q) {
    if (q < wave) {
      offT += P.wt[q][0];
      offQ += P.wt[q][1];
    }
    segT += P.wt[q][0];
    segQ += P.wt[q][1];
  }
  if (WRITE) {
    int32_t vt[kPrepPer], vq[kPrepPer], vg[kPrepPer], vh[kPrepPer];
    int64_t T8[kPrepPer], Q8[kPrepPer];
#pragma unroll
    for (int e = 0; e < kPrepPer; ++e) {
      const int p = p0 + e;
      const int64_t T = offT + lt[e], Q = offQ + lq[e];
      T8[e] = T;
      Q8[e] = Q;
      vt[e] = (int32_t)(uint32_t)(uint64_t)T;
      vq[e] = (int32_t)(uint32_t)(uint64_t)Q;
      const uint32_t Z = 2u * (uint32_t)(uint64_t)Q - (uint32_t)(2 * p + 1) * (uint32_t)(uint64_t)T;
      vg[e] = (int32_t)(0u - (uint32_t)(N - p) * (uint32_t)w[e] - Z);
      vh[e] = (int32_t)(Z - (uint32_t)(N - p - 1) * (uint32_t)w[e + 1]);
      if (p < N) lacc += T;
    }
    if (p0 + kPrepPer <= N) {
      // the thread's eight consecutive positions as 16-byte stores (p0 is a
      // multiple of 8 and every array 256-byte aligned): a quarter of the
      // store instructions of one word per position
      auto put4 = [](int32_t* d, int a, int b, int c, int e) { *(int4*)d = make_int4(a, b, c, e); };
      auto put2 = [](int64_t* d, int64_t a, int64_t b) {
        *(longlong2*)d = make_longlong2((long long)a, (long long)b);
      };
#pragma unroll
      for (int h = 0; h < kPrepPer; h += 4) {
        put4(s.xy + p0 + h, xy[h + 1], xy[h + 2], xy[h + 3], xy[h + 4]);
        put4(s.w + p0 + h, w[h], w[h + 1], w[h + 2], w[h + 3]);
        put4(s.t32 + p0 + h, vt[h], vt[h + 1], vt[h + 2], vt[h + 3]);
        put4(s.q32 + p0 + h, vq[h], vq[h + 1], vq[h + 2], vq[h + 3]);
        put4(s.g32 + p0 + h, vg[h], vg[h + 1], vg[h + 2], vg[h + 3]);
        put4(s.h32 + p0 + h, vh[h], vh[h + 1], vh[h + 2], vh[h + 3]);
      }
#pragma unroll
      for (int h = 0; h < kPrepPer; h += 2) {
        put2(s.t64 + p0 + h, T8[h], T8[h + 1]);
        put2(s.q64 + p0 + h, Q8[h], Q8[h + 1]);
      }
    } else {
#pragma unroll
      for (int e = 0; e < kPrepPer; ++e) {
        const int p = p0 + e;
        if (p < N) {
          s.xy[p] = xy[e + 1];
          s.w[p] = w[e];
          s.t64[p] = T8[e];
          s.q64[p] = Q8[e];
          s.t32[p] = vt[e];
          s.q32[p] = vq[e];
          s.g32[p] = vg[e];
          s.h32[p] = vh[e];
        }
      }
    }
  }
  __syncthreads();  // every thread has read wt
}

// The tour's T[N-1] and int32-path flags.
__device__ __forceinline__ void prep_flags(const SlotDev& s, int64_t Ttot) {
  // Moves computed mod 2^32 are exact while |delta| < 2^31.  Node-wise:
  // a 2-opt reversal moves each reversed node's arrival by at most
  // Ttot + 2 maxd, so |delta| <= N (Ttot + 6 maxd) + 3 Ttot; a swap changes
  // 4 edges (|delta| <= 4 N maxd); an or-opt-k moves k nodes by at most
  // Ttot each and every other node by at most (k + 4) maxd, so
  // |delta| <= 3 Ttot + 7 N maxd for k <= 3.
  const int64_t N = s.N, md = s.maxd;
  const int64_t b2 = N * (Ttot + 6 * md) + 3 * Ttot;
  const int64_t bso = 3 * Ttot + 8 * N * md;
  s.info[MLP_INFO_TTOT] = Ttot;
  s.info[MLP_INFO_PATH32] = (b2 < 2147483647LL ? MLP_P32_2OPT : 0) |
                            (bso < 2147483647LL ? MLP_P32_SWAP_OROPT : 0) |
                            (bso < (1LL << 26) ? MLP_PK_SWAP_OROPT : 0) |
                            (Ttot < (1LL << 24) ? MLP_T24 : 0);
}

// The whole tour's prefix arrays and info words by one workgroup (inlined,
// so that the tour and coordinate arrays keep their address space: LDS in a
// merge, global memory in k_prep).
__device__ __forceinline__ void prep_wg(const SlotDev& s, const int32_t* tour, PrepLds& P,
                                        const int32_t* xyc = nullptr) {
  const int N = s.N, tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
  int64_t lacc = 0;  // this thread's share of L = sum_p T[p]
  if (tid == 0) {
    P.carry[0] = 0;
    P.carry[1] = 0;
  }
  __syncthreads();
  for (int base = 0; base < N; base += kBlock * kPrepPer) {
    const int64_t cT = P.carry[0], cQ = P.carry[1];
    int64_t segT, segQ;
    prep_segment<true>(s, tour, xyc, base, cT, cQ, P, segT, segQ, lacc);
    if (tid == 0) {
      P.carry[0] = cT + segT;
      P.carry[1] = cQ + segQ;
    }
    __syncthreads();
  }
  const int64_t lw = wave_sum64(lacc);
  if (lane == 0) P.lsum[wave] = lw;
  __syncthreads();
  if (tid == 0) {
    int64_t L = 0;
    for (int q = 0; q < kBlock / kWave; ++q) L += P.lsum[q];
    s.info[MLP_INFO_LATENCY] = L;
    prep_flags(s, P.carry[0]);  // T[N-1]: positions past N-1 add nothing
  }
  __syncthreads();
}

}  // namespace

// ===========================================================================
// kernels
// ===========================================================================

__global__ __launch_bounds__(kBlock) void k_prep(const SlotDev* __restrict__ table,
                                                 SlotListInline list, const int32_t* list_dev) {
  __shared__ PrepLds P;
  const int id = list_at(list, list_dev, blockIdx.x);
  const SlotDev& s = table[id];
  prep_wg(s, s.tour, P);
  if (threadIdx.x == 0) {
    s.info[MLP_INFO_ROUND_MOVES] = 0;
    s.info[MLP_INFO_TOTAL_MOVES] = 0;
    s.info[MLP_INFO_ACTIVE] = 1;
    s.info[MLP_INFO_ROUNDS] = 0;
  }
}

// Start of a DVND run: every listed slot becomes active with no rounds.
__global__ __launch_bounds__(kBlock) void k_dvnd_start(const SlotDev* __restrict__ table,
                                                       SlotListInline list,
                                                       const int32_t* list_dev, int count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) {
    const SlotDev& s = table[list_at(list, list_dev, t)];
    s.info[MLP_INFO_ACTIVE] = 1;
    s.info[MLP_INFO_ROUNDS] = 0;
    s.info[MLP_INFO_ROUND_MOVES] = 0;
  }
}

namespace {

// End of a DVND round for one active slot (run by the round's last merge):
// count the round; a round that applied no move (or reached max_rounds > 0)
// stops the slot; the next round counts its moves from zero.
__device__ __forceinline__ void round_tail(const SlotDev& s, int max_rounds) {
  const int64_t r = s.info[MLP_INFO_ROUNDS] + 1;
  s.info[MLP_INFO_ROUNDS] = r;
  if (s.info[MLP_INFO_ROUND_MOVES] == 0 || (max_rounds > 0 && r >= max_rounds))
    s.info[MLP_INFO_ACTIVE] = 0;
  s.info[MLP_INFO_ROUND_MOVES] = 0;
}

}  // namespace

// Start of a DVND round: active slots count their moves from zero; the
// persistent-grid work counters of the round's launches are cleared (a
// kernel, not a memset, so the round can be captured in a graph).
__global__ __launch_bounds__(kBlock) void k_round_begin(const SlotDev* __restrict__ table,
                                                        SlotListInline list,
                                                        const int32_t* list_dev, int count,
                                                        int* counters, int ncounters) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) {
    const SlotDev& s = table[list_at(list, list_dev, t)];
    if (s.info[MLP_INFO_ACTIVE]) s.info[MLP_INFO_ROUND_MOVES] = 0;
  }
  if (t < ncounters) counters[t] = 0;
}

// End of a DVND round: an active slot that applied no move (or reached
// max_rounds > 0) stops.
__global__ __launch_bounds__(kBlock) void k_round_end(const SlotDev* __restrict__ table,
                                                      SlotListInline list,
                                                      const int32_t* list_dev, int count,
                                                      int max_rounds) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) {
    const SlotDev& s = table[list_at(list, list_dev, t)];
    if (s.info[MLP_INFO_ACTIVE]) {
      const int64_t r = s.info[MLP_INFO_ROUNDS] + 1;
      s.info[MLP_INFO_ROUNDS] = r;
      if (s.info[MLP_INFO_ROUND_MOVES] == 0 || (max_rounds > 0 && r >= max_rounds))
        s.info[MLP_INFO_ACTIVE] = 0;
    }
  }
}

// One workgroup per work item (persistent == nullptr) or a persistent grid
// draining the item counter.
__global__ __launch_bounds__(kBlock) void k_eval(const SlotDev* __restrict__ table,
                                                 EvalSegsInline inl, const EvalSeg* segs_dev,
                                                 int nsegs, int total, int rowmap,
                                                 int* counter, int skip_inactive) {
  if (!counter) {
    const int item = blockIdx.x;
    const EvalSeg sg = segs_dev ? find_seg(segs_dev, nsegs, item) : find_seg_inline(inl, nsegs, item);
    eval_item(table, sg, item, rowmap, skip_inactive != 0);
    return;
  }
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= total) break;
    const EvalSeg sg = segs_dev ? find_seg(segs_dev, nsegs, item) : find_seg_inline(inl, nsegs, item);
    eval_item(table, sg, item, rowmap, skip_inactive != 0);
  }
}

namespace {

#define MLP_CLAIM_OR(p, m) atomicOr((p), (m))
#include "mlp_claims.h"

// Positions a kept move rewrites, and their new sources (mlkswap.cu:1066-1106,
// mlk2opt.cu:725-768, mlkoropt.cu:886-950 restated as a position map).
constexpr int kShortMove = 64;
constexpr int kLongMoves = 512;
static_assert(MLP_MAX_N / (kShortMove + 1) < kLongMoves, "disjoint long spans fit the list");

__device__ __forceinline__ int move_span(int nb, int k, int i, int j) {
  if (nb == 0) return 2;
  if (nb == 1) return j - i + 1;
  return i < j ? j + k - i : i + k - j;
}

__device__ __forceinline__ void apply_span(int nb, int k, int i, int j, const int32_t* told,
                                           int32_t* tnew, int start, int stride) {
  if (nb == 0) {
    if (start == 0) {
      tnew[i] = told[j];
      tnew[j] = told[i];
    }
  } else if (nb == 1) {
    for (int p = i + start; p <= j; p += stride) tnew[p] = told[i + j - p];
  } else if (i < j) {
    for (int p = i + start; p < j + k; p += stride) tnew[p] = p < j ? told[p + k] : told[i + p - j];
  } else {
    for (int p = j + start; p < i + k; p += stride) tnew[p] = p < j + k ? told[i + p - j] : told[p - k];
  }
}

// Same-id conflict predicates of mlkernel.cu:51-71 without short-circuit
// branches.
template <int NB>
__device__ __forceinline__ bool conflicts_nb(int i1, int j1, int i2, int j2) {
  constexpr int K = NB >= 2 ? NB - 1 : 0;
  if (NB == 0)
    return ((int)((unsigned)(i1 - i2 + 1) <= 2u) | (int)((unsigned)(i1 - j2 + 1) <= 2u) |
            (int)((unsigned)(j1 - i2 + 1) <= 2u) | (int)((unsigned)(j1 - j2 + 1) <= 2u)) != 0;
  if (NB == 1)
    return ((int)(j1 < i2 - 1) | (int)(i1 > j2 + 1) | (int)(j2 < i1 - 1) | (int)(i2 > j1 + 1)) == 0;
  return ((int)(max(i1, j1) + K < min(i2, j2)) | (int)(min(i1, j1) > max(i2, j2) + K)) == 0;
}

// The conflict predicates by kind (0 swap, 1 2-opt, 2 or-opt-k with k
// given at run time).
template <int KIND>
__device__ __forceinline__ bool conflicts_kind(int k, int i1, int j1, int i2, int j2) {
  if constexpr (KIND < 2) return conflicts_nb<KIND>(i1, j1, i2, j2);
  return ((int)(max(i1, j1) + k < min(i2, j2)) | (int)(min(i1, j1) > max(i2, j2) + k)) == 0;
}

// Greedy over the M candidates sorted by (cost, row), on one wave, 64 at a
// time.  The bitmap of positions claimed by kept moves is an exact conflict
// test against every move kept so far (swap: positions i-1..i+1, j-1..j+1;
// 2-opt: [i, j+1]; or-opt: [min(i,j), max(i,j)+k]); within a batch the
// candidates that pass it are decided in order by a scalar loop over the
// kept ones.  Kept moves go to keys[q]
// as (i << 32 | j) and to the slot's kept arrays.  Returns their count.
// Candidates from index `tail` on are read from gtail (a sort whose keys
// exceed the LDS buffer); kept moves are written below `tail`.  The next
// batch's keys and ctx and the key after it are read while a batch is
// decided (kept moves overwrite only keys of decided batches).
template <int KIND>
__device__ __forceinline__ int greedy_wave(const SlotDev& s, int nb, int N, int rowmap, int M,
                                           int64_t* keys, const uint16_t* ctxL, uint32_t* bits,
                                           int lane, const int64_t* gtail, int tail) {
  const int K = KIND == 2 ? nb - 1 : 0;
  auto key_at = [&](int t) -> int64_t {
    return t < M ? (t < tail ? keys[t] : ldg(gtail + t)) : MLP_KEY_NONE;
  };
  auto ctx_at = [&](int t, int64_t key) -> int {
    return t < M ? (int)ctxL[(int)(key & MLP_KEY_MASK)] : 0;
  };
  // slot fields in registers: the loop's global stores would otherwise make
  // the compiler reload them (a global load on every batch)
  int64_t* const kcost = s.kept_cost[nb];
  int32_t* const kij = s.kept_ij[nb];
  int64_t skc = key_at(lane);
  int ctxc = ctx_at(lane, skc);
  int64_t skn = key_at(kWave + lane);
  int kept_total = 0;
  for (int base = 0; base < M; base += kWave) {
    const int t = base + lane;
    const bool valid = t < M;
    const int ctxn = ctx_at(t + kWave, skn);
    const int64_t sknn = key_at(t + 2 * kWave);
    int i = 0, j = 0;
    const int row = (int)(skc & MLP_KEY_MASK);
    const int64_t cost = skc >> MLP_KEY_SHIFT;
    if (valid) move_ij(nb, rowmap, N, row, ctxc, i, j);
    const int a = KIND == 1 ? i : min(i, j);
    const int b = KIND == 1 ? j + 1 : max(i, j) + K;
    bool pre = valid;
    if (valid) {
      if (KIND == 0) pre = !((int)claimed(bits, N, i) | (int)claimed(bits, N, j));
      else pre = !claims_any(bits, N, a, b);
    }
    // the first remaining candidate is kept and every candidate it
    // conflicts with leaves the batch (the predicates are symmetric), so the
    // loop runs once per kept move
    uint64_t m = __ballot(pre), kept = 0;
    while (m) {
      const int q = __builtin_ctzll(m);
      kept |= 1ull << q;
      const int iq = __builtin_amdgcn_readlane(i, q), jq = __builtin_amdgcn_readlane(j, q);
      const bool c = conflicts_kind<KIND>(K, iq, jq, i, j);
      m &= ~(__ballot(c) | (1ull << q));
    }
    if ((kept >> lane) & 1ull) {
      if (KIND == 0) {
        claims_set(bits, N, i - 1, i + 1);
        claims_set(bits, N, j - 1, j + 1);
      } else {
        claims_set(bits, N, a, b);
      }
      const int q = kept_total + __popcll(kept & ((1ull << lane) - 1ull));
      keys[q] = ((int64_t)i << 32) | (uint32_t)j;  // q <= t: already consumed
      kcost[q] = cost;
      kij[3 * q + 0] = i;
      kij[3 * q + 1] = j;
      kij[3 * q + 2] = row;
    }
    kept_total += __popcll(kept);
    skc = skn;
    ctxc = ctxn;
    skn = sknn;
  }
  return kept_total;
}


// Applies K kept moves (disjoint regions) to s.tour through a position ->
// source map: a move with a short region is applied by one thread, the few
// long ones (regions are disjoint, so at most N / kShortMove) by all threads.
// ij(q) gives the q-th kept move as (i << 32 | j); tnew ends as the new tour.
// loaded: told and tnew already hold the tour (copied at the kernel's start,
// its load overlapping the row-key loads).
template <typename IJ>
__device__ __forceinline__ void apply_kept(const SlotDev& s, int nb, int K, IJ ij, int32_t* told,
                                           int32_t* tnew, int& s_long, int* s_longq,
                                           bool loaded = false, bool copy_back = true) {
  const int N = s.N, tid = threadIdx.x, k = nb_k(nb);
  int32_t* const tour = s.tour;  // in a register across the loops' stores
  if (!loaded)
    for (int p = tid; p < N; p += kBlock) {
      const int c = tour[p];
      told[p] = c;
      tnew[p] = c;
    }
  if (tid == 0) s_long = 0;
  __syncthreads();
  for (int q = tid; q < K; q += kBlock) {
    const int64_t v = ij(q);
    const int i = (int)(v >> 32), j = (int)(uint32_t)v;
    if (move_span(nb, k, i, j) <= kShortMove) {
      apply_span(nb, k, i, j, told, tnew, 0, 1);
    } else {
      const int at = atomicAdd(&s_long, 1);
      if (at < kLongMoves) s_longq[at] = q;
    }
  }
  __syncthreads();
  const int nl = min(s_long, kLongMoves);
  for (int t = 0; t < nl; ++t) {
    const int64_t v = ij(s_longq[t]);
    apply_span(nb, k, (int)(v >> 32), (int)(uint32_t)v, told, tnew, tid, kBlock);
  }
  __syncthreads();
  if (copy_back)
    for (int p = tid; p < N; p += kBlock) tour[p] = tnew[p];
}

}  // namespace

// Bitonic sorts of 64-bit keys by the workgroup.  Stage kk (a power of two)
// merges blocks of kk keys, element i ascending when (i & kk) == 0 (the
// whole array flipped for a descending sort); its phase jj compares i with
// i + jj for (i & jj) == 0.  The array is cut into segments of 64 E keys
// (E = 2, 4 or 8: every wave of the workgroup holds at least one), held in
// the registers of one wave, key base + 64 e + lane in v[e] of that lane:
// phases with jj of 64 .. 32 E pair keys inside a lane, jj of 1..32 pair
// lanes (lane ^ jj); phases with jj >= 64 E go pair by pair through memory
// between barriers.  The stages up to 64 E take one segment pass, and each
// later stage one pass after its memory phases.
constexpr int kSortSeg = 8 * kWave;  // the largest segment

__device__ __forceinline__ int64_t xor_lane64(int64_t v, int m) {
  const int lo = __shfl_xor((int)(uint32_t)(uint64_t)v, m, kWave);
  const int hi = __shfl_xor((int)(uint32_t)((uint64_t)v >> 32), m, kWave);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// The direction bit (i & kk) of key base + 64 e + lane: the lane's bit when
// kk < 64 (the same for every e), otherwise a bit of base + 64 e (uniform
// over the wave).
template <int E, int JJ>
__device__ __forceinline__ void seg_phase(int64_t (&v)[E], int base, int lane, int kk, bool desc) {
  if constexpr (JJ >= kWave) {
    constexpr int D = JJ / kWave;  // kk > JJ >= 64
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if (e & D) continue;
      const bool asc = (((base + kWave * e) & kk) == 0) != desc;
      const int64_t a = v[e], b = v[e + D];
      // keys are distinct apart from KEY_NONE padding: (a > b) == asc is
      // the swap test of either direction (branch-free selects)
      const bool sw = (a > b) == asc;
      v[e] = sw ? b : a;
      v[e + D] = sw ? a : b;
    }
  } else {
    const bool lower = (lane & JJ) == 0;
    int64_t q[E];
#pragma unroll
    for (int e = 0; e < E; ++e) q[e] = xor_lane64(v[e], JJ);
    // the lower key of a pair keeps the minimum in an ascending block
    if (kk < kWave) {
      const bool tmin = lower == (((lane & kk) == 0) != desc);
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = ((q[e] < v[e]) == tmin) ? q[e] : v[e];
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const bool asc = (((base + kWave * e) & kk) == 0) != desc;
        v[e] = ((q[e] < v[e]) == (lower == asc)) ? q[e] : v[e];
      }
    }
  }
}

// phases min(jj_hi, 32 E) .. 1 of stage kk on one segment (jj_hi uniform)
template <int E>
__device__ __forceinline__ void seg_phases(int64_t (&v)[E], int base, int lane, int kk, int jj_hi,
                                           bool desc) {
  if constexpr (E >= 8) {
    if (jj_hi >= 256) seg_phase<E, 256>(v, base, lane, kk, desc);
  }
  if constexpr (E >= 4) {
    if (jj_hi >= 128) seg_phase<E, 128>(v, base, lane, kk, desc);
  }
  if (jj_hi >= 64) seg_phase<E, 64>(v, base, lane, kk, desc);
  if (jj_hi >= 32) seg_phase<E, 32>(v, base, lane, kk, desc);
  if (jj_hi >= 16) seg_phase<E, 16>(v, base, lane, kk, desc);
  if (jj_hi >= 8) seg_phase<E, 8>(v, base, lane, kk, desc);
  if (jj_hi >= 4) seg_phase<E, 4>(v, base, lane, kk, desc);
  if (jj_hi >= 2) seg_phase<E, 2>(v, base, lane, kk, desc);
  seg_phase<E, 1>(v, base, lane, kk, desc);
}

// Stages kk0 .. kk1 (powers of two) on every segment of k[0..P), phases
// below 64 E only (those of a larger stage follow its memory phases).
template <int E>
__device__ __forceinline__ void seg_pass(int64_t* k, int P, int kk0, int kk1, bool desc) {
  const int lane = threadIdx.x % kWave;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);  // uniform segment bases
  for (int base = wave * kWave * E; base < P; base += kBlock * E) {
    int64_t v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = k[base + kWave * e + lane];
    for (int kk = kk0; kk <= kk1; kk <<= 1) seg_phases<E>(v, base, lane, kk, kk >> 1, desc);
#pragma unroll
    for (int e = 0; e < E; ++e) k[base + kWave * e + lane] = v[e];
  }
}

// Phase jj (>= 64 E) of stage kk through memory: P/2 compare-swaps, up to
// four pairs per thread loaded before any is compared.
__device__ __forceinline__ void pair_phase(int64_t* k, int P, int kk, int jj, bool desc) {
  const int half = P >> 1, lj = 31 - __clz(jj);
  for (int p0 = threadIdx.x; p0 < half; p0 += 4 * kBlock) {
    int lo[4];
    int64_t va[4], vb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int pr = min(p0 + u * kBlock, half - 1);
      lo[u] = ((pr >> lj) << (lj + 1)) | (pr & (jj - 1));
      va[u] = k[lo[u]];
      vb[u] = k[lo[u] + jj];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool sw = (va[u] > vb[u]) == (((lo[u] & kk) == 0) != desc);
      if (p0 + u * kBlock < half) {
        k[lo[u]] = sw ? vb[u] : va[u];
        k[lo[u] + jj] = sw ? va[u] : vb[u];
      }
    }
  }
}

// Sort of k[0..P), P a power of two >= 256 E, ascending or descending.
template <int E>
__device__ __forceinline__ void sort_block(int64_t* k, int P, bool desc) {
  constexpr int SEG = kWave * E;
  seg_pass<E>(k, P, 2, min(P, SEG), desc);
  __syncthreads();
  for (int kk = 2 * SEG; kk <= P; kk <<= 1) {
    for (int jj = kk >> 1; jj >= SEG; jj >>= 1) {
      pair_phase(k, P, kk, jj, desc);
      __syncthreads();
    }
    seg_pass<E>(k, P, kk, kk, desc);
    __syncthreads();
  }
}

// Ascending sort of a bitonic sequence k[0..P) (the last stage of a sort),
// P >= 2048.
__device__ __forceinline__ void merge_block(int64_t* k, int P) {
  for (int jj = P >> 1; jj >= kSortSeg; jj >>= 1) {
    pair_phase(k, P, 2 * P, jj, false);
    __syncthreads();
  }
  seg_pass<8>(k, P, 2 * P, 2 * P, false);
  __syncthreads();
}

// Ascending sort of keys[0..M) (keys[M..P) padded with KEY_NONE, P the
// power of two at or above max(M, 512)), in segments of 128, 256 or 512
// keys so that each of the four waves holds one.
__device__ __forceinline__ void bitonic_keys(int64_t* keys, int M) {
  int P = 2 * kBlock;
  while (P < M) P <<= 1;
  for (int t = M + threadIdx.x; t < P; t += kBlock) keys[t] = MLP_KEY_NONE;
  __syncthreads();
  if (P <= 2 * kBlock) sort_block<2>(keys, P, false);
  else if (P <= 4 * kBlock) sort_block<4>(keys, P, false);
  else sort_block<8>(keys, P, false);
}

template <typename T>
__device__ __forceinline__ void stg(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}

// n keys from global memory into LDS, eight loads in flight per thread;
// those from index `valid` on read as KEY_NONE.
__device__ __forceinline__ void keys_to_lds(int64_t* sb, const int64_t* g, int n, int valid) {
  for (int t0 = threadIdx.x; t0 < n; t0 += 8 * kBlock) {
    int64_t v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = t0 + u * kBlock;
      v[u] = (t < n && t < valid) ? ldg(g + t) : MLP_KEY_NONE;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (t0 + u * kBlock < n) sb[t0 + u * kBlock] = v[u];
  }
}

__device__ __forceinline__ void keys_to_global(int64_t* g, const int64_t* sb, int n) {
  for (int t = threadIdx.x; t < n; t += kBlock) stg(g + t, sb[t]);
}

// A merge whose working arrays are in global memory still sorts up to
// kGmemSortCap improving rows in LDS; the rest of its dynamic LDS holds the
// conflict bitmap.
constexpr int kGmemSortCap = 16384;

// The merge's small shared variables (counts, the long-move queue, the
// prefix rebuild's scan carries).
struct MergeShared {
  int m, kept, lng;
  int longq[kLongMoves];
  PrepLds PL;
};

// 2.-3. of a merge: sort of the M improving (cost, row) keys == stable sort
// by cost of the row list, then the greedy filter on wave 0.  Returns the
// kept count; sk points at the sorted keys, whose first entries then hold the
// kept moves as (i << 32 | j).
template <bool GMEM>
__device__ __forceinline__ int sort_greedy(const SlotDev& s, int nb, int rowmap, int N, int M,
                                           int64_t* keys, int64_t* lds, const uint16_t* ctxL,
                                           uint32_t* bits, MergeShared& S, int64_t*& sk) {
  const int tid = threadIdx.x;
  const int64_t* gtail = nullptr;  // sorted keys beyond the LDS buffer
  int tail = INT32_MAX;
  // Up to one key per lane of a wave, keys in LDS: rank sort on wave 0, which
  // also runs the greedy next (the keys are distinct; a key's rank is the
  // number of smaller keys; every lane reads all keys before any lane
  // writes).  Otherwise a bitonic sort over the workgroup.
  if (!GMEM && M <= kWave) {
    if (tid < kWave) {
      const int64_t my = tid < M ? keys[tid] : 0;
      int rank = 0;
      for (int u = 0; u < M; ++u) rank += keys[u] < my ? 1 : 0;
      __builtin_amdgcn_wave_barrier();
      if (tid < M) keys[rank] = my;
      __builtin_amdgcn_wave_barrier();
    }
  } else if (!GMEM && M <= kBlock) {
    // one key per thread: the same rank sort over the workgroup (M reads of
    // LDS broadcasts per thread against the 36 barrier phases of a bitonic
    // sort of 256)
    const int64_t my = tid < M ? keys[tid] : 0;
    int rank = 0;
    if (tid < M)
      for (int u = 0; u < M; ++u) rank += keys[u] < my ? 1 : 0;
    __syncthreads();
    if (tid < M) keys[rank] = my;
    __syncthreads();
  } else if (GMEM && M <= kGmemSortCap) {
    // the slot's working arrays are in global memory, but its improving rows
    // fit the LDS sort buffer: sort them there, and let the greedy and the
    // apply read them from LDS
    int64_t* sb = lds;
    keys_to_lds(sb, keys, M, M);
    bitonic_keys(sb, M);
    sk = sb;
  } else if (GMEM && M <= 2 * kGmemSortCap) {
    // twice the LDS buffer (keys[M..2 cap) in the slot's memory are padded):
    // the lower half sorted ascending and the upper half descending in LDS,
    // the last stage's first phase in global memory, then each half's
    // remaining phases in LDS, the upper half first: the lower one stays in
    // LDS for the greedy (the kept moves, at most N/2, fit below it)
    int64_t* sb = lds;
    constexpr int H = kGmemSortCap;
    for (int h = 0; h < 2; ++h) {
      keys_to_lds(sb, keys + h * H, H, M - h * H);
      __syncthreads();
      sort_block<8>(sb, H, h == 1);
      keys_to_global(keys + h * H, sb, H);
      __syncthreads();
    }
    for (int t0 = tid; t0 < H; t0 += 4 * kBlock) {
      int64_t a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = ldg(keys + t0 + u * kBlock);
        b[u] = ldg(keys + t0 + u * kBlock + H);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        stg(keys + t0 + u * kBlock, min(a[u], b[u]));
        stg(keys + t0 + u * kBlock + H, max(a[u], b[u]));
      }
    }
    __syncthreads();
    for (int h = 1; h >= 0; --h) {
      keys_to_lds(sb, keys + h * H, H, H);
      __syncthreads();
      merge_block(sb, H);
      if (h == 1) {
        keys_to_global(keys + H, sb, H);
        __syncthreads();
      }
    }
    sk = sb;
    gtail = keys;
    tail = H;
  } else {
    bitonic_keys(keys, M);
  }
  // 3. greedy filter on wave 0
  if (tid < kWave) {
    int kept = 0;
    if (nb == 0) kept = greedy_wave<0>(s, nb, N, rowmap, M, sk, ctxL, bits, tid, gtail, tail);
    else if (nb == 1) kept = greedy_wave<1>(s, nb, N, rowmap, M, sk, ctxL, bits, tid, gtail, tail);
    else if (nb < MLP_NB_MAX) kept = greedy_wave<2>(s, nb, N, rowmap, M, sk, ctxL, bits, tid, gtail, tail);
    if (tid == 0) S.kept = kept;
  }
  __syncthreads();
  return S.kept;
}

// Rows [q0, q1) of a merge: the tour and its coordinates by client copied
// (told, tnew, xyc), and each improving row's (cost, row) key handed to
// put(key) with its ctx recorded.  Each thread issues the loads of eight
// rows (strided by kBlock) before it uses the first.
template <typename Put>
__device__ __forceinline__ void merge_rows(const SlotDev& s, const int64_t* rk, int R, int q0,
                                           int q1, int32_t* told, int32_t* tnew, int32_t* xyc,
                                           uint16_t* ctxL, Put put) {
  constexpr int P = 8;
  const int N = s.N, tid = threadIdx.x;
  const int32_t* const tour = s.tour;  // in registers across the loop's stores
  const int32_t* const xy = s.xy;
  const int rend = min(R, q1), nend = min(N, q1);
  for (int base = q0; base < q1; base += kBlock * P) {
    int64_t key[P];
    int c[P], xq[P];
#pragma unroll
    for (int e = 0; e < P; ++e) {
      const int q = base + e * kBlock + tid;
      key[e] = q < rend ? rk[q] : MLP_KEY_NONE;
      c[e] = q < nend ? tour[q] : 0;
      xq[e] = q < nend ? xy[q] : 0;
    }
#pragma unroll
    for (int e = 0; e < P; ++e) {
      const int q = base + e * kBlock + tid;
      if (q < nend) {
        told[q] = c[e];
        tnew[q] = c[e];
        xyc[c[e]] = xq[e];
      }
      if (key[e] != MLP_KEY_NONE && (key[e] >> MLP_KEY_SHIFT) < 0) {
        put((key[e] >> MLP_KEY_SHIFT) * (int64_t)(1 << MLP_KEY_SHIFT) + q);
        ctxL[q] = (uint16_t)(key[e] & MLP_KEY_MASK);
      }
    }
  }
}

// Merge + apply of one slot by one workgroup; lds holds the working arrays
// (mlp_merge_lds_bytes(pcap_l, ncap_l)).
__device__ __forceinline__ void merge_slot(const SlotDev& s, int nb, int rowmap, int pcap_l,
                                           int ncap_l, int skip_inactive, int tail, int64_t* lds,
                                           MergeShared& S) {
  if (skip_inactive && s.info[MLP_INFO_ACTIVE] == 0) return;
  const int N = s.N, tid = threadIdx.x;
  const int R = rows_of(nb, rowmap, N);
  const int pcap = pcap_l, ncap = ncap_l;  // capacities of the launch
  int64_t* keys = lds;
  int32_t* told = (int32_t*)(keys + pcap);
  int32_t* tnew = told + ncap;
  int32_t* xyc = tnew + ncap;  // packed coordinates by client
  const int nwords = claim_words(N);
  uint32_t* bits = (uint32_t*)(xyc + ncap);
  uint16_t* ctxL = (uint16_t*)((uint32_t*)(xyc + ncap) + claim_words(ncap));  // ctx of each improving row
  int64_t* sk = keys;  // the sorted keys (greedy, apply)

  if (tid == 0) S.m = 0;
  for (int w = tid; w < nwords; w += kBlock) bits[w] = 0;
  __syncthreads();
  // 1. improving rows -> (cost, row) keys.  Inside a DVND round the eval
  // wrote the buffer opposite info[PARITY0 + nb]; this merge commits it.
  // The tour (for the apply step) and its coordinates (for the prefix
  // rebuild, by client) are copied in the same loop, so their loads go out
  // with the row keys'.
  const int par = (int)(s.info[MLP_INFO_PARITY0 + nb] & 1) ^ (skip_inactive ? 1 : 0);
  const int64_t* rk = s.rows[nb] + (par ? s.rowcap : 0);
  if (skip_inactive) {
    __syncthreads();  // every thread has read the old parity
    if (tid == 0) s.info[MLP_INFO_PARITY0 + nb] = par;
  }
  merge_rows(s, rk, R, 0, max(R, N), told, tnew, xyc, ctxL,
             [&](int64_t k) { keys[atomicAdd(&S.m, 1)] = k; });
  __syncthreads();
  const int M = S.m;
  if (M == 0) {
    if (tid == 0) {
      s.info[MLP_INFO_KEPT0 + nb] = 0;
      if (tail) round_tail(s, tail - 1);
    }
    return;
  }
  const int K = sort_greedy<false>(s, nb, rowmap, N, M, keys, lds, ctxL, bits, S, sk);
  // 4. apply the kept moves, rebuild the prefix arrays
  apply_kept(s, nb, K, [&](int q) { return sk[q]; }, told, tnew, S.lng, S.longq, true);
  prep_wg(s, tnew, S.PL, xyc);
  if (tid == 0) {
    s.info[MLP_INFO_KEPT0 + nb] = K;
    s.info[MLP_INFO_ROUND_MOVES] += K;
    s.info[MLP_INFO_TOTAL_MOVES] += K;
    if (tail) round_tail(s, tail - 1);
  }
}

// One workgroup per slot.  Dynamic LDS: keys[P] int64 | told[N] | tnew[N] | bits | ctx[P].
// tail != 0: this is the last merge of a DVND round, and it also ends the
// round for each active slot (round_tail, max_rounds = tail - 1).
__global__ __launch_bounds__(kBlock) void k_merge_apply(const SlotDev* __restrict__ table,
                                                        SlotListInline list,
                                                        const int32_t* list_dev, int nb,
                                                        int rowmap, int pcap_l, int ncap_l,
                                                        int skip_inactive, int tail) {
  extern __shared__ int64_t lds[];
  __shared__ MergeShared S;
  const int id = list_at(list, list_dev, blockIdx.x);
  merge_slot(table[id], nb, rowmap, pcap_l, ncap_l, skip_inactive, tail, lds, S);
}

// ---------------------------------------------------------------------------
// Merge of tours whose working arrays exceed LDS, over many workgroups per
// slot (grid.y = slot list, grid.x = MLP_MG_CHUNK-position segments):
//   k_mg_compact  the improving rows as (cost, row) keys (gathered per
//                 workgroup in LDS, then appended with one atomic), the tour
//                 copied to told / tnew, the coordinates by client;
//   k_mg_select   one workgroup per slot: sort, greedy, the kept moves'
//                 spans written into tnew, the merge's and round's counters;
//   k_mg_sums     per segment of the new tour: sums of w and p w;
//   k_mg_rebuild  per segment: the prefix arrays with the preceding
//                 segments' sums as carries, the tour, the latency (one
//                 atomic add per segment), T[N-1] and the int32 flags.
// The one-workgroup merge runs these steps one after another over the whole
// tour; here the O(N) steps are spread over the GPU.
static_assert(MLP_MG_CHUNK == kBlock * kPrepPer, "one segment per workgroup");

struct MgArrays {
  int64_t* keys;
  int32_t* told;
  int32_t* tnew;
  int32_t* xyc;
  uint16_t* ctxL;
  int64_t* csum;  // [segments][2]
};

// The slot's global working arrays: the one-workgroup merge's layout
// (keys[pcap] | told[N] | tnew[N] | xyc[N] | bitmap | ctx[pcap]), then the
// segment sums.
__device__ __forceinline__ MgArrays mg_arrays(const SlotDev& s) {
  int pcap = 1;
  while (pcap < s.rowcap) pcap <<= 1;
  const int N = s.N;
  MgArrays a;
  a.keys = (int64_t*)s.mscratch;
  a.told = (int32_t*)(a.keys + pcap);
  a.tnew = a.told + N;
  a.xyc = a.tnew + N;
  a.ctxL = (uint16_t*)((uint32_t*)(a.xyc + N) + claim_words(N));
  a.csum = (int64_t*)(((uintptr_t)(a.ctxL + pcap) + 7) & ~(uintptr_t)7);
  return a;
}

__global__ __launch_bounds__(kBlock) void k_mg_compact(const SlotDev* __restrict__ table,
                                                       SlotListInline list,
                                                       const int32_t* list_dev, int nb, int rowmap,
                                                       int skip_inactive) {
  __shared__ int64_t lk[MLP_MG_CHUNK];
  __shared__ int lm;
  __shared__ int64_t lbase;
  const SlotDev& s = table[list_at(list, list_dev, blockIdx.y)];
  if (skip_inactive && s.info[MLP_INFO_ACTIVE] == 0) return;
  const int N = s.N, tid = threadIdx.x, R = rows_of(nb, rowmap, N);
  const int q0 = blockIdx.x * MLP_MG_CHUNK, q1 = min(q0 + MLP_MG_CHUNK, max(R, N));
  if (q0 >= q1) return;
  const MgArrays a = mg_arrays(s);
  // the eval of a DVND round wrote the buffer opposite info[PARITY0 + nb]
  // (k_mg_select commits it once every segment has read it)
  const int par = (int)(s.info[MLP_INFO_PARITY0 + nb] & 1) ^ (skip_inactive ? 1 : 0);
  const int64_t* rk = s.rows[nb] + (par ? s.rowcap : 0);
  if (tid == 0) lm = 0;
  __syncthreads();
  merge_rows(s, rk, R, q0, q1, a.told, a.tnew, a.xyc, a.ctxL,
             [&](int64_t k) { lk[atomicAdd(&lm, 1)] = k; });
  __syncthreads();
  const int m = lm;
  if (m == 0) return;
  if (tid == 0)
    lbase = (int64_t)atomicAdd((unsigned long long*)&s.info[MLP_INFO_MCOUNT], (unsigned long long)m);
  __syncthreads();
  for (int t = tid; t < m; t += kBlock) a.keys[lbase + t] = lk[t];
}

__global__ __launch_bounds__(kBlock) void k_mg_select(const SlotDev* __restrict__ table,
                                                      SlotListInline list, const int32_t* list_dev,
                                                      int nb, int rowmap, int skip_inactive,
                                                      int tail) {
  extern __shared__ int64_t lds[];
  __shared__ MergeShared S;
  const SlotDev& s = table[list_at(list, list_dev, blockIdx.x)];
  const int tid = threadIdx.x;
  if (skip_inactive && s.info[MLP_INFO_ACTIVE] == 0) {
    if (tid == 0) s.info[MLP_INFO_MG_APPLY] = 0;
    return;
  }
  const int N = s.N;
  const MgArrays a = mg_arrays(s);
  const int M = (int)s.info[MLP_INFO_MCOUNT];
  const int par = (int)(s.info[MLP_INFO_PARITY0 + nb] & 1) ^ (skip_inactive ? 1 : 0);
  uint32_t* bits = (uint32_t*)(lds + kGmemSortCap);
  const int nwords = claim_words(N);
  for (int w = tid; w < nwords; w += kBlock) bits[w] = 0;
  __syncthreads();  // every thread has read the count and the parity
  if (tid == 0) {
    s.info[MLP_INFO_MCOUNT] = 0;
    if (skip_inactive) s.info[MLP_INFO_PARITY0 + nb] = par;
  }
  if (M == 0) {
    if (tid == 0) {
      s.info[MLP_INFO_KEPT0 + nb] = 0;
      s.info[MLP_INFO_MG_APPLY] = 0;
      if (tail) round_tail(s, tail - 1);
    }
    return;
  }
  int64_t* sk = a.keys;
  const int K = sort_greedy<true>(s, nb, rowmap, N, M, a.keys, lds, a.ctxL, bits, S, sk);
  // tnew becomes the new tour; k_mg_rebuild copies it back
  apply_kept(s, nb, K, [&](int q) { return sk[q]; }, a.told, a.tnew, S.lng, S.longq, true, false);
  if (tid == 0) {
    s.info[MLP_INFO_KEPT0 + nb] = K;
    s.info[MLP_INFO_ROUND_MOVES] += K;
    s.info[MLP_INFO_TOTAL_MOVES] += K;
    s.info[MLP_INFO_MG_APPLY] = K;
    if (tail) round_tail(s, tail - 1);
  }
}

__global__ __launch_bounds__(kBlock) void k_mg_sums(const SlotDev* __restrict__ table,
                                                    SlotListInline list, const int32_t* list_dev) {
  __shared__ PrepLds P;
  const SlotDev& s = table[list_at(list, list_dev, blockIdx.y)];
  if (s.info[MLP_INFO_MG_APPLY] == 0) return;
  const int base = blockIdx.x * MLP_MG_CHUNK;
  if (base >= s.N) return;
  const MgArrays a = mg_arrays(s);
  int64_t segT, segQ, lacc = 0;
  prep_segment<false>(s, a.tnew, a.xyc, base, 0, 0, P, segT, segQ, lacc);
  if (threadIdx.x == 0) {
    a.csum[2 * blockIdx.x] = segT;
    a.csum[2 * blockIdx.x + 1] = segQ;
    if (blockIdx.x == 0) s.info[MLP_INFO_LATENCY] = 0;  // k_mg_rebuild adds the segments
  }
}

__global__ __launch_bounds__(kBlock) void k_mg_rebuild(const SlotDev* __restrict__ table,
                                                       SlotListInline list,
                                                       const int32_t* list_dev) {
  __shared__ PrepLds P;
  const SlotDev& s = table[list_at(list, list_dev, blockIdx.y)];
  if (s.info[MLP_INFO_MG_APPLY] == 0) return;
  const int N = s.N, base = blockIdx.x * MLP_MG_CHUNK;
  if (base >= N) return;
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
  const MgArrays a = mg_arrays(s);
  const int nseg = (N + MLP_MG_CHUNK - 1) / MLP_MG_CHUNK;
  int64_t cT = 0, cQ = 0, Ttot = 0;
  for (int c = 0; c < nseg; ++c) {
    const int64_t t = a.csum[2 * c], q = a.csum[2 * c + 1];
    if (c < (int)blockIdx.x) {
      cT += t;
      cQ += q;
    }
    Ttot += t;
  }
  int64_t segT, segQ, lacc = 0;
  prep_segment<true>(s, a.tnew, a.xyc, base, cT, cQ, P, segT, segQ, lacc);
  for (int p = base + tid; p < min(base + MLP_MG_CHUNK, N); p += kBlock) s.tour[p] = a.tnew[p];
  const int64_t lw = wave_sum64(lacc);
  if (lane == 0) P.lsum[wave] = lw;
  __syncthreads();
  if (tid == 0) {
    int64_t L = 0;
    for (int q = 0; q < kBlock / kWave; ++q) L += P.lsum[q];
    atomicAdd((unsigned long long*)&s.info[MLP_INFO_LATENCY], (unsigned long long)L);
    if (blockIdx.x == 0) prep_flags(s, Ttot);
  }
}

// ---------------------------------------------------------------------------
// TOTAL mode (kernelSwapTotal mlkswap.cu:334-526, kernel2OptTotal
// mlk2opt.cu:270-447, kernelOrOptTotal mlkoropt.cu:368-612, merged by
// mergeGPU mlkernel.cu:254-384 with testSwapTotal etc. :183-250).  The
// reference launches B threads per ADS row; thread t scans the row-local
// indices ctx = t, t + B, ... and keeps the first strictly smallest cost, so
// a row yields B entries (row * B + t).  Here one workgroup takes a row, its
// threads stride over the row's indices and fold each move's key
// (delta << 20 | ctx) into entry ctx mod B with an LDS atomic minimum: the
// smallest key of an entry is the reference's first strict minimum.
// ---------------------------------------------------------------------------
namespace {

template <int KIND, bool P32>
__device__ __forceinline__ void total_row(const SlotDev& s, int nb, int row, int B, int64_t* best) {
  const int N = s.N;
  const float dr = round_dr(s.round);
  const int k = nb_k(nb);
  int lo, hi;
  ctx_range(nb, 0, N, row, lo, hi);
  int slot = (lo + (int)threadIdx.x) % B;
  const int step = kBlock % B;
  for (int ctx = lo + (int)threadIdx.x; ctx < hi; ctx += kBlock) {
    int i, j;
    move_ij(nb, 0, N, row, ctx, i, j);
    int64_t d;
    if (KIND == 0) d = delta_swap<P32>(s, i, j, dr);
    else if (KIND == 1) d = delta_2opt<P32>(s, i, j, dr);
    else d = delta_oropt<P32>(s, k, i, j, dr);
    atomicMin((long long*)&best[slot], (long long)(d * (int64_t)(1 << MLP_KEY_SHIFT) + ctx));
    slot += step;
    if (slot >= B) slot -= B;
  }
}

template <bool P32>
__device__ __forceinline__ void total_row_dispatch(const SlotDev& s, int nb, int row, int B,
                                                   int64_t* best) {
  if (nb == 0) total_row<0, P32>(s, nb, row, B, best);
  else if (nb == 1) total_row<1, P32>(s, nb, row, B, best);
  else total_row<2, P32>(s, nb, row, B, best);
}

__device__ __forceinline__ bool int32_path(const SlotDev& s, int nb) {
  if (nb <= 4) return (s.info[MLP_INFO_PATH32] & (nb == 1 ? MLP_P32_2OPT : MLP_P32_SWAP_OROPT)) != 0;
  const int64_t k = nb_k(nb);  // or-opt-k, k > 3: |delta| <= k Ttot + (k + 4) N maxd
  return k * s.info[MLP_INFO_TTOT] + (k + 5) * (int64_t)s.N * s.maxd < 2147483647LL;
}

}  // namespace

// grid = rows of nb (ADS row map); dynamic LDS: B int64.
__global__ __launch_bounds__(kBlock) void k_total_eval(const SlotDev* __restrict__ table, int id,
                                                       int nb, int B, int64_t* __restrict__ out) {
  extern __shared__ int64_t best[];
  const SlotDev& s = table[id];
  const int row = blockIdx.x;
  for (int t = threadIdx.x; t < B; t += kBlock) best[t] = MLP_KEY_NONE;
  __syncthreads();
  if (int32_path(s, nb)) total_row_dispatch<true>(s, nb, row, B, best);
  else total_row_dispatch<false>(s, nb, row, B, best);
  __syncthreads();
  int64_t* o = out + (int64_t)row * B;
  for (int t = threadIdx.x; t < B; t += kBlock) o[t] = best[t];
}

// Sort keys of the entries: an improving entry's delta + bias (bias >= every
// |delta|), bias for the rest.  Values: row << 15 | ctx (N <= 32768), from
// which the merge recovers the move and the entry index row * B + ctx mod B
// without reading the entry again.  Keys are written in entry order, so a
// stable sort orders the entries by (cost, entry) like the reference's.
__global__ void k_total_keys(const int64_t* __restrict__ tk, int64_t M, int B, uint64_t bias,
                             uint64_t* __restrict__ ku, int32_t* __restrict__ ve) {
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < M;
       e += (int64_t)gridDim.x * kBlock) {
    const int64_t c = tk[e];
    uint64_t u = bias;
    int32_t v = 0;
    if (c != MLP_KEY_NONE) {
      const int64_t d = c >> MLP_KEY_SHIFT;
      if (d < 0) {
        u = d + (int64_t)bias > 0 ? (uint64_t)(d + (int64_t)bias) : 0;
        v = (int32_t)((e / B) << 15) | (int32_t)(c & MLP_KEY_MASK);
      }
    }
    ku[e] = u;
    ve[e] = v;
  }
}

namespace {

constexpr int kTotalPer = 8;                       // candidates per thread per chunk
constexpr int kTotalChunk = kTotalPer * kBlock;    // candidates decoded per chunk

struct TotalLds {
  int32_t ci[kTotalChunk], cj[kTotalChunk], ce[kTotalChunk];
  int64_t cc[kTotalChunk];
  int wc[kTotalPer][kBlock / kWave];
  int end, kept, nlong;
  int longq[kLongMoves];
};

// Greedy over S decoded candidates in order (one wave), continuing a kept
// list of `base` moves: same batch logic as greedy_wave.
template <int NB>
__device__ int greedy_list(const SlotDev& s, int S, const TotalLds& T, uint32_t* bits, int lane,
                           int base_kept) {
  constexpr int K = NB >= 2 ? NB - 1 : 0;
  // slot fields in registers (the loop's global stores would reload them)
  const int n = s.N;
  int64_t* const kcost = s.kept_cost[NB];
  int32_t* const kij = s.kept_ij[NB];
  int kept_total = base_kept;
  for (int base = 0; base < S; base += kWave) {
    const int t = base + lane;
    const bool valid = t < S;
    const int i = valid ? T.ci[t] : 0, j = valid ? T.cj[t] : 0;
    const int a = NB == 1 ? i : min(i, j);
    const int b = NB == 1 ? j + 1 : max(i, j) + K;
    bool pre = valid;
    if (valid) {
      if (NB == 0) pre = !claimed(bits, n, i) && !claimed(bits, n, j);
      else pre = !claims_any(bits, n, a, b);
    }
    // the first remaining candidate is kept and every candidate it
    // conflicts with leaves the batch (the predicates are symmetric), so the
    // loop runs once per kept move
    uint64_t m = __ballot(pre), kept = 0;
    while (m) {
      const int q = __builtin_ctzll(m);
      kept |= 1ull << q;
      const int iq = __builtin_amdgcn_readlane(i, q), jq = __builtin_amdgcn_readlane(j, q);
      const bool c = conflicts_nb<NB>(iq, jq, i, j);
      m &= ~(__ballot(c) | (1ull << q));
    }
    if ((kept >> lane) & 1ull) {
      if (NB == 0) {
        claims_set(bits, n, i - 1, i + 1);
        claims_set(bits, n, j - 1, j + 1);
      } else {
        claims_set(bits, n, a, b);
      }
      const int q = kept_total + __popcll(kept & ((1ull << lane) - 1ull));
      kcost[q] = T.cc[t];
      kij[3 * q + 0] = i;
      kij[3 * q + 1] = j;
      kij[3 * q + 2] = T.ce[t];
    }
    kept_total += __popcll(kept);
  }
  return kept_total;
}

}  // namespace

// Greedy merge of the sorted entries and apply, one workgroup.  The sorted
// list is consumed in chunks: all threads decode a chunk and drop the
// candidates that already conflict with a kept move (the kept set only
// grows, so they would be dropped anyway), keeping the survivors in order;
// wave 0 then runs the greedy over the survivors while the next chunk's keys
// are already being loaded.  The first non-improving key ends the list.
// Dynamic LDS: bits | told[N] | tnew[N] (told/tnew in the slot's merge
// scratch when GMEM).
template <bool GMEM>
__global__ __launch_bounds__(kBlock) void k_total_merge(const SlotDev* __restrict__ table, int id,
                                                        int nb, int B, int64_t M, uint64_t bias,
                                                        const uint64_t* __restrict__ ku,
                                                        const int32_t* __restrict__ ve) {
  extern __shared__ int64_t lds[];
  __shared__ TotalLds T;
  __shared__ PrepLds PL;
  const SlotDev& s = table[id];
  const int N = s.N, tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
  const int k = nb_k(nb);
  const int nwords = claim_words(N);
  uint32_t* bits = (uint32_t*)lds;
  int32_t* told = GMEM ? (int32_t*)s.mscratch : (int32_t*)(bits + nwords);
  int32_t* tnew = told + N;
  for (int w = tid; w < nwords; w += kBlock) bits[w] = 0;
  if (tid == 0) {
    T.end = 0;
    T.kept = 0;
  }
  const uint64_t lt = (1ull << lane) - 1ull;
  // raw keys / values of the next chunk (bias = past the end or not improving)
  uint64_t nk[kTotalPer];
  int32_t nv[kTotalPer];
#pragma unroll
  for (int u = 0; u < kTotalPer; ++u) {
    const int64_t q = (int64_t)u * kBlock + tid;
    nk[u] = q < M ? ku[q] : bias;
    nv[u] = q < M ? ve[q] : 0;
  }
  __syncthreads();
  for (int64_t base = 0; base < M; base += kTotalChunk) {
    uint64_t ck[kTotalPer];
    int32_t cv[kTotalPer];
#pragma unroll
    for (int u = 0; u < kTotalPer; ++u) {
      ck[u] = nk[u];
      cv[u] = nv[u];
      const int64_t q = base + kTotalChunk + (int64_t)u * kBlock + tid;
      nk[u] = q < M ? ku[q] : bias;
      nv[u] = q < M ? ve[q] : 0;
    }
    int ci[kTotalPer], cj[kTotalPer];
    uint64_t bal[kTotalPer];
    bool end = false;
#pragma unroll
    for (int u = 0; u < kTotalPer; ++u) {
      bool pass = false;
      ci[u] = cj[u] = 0;
      if (ck[u] < bias) {
        const int row = cv[u] >> 15, ctx = cv[u] & 0x7fff;
        int i, j;
        move_ij(nb, 0, N, row, ctx, i, j);
        ci[u] = i;
        cj[u] = j;
        if (nb == 0) pass = !claimed(bits, N, i) && !claimed(bits, N, j);
        else if (nb == 1) pass = !claims_any(bits, N, i, j + 1);
        else pass = !claims_any(bits, N, min(i, j), max(i, j) + k);
      } else {
        end = true;  // past the list or the first non-improving key
      }
      bal[u] = __ballot(pass);
      if (lane == 0) T.wc[u][wave] = __popcll(bal[u]);
    }
    if (end) T.end = 1;
    __syncthreads();
    int S = 0;
#pragma unroll
    for (int u = 0; u < kTotalPer; ++u) {
      int off = S;
      for (int w = 0; w < kBlock / kWave; ++w) {
        if (w < wave) off += T.wc[u][w];
        S += T.wc[u][w];
      }
      if ((bal[u] >> lane) & 1ull) {
        const int at = off + __popcll(bal[u] & lt);
        const int row = cv[u] >> 15, ctx = cv[u] & 0x7fff;
        T.ci[at] = ci[u];
        T.cj[at] = cj[u];
        T.ce[at] = row * B + ctx % B;
        T.cc[at] = (int64_t)ck[u] - (int64_t)bias;
      }
    }
    __syncthreads();
    // read before the barrier below: the next chunk's decode may set it
    const bool stop = T.end != 0;
    if (wave == 0 && S > 0) {
      int kept = T.kept;
#define MLP_TGREEDY(NB) \
  case NB: kept = greedy_list<NB>(s, S, T, bits, lane, kept); break
      switch (nb) {
        MLP_TGREEDY(0); MLP_TGREEDY(1); MLP_TGREEDY(2); MLP_TGREEDY(3); MLP_TGREEDY(4);
        MLP_TGREEDY(5); MLP_TGREEDY(6); MLP_TGREEDY(7); MLP_TGREEDY(8); MLP_TGREEDY(9);
        MLP_TGREEDY(10); MLP_TGREEDY(11); MLP_TGREEDY(12); MLP_TGREEDY(13); MLP_TGREEDY(14);
        default: break;
      }
#undef MLP_TGREEDY
      if (lane == 0) T.kept = kept;
    }
    __syncthreads();
    if (stop) break;
  }
  const int K = T.kept;
  const int32_t* kij = s.kept_ij[nb];
  apply_kept(
      s, nb, K,
      [&](int q) { return ((int64_t)kij[3 * q] << 32) | (uint32_t)kij[3 * q + 1]; }, told, tnew,
      T.nlong, T.longq);
  prep_wg(s, tnew, PL);
  if (tid == 0) {
    s.info[MLP_INFO_KEPT0 + nb] = K;
    s.info[MLP_INFO_ROUND_MOVES] += K;
    s.info[MLP_INFO_TOTAL_MOVES] += K;
  }
}

// MLMove64 words (mlads.h:190-206) of per-row keys: low word id:4 | i:14 |
// j:14, high word the cost as the reference's int32 (low 32 bits of the
// delta); a row without a move is COST_INFTY with move word 0
// (mlkswap.cu:175-176).
__device__ __forceinline__ uint64_t move64_of(int64_t key, int nb, int rowmap, int N, int row) {
  if (key == MLP_KEY_NONE) return (uint64_t)0x7fffffffu << 32;
  int i, j;
  move_ij(nb, rowmap, N, row, (int)(key & MLP_KEY_MASK), i, j);
  const uint32_t lo =
      ((uint32_t)(j & 0x3fff) << 18) | ((uint32_t)(i & 0x3fff) << 4) | (uint32_t)(nb & 15);
  return ((uint64_t)(uint32_t)(int32_t)(key >> MLP_KEY_SHIFT) << 32) | lo;
}

// Rows of the last evaluation of nb (the buffer info[PARITY0 + nb] names).
__global__ void k_rows_moves64(const SlotDev* __restrict__ table, int id, int nb, int rowmap,
                               int R, uint64_t* out) {
  const SlotDev& s = table[id];
  const int par = (int)(s.info[MLP_INFO_PARITY0 + nb] & 1);
  const int64_t* keys = s.rows[nb] + (par ? s.rowcap : 0);
  for (int r = blockIdx.x * kBlock + threadIdx.x; r < R; r += gridDim.x * kBlock)
    out[r] = move64_of(keys[r], nb, rowmap, s.N, r);
}

// Rows of the last evaluation of every neighborhood in mask (ids 0..4, in
// id order, rows_of(nb) entries each) as cost (INT64_MAX without a move),
// i and j: the records mlp_get_rows returns, written to device arrays.
__global__ void k_rows_out(const SlotDev* __restrict__ table, int id, int mask, int rowmap,
                           int total, int64_t* cost, int32_t* oi, int32_t* oj) {
  const SlotDev& s = table[id];
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < total; e += gridDim.x * kBlock) {
    int nb = 0, r = e;
    for (; nb < MLP_NB_COUNT; ++nb) {
      if (!((mask >> nb) & 1)) continue;
      const int R = max(rows_of(nb, rowmap, s.N), 0);
      if (r < R) break;
      r -= R;
    }
    if (nb >= MLP_NB_COUNT) continue;
    const int par = (int)(s.info[MLP_INFO_PARITY0 + nb] & 1);
    const int64_t key = s.rows[nb][(par ? s.rowcap : 0) + r];
    int i = 0, j = 0;
    int64_t c = INT64_MAX;
    if (key != MLP_KEY_NONE) {
      c = key >> MLP_KEY_SHIFT;
      move_ij(nb, rowmap, s.N, r, (int)(key & MLP_KEY_MASK), i, j);
    }
    cost[e] = c;
    oi[e] = i;
    oj[e] = j;
  }
}

// TOTAL entries row * B + t of nb (ADS row map).
__global__ void k_total_moves64(const int64_t* __restrict__ tk, int64_t M, int B, int nb, int N,
                                uint64_t* out) {
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < M;
       e += (int64_t)gridDim.x * kBlock)
    out[e] = move64_of(tk[e], nb, 0, N, (int)(e / B));
}

// The reference's ADS buffer: header, coords, [client|w], T, C (uint32 wrap).
// S (prefix of T, uint32) is produced first by k_ads_prefix into scratch.
__global__ void k_ads_prefix(const SlotDev* __restrict__ table, int id, uint32_t* S) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const SlotDev& s = table[id];
    uint32_t acc = 0;
    for (int p = 0; p < s.N; ++p) {
      acc += (uint32_t)s.t32[p];
      S[p] = acc;
    }
  }
}

__global__ void k_ads_fill(const SlotDev* __restrict__ table, int id, const uint32_t* S,
                           uint32_t* out) {
  const SlotDev& s = table[id];
  const int N = s.N;
  const int64_t total = (int64_t)N * N;
  uint32_t* hdr = out;
  uint32_t* coords = out + 32;
  uint32_t* sol = coords + N;
  uint32_t* T = sol + N;
  uint32_t* C = T + N;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (g < 32) {
    uint32_t v = 0;
    if (g == 0) v = (uint32_t)(128 + 4 * (int64_t)N * (N + 3));
    else if (g == 1 || g == 2) v = (uint32_t)N;
    else if (g == 3) v = (uint32_t)(uint64_t)s.info[MLP_INFO_LATENCY];
    else if (g == 4) v = 1;
    else if (g == 5) v = (uint32_t)s.round;
    hdr[g] = v;
  }
  for (int64_t p = g; p < N; p += stride) {
    coords[p] = (uint32_t)s.xy[p];
    sol[p] = p == 0 ? (uint32_t)s.tour[0] : (((uint32_t)s.tour[p] << 16) | ((uint32_t)s.w[p] & 0xffffu));
    T[p] = (uint32_t)s.t32[p];
  }
  for (int64_t e = g; e < total; e += stride) {
    const int i = (int)(e / N), j = (int)(e % N);
    uint32_t v;
    const uint32_t Ti = (uint32_t)s.t32[i];
    if (i < j) v = (S[j] - S[i]) - (uint32_t)(j - i) * Ti;
    else if (i > j) v = (uint32_t)(i - j) * Ti - (S[i - 1] - (j > 0 ? S[j - 1] : 0u));
    else v = 0;
    C[e] = v;
  }
}

__global__ void k_dist_table(int s0, int count, int round, int32_t* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < count) out[q] = dist_s(s0 + q, round_dr(round));
}

// Mismatches of sqrt_rn_int against the library's correctly rounded sqrtf
// over s in [s0, s0 + count).
__global__ void k_sqrt_selfcheck(uint32_t s0, uint32_t count, unsigned long long* bad) {
  const uint32_t base = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  unsigned long long nb = 0;
  for (uint32_t q = base; q < count; q += stride) {
    const float f = (float)(s0 + q);
    nb += (__float_as_uint(sqrt_rn_int(f)) != __float_as_uint(__builtin_sqrtf(f))) ? 1ull : 0ull;
  }
  if (nb) atomicAdd(bad, nb);
}

__global__ void k_dist_list(const int32_t* s, int count, int round, int32_t* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < count) out[q] = dist_s(s[q], round_dr(round));
}

__global__ void k_weights(const int32_t* xs, const int32_t* ys, int n, int round, int32_t* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < (int64_t)n * n) {
    const int a = (int)(e / n), b = (int)(e % n);
    out[e] = dist_xy((xs[a] << 16) | ys[a], (xs[b] << 16) | ys[b], round_dr(round));
  }
}

// Clears the row buffers a sweep segment is about to fold into (one block
// per segment).
__global__ __launch_bounds__(kBlock) void k_clear_rows(const SlotDev* __restrict__ table,
                                                       SweepSegsInline inl,
                                                       const SweepSeg* segs_dev, int nsegs) {
  if ((int)blockIdx.x >= nsegs) return;
  const SweepSeg sg = segs_dev ? segs_dev[blockIdx.x] : inl.s[blockIdx.x];
  const SlotDev& s = table[sg.slot];
  for (int nb = 0; nb < MLP_NB_COUNT; ++nb)
    if ((sg.mask >> nb) & 1) {
      int64_t* b = s.rows[nb] + (((sg.parity >> nb) & 1) ? s.rowcap : 0);
      for (int r = threadIdx.x; r < s.rowcap; r += kBlock) b[r] = MLP_KEY_NONE;
    }
}

// Fixed-cost probes with the one-tour sweep's launch shape and kernel
// argument size (a SweepSegsInline table): kind 0 does nothing, 1 touches the
// tile's LDS and synchronises, 2 adds the dependent loads a sweep tile starts
// with (slot-table entry, then info and xy of the slot) and one store per
// workgroup, 3 adds four 64-bit atomic minima per workgroup and a row reset
// store, as a tile's row folds do.  out holds at least 2 * 8192 int64.
__global__ __launch_bounds__(kBlock) void k_probe(const SlotDev* __restrict__ table,
                                                  SweepSegsInline inl, int kind, int64_t* out) {
  __shared__ int L[1536];
  if (kind == 0) {
    if (threadIdx.x == 1023) out[0] = inl.s[blockIdx.x & 31].blk0;
    return;
  }
  int v = threadIdx.x;
  if (kind >= 2) {
    const SlotDev& s = table[inl.s[0].slot];
    v += (int)ldg(s.info + MLP_INFO_PATH32) + ldg(s.xy + (int)((blockIdx.x * 4 + threadIdx.x) % s.N));
  }
  L[threadIdx.x] = v;
  __syncthreads();
  if (kind == 1) {
    if (threadIdx.x == 1023) out[0] = L[(threadIdx.x + 1) & 255];
    return;
  }
  if (kind == 2) {
    if (threadIdx.x == 0) out[blockIdx.x % 8192] = L[255];
    return;
  }
  if (threadIdx.x < 4) out[8192 + (blockIdx.x % 2048) * 4 + threadIdx.x] = MLP_KEY_NONE;
  if ((threadIdx.x & 63) == 0)
    atomicMin((long long*)(out + (blockIdx.x % 2048) * 4 + threadIdx.x / 64),
              (long long)(L[threadIdx.x] + (int)blockIdx.x));
}

__global__ void k_fill_i64(int64_t* p, int64_t count, int64_t v) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count;
       t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v;
}

__global__ void k_set_slot(SlotDev* table, int id, SlotDev d) {
  if (blockIdx.x == 0 && threadIdx.x == 0) table[id] = d;
}

// A new slot's block in one launch: every 64-bit word zero except the row
// buffers (KEY_NONE), and the slot's table entry.
__global__ void k_slot_init(SlotDev* table, int id, SlotDev d, int64_t* mem, int64_t words,
                            SlotInitRows rows) {
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < words;
       w += (int64_t)gridDim.x * blockDim.x) {
    int64_t v = 0;
    for (int nb = 0; nb < rows.nbs; ++nb)
      if ((uint64_t)(w - rows.start[nb]) < (uint64_t)rows.count[nb]) v = MLP_KEY_NONE;
    mem[w] = v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) table[id] = d;
}

// Offset of slot ids[q]'s tour in a list-ordered array of tours: the sum of
// the earlier slots' N (one block).
__device__ int64_t tour_offset(const SlotDev* __restrict__ table, const int32_t* ids, int q) {
  __shared__ int64_t s_off;
  if (threadIdx.x == 0) s_off = 0;
  __syncthreads();
  int64_t part = 0;
  for (int r = threadIdx.x; r < q; r += blockDim.x) part += table[ids[r]].N;
  if (part) atomicAdd((unsigned long long*)&s_off, (unsigned long long)part);
  __syncthreads();
  return s_off;
}

// Tours of several slots into one array, in list order (block q: slot ids[q]).
__global__ void k_gather_tours(const SlotDev* __restrict__ table, const int32_t* ids, int count,
                               int32_t* out) {
  const int q = blockIdx.x;
  if (q >= count) return;
  const int64_t off = tour_offset(table, ids, q);
  const SlotDev& s = table[ids[q]];
  for (int p = threadIdx.x; p < s.N; p += blockDim.x) out[off + p] = s.tour[p];
}

// The inverse: a list-ordered array of tours into the slots.
__global__ void k_scatter_tours(const SlotDev* __restrict__ table, const int32_t* ids, int count,
                                const int32_t* in) {
  const int q = blockIdx.x;
  if (q >= count) return;
  const int64_t off = tour_offset(table, ids, q);
  const SlotDev& s = table[ids[q]];
  for (int p = threadIdx.x; p < s.N; p += blockDim.x) s.tour[p] = in[off + p];
}

__global__ void k_gather_info(const SlotDev* __restrict__ table, const int32_t* ids, int count,
                              int word, int64_t* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) out[t] = table[ids[t]].info[word];
}

// ===========================================================================
// k_sweep: full-neighborhood sweep of the ADS row map.
//
// kind 0 items own a tile of R consecutive rows "by" x C columns j of swap
// and of each or-opt-k.  Every one of those moves reads distances
// d(tour[a], tour[c]) for a in [by, by+k+1] only, so the workgroup computes
// the tour-ordered distance vectors v_a[c], a in [b0, b0+R+3], once into LDS
// (uint16) and all four neighborhoods' rows read them.  Per move:
//   swap    delta = Cs(i) + Bs(j) + (N-i) v_{i-1}[j] + (N-i-1) v_{i+1}[j]
//                 + (N-j) v_i[j-1] + (N-j-1) v_i[j+1]
//   or-opt  forward  RF(i) + CF(j) + (N-j-k) v_{i+k-1}[j+k] + (N-j) v_i[j+k-1] - j S(i)
//           backward RB(i) + CB(j) + (N-j) v_i[j-1] + (N-j-k) v_{i+k-1}[j] - j S(i)
// with row constants (Cs, RF, RB, S) and column constants (Bs, CF, CB) in
// LDS.  These are the positional forms of the reference's ODC kernels
// (mlkswap.cu:770-785, mlkoropt.cu:704-736) regrouped; each equals the ADS
// concatenation value of mlkswap.cu:210-278 / mlkoropt.cu:187-300.
// kind 1 items own four 2-opt rows (one per wave) x C row-local indices: a
// row is two diagonals j - i = m, and delta = G[i] + H[j] + (N-i) e[i-1] +
// (N-j-1) e[i] + m (T[i] + T[j]) with e[p] = d(tour[p], tour[p+m+1]) in LDS.
// Partial per-row bests of a tile are folded into the row buffer with a
// 64-bit atomicMin on the (delta, index) key.  Tours whose |delta| bound does
// not fit int32 take the per-lane int64 path.
// ===========================================================================
namespace {


template <int R, int C>
struct TileLds {
  uint16_t v[R + 4][C + 4];  // v[a - b0][c - cbase] = d(tour[a], tour[c])
  uint16_t w[C + 4];         // w[c] staged at c - cbase (a distance: below 2^16)
  int32_t t[C + 4];          // T[c] (mod 2^32) staged at c - cbase
  int32_t rw[R + 8];         // row side: w, T and xy at positions b0 .. b0 + R + 7
  int32_t rt[R + 8];
  int32_t rxy[R + 8];
};

// 2-opt inside a fixed-row tile: the column constants h, the row constants g
// and the tile's per-diagonal minimum keys (diagonal D = j - i at slot
// D - (q0 - b0 - R)); dgx holds the second key of the middle diagonals that
// belong to two rows of the 2-opt row map.
template <int R, int C>
struct TwoLds {
  int32_t h[C + 4];
  int32_t rg[R + 8];
  int64_t dg[C + R];
  int64_t dgx[4];
};

struct NoTwoLds {};

// ODC row map inside a fixed-row tile: per-diagonal minimum keys of or-opt-k
// (k = 1..3) at slot (j - i) - (q0 - b0 - R - 1).  With packed keys (|delta|
// < 2^26) the slots are int32 at the start of each array: delta * 32 + the
// move's tile row (forward moves) or R - 1 - that row (backward moves), so
// the minimum is the lowest delta and then the lowest ODC index.
template <int R, int C>
struct OdcLds {
  int64_t dg[3][C + R];
};

// Tile kinds: plain, with 2-opt (TWO), ODC row map (ODC).
constexpr int kPlain = 0, kTwo = 1, kOdc = 2;

template <int R, int C, int X>
struct FixedLds {
  TileLds<R, C> t;
  std::conditional_t<X == kTwo, TwoLds<R, C>, NoTwoLds> two;
  std::conditional_t<X == kOdc, OdcLds<R, C>, NoTwoLds> odc;
};

template <int C>
struct DiagLds {
  uint16_t e[MLP_2OPT_ROWS][2][C + 2];
};

template <int R, int C, int X>
union SweepLds {
  FixedLds<R, C, X> f;
  DiagLds<C> d;
};

__device__ __forceinline__ uint32_t umul(uint32_t a, uint32_t b) { return __umul24(a, b); }
// signed 24-bit multiply, wrapped to 32 bits (|a|, |b| < 2^23)
__device__ __forceinline__ uint32_t smul(int a, int b) { return (uint32_t)__mul24(a, b); }

// Word 3 of a raw buffer descriptor on gfx9-family parts (untyped dword
// access, no swizzle).
constexpr int kRsrcWord3 = 0x00020000;

// Wave index as a wave-uniform value (kept in a scalar register, so that
// branches on it are uniform branches).
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
}

__device__ __forceinline__ int64_t pair_key(int32_t d, int32_t c) {
  return c < 0 ? MLP_KEY_NONE : (int64_t)d * (int64_t)(1 << MLP_KEY_SHIFT) + c;
}

__device__ __forceinline__ int64_t* row_buf(const SlotDev& s, int nb, int par) {
  return s.rows[nb] + (par ? s.rowcap : 0);
}

__device__ __forceinline__ void fold_key(int64_t* p, int64_t key) {
  if (key != MLP_KEY_NONE) atomicMin((long long*)p, (long long)key);
}

// fold_key for rows that many tiles fold into (the diagonals of the ODC
// or-opt rows and of 2-opt in fixed tiles, one key per tile the diagonal
// crosses): the row's current value is read first and the atomic is issued
// only by a key below it.  Values only decrease during the launch, so a stale
// read costs at most an unneeded atomic.  It pays when tiles of a tour run
// one after another (most keys then lose to the value already there); when
// they all run at once every read still finds the reset value, and the read
// only adds latency.
__device__ __forceinline__ void fold_key_filtered(int64_t* p, int64_t key) {
  if (key != MLP_KEY_NONE && key < __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMin((long long*)p, (long long)key);
}

__device__ __forceinline__ uint32_t lane_bcast(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

// Reset of the buffer holding the previous result (par ^ 1) by the tiles of
// column 0; the first tile also clears both tails beyond this row map.
__device__ __forceinline__ void reset_other(const SlotDev& s, int nb, int par, int r0, int r1, bool first,
                                            int rowmap = 0) {
  int64_t* o = row_buf(s, nb, par ^ 1);
  for (int r = r0 + (int)threadIdx.x; r < r1; r += kBlock) o[r] = MLP_KEY_NONE;
  if (first) {
    for (int r = rows_of(nb, rowmap, s.N) + (int)threadIdx.x; r < s.rowcap; r += kBlock) {
      o[r] = MLP_KEY_NONE;
      s.rows[nb][(par ? s.rowcap : 0) + r] = MLP_KEY_NONE;
    }
  }
}

// Value barrier: the compiler sees an opaque value, so a per-column term
// stays one shared value instead of being distributed into every row's
// products (which turns two 24-bit multiplies into one quarter-rate 32-bit
// multiply per row).
__device__ __forceinline__ uint32_t opaque(uint32_t v) {
  asm("" : "+v"(v));
  return v;
}

// Per-lane running minimum over increasing indices: strict < keeps the
// first (lowest) index on ties.  Selects, not branches.
__device__ __forceinline__ void take_min(int32_t& bd, int32_t& bc, uint32_t d, int c) {
  const bool lt = (int32_t)d < bd;
  bc = lt ? c : bc;
  bd = lt ? (int32_t)d : bd;
}

__device__ __forceinline__ void take_min_if(int32_t& bd, int32_t& bc, uint32_t d, int c, bool ok) {
  const bool lt = ok && (int32_t)d < bd;
  bc = lt ? c : bc;
  bd = lt ? (int32_t)d : bd;
}

template <bool B>
using Flag = std::integral_constant<bool, B>;

// Packed per-lane key of a tour whose |delta| < 2^26: delta * 32 + the
// lane's column iteration (below 32), so one 32-bit minimum keeps the lowest
// delta and, among equal ones, the lane's lowest column.
__device__ __forceinline__ int32_t pack_key(uint32_t d, int it) {
  return (int32_t)((d << 5) + (uint32_t)it);
}

// The row key from per-lane packed keys: one wave minimum, then the lowest
// lane holding it (its column is the lowest among equal deltas).  Column of
// lane l at iteration it: jb0 + it * 64 + l; ctx = column + off.
__device__ __forceinline__ int64_t packed_row_key(int32_t bk, int jb0, int off) {
  const int32_t m = wave_min_i32(bk);
  if (m == INT32_MAX) return MLP_KEY_NONE;
  const int l0 = (int)__builtin_ctzll(__builtin_amdgcn_ballot_w64(bk == m));
  const int ctx = jb0 + (m & 31) * kWave + l0 + off;
  return (int64_t)(m >> 5) * (int64_t)(1 << MLP_KEY_SHIFT) + ctx;
}

// Swap rows of a tile: each wave owns RW consecutive rows (r0 = wave * RW),
// lanes stride over the tile's columns j and evaluate all RW rows for their
// column.  Row r: i = b0+r+1, j in [i+2, N-2].  delta = Cs_r + Bs(j) +
// (N-i) v_{i-1}[j] + (N-i-1) v_{i+1}[j] + (N-j) v_i[j-1] + (N-j-1) v_i[j+1].
//
// TWO: the same loop also evaluates the 2-opt moves (i, j), j in [i+1, N-2],
// of rows i <= N-3 (mlk2opt.cu:168-219 in the closed form
// g[i] + h[j] + (N-i) v_{i-1}[j] + (N-j-1) v_i[j+1] + (j-i)(T[i] + T[j]),
// written A_i + B_j + j T[i] - i T[j] with A_i = g[i] - i T[i] and
// B_j = h[j] + j T[j]); its two distance products are swap's first and last.
// The 2-opt row map's rows are diagonals (row D-1 for j - i = D <= ceil(N/2),
// row N-2-D for D > N-2-ceil(N/2), both in between), so each move's key
// (delta << 20 | row-local index) is folded into the tile's per-diagonal
// minimum in LDS; the tile flushes one key per diagonal afterwards.  Needs
// the int32 2-opt path and T[N-1] < 2^24 (the j T[i] and i T[j] products in
// 24-bit multiplies).
template <int R, int C, int RW, bool TWO, typename TL, bool ODC = false>
__device__ __forceinline__ void tile_swap(const SlotDev& s, int b0, int q0, int q1, int cbase, TileLds<R, C>& L,
                          TL& T2, int parity, int wave, int lane, bool pk) {
  const int N = s.N;
  const int r0 = wave * RW;
  const int rv = min(RW, N - 4 - b0 - r0);  // rows with i <= N-4
  const int rv2 = TWO ? min(RW, N - 3 - b0 - r0) : rv;  // 2-opt rows: i <= N-3
  if (rv2 <= 0) return;
  if constexpr (ODC) pk = false;
  int32_t bd[RW], bc[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    bd[r] = INT32_MAX;
    bc[r] = -1;
  }
  uint32_t myCs = 0u, myA = 0u, myT = 0u;
  if (lane < rv2) {
    const int o = r0 + lane + 1;  // i - b0
    const int i = b0 + o;
    myCs = 0u - umul(N - i, (uint32_t)L.rw[o]) - umul(N - i - 1, (uint32_t)L.rw[o + 1]);
    if constexpr (TWO) {
      myT = (uint32_t)L.rt[o];
      myA = (uint32_t)T2.rg[o] - (uint32_t)i * myT;
    }
  }
  // ODC row map (mlkswap.cu:764-768): ADS row ar = b0 + r0 + r is the
  // first part of ODC row ar (index j - 3, lowest j first on ties) while
  // ar < rowsO, and the second part of ODC row N - 4 - ar (index N - 2 - j,
  // highest j first) while N - 4 - ar < rowsO
  const int rowsO = ((N + 1) >> 1) - 1;
  int32_t be[RW], bf[RW];
  bool up[RW], dn[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    be[r] = INT32_MAX;
    bf[r] = -1;
    up[r] = b0 + r0 + r < rowsO;
    dn[r] = N - 4 - (b0 + r0 + r) < rowsO;
  }
  uint32_t Cs[RW], A2[RW], T2i[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    Cs[r] = lane_bcast(myCs, r);
    if constexpr (TWO) {
      A2[r] = lane_bcast(myA, r);
      T2i[r] = lane_bcast(myT, r);
    }
  }
  const uint16_t(*V)[C + 4] = L.v + r0;
  const int jhi = min(q1, N - 1);
  const int jlo = max(q0, b0 + r0 + (TWO ? 2 : 3));
  const int jb0 = jlo - ((jlo - q0) & (kWave - 1));
  // 2-opt: diagonal slot of (r, j) is j - (b0 + r0 + r + 1) - dlo; row-map
  // split of the diagonals (rows2 = ceil(N/2)): D <= rows2 -> row D - 1 with
  // index j - 2; D > N - 2 - rows2 -> row N - 2 - D with index N - 2 - j
  const int rows2 = (N + 1) >> 1;
  const int dlo = q0 - b0 - R;
  const int mid0 = N - 1 - rows2;  // first diagonal that also has a part-0 row
  const bool mids = TWO && (q0 - b0 - R <= rows2) && (q1 - 1 - b0 - 1 >= mid0);
  // FR: every row of the wave is in range (no per-row guard in the loop);
  // PK: packed keys in bd (bc unused)
  auto cols = [&](auto fr, auto pkc) {
    constexpr bool FR = decltype(fr)::value;
    constexpr bool PK = decltype(pkc)::value;
    int it = 0;
    for (int jb = jb0; jb < jhi; jb += kWave, ++it) {
      const int j = jb + lane;
      const int jj = min(max(j, jlo), jhi - 1);
      const int x = jj - cbase;
      const uint32_t nj = (uint32_t)(N - jj);
      const uint32_t cs = opaque(0u - umul(nj, L.w[x]) - umul(nj - 1u, L.w[x + 1]));
      uint32_t Bj = 0u, Tj = 0u;
      int ctx1 = 0, ctx0 = 0;
      if constexpr (TWO) {
        Tj = (uint32_t)L.t[x];
        Bj = opaque((uint32_t)T2.h[x] + umul((uint32_t)jj, Tj));
        ctx1 = jj - 2;
        ctx0 = N - 2 - jj;
      }
      uint32_t c[RW + 2];
#pragma unroll
      for (int a = 0; a < RW + 2; ++a) c[a] = V[a][x];
      const bool full = jb >= b0 + r0 + RW + 2 && jb + kWave <= jhi;
#pragma unroll
      for (int r = 0; r < RW; ++r) {
        const uint32_t a1 = (uint32_t)(N - b0 - r0 - r - 1);
        const uint32_t v4 = V[r + 1][x + 1];
        const uint32_t X = umul(a1, c[r]) + umul(nj - 1u, v4);  // shared by swap and 2-opt
        if (FR || r < rv) {
          const uint32_t d = Cs[r] + cs + X + umul(a1 - 1u, c[r + 2]) +
                             umul(nj, V[r + 1][x - 1]);
          if constexpr (ODC) {
            const bool ok = full || (j < jhi && j >= b0 + r0 + r + 3);
            if (up[r]) take_min_if(bd[r], bc[r], d, j, ok);
            if (dn[r]) {  // <= keeps the later (higher) column of equal deltas
              const bool le = ok && (int32_t)d <= be[r];
              bf[r] = le ? j : bf[r];
              be[r] = le ? (int32_t)d : be[r];
            }
          } else if (PK) {
            const int32_t k = pack_key(d, it);
            if (full) bd[r] = min(bd[r], k);
            else bd[r] = (j < jhi && j >= b0 + r0 + r + 3) ? min(bd[r], k) : bd[r];
          } else if (full) {
            take_min(bd[r], bc[r], d, j);
          } else {
            take_min_if(bd[r], bc[r], d, j, j < jhi && j >= b0 + r0 + r + 3);
          }
        }
        if constexpr (TWO) {
          if (FR || r < rv2) {
            const int i = b0 + r0 + r + 1;
            const uint32_t d2 =
                A2[r] + Bj + X + umul((uint32_t)jj, T2i[r]) - umul((uint32_t)i, Tj);
            const int D = jj - i;
            const bool ok = full || (j < jhi && j >= i + 1);
            const int64_t base = (int64_t)(int32_t)d2 * (int64_t)(1 << MLP_KEY_SHIFT);
            if (ok) atomicMin((long long*)&T2.dg[D - dlo], (long long)(base + (D <= rows2 ? ctx1 : ctx0)));
            if (mids && ok && D >= mid0 && D <= rows2)
              atomicMin((long long*)&T2.dgx[D - mid0], (long long)(base + ctx0));
          }
        }
      }
    }
  };
  int64_t* out = row_buf(s, 0, parity & 1) + b0 + r0;
  if constexpr (ODC) {
    if (rv2 == RW) cols(Flag<true>{}, Flag<false>{});
    else cols(Flag<false>{}, Flag<false>{});
    int64_t* ob = row_buf(s, 0, parity & 1);
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int ar = b0 + r0 + r;
      if (r < rv && up[r]) {
        const int32_t m = wave_min_i32(bc[r] >= 0 ? bd[r] : INT32_MAX);
        const int32_t c = wave_min_i32((bc[r] >= 0 && bd[r] == m) ? bc[r] : INT32_MAX);
        if (lane == 0 && c != INT32_MAX)
          fold_key(ob + ar, (int64_t)m * (int64_t)(1 << MLP_KEY_SHIFT) + (c - 3));
      }
      if (r < rv && dn[r]) {
        int32_t m, c;
        wave_argmin_hi(be[r], bf[r], m, c);
        if (lane == 0 && c >= 0)
          fold_key(ob + (N - 4 - ar), (int64_t)m * (int64_t)(1 << MLP_KEY_SHIFT) + (N - 2 - c));
      }
    }
    return;
  }
  if (pk) {
    if (rv2 == RW) cols(Flag<true>{}, Flag<true>{});
    else cols(Flag<false>{}, Flag<true>{});
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int64_t key = packed_row_key(bd[r], jb0, 0);
      if (lane == 0 && r < rv) fold_key(out + r, key);
    }
    return;
  }
  if (rv2 == RW) cols(Flag<true>{}, Flag<false>{});
  else cols(Flag<false>{}, Flag<false>{});
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const int64_t key = wave_argmin(bd[r], bc[r]);
    if (lane == 0 && r < rv) fold_key(out + r, key);
  }
}

// Or-opt-K rows of a tile, each wave owning RW rows.  Row r (by = b0+r):
// forward i = by+1 for j >= by+2, backward i = by+2 for j <= by+1, j in
// [1, N-K-1]; row-local index j - 1.
// ODC: the ODC row map's rows are diagonals (mlkoropt.cu:696-723): a forward
// move (j > i) belongs to row j - i - 1 with index j - 2, a backward one with
// i - j >= 2 to row N - K - 1 - (i - j) with index N - K - 1 - i (backward
// moves by one position are not in that row map); each move's key is folded
// into the tile's per-diagonal minimum dg[(j - i) - (q0 - b0 - R - 1)] in LDS
// (an int32 packed key when pk).
template <int R, int C, int RW, int K, bool ODC = false>
__device__ __forceinline__ void tile_oropt(const SlotDev& s, int b0, int q0, int q1, int cbase, TileLds<R, C>& L,
                           int parity, int wave, int lane, bool pk, int64_t* dg = nullptr) {
  const int N = s.N;
  const float dr = round_dr(s.round);
  const int r0 = wave * RW;
  const int rv = min(RW, N - K - 2 - b0 - r0);  // rows with by <= N-K-3
  if (rv <= 0) return;
  const int bb = b0 + r0;  // first row of this wave
  int32_t bd[RW], bc[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    bd[r] = INT32_MAX;
    bc[r] = -1;
  }
  uint32_t mRF = 0u, mRB = 0u;
  int mSF = 0, mSB = 0;
  if (lane < rv) {
    const int o = r0 + lane + 1;  // iF - b0; iB = iF + 1
    const int iF = b0 + o, iB = iF + 1;
    const uint32_t SF = (uint32_t)L.rt[o + K - 1] - (uint32_t)L.rt[o];
    const uint32_t SB = (uint32_t)L.rt[o + K] - (uint32_t)L.rt[o + 1];
    const uint32_t dF = (uint32_t)dist_xy(L.rxy[o - 1], L.rxy[o + K], dr);
    const uint32_t dB = (uint32_t)dist_xy(L.rxy[o], L.rxy[o + 1 + K], dr);
    mRF = (uint32_t)(N - iF) * (dF - (uint32_t)L.rw[o]) -
          (uint32_t)(N - iF - K) * (uint32_t)L.rw[o + K] + (uint32_t)iF * SF -
          (uint32_t)K * (uint32_t)L.rt[o + K];
    mRB = (uint32_t)(N - iB - K) * (dB - (uint32_t)L.rw[o + 1 + K]) -
          (uint32_t)(N - iB) * (uint32_t)L.rw[o + 1] + (uint32_t)iB * SB -
          (uint32_t)K * (uint32_t)L.rt[o];
    mSF = -(int)SF;
    mSB = -(int)SB;
  }
  uint32_t RF[RW], RB[RW];
  int nSF[RW], nSB[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    RF[r] = lane_bcast(mRF, r);
    RB[r] = lane_bcast(mRB, r);
    nSF[r] = (int)lane_bcast((uint32_t)mSF, r);
    nSB[r] = (int)lane_bcast((uint32_t)mSB, r);
  }
  const uint16_t(*V)[C + 4] = L.v + r0;
  const int jhi = min(q1, N - K);
  const int slo = q0 - b0 - R - 1;
  // PK: packed keys in bd (bc unused)
  int32_t* dg32 = reinterpret_cast<int32_t*>(dg);
  auto upd = [&](auto pkc, int r, uint32_t d, int it, int j, bool ok, bool fwd) {
    if constexpr (ODC && decltype(pkc)::value) {
      const int S = j - bb - r - (fwd ? 1 : 2);  // j - i
      if (ok && S != -1)
        atomicMin(&dg32[S - slo], pack_key(d, fwd ? r0 + r : R - 1 - (r0 + r)));
    } else if constexpr (ODC) {
      const int i = fwd ? bb + r + 1 : bb + r + 2;
      const int S = j - i;
      if (ok && S != -1)
        atomicMin((long long*)&dg[S - slo],
                  (long long)((int64_t)(int32_t)d * (int64_t)(1 << MLP_KEY_SHIFT) +
                              (S >= 1 ? j - 2 : N - K - 1 - i)));
    } else if constexpr (decltype(pkc)::value) {
      bd[r] = ok ? min(bd[r], pack_key(d, it)) : bd[r];
    } else {
      take_min_if(bd[r], bc[r], d, j - 1, ok);
    }
  };
  auto cols = [&](auto fr, auto pkc) {
    constexpr bool FR = decltype(fr)::value;
    int it = 0;
    for (int jb = q0; jb < jhi; jb += kWave, ++it) {
      const int j = jb + lane;
      const int jj = min(j, jhi - 1);
      const int x = jj - cbase;
      const uint32_t nj = (uint32_t)(N - jj);
      const bool tail = jb + kWave > jhi;
      const bool allF = !tail && jb >= bb + RW + 1;
      const bool allB = !tail && jb + kWave - 1 <= bb + 1;
      if (allF) {
        const uint32_t CF = opaque(0u - umul(nj - K, L.w[x + K]) + (uint32_t)K * (uint32_t)L.t[x + K - 1]);
#pragma unroll
        for (int r = 0; r < RW; ++r) {
          if (FR || r < rv) {
            uint32_t d = RF[r] + CF + umul(nj - K, V[r + K][x + K]) + umul(nj, V[r + 1][x + K - 1]);
            if (K > 1) d += smul(jj, nSF[r]);
            upd(pkc, r, d, it, j, true, true);
          }
        }
      } else if (allB) {
        const uint32_t CB = opaque(0u - umul(nj, L.w[x]) + (uint32_t)K * (uint32_t)L.t[x]);
#pragma unroll
        for (int r = 0; r < RW; ++r) {
          if (FR || r < rv) {
            uint32_t d = RB[r] + CB + umul(nj, V[r + 2][x - 1]) + umul(nj - K, V[r + K + 1][x]);
            if (K > 1) d += smul(jj, nSB[r]);
            upd(pkc, r, d, it, j, true, false);
          }
        }
      } else {
        const uint32_t CF = opaque(0u - umul(nj - K, L.w[x + K]) + (uint32_t)K * (uint32_t)L.t[x + K - 1]);
        const uint32_t CB = opaque(0u - umul(nj, L.w[x]) + (uint32_t)K * (uint32_t)L.t[x]);
#pragma unroll
        for (int r = 0; r < RW; ++r) {
          if (FR || r < rv) {
            const bool fwd = jj >= bb + r + 2;
            uint32_t dF = RF[r] + CF + umul(nj - K, V[r + K][x + K]) + umul(nj, V[r + 1][x + K - 1]);
            uint32_t dB = RB[r] + CB + umul(nj, V[r + 2][x - 1]) + umul(nj - K, V[r + K + 1][x]);
            if (K > 1) {
              dF += smul(jj, nSF[r]);
              dB += smul(jj, nSB[r]);
            }
            upd(pkc, r, fwd ? dF : dB, it, j, j < jhi, fwd);
          }
        }
      }
    }
  };
  int64_t* out = row_buf(s, K + 1, (parity >> (K + 1)) & 1) + bb;
  if constexpr (ODC) {  // the tile flushes the diagonals afterwards
    if (pk) {
      if (rv == RW) cols(Flag<true>{}, Flag<true>{});
      else cols(Flag<false>{}, Flag<true>{});
    } else {
      if (rv == RW) cols(Flag<true>{}, Flag<false>{});
      else cols(Flag<false>{}, Flag<false>{});
    }
    return;
  }
  if (pk) {
    if (rv == RW) cols(Flag<true>{}, Flag<true>{});
    else cols(Flag<false>{}, Flag<true>{});
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int64_t key = packed_row_key(bd[r], q0, -1);  // ctx = j - 1
      if (lane == 0 && r < rv) fold_key(out + r, key);
    }
    return;
  }
  if (rv == RW) cols(Flag<true>{}, Flag<false>{});
  else cols(Flag<false>{}, Flag<false>{});
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const int64_t key = wave_argmin(bd[r], bc[r]);
    if (lane == 0 && r < rv) fold_key(out + r, key);
  }
}

// int64 fallback of a fixed-i tile (column tile 0 only): one wave per
// (neighborhood, row) over the whole row.
template <int R, int RM = 0>
__device__ __forceinline__ void sweep_fixed_wide(const SlotDev& s, int rb, int ct, int mask, int parity) {
  const int N = s.N, wave = wave_id(), lane = threadIdx.x % kWave;
  const int b0 = rb * R;
  if (ct != 0) return;
#pragma unroll
  for (int nb = 0; nb < MLP_NB_COUNT; ++nb)
    if (nb != 1 && ((mask >> nb) & 1))
      reset_other(s, nb, (parity >> nb) & 1, b0, min(b0 + R, rows_of(nb, RM, N)), rb == 0, RM);
  for (int sg = wave; sg < 4 * R; sg += kBlock / kWave) {
    const int g = sg / R, by = b0 + sg % R;
    const int nb = g == 0 ? 0 : g + 1;
    if (!((mask >> nb) & 1) || by >= rows_of(nb, RM, N)) continue;
    int64_t* out = row_buf(s, nb, (parity >> nb) & 1) + by;
    if (nb == 0) eval_row<0, false>(s, nb, RM, by, lane, out);
    else eval_row<2, false>(s, nb, RM, by, lane, out);
  }
}

// 2-opt rows [b0, b0 + R) of a fixed-row tile whose launch includes 2-opt
// but whose tour does not take the in-tile path (64-bit deltas, or
// T[N-1] >= 2^24): column tile 0 resets them and evaluates each diagonal row
// with one wave.
template <int R>
__device__ __forceinline__ void two_rows_wide(const SlotDev& s, int rb, int ct, int parity,
                                              bool p32) {
  if (ct != 0) return;
  const int N = s.N, wave = wave_id(), lane = threadIdx.x % kWave;
  const int rows2 = (N + 1) >> 1, b0 = rb * R;
  const int par = (parity >> 1) & 1;
  reset_other(s, 1, par, b0, min(b0 + R, rows2), rb == 0);
  for (int by = b0 + wave; by < min(b0 + R, rows2); by += kBlock / kWave) {
    if (p32) eval_row<1, true>(s, 1, 0, by, lane, row_buf(s, 1, par) + by);
    else eval_row<1, false>(s, 1, 0, by, lane, row_buf(s, 1, par) + by);
  }
}

template <int R, int C, int X>
__device__ __forceinline__ void sweep_fixed(const SlotDev& s, int rb, int ct, int mask, int parity,
                            FixedLds<R, C, X>& F, bool filtered) {
  constexpr bool TWO = X == kTwo;
  constexpr bool ODC = X == kOdc;  // ODC row map for swap and or-opt
  constexpr int RM = ODC ? 1 : 0;
  constexpr int RW = R / (kBlock / kWave);
  constexpr int FI = (C + 4 + kBlock - 1) / kBlock;  // fill iterations per thread
  TileLds<R, C>& L = F.t;
  const int N = s.N, tid = threadIdx.x, wave = wave_id(), lane = tid % kWave;
  const float dr = round_dr(s.round);
  const int b0 = rb * R;
  const int q0 = 1 + ct * C;                 // move j in [q0, q1)
  const int q1 = min(q0 + C, N - 1);
  const int cbase = q0 - 1;                  // LDS column 0 <-> tour position q0 - 1
  const int cend = min(q1 + 3, N);
  // Swap and 2-opt moves have j >= i + 1 >= b0 + 2 and read columns from
  // j - 1 on: without or-opt the block starts at position b0 + 1 (the
  // columns left of it are not filled and never used).
  const int cfill = (mask & 0x1c) ? cbase : max(cbase, min(b0 + 1, cend - 1));
  // every global load of the tile is issued before the first wait
  const int64_t pflags = ldg(s.info + MLP_INFO_PATH32);
  const bool p32 = (pflags & MLP_P32_SWAP_OROPT) != 0;
  const bool pk = (pflags & MLP_PK_SWAP_OROPT) != 0;  // |delta| < 2^26: packed keys
  // 2-opt in this tile: int32 2-opt deltas and 24-bit prefix times
  const bool two = TWO && (mask & 2) && p32 && (pflags & MLP_P32_2OPT) && (pflags & MLP_T24);
  int xa[R + 4];
#pragma unroll
  for (int va = 0; va < R + 4; ++va) xa[va] = ldg(s.xy + min(b0 + va, N - 1));
  int xc[FI], wv[FI], tv[FI], hv[FI];
#pragma unroll
  for (int it = 0; it < FI; ++it) {
    const int c = min(cfill + tid + it * kBlock, cend - 1);
    xc[it] = ldg(s.xy + c);
    wv[it] = ldg(s.w + c);
    tv[it] = ldg(s.t32 + c);
    if (TWO) hv[it] = ldg(s.h32 + c);
  }
  int rwv = 0, rtv = 0, rxv = 0, rgv = 0;
  if (tid < R + 8) {
    const int p = min(b0 + tid, N - 1);
    rwv = ldg(s.w + p);
    rtv = ldg(s.t32 + p);
    rxv = ldg(s.xy + p);
    if (TWO) rgv = ldg(s.g32 + p);
  }
  if (TWO && (mask & 2) && !two)
    two_rows_wide<R>(s, rb, ct, parity, (pflags & MLP_P32_2OPT) != 0);
  if (!p32) {
    sweep_fixed_wide<R, RM>(s, rb, ct, mask, parity);
    return;
  }
  if (ct == 0) {
#pragma unroll
    for (int nb = 0; nb < MLP_NB_COUNT; ++nb)
      if (nb != 1 && ((mask >> nb) & 1))
        reset_other(s, nb, (parity >> nb) & 1, b0, min(b0 + R, rows_of(nb, RM, N)), rb == 0, RM);
    if (two) reset_other(s, 1, (parity >> 1) & 1, b0, min(b0 + R, (N + 1) >> 1), rb == 0);
  }
  // swap alone covers j >= i + 2 only: a tile left of that diagonal is empty
  if (mask == 1 && q1 <= b0 + 3) return;
#pragma unroll
  for (int it = 0; it < FI; ++it) {
    const int c = cfill + tid + it * kBlock;
    if (c < cend) {
#pragma unroll
      for (int va = 0; va < R + 4; ++va) L.v[va][c - cbase] = (uint16_t)dist_xy(xa[va], xc[it], dr);
      L.w[c - cbase] = (uint16_t)wv[it];
      L.t[c - cbase] = tv[it];
      if constexpr (TWO) F.two.h[c - cbase] = hv[it];
    }
  }
  if (tid < R + 8) {
    L.rw[tid] = rwv;
    L.rt[tid] = rtv;
    L.rxy[tid] = rxv;
    if constexpr (TWO) F.two.rg[tid] = rgv;
  }
  if constexpr (TWO) {
    for (int q = tid; q < C + R; q += kBlock) F.two.dg[q] = MLP_KEY_NONE;
    if (tid < 4) F.two.dgx[tid] = MLP_KEY_NONE;
  }
  if constexpr (ODC) {
    if (pk)
      for (int q = tid; q < 3 * (C + R); q += kBlock)
        reinterpret_cast<int32_t*>(F.odc.dg[q / (C + R)])[q % (C + R)] = INT32_MAX;
    else
      for (int q = tid; q < 3 * (C + R); q += kBlock) F.odc.dg[q / (C + R)][q % (C + R)] = MLP_KEY_NONE;
  }
  // the or-opt diagonals of an ODC tile (unused otherwise)
  int64_t* dg1 = nullptr;
  int64_t* dg2 = nullptr;
  int64_t* dg3 = nullptr;
  if constexpr (ODC) {
    dg1 = F.odc.dg[0];
    dg2 = F.odc.dg[1];
    dg3 = F.odc.dg[2];
  }
  __syncthreads();
  if (!TWO && R == kBlock / kWave && __popc(mask & 0x1d) >= 2) {
    // Several neighborhoods on a four-row tile: wave w takes the w-th
    // neighborhood of the mask over all rows, so the neighborhoods run side
    // by side instead of one after another (the latency of a one-tour launch).
    int nb = -1, q = wave;
#pragma unroll
    for (int b = 0; b < MLP_NB_COUNT; ++b)
      if (b != 1 && ((mask >> b) & 1)) {
        if (q == 0) nb = b;
        --q;
      }
    if (nb == 0)
      tile_swap<R, C, R, false, decltype(F.two), ODC>(s, b0, q0, q1, cbase, L, F.two, parity, 0,
                                                      lane, pk);
    else if (nb == 2) tile_oropt<R, C, R, 1, ODC>(s, b0, q0, q1, cbase, L, parity, 0, lane, pk, dg1);
    else if (nb == 3) tile_oropt<R, C, R, 2, ODC>(s, b0, q0, q1, cbase, L, parity, 0, lane, pk, dg2);
    else if (nb == 4) tile_oropt<R, C, R, 3, ODC>(s, b0, q0, q1, cbase, L, parity, 0, lane, pk, dg3);
    if constexpr (!ODC) return;
  } else {
    if (two) tile_swap<R, C, RW, TWO>(s, b0, q0, q1, cbase, L, F.two, parity, wave, lane, pk);
    else if (mask & 1)
      tile_swap<R, C, RW, false, decltype(F.two), ODC>(s, b0, q0, q1, cbase, L, F.two, parity,
                                                       wave, lane, pk);
    if (mask & 4) tile_oropt<R, C, RW, 1, ODC>(s, b0, q0, q1, cbase, L, parity, wave, lane, pk, dg1);
    if (mask & 8) tile_oropt<R, C, RW, 2, ODC>(s, b0, q0, q1, cbase, L, parity, wave, lane, pk, dg2);
    if (mask & 16) tile_oropt<R, C, RW, 3, ODC>(s, b0, q0, q1, cbase, L, parity, wave, lane, pk, dg3);
  }
  if constexpr (ODC) {
    // one key per or-opt diagonal of the tile into its ODC row
    __syncthreads();
    const int slo = q0 - b0 - R - 1;
    for (int q = tid; q < 3 * (C + R); q += kBlock) {
      const int k = q / (C + R) + 1, sl = q % (C + R);
      if (!((mask >> (k + 1)) & 1)) continue;
      const int S = slo + sl;  // j - i
      int64_t key;
      if (pk) {
        // packed: delta * 32 + tile row (forward) or R - 1 - tile row
        // (backward); ODC index j - 2 = S + i - 2 or N - k - 1 - i
        const int32_t kp = reinterpret_cast<const int32_t*>(F.odc.dg[k - 1])[sl];
        if (kp == INT32_MAX) continue;
        const int o = kp & 31;
        const int idx = S >= 1 ? S + b0 + o - 1 : N - k - 1 - (b0 + (R - 1 - o) + 2);
        key = (int64_t)(kp >> 5) * (int64_t)(1 << MLP_KEY_SHIFT) + idx;
      } else {
        key = F.odc.dg[k - 1][sl];
        if (key == MLP_KEY_NONE) continue;
      }
      int64_t* p = row_buf(s, k + 1, (parity >> (k + 1)) & 1) + (S >= 1 ? S - 1 : N - k - 1 + S);
      if (filtered) fold_key_filtered(p, key);
      else fold_key(p, key);
    }
  }
  if constexpr (TWO) {
    if (two) {
      // one key per diagonal of the tile into the 2-opt row buffer
      __syncthreads();
      const int rows2 = (N + 1) >> 1, dlo = q0 - b0 - R, mid0 = N - 1 - rows2;
      int64_t* out = row_buf(s, 1, (parity >> 1) & 1);
      for (int q = tid; q < C + R; q += kBlock) {
        const int64_t key = F.two.dg[q];
        const int D = dlo + q;
        int64_t* p = out + (D <= rows2 ? D - 1 : N - 2 - D);
        if (key == MLP_KEY_NONE) continue;
        if (filtered) fold_key_filtered(p, key);
        else fold_key(p, key);
      }
      if (tid < 4 && F.two.dgx[tid] != MLP_KEY_NONE)
        fold_key(out + (N - 2 - (mid0 + tid)), F.two.dgx[tid]);
    }
  }
}

__device__ __forceinline__ void sweep_2opt_wide(const SlotDev& s, int rb, int ct, int parity) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int rows = (s.N + 1) >> 1;
  const int par = (parity >> 1) & 1;
  if (ct != 0) return;
  reset_other(s, 1, par, rb * MLP_2OPT_ROWS, min(rb * MLP_2OPT_ROWS + MLP_2OPT_ROWS, rows), rb == 0);
  const int by = rb * MLP_2OPT_ROWS + wave;
  if (by < rows) eval_row<1, false>(s, 1, 0, by, lane, row_buf(s, 1, par) + by);
}

template <int C>
__device__ __forceinline__ void sweep_2opt(const SlotDev& s, int rb, int ct, int parity, DiagLds<C>& L) {
  constexpr int FE = (C + 1 + kWave - 1) / kWave;  // e entries per lane and part
  constexpr int FT = (C + 2 + kWave - 1) / kWave;  // e entries per lane, both parts
  // Wide tiles hold both parts of every row of a tour up to C positions:
  // their two parts' entries are filled as one flat range (one distance per
  // entry; 200 tours n=1000, 16x1024: 281 -> 265 us), narrow tiles fill each
  // part on its own (the flat form measured 3% slower at 4x256)
  constexpr bool FLAT = C >= 1024;
  constexpr int FP = FLAT ? 1 : 2, FI = FLAT ? FT : FE;
  // compute iterations per load batch (a part of a tile holds up to C + 1
  // entries: two for the 128-column tiles, whose batches would otherwise be
  // half empty)
  constexpr int CU = C >= 4 * kWave ? 4 : 2;
  const int N = s.N, tid = threadIdx.x, wave = wave_id(), lane = tid % kWave;
  const float dr = round_dr(s.round);
  const int rows = (N + 1) >> 1;
  const int par = (parity >> 1) & 1;
  const bool p32 = (ldg(s.info + MLP_INFO_PATH32) & MLP_P32_2OPT) != 0;
  const int by = rb * MLP_2OPT_ROWS + wave;
  const bool live = by < rows;
  const int c0 = ct * C, c1 = min(c0 + C, N - 3);  // row-local indices [c0, c1)
  // part 0: ctx < by  -> m = N-2-by, i = by - ctx  in [by - cb1 + 1, by - c0]
  // part 1: ctx >= by -> m = by+1,   i = ctx-by+1  in [ca0 - by + 1, c1 - by]
  // ilo / ihi / m are uniform over the wave (one row per wave).
  int ilo[2], ihi[2], mm[2];
  const int cb1 = min(c1, by), ca0 = max(c0, by);
  mm[0] = N - 2 - by;
  ilo[0] = by - cb1 + 1;
  ihi[0] = (live && c0 < cb1) ? by - c0 : ilo[0] - 1;
  mm[1] = by + 1;
  ilo[1] = ca0 - by + 1;
  ihi[1] = (live && ca0 < c1) ? c1 - by : ilo[1] - 1;
  // e[x] = d(tour[ilo-1+x], tour[ilo+x+m]) for x in [0, ihi-ilo+1]; flat:
  // both parts' entries (at most C + 2 together) as one range, part 0's first
  const int ne0 = ihi[0] >= ilo[0] ? ihi[0] - ilo[0] + 2 : 0;
  const int ne1 = ihi[1] >= ilo[1] ? ihi[1] - ilo[1] + 2 : 0;
  const int net = ne0 + ne1;
  int pa[FP][FI], pb[FP][FI];
  if constexpr (FLAT) {
    if (net > 0) {
#pragma unroll
      for (int it = 0; it < FI; ++it) {
        const int x = min(lane + it * kWave, net - 1);
        const bool p1 = x >= ne0;
        const int p = (p1 ? ilo[1] - 1 - ne0 : ilo[0] - 1) + x;
        pa[0][it] = ldg(s.xy + p);
        pb[0][it] = ldg(s.xy + p + (p1 ? mm[1] : mm[0]) + 1);
      }
    }
  } else {
#pragma unroll
    for (int part = 0; part < FP; ++part) {
      const int ne = part ? ne1 : ne0;
      if (ne >= 2) {
#pragma unroll
        for (int it = 0; it < FI; ++it) {
          const int p = ilo[part] - 1 + min(lane + it * kWave, ne - 1);
          pa[part][it] = ldg(s.xy + p);
          pb[part][it] = ldg(s.xy + p + mm[part] + 1);
        }
      }
    }
  }
  // Tours whose 2-opt deltas may leave int32: the wide tiles keep the
  // diagonal tiling (shared distances, column tiles) with 64-bit prefix
  // arithmetic; the small ones, whose register budget sets their occupancy,
  // take one wave per row over the whole row.
  if (!p32 && C < 1024) {
    sweep_2opt_wide(s, rb, ct, parity);
    return;
  }
  if (ct == 0)
    reset_other(s, 1, par, rb * MLP_2OPT_ROWS, min(rb * MLP_2OPT_ROWS + MLP_2OPT_ROWS, rows), rb == 0);
  if constexpr (FLAT) {
    if (net > 0) {
#pragma unroll
      for (int it = 0; it < FI; ++it) {
        const int x = lane + it * kWave;
        if (x < net) {
          const bool p1 = x >= ne0;
          L.e[wave][p1 ? 1 : 0][p1 ? x - ne0 : x] = (uint16_t)dist_xy(pa[0][it], pb[0][it], dr);
        }
      }
    }
  } else {
#pragma unroll
    for (int part = 0; part < FP; ++part) {
      const int ne = part ? ne1 : ne0;
      uint16_t* e = L.e[wave][part];
      if (ne >= 2) {
#pragma unroll
        for (int it = 0; it < FI; ++it) {
          const int x = lane + it * kWave;
          if (x < ne) e[x] = (uint16_t)dist_xy(pa[part][it], pb[part][it], dr);
        }
      }
    }
  }
  __syncthreads();
  // A tour above the int32 bound can still have rows whose moves all stay
  // below 2^31: reversing positions i..j moves each reversed node's arrival
  // by at most (T[j] - T[i]) + maxd and every later node's by at most
  // 2 maxd, so |delta| <= (m + 1)(T[j] - T[i] + maxd) + 2 (N - 1 - j) maxd
  // for j = i + m; such a row takes the int32 path (the bound is uniform
  // over the wave, as the wave owns the row).
  bool w32 = p32;
  if (!p32) {
    const int64_t md = s.maxd;
    bool ok = true;
#pragma unroll
    for (int part = 0; part < 2; ++part)
      if (ihi[part] >= ilo[part]) {
        const int64_t span = ldg(s.t64 + ihi[part] + mm[part]) - ldg(s.t64 + ilo[part]);
        const int64_t b = (int64_t)(mm[part] + 1) * (span + md) +
                          2 * (int64_t)(N - 1 - ilo[part] - mm[part]) * md;
        ok = ok && b < 2147483647LL;
      }
    w32 = __builtin_amdgcn_readfirstlane((int)ok) != 0;
  }
  if (!w32) {
    // delta = (N-i)(e[x] - w[i]) + (N-j-1)(e[x+1] - w[j+1]) + 2(Q[j] - Q[i])
    //         - (i+j+1)(T[j] - T[i])   (the ADS form of mlk2opt.cu, 64-bit)
    // each lane keeps the smallest key delta * 2^20 + ctx of its moves
    int64_t best = MLP_KEY_NONE;
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      const uint16_t* e = L.e[wave][part];
      const int m = mm[part], lo = ilo[part], hi = ihi[part];
      for (int i0 = lo; i0 <= hi; i0 += CU * kWave) {
        int32_t wi[CU], wj[CU];
        int64_t qi[CU], qj[CU], ti[CU], tj[CU];
#pragma unroll
        for (int u = 0; u < CU; ++u) {
          const int i = min(i0 + lane + u * kWave, hi);
          const int j = i + m;
          wi[u] = ldg(s.w + i);
          wj[u] = ldg(s.w + j + 1);
          qi[u] = ldg(s.q64 + i);
          qj[u] = ldg(s.q64 + j);
          ti[u] = ldg(s.t64 + i);
          tj[u] = ldg(s.t64 + j);
        }
#pragma unroll
        for (int u = 0; u < CU; ++u) {
          const int iu = i0 + lane + u * kWave;
          const int i = min(iu, hi);
          const int j = i + m;
          const int x = i - lo;
          const int t1 = __mul24(N - i, (int)e[x] - wi[u]);
          const int t2 = __mul24(N - j - 1, (int)e[x + 1] - wj[u]);
          const int64_t d = (int64_t)t1 + t2 + 2 * (qj[u] - qi[u]) - (int64_t)(i + j + 1) * (tj[u] - ti[u]);
          const int ctx = part == 0 ? by - i : i + by - 1;
          const int64_t key = d * (int64_t)(1 << MLP_KEY_SHIFT) + ctx;
          best = (iu <= hi && key < best) ? key : best;
        }
      }
    }
    best = wave_min64(best);
    if (live && lane == 0) fold_key(row_buf(s, 1, par) + by, best);
    return;
  }
  // The prefix arrays are read from global memory in batches of CU x 64
  // positions; the batches of both parts form one sequence and batch k + 1
  // is loaded while batch k is evaluated, so each batch's memory latency
  // hides behind the previous batch's arithmetic.
  int32_t bd = INT32_MAX, bc = -1;
  const int nb0 = ihi[0] >= ilo[0] ? (ihi[0] - ilo[0]) / (CU * kWave) + 1 : 0;
  const int nb1 = ihi[1] >= ilo[1] ? (ihi[1] - ilo[1]) / (CU * kWave) + 1 : 0;
  uint32_t gi[CU], hj[CU], ti[CU], tj[CU];
  // raw buffer loads: one scalar descriptor per array and a 32-bit byte
  // offset per lane (no 64-bit address arithmetic); reads beyond 4N bytes
  // would return 0, none is issued (indices are clamped to the row)
  const int nbytes = 4 * N;
  const __amdgpu_buffer_rsrc_t rG = __builtin_amdgcn_make_buffer_rsrc((void*)s.g32, 0, nbytes, kRsrcWord3);
  const __amdgpu_buffer_rsrc_t rH = __builtin_amdgcn_make_buffer_rsrc((void*)s.h32, 0, nbytes, kRsrcWord3);
  const __amdgpu_buffer_rsrc_t rT = __builtin_amdgcn_make_buffer_rsrc((void*)s.t32, 0, nbytes, kRsrcWord3);
  auto load_batch = [&](int k, uint32_t* g, uint32_t* h, uint32_t* a, uint32_t* b) {
    const int part = k < nb0 ? 0 : 1;
    const int lo = ilo[part], hi = ihi[part], m = mm[part];
    const int i0 = lo + (k - (part ? nb0 : 0)) * CU * kWave;
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const uint32_t i = (uint32_t)min(i0 + lane + u * kWave, hi);
      const uint32_t jm = i + (uint32_t)m;
      g[u] = __builtin_amdgcn_raw_buffer_load_b32(rG, 4u * i, 0, 0);
      h[u] = __builtin_amdgcn_raw_buffer_load_b32(rH, 4u * jm, 0, 0);
      a[u] = __builtin_amdgcn_raw_buffer_load_b32(rT, 4u * i, 0, 0);
      b[u] = __builtin_amdgcn_raw_buffer_load_b32(rT, 4u * jm, 0, 0);
    }
  };
  // A lane meets its indices in increasing i: ctx = by - i falls in part
  // 0 and ctx = i + by - 1 rises in part 1 (above every part-0 ctx), so
  // the lowest ctx among equal costs is the last equal one in part 0
  // (<=) and the first in part 1 (<).
  // (the part is a template constant: as a run-time value the tie-break's
  // two compares, the ctx select and the batch offset were computed per lane)
  auto eval_part = [&](auto pc, int k, const uint32_t* g, const uint32_t* h, const uint32_t* a,
                       const uint32_t* b) {
    constexpr int part = decltype(pc)::value ? 1 : 0;
    const uint16_t* e = L.e[wave][part];
    const int m = mm[part], lo = ilo[part], hi = ihi[part];
    const int i0 = lo + (k - (part ? nb0 : 0)) * CU * kWave;
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const int i = i0 + lane + u * kWave;
      const int ic = min(i, hi);
      const int j = ic + m;
      const int x = ic - lo;
      uint32_t d = g[u] + h[u];
      d += umul((uint32_t)(N - ic), e[x]) + umul((uint32_t)(N - j - 1), e[x + 1]);
      d += (uint32_t)m * (a[u] + b[u]);
      const int ctx = part == 0 ? by - ic : ic + by - 1;
      const bool take = i <= hi && (part == 0 ? (int32_t)d <= bd : (int32_t)d < bd);
      bd = take ? (int32_t)d : bd;
      bc = take ? ctx : bc;
    }
  };
  auto eval_batch = [&](int k, const uint32_t* g, const uint32_t* h, const uint32_t* a,
                        const uint32_t* b) {
    if (k < nb0) eval_part(Flag<false>{}, k, g, h, a, b);
    else eval_part(Flag<true>{}, k, g, h, a, b);
  };
  // two register sets, alternating: batch k + 1 loads while batch k runs
  const int nbat = nb0 + nb1;
  uint32_t gj[CU], hk[CU], tk[CU], tl[CU];
  if (nbat > 0) load_batch(0, gi, hj, ti, tj);
  for (int k = 0; k < nbat; k += 2) {
    if (k + 1 < nbat) load_batch(k + 1, gj, hk, tk, tl);
    eval_batch(k, gi, hj, ti, tj);
    if (k + 1 >= nbat) break;
    if (k + 2 < nbat) load_batch(k + 2, gi, hj, ti, tj);
    eval_batch(k + 1, gj, hk, tk, tl);
  }
  const int64_t key = wave_argmin(bd, bc);
  if (live && lane == 0) fold_key(row_buf(s, 1, par) + by, key);
}

__device__ __forceinline__ SweepSeg find_sweep(const SweepSeg* segs, int nsegs, int item) {
  int lo = 0, hi = nsegs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].blk0 <= item) lo = mid;
    else hi = mid - 1;
  }
  return segs[lo];
}

__device__ __forceinline__ SweepSeg find_sweep_inline(const SweepSegsInline& inl, int nsegs,
                                                      int item) {
  return find_inline<SweepSeg>(inl, nsegs, item);
}

// A launch whose slots all have the same segments (flags bits 8..31: the
// items per slot; bit 2: two segments per slot, the first starting at item
// 0): the segment of an item by division instead of a binary search over
// device memory.
__device__ __forceinline__ SweepSeg find_sweep_dev(const SweepSeg* segs, int nsegs, int item,
                                                   int flags) {
  const int group = (int)((uint32_t)flags >> 8);
  if (group == 0) return find_sweep(segs, nsegs, item);
  const int g = item / group;
  if (!(flags & 4)) return segs[g];
  return segs[2 * g + (item - g * group >= segs[1].blk0 ? 1 : 0)];
}

template <int R, int C, int X>
__device__ __forceinline__ void sweep_item(const SlotDev* __restrict__ table, const SweepSeg& sg, int item,
                           SweepLds<R, C, X>& L, bool dvnd, bool filtered) {
  const SlotDev& s = table[sg.slot];
  if (dvnd && s.info[MLP_INFO_ACTIVE] == 0) return;
  // bit nb of par: the buffer this sweep writes (the one not holding the last result)
  int par = sg.parity;
  if (dvnd) {
    par = 0;
#pragma unroll
    for (int nb = 0; nb < MLP_NB_COUNT; ++nb)
      if ((sg.mask >> nb) & 1) par |= (1 - (int)(s.info[MLP_INFO_PARITY0 + nb] & 1)) << nb;
  }
  const int b = item - sg.blk0;
  const int rb = b / sg.ncols, ct = b % sg.ncols;
  if (!dvnd && b == 0 && threadIdx.x == 0) {
#pragma unroll
    for (int nb = 0; nb < MLP_NB_COUNT; ++nb)
      if ((sg.mask >> nb) & 1) s.info[MLP_INFO_PARITY0 + nb] = (par >> nb) & 1;
  }
  if (sg.kind == 0) sweep_fixed<R, C, X>(s, rb, ct, sg.mask, par, L.f, filtered);
  else sweep_2opt<C>(s, rb, ct, par, L.d);
}

}  // namespace

// Resident waves per SIMD the register budget is held to (at most 72
// VGPRs for seven): the small shapes serve one-tour launches, whose latency
// depends on every tile being resident at once, and many-tour launches,
// which gain from the occupancy.  The large shapes are held to three
// workgroups per CU by their LDS tiles anyway.
template <int R, int C, int X>
__global__ __launch_bounds__(kBlock, (R * C <= 2048 ? 7 : 1)) void k_sweep(
    const SlotDev* __restrict__ table,
                                                  SweepSegsInline inl, const SweepSeg* segs_dev,
                                                  int nsegs, int total, int* counter,
                                                  int flags) {
  // flags bit 0: a DVND launch (inactive slots skipped, parities read from
  // info); bit 1: filtered folds of the diagonal rows (many more tiles than
  // resident workgroups, so most tiles fold after others of their tour);
  // bits 2, 8..31: equal segments per slot (find_sweep_dev)
  __shared__ SweepLds<R, C, X> L;
  const bool dvnd = (flags & 1) != 0, filtered = (flags & 2) != 0;
  if (!counter) {
    const int item = blockIdx.x;
    const SweepSeg sg = segs_dev ? find_sweep_dev(segs_dev, nsegs, item, flags)
                                  : find_sweep_inline(inl, nsegs, item);
    sweep_item<R, C, X>(table, sg, item, L, dvnd, filtered);
    return;
  }
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= total) break;
    const SweepSeg sg = segs_dev ? find_sweep_dev(segs_dev, nsegs, item, flags)
                                  : find_sweep_inline(inl, nsegs, item);
    sweep_item<R, C, X>(table, sg, item, L, dvnd, filtered);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// DVND rounds as one persistent launch (FusedPlan): each workgroup takes the
// next item of the queue, waits until that item's own dependencies are done
// (its slot's previous merge; for a merge, all tiles of its eval phase), runs
// it, and publishes its completion.  Slots advance independently inside the
// launch, so short tours do not wait for the longest tour of the batch at
// every neighborhood.  Waits poll with a sleep and give up after about
// 2^22 polls (the flag sync[2] is then set and reported by the host), so the
// grid always exits.
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ int sync_load(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread 0 waits until *p >= target, then acquires at agent scope; the
// barrier orders every thread's later reads after that acquire.  One thread
// fences for the workgroup: an agent-scope fence is a cache invalidation (and
// a release a cache write-back) of the whole XCD, and the workgroup's waves
// share one CU's caches.
__device__ __forceinline__ void wait_at_least(const int32_t* p, int target, int32_t* timeout) {
  if (threadIdx.x == 0) {
    int polls = 0;
    while (sync_load(p) < target) {
      __builtin_amdgcn_s_sleep(8);
      if (++polls > (1 << 22)) {
        atomicExch(timeout, 1);
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __builtin_amdgcn_s_waitcnt(0);  // the invalidation has completed
  }
  __syncthreads();
}

// Every wave waits for its own stores and atomics to complete, the barrier
// gathers the workgroup, and thread 0 releases at agent scope and counts the
// item.
__device__ __forceinline__ void signal_done(int32_t* p) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_s_waitcnt(0);  // the write-back has completed
    __hip_atomic_fetch_add(p, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace

template <int R, int C>
__global__ __launch_bounds__(kBlock) void k_dvnd_fused(const SlotDev* __restrict__ table,
                                                       FusedPlan P) {
  extern __shared__ int64_t dyn[];
  __shared__ MergeShared MS;
  __shared__ int s_item;
  SweepLds<R, C, kPlain>& L = *reinterpret_cast<SweepLds<R, C, kPlain>*>(dyn);
  int32_t* work = P.sync;
  int32_t* exits = P.sync + 1;
  int32_t* timeout = P.sync + 2;
  int32_t* done = P.sync + 4;
  int32_t* merged = done + P.count;
  const int total = P.per_round * P.rounds;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(work, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= total) break;
    const int rr = item / P.per_round, x = item - rr * P.per_round;
    int ph = 0;
    while (ph < 9 && x >= P.phase[ph + 1]) ++ph;
    const int nb = ph >> 1;
    if (ph & 1) {  // merge + apply of slot q
      const int q = x - P.phase[ph];
      int all = 0, upto = 0;
      for (int b = 0; b < MLP_NB_COUNT; ++b) {
        const int t = P.prefix[b * (P.count + 1) + q + 1] - P.prefix[b * (P.count + 1) + q];
        all += t;
        if (b <= nb) upto += t;
      }
      wait_at_least(merged + q, rr * MLP_NB_COUNT + nb, timeout);
      wait_at_least(done + q, rr * all + upto, timeout);
      merge_slot(table[P.slots[q]], nb, 0, P.pcap, P.ncap, 1,
                        nb == MLP_NB_COUNT - 1 ? P.tail : 0, dyn, MS);
      signal_done(merged + q);
    } else {  // up to P.group tiles of the evaluation of nb for slot q
      const int e = x - P.phase[ph];
      const int32_t* pre = P.prefix + nb * (P.count + 1);
      int lo = 0, hi = P.count - 1;
      while (lo < hi) {  // the slot whose tiles hold tile e
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= e) lo = mid;
        else hi = mid - 1;
      }
      const int q = lo;
      wait_at_least(merged + q, rr * MLP_NB_COUNT + nb, timeout);
      SweepSeg sg;
      sg.slot = P.slots[q];
      sg.kind = nb == 1 ? 1 : 0;
      sg.mask = 1 << nb;
      sg.blk0 = 0;
      sg.ncols = P.ncols[2 * q + (nb == 1 ? 1 : 0)];
      sg.parity = 0;
      const int g = e - pre[q];
      const int t1 = min((g + 1) * P.group, P.tiles[nb * P.count + q]);
      for (int t = g * P.group; t < t1; ++t) {
        sweep_item<R, C, kPlain>(table, sg, t, L, true, true);
        __syncthreads();  // the next tile reuses the LDS
      }
      signal_done(done + q);
    }
    __syncthreads();
  }
  // the last workgroup out clears the counters for the next launch
  if (threadIdx.x == 0) s_item = atomicAdd(exits, 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (s_item) {
    __threadfence();
    for (int q = threadIdx.x; q < 2 * P.count; q += kBlock) done[q] = 0;
    if (threadIdx.x == 0) {
      *work = 0;
      *exits = 0;
    }
  }
}

template __global__ void k_dvnd_fused<MLP_FUSED_R, MLP_FUSED_C>(const SlotDev*, FusedPlan);

#define MLP_SWEEP_INST(R, C, X)                                                       \
  template __global__ void k_sweep<R, C, X>(const SlotDev*, SweepSegsInline, const SweepSeg*,   \
                                            int, int, int*, int);
MLP_SWEEP_INST(4, 256, kPlain)
MLP_SWEEP_INST(8, 512, kPlain)
MLP_SWEEP_INST(16, 1024, kPlain)
MLP_SWEEP_INST(8, 256, kPlain)
MLP_SWEEP_INST(4, 512, kPlain)
// shapes of eight or more rows also come with 2-opt evaluated in the fixed tiles
MLP_SWEEP_INST(8, 512, kTwo)
MLP_SWEEP_INST(12, 1024, kTwo)
MLP_SWEEP_INST(8, 256, kTwo)
// the ODC row map: 4 x 256, 8 x 512 (also for 16 x 1024), 4 x 512, and
// 16 x 128 for 8 x 256 (its or-opt diagonals cross half as many tiles)
MLP_SWEEP_INST(4, 256, kOdc)
MLP_SWEEP_INST(8, 512, kOdc)
MLP_SWEEP_INST(16, 128, kOdc)
MLP_SWEEP_INST(4, 512, kOdc)

// ===========================================================================
// host-side launch wrappers (called by the C ABI runtime in mlp_runtime.cpp)
// ===========================================================================
#include "mlp_launch.h"

hipError_t mlp_launch_prep(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                           const int32_t* list_dev, int count) {
  hipLaunchKernelGGL(k_prep, dim3(count), dim3(kBlock), 0, st, table, inl, list_dev);
  return hipGetLastError();
}

hipError_t mlp_launch_dvnd_start(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                                 const int32_t* list_dev, int count) {
  hipLaunchKernelGGL(k_dvnd_start, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                     table, inl, list_dev, count);
  return hipGetLastError();
}

hipError_t mlp_launch_round_begin(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                                  const int32_t* list_dev, int count, int* counters,
                                  int ncounters) {
  const int n = std::max(count, ncounters);
  hipLaunchKernelGGL(k_round_begin, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, table,
                     inl, list_dev, count, counters, ncounters);
  return hipGetLastError();
}

hipError_t mlp_launch_round_end(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                                const int32_t* list_dev, int count, int max_rounds) {
  hipLaunchKernelGGL(k_round_end, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                     table, inl, list_dev, count, max_rounds);
  return hipGetLastError();
}

hipError_t mlp_launch_eval(hipStream_t st, const SlotDev* table, const EvalSegsInline& inl,
                           const EvalSeg* segs_dev, int nsegs, int total, int rowmap,
                           int* counter, int grid, int skip_inactive) {
  hipLaunchKernelGGL(k_eval, dim3(grid), dim3(kBlock), 0, st, table, inl, segs_dev, nsegs, total,
                     rowmap, counter, skip_inactive);
  return hipGetLastError();
}

hipError_t mlp_launch_sweep(hipStream_t st, int tile, const SlotDev* table,
                            const SweepSegsInline& inl, const SweepSeg* segs_dev, int nsegs,
                            int total, int* counter, int grid, int flags) {
  // MLP_TILE_TWO: fixed-row tiles with 2-opt (mask bit 1), the shapes' TWO
  // variants; MLP_TILE_ODC: the ODC row map (16 x 1024 runs as 8 x 512,
  // 8 x 256 as 16 x 128)
#define MLP_SWEEP_LAUNCH(R, C, X)                                                              \
  hipLaunchKernelGGL((k_sweep<R, C, X>), dim3(grid), dim3(kBlock), 0, st, table, inl,          \
                     segs_dev, nsegs, total, counter, flags)
  const bool two = (tile & MLP_TILE_TWO) != 0, odc = (tile & MLP_TILE_ODC) != 0;
  switch (tile & 0xff) {
    case MLP_TILE_4x256:
      if (odc) MLP_SWEEP_LAUNCH(4, 256, kOdc);
      else MLP_SWEEP_LAUNCH(4, 256, kPlain);
      break;
    case MLP_TILE_8x512:
      if (two) MLP_SWEEP_LAUNCH(8, 512, kTwo);
      else if (odc) MLP_SWEEP_LAUNCH(8, 512, kOdc);
      else MLP_SWEEP_LAUNCH(8, 512, kPlain);
      break;
    case MLP_TILE_8x256:
      if (two) MLP_SWEEP_LAUNCH(8, 256, kTwo);
      else if (odc) MLP_SWEEP_LAUNCH(16, 128, kOdc);
      else MLP_SWEEP_LAUNCH(8, 256, kPlain);
      break;
    case MLP_TILE_4x512:
      if (odc) MLP_SWEEP_LAUNCH(4, 512, kOdc);
      else MLP_SWEEP_LAUNCH(4, 512, kPlain);
      break;
    default:
      if (two) MLP_SWEEP_LAUNCH(12, 1024, kTwo);
      else if (odc) MLP_SWEEP_LAUNCH(8, 512, kOdc);
      else MLP_SWEEP_LAUNCH(16, 1024, kPlain);
      break;
  }
  return hipGetLastError();
}

void mlp_tile_shape(int tile, int* rows, int* cols) {
  static const int kRows[MLP_TILE_COUNT] = {4, 8, 16, 8, 4};
  static const int kCols[MLP_TILE_COUNT] = {256, 512, 1024, 256, 512};
  const int t = tile & 0xff;
  *rows = kRows[t];
  *cols = kCols[t];
  // with 2-opt in the tile the 16 x 1024 shape runs 12 rows (its LDS then
  // still fits three workgroups per CU)
  if ((tile & MLP_TILE_TWO) && t == MLP_TILE_16x1024) *rows = 12;
  if ((tile & MLP_TILE_ODC) && t == MLP_TILE_8x256) {
    *rows = 16;
    *cols = 128;
  }
  if ((tile & MLP_TILE_ODC) && t == MLP_TILE_16x1024) {
    *rows = 8;
    *cols = 512;
  }
}

// keys[pcap] int64 | told[ncap] | tnew[ncap] | bits | ctx[pcap] uint16
size_t mlp_merge_lds_bytes(int pcap, int ncap) {
  return (size_t)pcap * 8 + (size_t)ncap * 12 + (size_t)claim_words(ncap) * 4 + (size_t)pcap * 2;
}

// Dynamic LDS of a merge whose working arrays are in global memory: the
// sort buffer and the conflict bitmap of the largest N.
constexpr size_t kGmemMergeLds = kGmemSortCap * 8 + claim_words(MLP_MAX_N) * 4;

hipError_t mlp_merge_setup(size_t max_dyn_lds) {
  hipError_t e = hipFuncSetAttribute((const void*)k_merge_apply,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_dyn_lds);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_mg_select, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kGmemMergeLds);
  if (e != hipSuccess) (void)hipGetLastError();  // do not leave it for the next launch check
  return e;
}

size_t mlp_fused_lds_bytes(int pcap, int ncap) {
  return std::max(sizeof(SweepLds<MLP_FUSED_R, MLP_FUSED_C, kPlain>), mlp_merge_lds_bytes(pcap, ncap));
}

hipError_t mlp_fused_setup(size_t max_dyn_lds) {
  hipError_t e = hipFuncSetAttribute((const void*)k_dvnd_fused<MLP_FUSED_R, MLP_FUSED_C>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_dyn_lds);
  if (e != hipSuccess) (void)hipGetLastError();
  return e;
}

hipError_t mlp_launch_fused(hipStream_t st, const SlotDev* table, const FusedPlan& plan, int grid) {
  hipLaunchKernelGGL((k_dvnd_fused<MLP_FUSED_R, MLP_FUSED_C>), dim3(grid), dim3(kBlock),
                     mlp_fused_lds_bytes(plan.pcap, plan.ncap), st, table, plan);
  return hipGetLastError();
}

hipError_t mlp_launch_merge_apply(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                                  const int32_t* list_dev, int count, int nb, int rowmap,
                                  int pcap, int ncap, int skip_inactive, int tail) {
  hipLaunchKernelGGL(k_merge_apply, dim3(count), dim3(kBlock), mlp_merge_lds_bytes(pcap, ncap), st,
                     table, inl, list_dev, nb, rowmap, pcap, ncap, skip_inactive, tail);
  return hipGetLastError();
}

hipError_t mlp_launch_merge_gmem(hipStream_t st, const SlotDev* table, const SlotListInline& inl,
                                 const int32_t* list_dev, int count, int nb, int rowmap,
                                 int skip_inactive, int tail, int max_n) {
  const dim3 seg((max_n + MLP_MG_CHUNK - 1) / MLP_MG_CHUNK, count);
  hipLaunchKernelGGL(k_mg_compact, seg, dim3(kBlock), 0, st, table, inl, list_dev, nb, rowmap,
                     skip_inactive);
  hipLaunchKernelGGL(k_mg_select, dim3(count), dim3(kBlock), kGmemMergeLds, st, table, inl,
                     list_dev, nb, rowmap, skip_inactive, tail);
  hipLaunchKernelGGL(k_mg_sums, seg, dim3(kBlock), 0, st, table, inl, list_dev);
  hipLaunchKernelGGL(k_mg_rebuild, seg, dim3(kBlock), 0, st, table, inl, list_dev);
  return hipGetLastError();
}

size_t mlp_total_merge_lds_bytes(int N) {
  return (size_t)claim_words(N) * 4 + (size_t)N * 8;
}

// The merge kernel's own static LDS (chunk arrays, prefix scratch) is below
// kTotalStaticLds; its dynamic part may use the rest.
constexpr size_t kTotalStaticLds = 48 * 1024;
static_assert(sizeof(TotalLds) + sizeof(PrepLds) <= kTotalStaticLds, "static LDS estimate");

size_t mlp_total_static_lds() { return kTotalStaticLds; }

hipError_t mlp_total_setup(size_t max_dyn_lds) {
  hipError_t e = hipFuncSetAttribute((const void*)k_total_eval,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_dyn_lds);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_total_merge<false>,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(max_dyn_lds - kTotalStaticLds));
  if (e != hipSuccess) (void)hipGetLastError();  // do not leave it for the next launch check
  return e;
}

hipError_t mlp_launch_total_eval(hipStream_t st, const SlotDev* table, int id, int nb, int rows,
                                 int B, int64_t* out) {
  hipLaunchKernelGGL(k_total_eval, dim3(rows), dim3(kBlock), (size_t)B * 8, st, table, id, nb, B,
                     out);
  return hipGetLastError();
}

hipError_t mlp_launch_total_keys(hipStream_t st, const int64_t* tk, int64_t M, int B,
                                 uint64_t bias, uint64_t* ku, int32_t* ve) {
  const int64_t blocks = std::min<int64_t>((M + kBlock - 1) / kBlock, 4096);
  hipLaunchKernelGGL(k_total_keys, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(kBlock), 0,
                     st, tk, M, B, bias, ku, ve);
  return hipGetLastError();
}

hipError_t mlp_launch_total_merge(hipStream_t st, const SlotDev* table, int id, int nb, int N,
                                  int B, int64_t M, uint64_t bias, const uint64_t* ku,
                                  const int32_t* ve, int gmem) {
  if (gmem)
    hipLaunchKernelGGL(k_total_merge<true>, dim3(1), dim3(kBlock), (size_t)claim_words(N) * 4,
                       st, table, id, nb, B, M, bias, ku, ve);
  else
    hipLaunchKernelGGL(k_total_merge<false>, dim3(1), dim3(kBlock), mlp_total_merge_lds_bytes(N),
                       st, table, id, nb, B, M, bias, ku, ve);
  return hipGetLastError();
}

hipError_t mlp_launch_ads(hipStream_t st, const SlotDev* table, int id, int N, uint32_t* S,
                          uint32_t* out) {
  hipLaunchKernelGGL(k_ads_prefix, dim3(1), dim3(64), 0, st, table, id, S);
  const int64_t total = (int64_t)N * N;
  int grid = (int)std::min<int64_t>((total + kBlock - 1) / kBlock, 65536);
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_ads_fill, dim3(grid), dim3(kBlock), 0, st, table, id, S, out);
  return hipGetLastError();
}

hipError_t mlp_launch_rows_moves64(hipStream_t st, const SlotDev* table, int id, int nb,
                                  int rowmap, int R, uint64_t* out) {
  const int grid = std::max(1, std::min((R + kBlock - 1) / kBlock, 1024));
  hipLaunchKernelGGL(k_rows_moves64, dim3(grid), dim3(kBlock), 0, st, table, id, nb, rowmap, R,
                     out);
  return hipGetLastError();
}

hipError_t mlp_launch_rows_out(hipStream_t st, const SlotDev* table, int id, int mask, int rowmap,
                              int total, int64_t* cost, int32_t* i, int32_t* j) {
  const int grid = std::max(1, std::min((total + kBlock - 1) / kBlock, 1024));
  hipLaunchKernelGGL(k_rows_out, dim3(grid), dim3(kBlock), 0, st, table, id, mask, rowmap, total,
                     cost, i, j);
  return hipGetLastError();
}

hipError_t mlp_launch_total_moves64(hipStream_t st, const int64_t* tk, int64_t M, int B, int nb,
                                   int N, uint64_t* out) {
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((M + kBlock - 1) / kBlock, 8192));
  hipLaunchKernelGGL(k_total_moves64, dim3(grid), dim3(kBlock), 0, st, tk, M, B, nb, N, out);
  return hipGetLastError();
}

hipError_t mlp_launch_dist_table(hipStream_t st, int s0, int count, int round, int32_t* out) {
  hipLaunchKernelGGL(k_dist_table, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s0,
                     count, round, out);
  return hipGetLastError();
}

hipError_t mlp_launch_dist_list(hipStream_t st, const int32_t* s, int count, int round,
                                int32_t* out) {
  hipLaunchKernelGGL(k_dist_list, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s,
                     count, round, out);
  return hipGetLastError();
}

hipError_t mlp_launch_sqrt_selfcheck(hipStream_t st, uint32_t s0, uint32_t count,
                                     unsigned long long* bad) {
  hipLaunchKernelGGL(k_sqrt_selfcheck, dim3(4096), dim3(kBlock), 0, st, s0, count, bad);
  return hipGetLastError();
}

hipError_t mlp_launch_weights(hipStream_t st, const int32_t* xs, const int32_t* ys, int n,
                              int round, int32_t* out) {
  const int64_t total = (int64_t)n * n;
  hipLaunchKernelGGL(k_weights, dim3((unsigned)((total + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     st, xs, ys, n, round, out);
  return hipGetLastError();
}

hipError_t mlp_launch_clear_rows(hipStream_t st, const SlotDev* table, const SweepSegsInline& inl,
                                 const SweepSeg* segs_dev, int nsegs) {
  hipLaunchKernelGGL(k_clear_rows, dim3(nsegs), dim3(kBlock), 0, st, table, inl, segs_dev, nsegs);
  return hipGetLastError();
}

hipError_t mlp_launch_probe(hipStream_t st, const SlotDev* table, int slot, int kind, int grid,
                            int64_t* out) {
  SweepSegsInline inl{};
  inl.s[0].slot = slot;
  hipLaunchKernelGGL(k_probe, dim3(grid), dim3(kBlock), 0, st, table, inl, kind, out);
  return hipGetLastError();
}

hipError_t mlp_launch_fill_i64(hipStream_t st, int64_t* p, int64_t count, int64_t v) {
  const int64_t blocks = std::min<int64_t>((count + kBlock - 1) / kBlock, 1024);
  hipLaunchKernelGGL(k_fill_i64, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(kBlock), 0, st,
                     p, count, v);
  return hipGetLastError();
}

hipError_t mlp_launch_set_slot(hipStream_t st, SlotDev* table, int id, const SlotDev& d) {
  hipLaunchKernelGGL(k_set_slot, dim3(1), dim3(64), 0, st, table, id, d);
  return hipGetLastError();
}

hipError_t mlp_launch_slot_init(hipStream_t st, SlotDev* table, int id, const SlotDev& d,
                                int64_t* mem, int64_t words, const SlotInitRows& rows) {
  const int64_t blocks = std::min<int64_t>((words + kBlock - 1) / kBlock, 1024);
  hipLaunchKernelGGL(k_slot_init, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(kBlock), 0, st,
                     table, id, d, mem, words, rows);
  return hipGetLastError();
}

hipError_t mlp_launch_gather_tours(hipStream_t st, const SlotDev* table, const int32_t* ids,
                                   int count, int32_t* out) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_tours, dim3(count), dim3(kBlock), 0, st, table, ids, count, out);
  return hipGetLastError();
}

hipError_t mlp_launch_scatter_tours(hipStream_t st, const SlotDev* table, const int32_t* ids,
                                    int count, const int32_t* in) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_tours, dim3(count), dim3(kBlock), 0, st, table, ids, count, in);
  return hipGetLastError();
}

hipError_t mlp_launch_gather_info(hipStream_t st, const SlotDev* table, const int32_t* ids,
                                  int count, int word, int64_t* out) {
  hipLaunchKernelGGL(k_gather_info, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                     table, ids, count, word, out);
  return hipGetLastError();
}

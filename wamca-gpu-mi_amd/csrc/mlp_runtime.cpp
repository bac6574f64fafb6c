// mlp_runtime.cpp -- host runtime behind the C ABI in include/mlp_gpu.h.
//
// Owns the HIP stream, the instance and slot tables, the device copy of the
// slot table read by the kernels, and the caches of device-side work lists.
// Every entry point returns a status and records a message; nothing throws
// across the ABI and nothing calls exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <map>
#include <mutex>
#include <vector>

#include "mlp_gpu.h"
#include "mlp_internal.h"
#include "mlp_launch.h"

namespace {

struct Instance {
  bool alive = false;
  int n = 0, round = 0, maxd = 0;
  int32_t* xs = nullptr;  // device, in one pool block with ys
  int32_t* ys = nullptr;
  size_t cls = 0;         // pool class of the block
};

struct Slot {
  bool alive = false;
  int inst = -1;
  int N = 0;
  void* mem = nullptr;
  size_t cls = 0;  // pool class of mem
  SlotDev dev{};
  int nbs = MLP_NB_COUNT;  // neighborhoods with buffers
  int eval_rowmap[MLP_NB_MAX] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  int parity[MLP_NB_MAX] = {};  // row buffer of the last eval
  bool parity_stale = false;    // DVND rounds flipped parities on the device
  bool merged[MLP_NB_MAX] = {};
  // TOTAL mode: entries (delta << 20 | ctx) of the last mlp_eval_total
  int64_t* total = nullptr;
  size_t total_cls = 0;
  int64_t total_cap = 0;  // entries the block holds
  int total_nb = -1, total_block = 0, total_rows = 0;
  bool total_merged = false;
};

#define MLP_TUNE_BUCKETS 48

struct DevList {
  void* p = nullptr;
};

int rows_of(int nb, int rowmap, int N) {
  if (nb == 0) return rowmap == 0 ? N - 4 : ((N + 1) / 2) - 1;
  if (nb == 1) return (N + 1) / 2;
  return N - (nb - 1) - 2;
}

void move_ij(int nb, int rowmap, int N, int row, int ctx, int& i, int& j) {
  if (nb == 0) {
    if (rowmap == 0) { i = row + 1; j = ctx; }
    else if (ctx >= row) { i = row + 1; j = ctx + 3; }
    else { i = N - row - 3; j = N - ctx - 2; }
  } else if (nb == 1) {
    if (ctx >= row) { i = ctx - row + 1; j = ctx + 2; }
    else { i = row - ctx; j = N - ctx - 2; }
  } else {
    const int k = nb - 1;
    if (rowmap == 0) { i = ctx > row ? row + 1 : row + 2; j = ctx + 1; }
    else if (ctx >= row) { i = ctx - row + 1; j = ctx + 2; }
    else { i = N - ctx - k - 1; j = row - ctx; }
  }
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Entries of one row buffer: the most rows any neighborhood has in any row map.
int row_cap(int N) {
  int r = 1;
  for (int nb = 0; nb < 3; ++nb) r = std::max(r, std::max(rows_of(nb, 0, N), rows_of(nb, 1, N)));
  return r;
}

// The merge's working arrays for one slot of N positions (capacity of the
// sorted-key array: the power of two at or above the row capacity).
int merge_pcap(int N) {
  int p = 1;
  while (p < row_cap(N)) p <<= 1;
  return p;
}
size_t merge_bytes(int N) { return mlp_merge_lds_bytes(merge_pcap(N), N); }
// The merge's working arrays in the slot's device memory: that layout, then
// the many-workgroup merge's per-segment sums.
size_t merge_gmem_bytes(int N) {
  return merge_bytes(N) + 8 + 16 * (size_t)((N + MLP_MG_CHUNK - 1) / MLP_MG_CHUNK);
}

// Byte layout of one slot's device block.
struct SlotLayout {
  size_t tour, xy, w, t32, q32, g32, h32, t64, q64, info, rows[MLP_NB_MAX], kc[MLP_NB_MAX],
      kij[MLP_NB_MAX], ms, total;
};

SlotLayout slot_layout(int N, int nbs) {
  SlotLayout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = align_up(off + bytes, 256);
    return at;
  };
  L.tour = take(4 * (size_t)N);
  L.xy = take(4 * (size_t)N);
  L.w = take(4 * (size_t)N);
  L.t32 = take(4 * (size_t)N);
  L.q32 = take(4 * (size_t)N);
  L.g32 = take(4 * (size_t)N);
  L.h32 = take(4 * (size_t)N);
  L.t64 = take(8 * (size_t)N);
  L.q64 = take(8 * (size_t)N);
  L.info = take(8 * MLP_INFO_WORDS);
  for (int nb = 0; nb < nbs; ++nb) {
    const size_t r = (size_t)row_cap(N);
    L.rows[nb] = take(16 * r);
    L.kc[nb] = take(8 * r);
    L.kij[nb] = take(12 * r);
  }
  L.ms = take(merge_gmem_bytes(N));  // merge working arrays in global memory
  L.total = off;
  return L;
}

}  // namespace

// Working arrays of mlp_eval_static (one slot's ADS prefix arrays, info words
// and row buffers, behind its slot-table entry): a file-scope device array of
// a fixed size, so that entry point allocates nothing.  Every GPU holds its
// own copy until the process ends; its device address is looked up once per
// device (mlp_ctx_create) and only that address is handed to kernels.
#define MLP_STATIC_SCRATCH_BYTES (4 << 20)
static __device__ __align__(256) char scratch[MLP_STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_set[kMaxDevices] = {};

// The calling thread's current device must be `device`.
void* scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_set[device]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;  // retried by the next context on this device
    }
    g_scratch_ptr[device] = ptr;
    g_scratch_set[device] = true;
  }
  return g_scratch_ptr[device];
}
}  // namespace

struct mlp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // non-blocking, for uploads while `stream` is captured
  int cu_count = 0;
  int64_t hbm = 0;
  char name[256] = {0};
  std::string err;
  std::vector<Instance> insts;
  std::vector<Slot> slots;
  SlotDev* d_table = nullptr;
  int table_cap = 0;
  int* d_counter = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // timed region
  bool region_open = false;
  int32_t* d_tstage = nullptr;  // device array of several slots' tours (uploads)
  size_t d_tstage_cap = 0;
  int64_t* d_i64 = nullptr;  // small device scratch for gathers
  int64_t* h_i64 = nullptr;  // pinned
  size_t i64_cap = 0;
  std::unordered_map<std::string, DevList> lists;
  // pinned staging of list uploads: a bump area reused after the stream drains
  char* h_lists = nullptr;
  size_t h_lists_cap = 0, h_lists_off = 0;
  size_t max_merge_lds = 0;
  size_t max_total_lds = 0;  // dynamic LDS limit of the TOTAL-mode kernels
  int tile_override = -1;  // -1 auto, else an MLP_TILE_* shape
  bool merge_force_global = false;  // merge working arrays in global memory for every size
  int8_t tune[32][MLP_TUNE_BUCKETS];  // best shape per (mask, log2 moves), -1 untuned
  // device memory pool: released blocks stay on per-class free lists (no
  // hipFree, so no implicit device synchronisation, while jobs come and go)
  std::map<size_t, std::vector<void*>> pool_free;
  std::vector<void*> pool_all;
  size_t pool_reserved = 0, pool_live = 0;
  // DVND rounds: launches are enqueued `dvnd_chunk` rounds at a time between
  // convergence checks; with dvnd_graph the chunk is captured once per active
  // set and replayed as a HIP graph; dvnd_persistent drains work counters.
  int nbs = MLP_NB_COUNT;       // neighborhoods of slots created from now on
  int dvnd_nbs = MLP_NB_COUNT;  // neighborhoods of a DVND round (ids 0 .. dvnd_nbs-1)
  int dvnd_chunk = 4;
  int dvnd_graph = 1;
  int dvnd_persistent = 0;
  int dvnd_fused = 0;         // a round is one persistent launch (k_dvnd_fused)
  int32_t* d_sync = nullptr;  // its queue and dependency counters (zero between launches)
  int sync_cap = 0;           // slots the counters cover
  bool capturing = false;     // the stream is being captured into a graph
  bool user_capture = false;  // ... by mlp_graph_begin
  std::vector<hipGraphExec_t> user_graphs;
  // host bookkeeping a graph's replay repeats: (slot, merge?, mask or nb, row map)
  struct GraphOp {
    int32_t slot, merge;
    uint32_t mask;
    int32_t rowmap;
  };
  std::vector<std::vector<GraphOp>> user_graph_ops;
  std::vector<GraphOp> capture_ops;
  // cached DVND chunk graphs, one per slot group of a DVND call
  hipGraphExec_t graph_exec[MLP_DVND_GROUPS] = {};
  std::string graph_key[MLP_DVND_GROUPS];
  // DVND slot groups: the largest slots on the context stream, the others
  // on streams of their own, so one group's merges overlap another's sweeps
  int dvnd_groups = 2;
  hipStream_t stream_main = nullptr, stream_grp[MLP_DVND_GROUPS] = {};  // [0] unused
  hipEvent_t ev_fork = nullptr, ev_join[MLP_DVND_GROUPS] = {};
  hipEvent_t ev_chunk[MLP_DVND_GROUPS] = {};  // end of a group's chunk (mlp_dvnd)
  void* scratch = nullptr;  // device address of this GPU's copy of the static scratch array
  // caller-owned device buffers (mlp_dev_alloc) and event-pair timers
  std::vector<void*> dev_bufs;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> timers;
  // per-launch timing of k_eval (HIP events on the context stream)
  bool timing = false;
  std::vector<hipEvent_t> tev;  // pairs
  size_t tcount = 0;
};

namespace {

int fail(mlp_ctx* c, int st, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return st;
}

#define HCK(call)                                                                             \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(ctx, MLP_EDEVICE, "%s failed: %s", #call, hipGetErrorString(e_));          \
  } while (0)

// Calls that synchronise or copy to the host are refused inside a capture.
#define NO_CAPTURE(ctx)                                                                       \
  do {                                                                                        \
    if ((ctx) && (ctx)->capturing)                                                            \
      return fail((ctx), MLP_ESTATE, "not allowed while a graph capture is open");            \
  } while (0)

bool slot_ok(mlp_ctx* ctx, int s) {
  return s >= 0 && s < (int)ctx->slots.size() && ctx->slots[s].alive;
}

// The context stream and the DVND group streams drained.
hipError_t sync_streams(mlp_ctx* ctx) {
  hipError_t e = hipStreamSynchronize(ctx->stream_main);
  for (int g = 1; g < MLP_DVND_GROUPS && e == hipSuccess; ++g)
    if (ctx->stream_grp[g]) e = hipStreamSynchronize(ctx->stream_grp[g]);
  return e;
}

// Pinned staging space for an upload ordered on the context stream: a bump
// area, reused (after the streams drain) when it is full.
int stage_alloc(mlp_ctx* ctx, size_t bytes, char** h) {
  if (ctx->h_lists_off + bytes > ctx->h_lists_cap) {
    // every copy out of the staging area is done once the streams drain
    HCK(sync_streams(ctx));
    ctx->h_lists_off = 0;
    if (bytes > ctx->h_lists_cap) {
      if (ctx->h_lists) hipHostFree(ctx->h_lists);
      ctx->h_lists = nullptr;
      ctx->h_lists_cap = 0;
      size_t cap = (size_t)1 << 20;
      while (cap < bytes) cap *= 2;
      HCK(hipHostMalloc((void**)&ctx->h_lists, cap, hipHostMallocDefault));
      ctx->h_lists_cap = cap;
    }
  }
  *h = ctx->h_lists + ctx->h_lists_off;
  ctx->h_lists_off += (bytes + 255) & ~(size_t)255;
  return MLP_OK;
}

// Device copy of an int32 list, uploaded once per distinct content.  The
// upload is ordered on the context stream from pinned staging, so a new list
// does not wait for the kernels already queued (a pageable hipMemcpy would
// drain the stream first).  Inside a graph capture a missing list is uploaded
// on a side stream (the captured stream is not touched), so it exists before
// any replay.
int device_list(mlp_ctx* ctx, const int32_t* v, int count, const int32_t** out) {
  std::string key((const char*)v, (size_t)count * 4);
  auto it = ctx->lists.find(key);
  if (it != ctx->lists.end()) {
    *out = (const int32_t*)it->second.p;
    return MLP_OK;
  }
  if (ctx->lists.size() >= 8192 && !ctx->capturing) {
    HCK(sync_streams(ctx));
    for (int g = 0; g < MLP_DVND_GROUPS; ++g)
      if (ctx->graph_exec[g]) {  // its launches point into the lists
        HCK(hipGraphExecDestroy(ctx->graph_exec[g]));
        ctx->graph_exec[g] = nullptr;
        ctx->graph_key[g].clear();
      }
    bool user_graphs = false;
    for (hipGraphExec_t g : ctx->user_graphs) user_graphs = user_graphs || g;
    if (!user_graphs) {  // recorded graphs may point into the lists: keep them
      for (auto& kv : ctx->lists) hipFree(kv.second.p);
      ctx->lists.clear();
    }
  }
  void* p = nullptr;
  const size_t bytes = (size_t)count * 4;
  HCK(hipMalloc(&p, std::max<size_t>(4, bytes)));
  if (ctx->capturing) {
    HCK(hipMemcpyAsync(p, v, bytes, hipMemcpyHostToDevice, ctx->side));
    HCK(hipStreamSynchronize(ctx->side));
  } else if (bytes > 0) {
    char* h = nullptr;
    const int st = stage_alloc(ctx, bytes, &h);
    if (st) return st;
    memcpy(h, v, bytes);
    HCK(hipMemcpyAsync(p, h, bytes, hipMemcpyHostToDevice, ctx->stream));
  }
  ctx->lists.emplace(std::move(key), DevList{p});
  *out = (const int32_t*)p;
  return MLP_OK;
}

// Slot-id list for a per-slot launch: inline kernel argument when small.
int slot_list(mlp_ctx* ctx, const int32_t* slots, int count, SlotListInline& inl,
              const int32_t** dev) {
  *dev = nullptr;
  if (count <= MLP_SLOTS_INLINE) {
    for (int q = 0; q < count; ++q) inl.id[q] = slots[q];
    return MLP_OK;
  }
  return device_list(ctx, slots, count, dev);
}

int ensure_i64(mlp_ctx* ctx, size_t count) {
  if (count <= ctx->i64_cap) return MLP_OK;
  const size_t cap = std::max<size_t>({count, 2 * ctx->i64_cap, 1024});
  HCK(hipStreamSynchronize(ctx->stream));
  if (ctx->d_i64) hipFree(ctx->d_i64);
  if (ctx->h_i64) hipHostFree(ctx->h_i64);
  ctx->d_i64 = nullptr;
  ctx->h_i64 = nullptr;
  ctx->i64_cap = 0;
  HCK(hipMalloc(&ctx->d_i64, cap * 8));
  HCK(hipHostMalloc(&ctx->h_i64, cap * 8, hipHostMallocDefault));
  ctx->i64_cap = cap;
  return MLP_OK;
}

int check_slots(mlp_ctx* ctx, const int32_t* slots, int count) {
  if (!ctx) return MLP_EINVAL;
  if (count < 0 || (count > 0 && !slots)) return fail(ctx, MLP_EINVAL, "bad slot list");
  for (int q = 0; q < count; ++q)
    if (!slot_ok(ctx, slots[q])) return fail(ctx, MLP_EINVAL, "slot %d does not exist", slots[q]);
  return MLP_OK;
}

int grow_table(mlp_ctx* ctx, int need) {
  if (need <= ctx->table_cap) return MLP_OK;
  int cap = std::max(64, ctx->table_cap);
  while (cap < need) cap *= 2;
  SlotDev* t = nullptr;
  HCK(hipStreamSynchronize(ctx->stream));
  HCK(hipMalloc(&t, sizeof(SlotDev) * (size_t)cap));
  HCK(hipMemset(t, 0, sizeof(SlotDev) * (size_t)cap));
  if (ctx->d_table) {
    HCK(hipMemcpy(t, ctx->d_table, sizeof(SlotDev) * (size_t)ctx->table_cap,
                  hipMemcpyDeviceToDevice));
    HCK(hipFree(ctx->d_table));
  }
  ctx->d_table = t;
  ctx->table_cap = cap;
  return MLP_OK;
}

size_t pool_class(size_t bytes) {
  size_t c = 4096;
  while (c < bytes) c <<= 1;
  return c;
}

int pool_alloc(mlp_ctx* ctx, size_t bytes, void** out, size_t* cls) {
  const size_t c = pool_class(bytes);
  auto it = ctx->pool_free.find(c);
  if (it != ctx->pool_free.end() && !it->second.empty()) {
    *out = it->second.back();
    it->second.pop_back();
  } else {
    void* p = nullptr;
    if (hipMalloc(&p, c) != hipSuccess) {
      (void)hipGetLastError();
      return fail(ctx, MLP_ENOMEM, "device allocation of %zu bytes failed", c);
    }
    ctx->pool_all.push_back(p);
    ctx->pool_reserved += c;
    *out = p;
  }
  ctx->pool_live += c;
  *cls = c;
  return MLP_OK;
}

// Stream-ordered release: work already queued on the context stream may still
// use the block; a later allocation reuses it only for work queued after.
void pool_release(mlp_ctx* ctx, void* p, size_t cls) {
  if (!p) return;
  ctx->pool_free[cls].push_back(p);
  ctx->pool_live -= cls;
}

}  // namespace

extern "C" {

const char* mlp_strerror(int st) {
  switch (st) {
    case MLP_OK: return "ok";
    case MLP_EINVAL: return "invalid argument";
    case MLP_ENOMEM: return "out of memory";
    case MLP_EDEVICE: return "HIP runtime error";
    case MLP_ERANGE: return "value out of supported range";
    case MLP_ESTATE: return "call out of order";
    default: return "unknown status";
  }
}

const char* mlp_ctx_error(const mlp_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int mlp_ctx_create(int device, mlp_ctx** out) {
  if (!out) return MLP_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MLP_EDEVICE;
  if (device < 0 || device >= ndev) return MLP_EINVAL;
  mlp_ctx* ctx = new (std::nothrow) mlp_ctx();
  if (!ctx) return MLP_ENOMEM;
  memset(ctx->tune, -1, sizeof ctx->tune);
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreate(&ctx->stream) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess ||
      hipEventCreate(&ctx->ev1) != hipSuccess || hipMalloc(&ctx->d_counter, 64) != hipSuccess) {
    delete ctx;
    return MLP_EDEVICE;
  }
  ctx->stream_main = ctx->stream;
  ctx->scratch = scratch_address(device);  // device is current
  ctx->cu_count = prop.multiProcessorCount;
  ctx->hbm = (int64_t)prop.totalGlobalMem;
  snprintf(ctx->name, sizeof ctx->name, "%s", prop.gcnArchName);
  // dynamic LDS of the merge kernel: up to the device's per-block limit
  int lds_attr = 0;
  if (hipDeviceGetAttribute(&lds_attr, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
      hipSuccess)
    lds_attr = 64 * 1024;
  size_t lds = (size_t)lds_attr;
  if (lds > 160 * 1024) lds = 160 * 1024;
  ctx->max_merge_lds = lds > 8192 ? lds - 8192 : lds;
  if (mlp_merge_setup(ctx->max_merge_lds) != hipSuccess) ctx->max_merge_lds = 64 * 1024 - 8192;
  (void)mlp_fused_setup(ctx->max_merge_lds);
  ctx->max_total_lds = ctx->max_merge_lds;
  if (mlp_total_setup(ctx->max_total_lds) != hipSuccess) ctx->max_total_lds = 64 * 1024 - 8192;
  (void)hipGetLastError();
  *out = ctx;
  return MLP_OK;
}

int mlp_ctx_destroy(mlp_ctx* ctx) {
  if (ctx && ctx->capturing) {  // abandon an open capture
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(ctx->stream, &g) == hipSuccess && g) hipGraphDestroy(g);
    ctx->capturing = ctx->user_capture = false;
  }
  if (!ctx) return MLP_EINVAL;
  hipSetDevice(ctx->device);
  ctx->stream = ctx->stream_main;
  (void)sync_streams(ctx);  // every stream's work done before its memory goes
  for (void* p : ctx->pool_all) hipFree(p);
  for (auto& kv : ctx->lists) hipFree(kv.second.p);
  if (ctx->d_table) hipFree(ctx->d_table);
  for (hipGraphExec_t g : ctx->graph_exec)
    if (g) hipGraphExecDestroy(g);
  for (hipGraphExec_t g : ctx->user_graphs)
    if (g) hipGraphExecDestroy(g);
  if (ctx->d_counter) hipFree(ctx->d_counter);
  if (ctx->d_i64) hipFree(ctx->d_i64);
  if (ctx->h_i64) hipHostFree(ctx->h_i64);
  if (ctx->d_tstage) hipFree(ctx->d_tstage);
  if (ctx->d_sync) hipFree(ctx->d_sync);
  if (ctx->h_lists) hipHostFree(ctx->h_lists);
  hipEventDestroy(ctx->ev0);
  hipEventDestroy(ctx->ev1);
  for (hipEvent_t e : ctx->tev) hipEventDestroy(e);
  for (void* p : ctx->dev_bufs) hipFree(p);
  for (auto& t : ctx->timers)
    if (t.first) {
      hipEventDestroy(t.first);
      hipEventDestroy(t.second);
    }
  hipStreamDestroy(ctx->stream_main);
  if (ctx->side) hipStreamDestroy(ctx->side);
  for (int g = 0; g < MLP_DVND_GROUPS; ++g) {
    if (ctx->stream_grp[g]) hipStreamDestroy(ctx->stream_grp[g]);
    if (ctx->ev_join[g]) hipEventDestroy(ctx->ev_join[g]);
    if (ctx->ev_chunk[g]) hipEventDestroy(ctx->ev_chunk[g]);
  }
  if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
  delete ctx;
  return MLP_OK;
}

int mlp_ctx_device(mlp_ctx* ctx, int32_t* device) {
  if (!ctx || !device) return MLP_EINVAL;
  *device = ctx->device;
  return MLP_OK;
}

int mlp_query(mlp_ctx* ctx, int32_t* idle) {
  if (!ctx || !idle) return MLP_EINVAL;
  const hipError_t e = hipStreamQuery(ctx->stream);
  if (e == hipSuccess) {
    *idle = 1;
  } else if (e == hipErrorNotReady) {
    (void)hipGetLastError();
    *idle = 0;
  } else {
    return fail(ctx, MLP_EDEVICE, "stream query: %s", hipGetErrorString(e));
  }
  return MLP_OK;
}

int mlp_sync(mlp_ctx* ctx) {
  NO_CAPTURE(ctx);
  if (!ctx) return MLP_EINVAL;
  HCK(hipStreamSynchronize(ctx->stream));
  return MLP_OK;
}

int mlp_pool_stats(mlp_ctx* ctx, int64_t* live_bytes, int64_t* reserved_bytes) {
  if (!ctx) return MLP_EINVAL;
  if (live_bytes) *live_bytes = (int64_t)ctx->pool_live;
  if (reserved_bytes) *reserved_bytes = (int64_t)ctx->pool_reserved;
  return MLP_OK;
}

int mlp_device_info(mlp_ctx* ctx, int32_t* cu_count, int64_t* hbm_bytes, char* name,
                    int32_t name_len) {
  if (!ctx) return MLP_EINVAL;
  if (cu_count) *cu_count = ctx->cu_count;
  if (hbm_bytes) *hbm_bytes = ctx->hbm;
  if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", ctx->name);
  return MLP_OK;
}

// Instance validation shared by instance creation and mlp_check_job.  The
// coordinates may be any int32 values whose span is at most 32767 in x and in
// y; an axis whose values leave [0, 32767] is translated to [0, span]
// (distances depend on differences only, so the reference's raw-coordinate
// distances, mlproblem.cpp:153-166, are unchanged), which keeps dx^2 + dy^2
// within int32 and the device's packed 16-bit coordinates exact.  *xmin_out
// and *ymin_out receive the translation (0 for an axis inside [0, 32767]),
// *maxd a bound of every distance.
static int check_instance(mlp_ctx* ctx, const int32_t* xs, const int32_t* ys, int32_t n,
                          int32_t round, int32_t* xmin_out, int32_t* ymin_out, int* maxd) {
  if (!xs || !ys || n < 2) return fail(ctx, MLP_EINVAL, "need at least 2 clients");
  if (n + 1 > MLP_MAX_N) return fail(ctx, MLP_ERANGE, "n = %d exceeds %d", n, MLP_MAX_N - 1);
  if (round != 0 && round != 1) return fail(ctx, MLP_EINVAL, "round must be 0 or 1");
  int64_t xmin = INT64_MAX, xmax = INT64_MIN, ymin = INT64_MAX, ymax = INT64_MIN;
  for (int c = 0; c < n; ++c) {
    xmin = std::min<int64_t>(xmin, xs[c]);
    xmax = std::max<int64_t>(xmax, xs[c]);
    ymin = std::min<int64_t>(ymin, ys[c]);
    ymax = std::max<int64_t>(ymax, ys[c]);
  }
  if (xmax - xmin > 32767 || ymax - ymin > 32767)
    return fail(ctx, MLP_ERANGE, "coordinate span %lld x %lld exceeds 32767",
                (long long)(xmax - xmin), (long long)(ymax - ymin));
  const double dx = (double)(xmax - xmin), dy = (double)(ymax - ymin);
  const int md = (int)std::ceil(std::sqrt(dx * dx + dy * dy)) + 1;
  // move keys hold delta * 2^20 + index in 64 bits: every |delta| is below
  // N (N + 32) maxd, which must stay below 2^43
  const uint64_t N = (uint64_t)n + 1;
  if (N * (N + 32) * (uint64_t)md >= (1ull << 43))
    return fail(ctx, MLP_ERANGE, "n = %d with a coordinate span of %d exceeds the 64-bit move keys",
                n, md);
  if (xmin_out) *xmin_out = (xmin < 0 || xmax > 32767) ? (int32_t)xmin : 0;
  if (ymin_out) *ymin_out = (ymin < 0 || ymax > 32767) ? (int32_t)ymin : 0;
  if (maxd) *maxd = md;
  return MLP_OK;
}

// A tour of N positions [0, perm(1..N-2), 0].
static int check_tour(mlp_ctx* ctx, const int32_t* t, int N, int q, std::vector<char>& seen) {
  const int n = N - 1;
  if (t[0] != 0 || t[N - 1] != 0)
    return fail(ctx, MLP_EINVAL, "tour %d must start and end at the depot 0", q);
  seen.assign((size_t)n, 0);
  for (int p = 1; p < N - 1; ++p) {
    if (t[p] <= 0 || t[p] >= n || seen[(size_t)t[p]])
      return fail(ctx, MLP_EINVAL, "tour %d is not a permutation (position %d)", q, p);
    seen[(size_t)t[p]] = 1;
  }
  return MLP_OK;
}

static int instance_create(mlp_ctx* ctx, const int32_t* xs, const int32_t* ys, int32_t n,
                           int32_t round, int32_t* inst, bool async) {
  if (!ctx || !inst) return MLP_EINVAL;
  Instance in;
  int32_t x0 = 0, y0 = 0;
  int st = check_instance(ctx, xs, ys, n, round, &x0, &y0, &in.maxd);
  if (st) return st;
  in.n = n;
  in.round = round;
  HCK(hipSetDevice(ctx->device));
  const size_t half = align_up(4 * (size_t)n, 256);
  void* mem = nullptr;
  st = pool_alloc(ctx, 2 * half, &mem, &in.cls);
  if (st) return st;
  in.xs = (int32_t*)mem;
  in.ys = (int32_t*)((char*)mem + half);
  // both coordinate arrays, translated to [0, span], in one copy
  auto fill = [&](char* h) {
    int32_t* hx = (int32_t*)h;
    int32_t* hy = (int32_t*)(h + half);
    for (int c = 0; c < n; ++c) {
      hx[c] = xs[c] - x0;
      hy[c] = ys[c] - y0;
    }
  };
  if (async) {
    // from pinned staging, ordered on the stream (the caller's arrays may be
    // released at once)
    char* h = nullptr;
    st = stage_alloc(ctx, 2 * half, &h);
    if (st) {
      pool_release(ctx, mem, in.cls);
      return st;
    }
    fill(h);
    HCK(hipMemcpyAsync(mem, h, 2 * half, hipMemcpyHostToDevice, ctx->stream));
  } else {
    std::vector<char> h(2 * half);
    fill(h.data());
    HCK(hipMemcpy(mem, h.data(), 2 * half, hipMemcpyHostToDevice));
  }
  in.alive = true;
  for (size_t q = 0; q < ctx->insts.size(); ++q)
    if (!ctx->insts[q].alive) {
      ctx->insts[q] = in;
      *inst = (int32_t)q;
      return MLP_OK;
    }
  ctx->insts.push_back(in);
  *inst = (int32_t)ctx->insts.size() - 1;
  return MLP_OK;
}

int mlp_instance_create(mlp_ctx* ctx, const int32_t* xs, const int32_t* ys, int32_t n,
                        int32_t round, int32_t* inst) {
  NO_CAPTURE(ctx);
  return instance_create(ctx, xs, ys, n, round, inst, false);
}

int mlp_instance_create_async(mlp_ctx* ctx, const int32_t* xs, const int32_t* ys, int32_t n,
                              int32_t round, int32_t* inst) {
  NO_CAPTURE(ctx);
  return instance_create(ctx, xs, ys, n, round, inst, true);
}

int mlp_check_job(mlp_ctx* ctx, const int32_t* xs, const int32_t* ys, int32_t n, int32_t round,
                  const int32_t* tour) {
  int st = check_instance(ctx, xs, ys, n, round, nullptr, nullptr, nullptr);
  if (st || !tour) return st ? st : fail(ctx, MLP_EINVAL, "tour is NULL");
  std::vector<char> seen;
  return check_tour(ctx, tour, n + 1, 0, seen);
}

int mlp_mem_info(mlp_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes) {
  if (!ctx) return MLP_EINVAL;
  size_t fr = 0, tot = 0;
  HCK(hipSetDevice(ctx->device));
  HCK(hipMemGetInfo(&fr, &tot));
  if (free_bytes) *free_bytes = (int64_t)fr;
  if (total_bytes) *total_bytes = (int64_t)tot;
  return MLP_OK;
}

int mlp_instance_destroy(mlp_ctx* ctx, int32_t inst) {
  NO_CAPTURE(ctx);
  if (!ctx) return MLP_EINVAL;
  if (inst < 0 || inst >= (int)ctx->insts.size() || !ctx->insts[inst].alive)
    return fail(ctx, MLP_EINVAL, "instance %d does not exist", inst);
  for (auto& s : ctx->slots)
    if (s.alive && s.inst == inst)
      return fail(ctx, MLP_ESTATE, "instance %d still has slots", inst);
  Instance& in = ctx->insts[inst];
  pool_release(ctx, in.xs, in.cls);
  in = Instance();
  return MLP_OK;
}

int mlp_instance_weights(mlp_ctx* ctx, int32_t inst, int32_t* w_out) {
  NO_CAPTURE(ctx);
  if (!ctx || !w_out) return MLP_EINVAL;
  if (inst < 0 || inst >= (int)ctx->insts.size() || !ctx->insts[inst].alive)
    return fail(ctx, MLP_EINVAL, "instance %d does not exist", inst);
  const Instance& in = ctx->insts[inst];
  const size_t bytes = 4 * (size_t)in.n * in.n;
  int32_t* d = nullptr;
  HCK(hipMalloc(&d, bytes));
  hipError_t e = mlp_launch_weights(ctx->stream, in.xs, in.ys, in.n, in.round, d);
  if (e == hipSuccess) e = hipMemcpyAsync(w_out, d, bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return fail(ctx, MLP_EDEVICE, "weights: %s", hipGetErrorString(e));
  return MLP_OK;
}

int64_t mlp_slot_footprint(int32_t n) {
  if (n < 2 || n + 1 > MLP_MAX_N) return -1;
  return (int64_t)slot_layout(n + 1, MLP_NB_COUNT).total;
}

int mlp_slot_create(mlp_ctx* ctx, int32_t inst, int32_t* slot) {
  NO_CAPTURE(ctx);
  if (!ctx || !slot) return MLP_EINVAL;
  if (inst < 0 || inst >= (int)ctx->insts.size() || !ctx->insts[inst].alive)
    return fail(ctx, MLP_EINVAL, "instance %d does not exist", inst);
  const Instance& in = ctx->insts[inst];
  const int N = in.n + 1;
  const SlotLayout L = slot_layout(N, ctx->nbs);
  int id = -1;
  for (size_t q = 0; q < ctx->slots.size(); ++q)
    if (!ctx->slots[q].alive) {
      id = (int)q;
      break;
    }
  if (id < 0) {
    id = (int)ctx->slots.size();
    ctx->slots.emplace_back();
  }
  int st = grow_table(ctx, id + 1);
  if (st) return st;
  Slot s;
  s.inst = inst;
  s.N = N;
  s.nbs = ctx->nbs;
  st = pool_alloc(ctx, L.total, &s.mem, &s.cls);
  if (st) return st;
  char* b = (char*)s.mem;
  SlotDev& d = s.dev;
  d.N = N;
  d.n = in.n;
  d.round = in.round;
  d.maxd = in.maxd;
  d.rowcap = row_cap(N);
  d.xs = in.xs;
  d.ys = in.ys;
  d.tour = (int32_t*)(b + L.tour);
  d.xy = (int32_t*)(b + L.xy);
  d.w = (int32_t*)(b + L.w);
  d.t32 = (int32_t*)(b + L.t32);
  d.q32 = (int32_t*)(b + L.q32);
  d.g32 = (int32_t*)(b + L.g32);
  d.h32 = (int32_t*)(b + L.h32);
  d.t64 = (int64_t*)(b + L.t64);
  d.q64 = (int64_t*)(b + L.q64);
  d.info = (int64_t*)(b + L.info);
  d.nbs = ctx->nbs;
  for (int nb = 0; nb < d.nbs; ++nb) {
    d.rows[nb] = (int64_t*)(b + L.rows[nb]);
    d.kept_cost[nb] = (int64_t*)(b + L.kc[nb]);
    d.kept_ij[nb] = (int32_t*)(b + L.kij[nb]);
  }
  d.mscratch = b + L.ms;
  // one launch: the block zeroed, the row buffers KEY_NONE, the table entry
  SlotInitRows rows{};
  rows.nbs = d.nbs;
  for (int nb = 0; nb < d.nbs; ++nb) {
    rows.start[nb] = (int64_t)(L.rows[nb] / 8);
    rows.count[nb] = 2 * (int64_t)d.rowcap;
  }
  HCK(mlp_launch_slot_init(ctx->stream, ctx->d_table, id, d, (int64_t*)s.mem,
                           (int64_t)(L.total / 8), rows));
  s.alive = true;
  ctx->slots[id] = s;
  *slot = id;
  return MLP_OK;
}

int mlp_slot_destroy(mlp_ctx* ctx, int32_t slot) {
  NO_CAPTURE(ctx);
  if (!ctx) return MLP_EINVAL;
  if (!slot_ok(ctx, slot)) return fail(ctx, MLP_EINVAL, "slot %d does not exist", slot);
  pool_release(ctx, ctx->slots[slot].mem, ctx->slots[slot].cls);
  pool_release(ctx, ctx->slots[slot].total, ctx->slots[slot].total_cls);
  ctx->slots[slot] = Slot();
  return MLP_OK;
}

int mlp_set_tours(mlp_ctx* ctx, const int32_t* slots, int32_t count, const int32_t* tours) {
  NO_CAPTURE(ctx);
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  if (count == 0) return MLP_OK;
  if (!tours) return fail(ctx, MLP_EINVAL, "tours is NULL");
  size_t total = 0;
  std::vector<char> seen;
  for (int q = 0; q < count; ++q) {
    const Slot& s = ctx->slots[slots[q]];
    const int32_t* t = tours + total;
    st = check_tour(ctx, t, s.N, q, seen);
    if (st) return st;
    total += (size_t)s.N;
  }
  // all tours in one copy from pinned staging, ordered on the stream (no
  // wait for queued kernels); several slots' tours land in a device array
  // that one kernel scatters into the slots
  char* h = nullptr;
  st = stage_alloc(ctx, total * 4, &h);
  if (st) return st;
  memcpy(h, tours, total * 4);
  if (count == 1) {
    HCK(hipMemcpyAsync(ctx->slots[slots[0]].dev.tour, h, total * 4, hipMemcpyHostToDevice,
                       ctx->stream));
  } else {
    if (total > ctx->d_tstage_cap) {
      const size_t cap = std::max<size_t>(total, 2 * ctx->d_tstage_cap + 4096);
      HCK(hipStreamSynchronize(ctx->stream));
      if (ctx->d_tstage) hipFree(ctx->d_tstage);
      ctx->d_tstage = nullptr;
      ctx->d_tstage_cap = 0;
      HCK(hipMalloc((void**)&ctx->d_tstage, cap * 4));
      ctx->d_tstage_cap = cap;
    }
    HCK(hipMemcpyAsync(ctx->d_tstage, h, total * 4, hipMemcpyHostToDevice, ctx->stream));
    const int32_t* dev = nullptr;
    st = device_list(ctx, slots, count, &dev);
    if (st) return st;
    HCK(mlp_launch_scatter_tours(ctx->stream, ctx->d_table, dev, count, ctx->d_tstage));
  }
  for (int q = 0; q < count; ++q) {
    Slot& s = ctx->slots[slots[q]];
    for (int nb = 0; nb < MLP_NB_MAX; ++nb) {
      s.eval_rowmap[nb] = -1;
      s.merged[nb] = false;
    }
    s.total_nb = -1;
    s.total_merged = false;
  }
  for (int q0 = 0; q0 < count; q0 += 65535) {
    const int c = std::min(65535, count - q0);
    SlotListInline inl;
    const int32_t* dev = nullptr;
    st = slot_list(ctx, slots + q0, c, inl, &dev);
    if (st) return st;
    HCK(mlp_launch_prep(ctx->stream, ctx->d_table, inl, dev, c));
  }
  return MLP_OK;
}

int mlp_get_tours(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t* tours) {
  NO_CAPTURE(ctx);
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  if (count && !tours) return fail(ctx, MLP_EINVAL, "tours is NULL");
  if (count == 0) return MLP_OK;
  size_t total = 0;
  for (int q = 0; q < count; ++q) total += (size_t)ctx->slots[slots[q]].N;
  // one copy into pinned memory: a single slot's tour directly, several
  // gathered into one device array first
  st = ensure_i64(ctx, (total + 1) / 2);
  if (st) return st;
  if (count == 1) {
    HCK(hipMemcpyAsync(ctx->h_i64, ctx->slots[slots[0]].dev.tour, 4 * total, hipMemcpyDeviceToHost,
                       ctx->stream));
  } else {
    const int32_t* dev = nullptr;
    st = device_list(ctx, slots, count, &dev);
    if (st) return st;
    HCK(mlp_launch_gather_tours(ctx->stream, ctx->d_table, dev, count, (int32_t*)ctx->d_i64));
    HCK(hipMemcpyAsync(ctx->h_i64, ctx->d_i64, 4 * total, hipMemcpyDeviceToHost, ctx->stream));
  }
  HCK(hipStreamSynchronize(ctx->stream));
  memcpy(tours, ctx->h_i64, 4 * total);
  return MLP_OK;
}

// Info words of several slots: word w0 into out0 and, when out1 is given,
// word w1 into out1, with one read-back.
static int gather_info(mlp_ctx* ctx, const int32_t* slots, int count, int w0, int64_t* out0,
                       int w1 = 0, int64_t* out1 = nullptr) {
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  if (count == 0) return MLP_OK;
  const size_t words = (size_t)count * (out1 ? 2 : 1);
  st = ensure_i64(ctx, words);
  if (st) return st;
  const int32_t* dev = nullptr;
  st = device_list(ctx, slots, count, &dev);
  if (st) return st;
  HCK(mlp_launch_gather_info(ctx->stream, ctx->d_table, dev, count, w0, ctx->d_i64));
  if (out1)
    HCK(mlp_launch_gather_info(ctx->stream, ctx->d_table, dev, count, w1, ctx->d_i64 + count));
  HCK(hipMemcpyAsync(ctx->h_i64, ctx->d_i64, 8 * words, hipMemcpyDeviceToHost, ctx->stream));
  HCK(hipStreamSynchronize(ctx->stream));
  memcpy(out0, ctx->h_i64, 8 * (size_t)count);
  if (out1) memcpy(out1, ctx->h_i64 + count, 8 * (size_t)count);
  return MLP_OK;
}

int mlp_get_latency(mlp_ctx* ctx, const int32_t* slots, int32_t count, int64_t* latency) {
  NO_CAPTURE(ctx);
  if (!latency && count) return fail(ctx, MLP_EINVAL, "latency is NULL");
  return gather_info(ctx, slots, count, MLP_INFO_LATENCY, latency);
}

namespace {

// Launches one eval; between mlp_timing_begin and mlp_timing_end the launch
// is bracketed by its own pair of HIP events.
int timed_launch(mlp_ctx* ctx, hipError_t (*launch)(mlp_ctx*, void*), void* arg) {
  if (ctx->capturing || !ctx->timing) {
    HCK(launch(ctx, arg));
    return MLP_OK;
  }
  if (ctx->tcount >= ctx->tev.size() / 2) return fail(ctx, MLP_ERANGE, "timing ring full");
  hipEvent_t e0 = ctx->tev[2 * ctx->tcount], e1 = ctx->tev[2 * ctx->tcount + 1];
  ctx->tcount++;
  HCK(hipEventRecord(e0, ctx->stream));
  HCK(launch(ctx, arg));
  HCK(hipEventRecord(e1, ctx->stream));
  return MLP_OK;
}

struct EvalArgs {
  EvalSegsInline inl;
  const EvalSeg* dev;
  SweepSegsInline sinl;
  const SweepSeg* sdev;
  int nsegs, total, rowmap, grid, skip;
  int* counter;
};

hipError_t launch_eval_cb(mlp_ctx* ctx, void* p) {
  EvalArgs& a = *(EvalArgs*)p;
  return mlp_launch_eval(ctx->stream, ctx->d_table, a.inl, a.dev, a.nsegs, a.total, a.rowmap,
                         a.counter, a.grid, a.skip);
}

// Resident workgroups per CU of each sweep tile shape (persistent grids).
const int kTileWgPerCu[MLP_TILE_COUNT] = {5, 4, 3, 4, 5};

// Moves one evaluation of `mask` over one tour of N positions (ADS row map).
int64_t tour_moves(int64_t N, uint32_t mask) {
  int64_t m = 0;
  if (mask & 1u) m += std::max<int64_t>(0, (N - 4) * (N - 3) / 2);
  if (mask & 2u) m += std::max<int64_t>(0, ((N + 1) / 2) * (N - 3));
  for (int k = 1; k <= 3; ++k)
    if (mask & (1u << (k + 1))) m += std::max<int64_t>(0, (N - k - 2) * (N - k - 1));
  return m;
}

// Moves one evaluation of `mask` over the slots covers (ADS row map).
int64_t launch_moves(const mlp_ctx* ctx, const int32_t* slots, int count, uint32_t mask) {
  int64_t m = 0;
  for (int q = 0; q < count; ++q) m += tour_moves(ctx->slots[slots[q]].N, mask);
  return m;
}

int work_bucket(int64_t moves) {
  int b = 0;
  while (b < MLP_TUNE_BUCKETS - 1 && (2LL << b) <= moves) ++b;
  return b;
}

// Shape for a launch: the autotuned table entry of (mask, work bucket) or the
// nearest tuned bucket of that mask; untuned masks use the launch size (small
// tiles give a small launch enough workgroups, large tiles evaluate fewer
// distances per move).
int auto_tile_moves(const mlp_ctx* ctx, int64_t moves, int count, uint32_t mask) {
  const int b = work_bucket(moves);
  const int8_t* row = ctx->tune[mask & 0x1fu];
  for (int d = 0; d < MLP_TUNE_BUCKETS; ++d) {
    if (b - d >= 0 && row[b - d] >= 0) return row[b - d];
    if (b + d < MLP_TUNE_BUCKETS && row[b + d] >= 0) return row[b + d];
  }
  if (count <= 2) return MLP_TILE_4x256;
  if (count <= 16) return MLP_TILE_8x512;
  return MLP_TILE_16x1024;
}

int auto_tile(const mlp_ctx* ctx, const int32_t* slots, int count, uint32_t mask) {
  return auto_tile_moves(ctx, launch_moves(ctx, slots, count, mask), count, mask);
}

// Forget the cached DVND chunk graph (its launches embed the neighborhood
// set, tile choices and merge mode of the moment it was captured).
int drop_dvnd_graph(mlp_ctx* ctx) {
  for (int g = 0; g < MLP_DVND_GROUPS; ++g) {
    if (ctx->graph_exec[g]) {
      if (!ctx->capturing) HCK(sync_streams(ctx));
      HCK(hipGraphExecDestroy(ctx->graph_exec[g]));
      ctx->graph_exec[g] = nullptr;
    }
    ctx->graph_key[g].clear();
  }
  return MLP_OK;
}

// A work counter cleared by a kernel (graph-capturable, unlike a memset).
hipError_t zero_counter(mlp_ctx* ctx, int* counter) {
  SlotListInline none{};
  return mlp_launch_round_begin(ctx->stream, ctx->d_table, none, nullptr, 0, counter, 1);
}

hipError_t launch_sweep_cb(mlp_ctx* ctx, void* p) {
  EvalArgs& a = *(EvalArgs*)p;
  return mlp_launch_sweep(ctx->stream, a.rowmap, ctx->d_table, a.sinl, a.sdev, a.nsegs, a.total,
                          a.counter, a.grid, a.skip);
}

}  // namespace

// One evaluation launch.  counter == nullptr: one workgroup per work item
// (persistent == 0) or the context's standalone counter cleared by a memset;
// a DVND round passes its own counter, cleared by k_round_begin.
// Every neighborhood of the mask has buffers in every slot.
static int check_mask(mlp_ctx* ctx, const int32_t* slots, int32_t count, uint32_t nb_mask) {
  if (nb_mask >> MLP_NB_MAX) return fail(ctx, MLP_EINVAL, "bad neighborhood mask");
  for (int q = 0; q < count; ++q)
    if (nb_mask >> ctx->slots[slots[q]].nbs)
      return fail(ctx, MLP_EINVAL, "slot %d has buffers for neighborhoods 0..%d only", slots[q],
                  ctx->slots[slots[q]].nbs - 1);
  return MLP_OK;
}

static int eval_impl(mlp_ctx* ctx, const int32_t* slots, int32_t count, uint32_t nb_mask,
                     int32_t rowmap, int32_t persistent, int skip, int* round_counter) {
  if (ctx->user_capture)
    for (int q = 0; q < count; ++q)
      ctx->capture_ops.push_back({slots[q], 0, nb_mask, rowmap});
  for (int q = 0; q < count; ++q) {
    Slot& s = ctx->slots[slots[q]];
    if (skip) {
      s.parity_stale = true;  // the DVND merges commit parities on the device
    } else if (s.parity_stale) {
      int64_t par[MLP_NB_MAX];
      HCK(hipStreamSynchronize(ctx->stream));
      HCK(hipMemcpy(par, s.dev.info + MLP_INFO_PARITY0, sizeof par, hipMemcpyDeviceToHost));
      for (int nb = 0; nb < MLP_NB_MAX; ++nb) s.parity[nb] = (int)(par[nb] & 1);
      s.parity_stale = false;
    }
    for (int nb = 0; nb < MLP_NB_MAX; ++nb)
      if (nb_mask & (1u << nb)) {
        s.eval_rowmap[nb] = rowmap;
        s.merged[nb] = false;
      }
  }
  EvalArgs a{};
  a.skip = skip;
  int st = MLP_OK;
  // the five tiled neighborhoods go to k_sweep; or-opt-k for k > 3
  // (ids 5..14) to the one-wave-per-row k_eval below
  if (nb_mask & 0x1fu) {
    // fixed-i tiles (swap + or-opt) and 2-opt diagonal tiles; under the ODC
    // row map the fixed tiles fold each move into its ODC row (tiles of at
    // most eight rows: 16 x 1024 runs as 8 x 512)
    const bool odc = rowmap == 1;
    int tile = ctx->tile_override;
    if (tile < 0) {
      tile = auto_tile(ctx, slots, count, nb_mask);
      // the tuned table is the ADS map's; ODC tiles of tours from 512
      // positions fold faster with 16 rows (16 x 128)
      int maxn = 0;
      for (int q = 0; q < count; ++q) maxn = std::max(maxn, ctx->slots[slots[q]].N);
      if (odc && maxn >= 512 && (nb_mask & 0x1cu)) tile = MLP_TILE_8x256;
    }
    if (odc && tile == MLP_TILE_16x1024) tile = MLP_TILE_8x512;
    int TR, TC;
    mlp_tile_shape(tile | (odc ? MLP_TILE_ODC : 0), &TR, &TC);
    // with swap in the launch, tiles of eight or more rows evaluate 2-opt
    // too (sharing swap's distance products) instead of 2-opt diagonal tiles
    const bool two = !odc && (nb_mask & 3u) == 3u && TR >= 8;
    if (two) mlp_tile_shape(tile | MLP_TILE_TWO, &TR, &TC);
    std::vector<SweepSeg> segs;
    segs.reserve((size_t)count * 2);
    int total = 0;
    const uint32_t fixed = (nb_mask & 0x1du) | (two ? 2u : 0u);
    for (int q = 0; q < count; ++q) {
      Slot& s = ctx->slots[slots[q]];
      int par = 0;
      if (!skip)
        for (int nb = 0; nb < MLP_NB_COUNT; ++nb)
          if (nb_mask & (1u << nb)) {
            s.parity[nb] ^= 1;
            par |= s.parity[nb] << nb;
          }
      if (fixed) {
        int rmax = 0;
        for (int nb : {0, 2, 3, 4})
          if (fixed & (1u << nb))  // tiles of column 0 also reset the ODC rows
            rmax = std::max({rmax, rows_of(nb, 0, s.N), odc ? rows_of(nb, 1, s.N) : 0});
        // 2-opt rows i = 1 .. N-3 and diagonal rows 0 .. ceil(N/2) - 1
        if (two) rmax = std::max({rmax, s.N - 3, rows_of(1, 0, s.N)});
        if (rmax > 0) {
          const int ncols = (s.N - 2 + TC - 1) / TC;
          segs.push_back(SweepSeg{slots[q], 0, (int)fixed, total, ncols, par});
          total += (rmax + TR - 1) / TR * ncols;
        }
      }
      if ((nb_mask & 2u) && !two) {
        const int r = rows_of(1, 0, s.N);
        if (r > 0 && s.N >= 4) {
          const int ncols = (s.N - 3 + TC - 1) / TC;
          segs.push_back(SweepSeg{slots[q], 1, 2, total, ncols, par});
          total += (r + MLP_2OPT_ROWS - 1) / MLP_2OPT_ROWS * ncols;
        }
      }
    }
    if (!segs.empty()) {
    a.nsegs = (int)segs.size();
    a.total = total;
    a.rowmap = tile;
    if (a.nsegs <= MLP_SEGS_INLINE) {
      for (int q = 0; q < a.nsegs; ++q) a.sinl.s[q] = segs[(size_t)q];
    } else {
      st = device_list(ctx, (const int32_t*)segs.data(), a.nsegs * MLP_SWEEPSEG_INTS,
                       (const int32_t**)&a.sdev);
      if (st) return st;
      // every slot with the same segments: the kernel finds an item's
      // segment by division
      const int sps = (a.nsegs == count) ? 1 : (a.nsegs == 2 * count ? 2 : 0);
      if (sps) {
        const int group = total / count, size0 = sps == 2 ? segs[1].blk0 : group;
        bool eq = total == group * count && group < (1 << 23);
        for (int k = 0; k < count && eq; ++k) {
          eq = segs[(size_t)k * sps].blk0 == k * group;
          if (sps == 2) eq = eq && segs[(size_t)k * 2 + 1].blk0 == k * group + size0;
        }
        if (eq) a.skip |= (sps == 2 ? 4 : 0) | (group << 8);
      }
    }
    a.grid = total;
    if (two) a.rowmap |= MLP_TILE_TWO;
    if (odc) a.rowmap |= MLP_TILE_ODC;
    // tiles of a tour run mostly one after another: filtered diagonal folds
    if (total > 2 * ctx->cu_count * kTileWgPerCu[tile]) a.skip |= 2;
    if (ctx->user_capture)  // replays fold into the same buffers: clear them in the graph
      HCK(mlp_launch_clear_rows(ctx->stream, ctx->d_table, a.sinl, a.sdev, a.nsegs));
    if (persistent) {
      a.grid = std::min(total, ctx->cu_count * kTileWgPerCu[tile]);
      if (round_counter) {
        a.counter = round_counter;
      } else {
        a.counter = ctx->d_counter;
        HCK(zero_counter(ctx, a.counter));
      }
    }
    st = timed_launch(ctx, launch_sweep_cb, &a);
    if (st) return st;
    }
    nb_mask &= ~0x1fu;
    if (!nb_mask) return MLP_OK;
    a = EvalArgs{};
    a.skip = skip;
  }
  std::vector<EvalSeg> segs;
  segs.reserve((size_t)count * 5);
  int total = 0;
  for (int q = 0; q < count; ++q) {
    Slot& s = ctx->slots[slots[q]];
    for (int nb = 0; nb < MLP_NB_MAX; ++nb) {
      if (!(nb_mask & (1u << nb))) continue;
      const int R = rows_of(nb, rowmap, s.N);
      if (R <= 0) continue;
      if (!skip) s.parity[nb] ^= 1;
      segs.push_back(EvalSeg{slots[q], nb, R, total, s.parity[nb]});
      total += (R + MLP_EVAL_ROWS_PER_WG - 1) / MLP_EVAL_ROWS_PER_WG;
    }
  }
  if (segs.empty()) return MLP_OK;
  a.nsegs = (int)segs.size();
  a.total = total;
  a.rowmap = rowmap;
  if (a.nsegs <= MLP_SEGS_INLINE) {
    for (int q = 0; q < a.nsegs; ++q) a.inl.s[q] = segs[(size_t)q];
  } else {
    st = device_list(ctx, (const int32_t*)segs.data(), a.nsegs * MLP_EVALSEG_INTS,
                     (const int32_t**)&a.dev);
    if (st) return st;
  }
  a.grid = total;
  if (persistent) {
    a.grid = std::min(total, ctx->cu_count * 8);
    if (round_counter) {
      a.counter = round_counter;
    } else {
      a.counter = ctx->d_counter;
      HCK(zero_counter(ctx, a.counter));
    }
  }
  return timed_launch(ctx, launch_eval_cb, &a);
}

int mlp_eval(mlp_ctx* ctx, const int32_t* slots, int32_t count, uint32_t nb_mask,
             int32_t rowmap, int32_t persistent) {
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  if (rowmap != 0 && rowmap != 1) return fail(ctx, MLP_EINVAL, "rowmap must be 0 or 1");
  st = check_mask(ctx, slots, count, nb_mask);
  if (st) return st;
  return eval_impl(ctx, slots, count, nb_mask, rowmap, persistent, 0, nullptr);
}

int mlp_graph_begin(mlp_ctx* ctx) {
  if (!ctx) return MLP_EINVAL;
  if (ctx->capturing) return fail(ctx, MLP_ESTATE, "a capture is already open");
  HCK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
  ctx->capturing = true;
  ctx->user_capture = true;
  ctx->capture_ops.clear();
  return MLP_OK;
}

int mlp_graph_end(mlp_ctx* ctx, int32_t* graph) {
  if (!ctx || !graph) return MLP_EINVAL;
  if (!ctx->user_capture) return fail(ctx, MLP_ESTATE, "no capture open");
  ctx->capturing = false;
  ctx->user_capture = false;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
  if (ec != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return fail(ctx, MLP_EDEVICE, "graph capture: %s", hipGetErrorString(ec));
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ei != hipSuccess) return fail(ctx, MLP_EDEVICE, "graph instantiate: %s", hipGetErrorString(ei));
  // the recorded evals changed the host's parity choice without running
  for (const auto& op : ctx->capture_ops)
    if (slot_ok(ctx, op.slot)) ctx->slots[op.slot].parity_stale = true;
  for (size_t q = 0; q < ctx->user_graphs.size(); ++q)
    if (!ctx->user_graphs[q]) {
      ctx->user_graphs[q] = ex;
      ctx->user_graph_ops[q] = ctx->capture_ops;
      *graph = (int32_t)q;
      return MLP_OK;
    }
  ctx->user_graphs.push_back(ex);
  ctx->user_graph_ops.push_back(ctx->capture_ops);
  *graph = (int32_t)ctx->user_graphs.size() - 1;
  return MLP_OK;
}

int mlp_graph_launch(mlp_ctx* ctx, int32_t graph) {
  if (!ctx) return MLP_EINVAL;
  if (ctx->capturing) return fail(ctx, MLP_ESTATE, "capture open");
  if (graph < 0 || graph >= (int)ctx->user_graphs.size() || !ctx->user_graphs[graph])
    return fail(ctx, MLP_EINVAL, "graph %d does not exist", graph);
  HCK(hipGraphLaunch(ctx->user_graphs[graph], ctx->stream));
  // replays record row-buffer parities the host did not choose; repeat the
  // evaluated / merged bookkeeping of the recorded calls
  for (const auto& op : ctx->user_graph_ops[graph]) {
    if (!slot_ok(ctx, op.slot)) continue;
    Slot& s = ctx->slots[op.slot];
    s.parity_stale = true;
    if (op.merge) {
      s.merged[op.mask] = true;
    } else {
      for (int nb = 0; nb < MLP_NB_MAX; ++nb)
        if (op.mask & (1u << nb)) {
          s.eval_rowmap[nb] = op.rowmap;
          s.merged[nb] = false;
        }
    }
  }
  return MLP_OK;
}

int mlp_graph_destroy(mlp_ctx* ctx, int32_t graph) {
  if (!ctx) return MLP_EINVAL;
  if (graph < 0 || graph >= (int)ctx->user_graphs.size() || !ctx->user_graphs[graph])
    return fail(ctx, MLP_EINVAL, "graph %d does not exist", graph);
  HCK(hipStreamSynchronize(ctx->stream));
  HCK(hipGraphExecDestroy(ctx->user_graphs[graph]));
  ctx->user_graphs[graph] = nullptr;
  return MLP_OK;
}

int mlp_set_neighborhoods(mlp_ctx* ctx, int32_t count) {
  if (!ctx) return MLP_EINVAL;
  if (count < MLP_NB_COUNT || count > MLP_NB_MAX)
    return fail(ctx, MLP_EINVAL, "neighborhood count must be in [%d, %d]", MLP_NB_COUNT, MLP_NB_MAX);
  ctx->nbs = count;
  return drop_dvnd_graph(ctx);
}

int mlp_dvnd_neighborhoods(mlp_ctx* ctx, int32_t count) {
  if (!ctx) return MLP_EINVAL;
  if (count < 1 || count > MLP_NB_MAX)
    return fail(ctx, MLP_EINVAL, "neighborhood count must be in [1, %d]", MLP_NB_MAX);
  ctx->dvnd_nbs = count;
  return drop_dvnd_graph(ctx);
}

int mlp_slot_bytes(mlp_ctx* ctx, int32_t n, int64_t* bytes) {
  if (!ctx || !bytes) return MLP_EINVAL;
  if (n < 2 || n + 1 > MLP_MAX_N) return fail(ctx, MLP_ERANGE, "n = %d out of range", n);
  *bytes = (int64_t)slot_layout(n + 1, ctx->nbs).total;
  return MLP_OK;
}

int mlp_merge_force_global(mlp_ctx* ctx, int32_t on) {
  if (!ctx) return MLP_EINVAL;
  ctx->merge_force_global = on != 0;
  return drop_dvnd_graph(ctx);
}

int mlp_autotune(mlp_ctx* ctx, const int32_t* slots, int32_t count, uint32_t nb_mask,
                 int32_t reps, int32_t* best_tile, float* us_per_launch) {
  NO_CAPTURE(ctx);
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  if (nb_mask == 0 || (nb_mask & ~0x1fu)) return fail(ctx, MLP_EINVAL, "bad neighborhood mask");
  if (count == 0 || reps <= 0) return fail(ctx, MLP_EINVAL, "nothing to tune");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HCK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return fail(ctx, MLP_EDEVICE, "event");
  }
  const int saved = ctx->tile_override;
  float best = 0;
  int bt = -1;
  for (int t = 0; t < MLP_TILE_COUNT && st == MLP_OK; ++t) {
    ctx->tile_override = t;
    st = eval_impl(ctx, slots, count, nb_mask, 0, 0, 0, nullptr);  // warm-up
    if (st) break;
    hipError_t e = hipEventRecord(e0, ctx->stream);
    for (int r = 0; r < reps && st == MLP_OK && e == hipSuccess; ++r)
      st = eval_impl(ctx, slots, count, nb_mask, 0, 0, 0, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) st = fail(ctx, MLP_EDEVICE, "autotune: %s", hipGetErrorString(e));
    const float us = ms * 1e3f / (float)reps;
    if (us_per_launch) us_per_launch[t] = us;
    if (bt < 0 || us < best) {
      best = us;
      bt = t;
    }
  }
  ctx->tile_override = saved;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (st) return st;
  ctx->tune[nb_mask][work_bucket(launch_moves(ctx, slots, count, nb_mask))] = (int8_t)bt;
  if (best_tile) *best_tile = bt;
  return drop_dvnd_graph(ctx);
}

int mlp_tune_get(mlp_ctx* ctx, int8_t* table) {
  if (!ctx || !table) return MLP_EINVAL;
  memcpy(table, ctx->tune, sizeof ctx->tune);
  return MLP_OK;
}

int mlp_tune_set(mlp_ctx* ctx, const int8_t* table) {
  if (!ctx || !table) return MLP_EINVAL;
  for (size_t q = 0; q < sizeof ctx->tune; ++q)
    if (table[q] < -1 || table[q] >= MLP_TILE_COUNT)
      return fail(ctx, MLP_EINVAL, "tune entry %zu = %d is not a tile shape", q, (int)table[q]);
  memcpy(ctx->tune, table, sizeof ctx->tune);
  return drop_dvnd_graph(ctx);
}

int mlp_set_tile(mlp_ctx* ctx, int32_t tile) {
  if (!ctx) return MLP_EINVAL;
  if (tile < -1 || tile >= MLP_TILE_COUNT)
    return fail(ctx, MLP_EINVAL, "tile must be -1 (auto) or a shape 0..%d", MLP_TILE_COUNT - 1);
  ctx->tile_override = tile;
  return MLP_OK;
}

int mlp_timing_begin(mlp_ctx* ctx, int32_t capacity) {
  NO_CAPTURE(ctx);
  if (!ctx || capacity <= 0) return MLP_EINVAL;
  HCK(hipStreamSynchronize(ctx->stream));
  while (ctx->tev.size() < 2 * (size_t)capacity) {
    hipEvent_t e;
    HCK(hipEventCreate(&e));
    ctx->tev.push_back(e);
  }
  ctx->tcount = 0;
  ctx->timing = true;
  return MLP_OK;
}

int mlp_timing_end(mlp_ctx* ctx, double* total_ms, int32_t* launches) {
  NO_CAPTURE(ctx);
  if (!ctx) return MLP_EINVAL;
  HCK(hipStreamSynchronize(ctx->stream));
  double tot = 0;
  for (size_t q = 0; q < ctx->tcount; ++q) {
    float ms = 0;
    HCK(hipEventElapsedTime(&ms, ctx->tev[2 * q], ctx->tev[2 * q + 1]));
    tot += ms;
  }
  if (total_ms) *total_ms = tot;
  if (launches) *launches = (int32_t)ctx->tcount;
  ctx->timing = false;
  ctx->tcount = 0;
  return MLP_OK;
}

int mlp_region_begin(mlp_ctx* ctx) {
  if (!ctx) return MLP_EINVAL;
  HCK(hipEventRecord(ctx->ev0, ctx->stream));
  ctx->region_open = true;
  return MLP_OK;
}

int mlp_region_end(mlp_ctx* ctx, float* ms) {
  NO_CAPTURE(ctx);
  if (!ctx || !ms) return MLP_EINVAL;
  if (!ctx->region_open) return fail(ctx, MLP_ESTATE, "no timed region open");
  HCK(hipEventRecord(ctx->ev1, ctx->stream));
  HCK(hipEventSynchronize(ctx->ev1));
  HCK(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  ctx->region_open = false;
  return MLP_OK;
}

int mlp_eval_each(mlp_ctx* ctx, const int32_t* slots, int32_t count, uint32_t nb_mask,
                  int32_t rowmap) {
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  if (rowmap != 0 && rowmap != 1) return fail(ctx, MLP_EINVAL, "rowmap must be 0 or 1");
  st = check_mask(ctx, slots, count, nb_mask);
  if (st) return st;
  for (int q = 0; q < count; ++q) {
    st = eval_impl(ctx, slots + q, 1, nb_mask, rowmap, 0, 0, nullptr);
    if (st) return st;
  }
  return MLP_OK;
}

int mlp_launch_probe(mlp_ctx* ctx, int32_t slot, int32_t kind, int32_t grid, int32_t reps) {
  NO_CAPTURE(ctx);
  if (!ctx) return MLP_EINVAL;
  if (!slot_ok(ctx, slot)) return fail(ctx, MLP_EINVAL, "slot %d does not exist", slot);
  if (kind < 0 || kind > 3 || grid < 1 || grid > 65535 || reps < 0)
    return fail(ctx, MLP_EINVAL, "bad probe arguments");
  int st = ensure_i64(ctx, 2 * 8192);
  if (st) return st;
  for (int r = 0; r < reps; ++r)
    HCK(mlp_launch_probe(ctx->stream, ctx->d_table, slot, kind, grid, ctx->d_i64));
  return MLP_OK;
}

int mlp_copy_probe(mlp_ctx* ctx, int64_t bytes, int32_t reps) {
  NO_CAPTURE(ctx);
  if (!ctx || bytes <= 0 || reps <= 0) return MLP_EINVAL;
  void *a = nullptr, *b = nullptr;
  HCK(hipMalloc(&a, (size_t)bytes));
  if (hipMalloc(&b, (size_t)bytes) != hipSuccess) {
    hipFree(a);
    return fail(ctx, MLP_ENOMEM, "copy probe buffers");
  }
  hipError_t e = hipMemsetAsync(a, 1, (size_t)bytes, ctx->stream);
  for (int r = 0; r < reps && e == hipSuccess; ++r)
    e = hipMemcpyAsync(b, a, (size_t)bytes, hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(a);
  hipFree(b);
  if (e != hipSuccess) return fail(ctx, MLP_EDEVICE, "copy probe: %s", hipGetErrorString(e));
  return MLP_OK;
}

int mlp_get_rows(mlp_ctx* ctx, int32_t slot, int32_t nb, int64_t* cost, int32_t* i, int32_t* j,
                 int32_t* rows) {
  NO_CAPTURE(ctx);
  if (!ctx) return MLP_EINVAL;
  if (!slot_ok(ctx, slot)) return fail(ctx, MLP_EINVAL, "slot %d does not exist", slot);
  if (nb < 0 || nb >= MLP_NB_MAX) return fail(ctx, MLP_EINVAL, "bad neighborhood %d", nb);
  const Slot& s = ctx->slots[slot];
  const int rm = s.eval_rowmap[nb];
  if (rm < 0) return fail(ctx, MLP_ESTATE, "neighborhood %d of slot %d not evaluated", nb, slot);
  const int R = std::max(0, rows_of(nb, rm, s.N));
  if (rows) *rows = R;
  if (R == 0) return MLP_OK;
  std::vector<int64_t> keys((size_t)R);
  HCK(hipStreamSynchronize(ctx->stream));
  int64_t par = 0;
  HCK(hipMemcpy(&par, s.dev.info + MLP_INFO_PARITY0 + nb, 8, hipMemcpyDeviceToHost));
  HCK(hipMemcpy(keys.data(), s.dev.rows[nb] + (par ? s.dev.rowcap : 0), 8 * (size_t)R,
                hipMemcpyDeviceToHost));
  for (int r = 0; r < R; ++r) {
    const int64_t k = keys[(size_t)r];
    int ii = 0, jj = 0;
    int64_t c = INT64_MAX;
    if (k != MLP_KEY_NONE) {
      c = k >> MLP_KEY_SHIFT;
      move_ij(nb, rm, s.N, r, (int)(k & MLP_KEY_MASK), ii, jj);
    }
    if (cost) cost[r] = c;
    if (i) i[r] = ii;
    if (j) j[r] = jj;
  }
  return MLP_OK;
}

int mlp_get_moves64(mlp_ctx* ctx, int32_t slot, int32_t nb, uint64_t* moves, int32_t* rows) {
  NO_CAPTURE(ctx);
  if (!ctx || !moves) return MLP_EINVAL;
  if (!slot_ok(ctx, slot)) return fail(ctx, MLP_EINVAL, "slot %d does not exist", slot);
  if (ctx->slots[slot].N > 16384) return fail(ctx, MLP_ERANGE, "N > 16384 (14-bit i/j of MLMove64)");
  std::vector<int64_t> c(MLP_MAX_N);
  std::vector<int32_t> ii(MLP_MAX_N), jj(MLP_MAX_N);
  int R = 0;
  int st = mlp_get_rows(ctx, slot, nb, c.data(), ii.data(), jj.data(), &R);
  if (st) return st;
  for (int r = 0; r < R; ++r) {
    int64_t cost = c[(size_t)r];
    int32_t c32;
    uint32_t lo;
    if (cost == INT64_MAX) {  // empty row: COST_INFTY, move word 0 (mlkswap.cu:175-176)
      c32 = 0x7fffffff;
      lo = 0;
    } else {
      if (cost < INT32_MIN || cost >= INT32_MAX)
        return fail(ctx, MLP_ERANGE, "row %d cost %lld does not fit int32", r, (long long)cost);
      c32 = (int32_t)cost;
      lo = ((uint32_t)jj[(size_t)r] << 18) | ((uint32_t)ii[(size_t)r] << 4) | (uint32_t)nb;
    }
    moves[r] = ((uint64_t)(uint32_t)c32 << 32) | lo;
  }
  if (rows) *rows = R;
  return MLP_OK;
}

// tail > 0: the last merge of a DVND round, which also ends the round
// (max_rounds = tail - 1).
static int merge_impl(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t nb, int skip,
                      int tail = 0) {
  if (count == 0) return MLP_OK;
  int rowmap = -1, pcap = 1, ncap = 1;
  std::vector<int32_t> in_lds, in_gmem;
  for (int q = 0; q < count; ++q) {
    Slot& s = ctx->slots[slots[q]];
    const int rm = s.eval_rowmap[nb];
    if (rm < 0)
      return fail(ctx, MLP_ESTATE, "neighborhood %d of slot %d not evaluated", nb, slots[q]);
    if (s.merged[nb])
      return fail(ctx, MLP_ESTATE, "neighborhood %d of slot %d already merged", nb, slots[q]);
    if (rowmap >= 0 && rm != rowmap)
      return fail(ctx, MLP_EINVAL, "slots evaluated with different row maps");
    rowmap = rm;
    // a slot whose working arrays fit LDS joins the LDS launch
    if (!ctx->merge_force_global && merge_bytes(s.N) <= ctx->max_merge_lds) {
      in_lds.push_back(slots[q]);
      pcap = std::max(pcap, merge_pcap(s.N));
      ncap = std::max(ncap, s.N);
    } else {
      in_gmem.push_back(slots[q]);
    }
  }
  if (!in_lds.empty() && mlp_merge_lds_bytes(pcap, ncap) > ctx->max_merge_lds) {
    // mixed sizes whose joint capacities exceed LDS: largest ones go global
    in_gmem.insert(in_gmem.end(), in_lds.begin(), in_lds.end());
    in_lds.clear();
  }
  for (int q = 0; q < count; ++q) ctx->slots[slots[q]].merged[nb] = true;
  if (ctx->user_capture)
    for (int q = 0; q < count; ++q) ctx->capture_ops.push_back({slots[q], 1, (uint32_t)nb, rowmap});
  int gmem_n = 1;
  for (int32_t id : in_gmem) gmem_n = std::max(gmem_n, ctx->slots[id].N);
  for (int g = 0; g < 2; ++g) {
    const std::vector<int32_t>& list = g ? in_gmem : in_lds;
    for (size_t q0 = 0; q0 < list.size(); q0 += 65535) {
      const int c = (int)std::min<size_t>(65535, list.size() - q0);
      SlotListInline inl;
      const int32_t* dev = nullptr;
      int st = slot_list(ctx, list.data() + q0, c, inl, &dev);
      if (st) return st;
      if (g)  // working arrays in device memory: the O(N) steps over many workgroups
        HCK(mlp_launch_merge_gmem(ctx->stream, ctx->d_table, inl, dev, c, nb, rowmap, skip, tail,
                                  gmem_n));
      else
        HCK(mlp_launch_merge_apply(ctx->stream, ctx->d_table, inl, dev, c, nb, rowmap, pcap, ncap,
                                   skip, tail));
    }
  }
  return MLP_OK;
}

int mlp_merge_apply(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t nb) {
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  if (nb < 0 || nb >= MLP_NB_MAX) return fail(ctx, MLP_EINVAL, "bad neighborhood %d", nb);
  return merge_impl(ctx, slots, count, nb, 0);
}

int mlp_get_applied(mlp_ctx* ctx, int32_t slot, int32_t nb, mlp_move* moves, int32_t cap,
                    int32_t* count) {
  NO_CAPTURE(ctx);
  if (!ctx || !count) return MLP_EINVAL;
  if (!slot_ok(ctx, slot)) return fail(ctx, MLP_EINVAL, "slot %d does not exist", slot);
  if (nb < 0 || nb >= MLP_NB_MAX) return fail(ctx, MLP_EINVAL, "bad neighborhood %d", nb);
  const Slot& s = ctx->slots[slot];
  if (!s.merged[nb]) return fail(ctx, MLP_ESTATE, "neighborhood %d of slot %d not merged", nb, slot);
  HCK(hipStreamSynchronize(ctx->stream));
  int64_t K = 0;
  HCK(hipMemcpy(&K, s.dev.info + MLP_INFO_KEPT0 + nb, 8, hipMemcpyDeviceToHost));
  *count = (int32_t)K;
  const int take = (int)std::min<int64_t>(K, cap);
  if (take > 0 && moves) {
    std::vector<int64_t> c((size_t)take);
    std::vector<int32_t> ij(3 * (size_t)take);
    HCK(hipMemcpy(c.data(), s.dev.kept_cost[nb], 8 * (size_t)take, hipMemcpyDeviceToHost));
    HCK(hipMemcpy(ij.data(), s.dev.kept_ij[nb], 12 * (size_t)take, hipMemcpyDeviceToHost));
    for (int q = 0; q < take; ++q)
      moves[q] = mlp_move{c[(size_t)q], nb, ij[3 * (size_t)q], ij[3 * (size_t)q + 1],
                          ij[3 * (size_t)q + 2]};
  }
  return MLP_OK;
}

// ---- TOTAL mode ------------------------------------------------------------

int mlp_eval_total(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t nb,
                   int32_t block) {
  NO_CAPTURE(ctx);
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  if (nb < 0 || nb >= MLP_NB_MAX) return fail(ctx, MLP_EINVAL, "bad neighborhood %d", nb);
  if (block < 1 || block > MLP_TOTAL_MAX_BLOCK)
    return fail(ctx, MLP_EINVAL, "block %d outside 1..%d", block, MLP_TOTAL_MAX_BLOCK);
  if ((size_t)block * 8 > ctx->max_total_lds)
    return fail(ctx, MLP_ERANGE, "block %d needs %zu bytes of LDS (limit %zu)", block,
                (size_t)block * 8, ctx->max_total_lds);
  for (int q = 0; q < count; ++q)
    if (nb >= ctx->slots[slots[q]].nbs)
      return fail(ctx, MLP_EINVAL, "slot %d has no buffers for neighborhood %d", slots[q], nb);
  for (int q = 0; q < count; ++q) {
    const int id = slots[q];
    Slot& s = ctx->slots[id];
    const int R = std::max(0, rows_of(nb, 0, s.N));
    const int64_t M = (int64_t)R * block;
    if (M > s.total_cap) {
      pool_release(ctx, s.total, s.total_cls);
      s.total = nullptr;
      s.total_cap = 0;
      void* p = nullptr;
      st = pool_alloc(ctx, 8 * (size_t)M, &p, &s.total_cls);
      if (st) return st;
      s.total = (int64_t*)p;
      s.total_cap = (int64_t)(s.total_cls / 8);
    }
    s.total_nb = nb;
    s.total_block = block;
    s.total_rows = R;
    s.total_merged = false;
    if (R > 0) HCK(mlp_launch_total_eval(ctx->stream, ctx->d_table, id, nb, R, block, s.total));
  }
  return MLP_OK;
}

int mlp_get_total(mlp_ctx* ctx, int32_t slot, int64_t* cost, int32_t* i, int32_t* j, int64_t cap,
                  int64_t* entries) {
  NO_CAPTURE(ctx);
  if (!ctx) return MLP_EINVAL;
  if (!slot_ok(ctx, slot)) return fail(ctx, MLP_EINVAL, "slot %d does not exist", slot);
  const Slot& s = ctx->slots[slot];
  if (s.total_nb < 0) return fail(ctx, MLP_ESTATE, "slot %d has no TOTAL evaluation", slot);
  const int64_t M = (int64_t)s.total_rows * s.total_block;
  if (entries) *entries = M;
  const int64_t take = std::min<int64_t>(M, std::max<int64_t>(cap, 0));
  if (take == 0) return MLP_OK;
  std::vector<int64_t> keys((size_t)take);
  HCK(hipStreamSynchronize(ctx->stream));
  HCK(hipMemcpy(keys.data(), s.total, 8 * (size_t)take, hipMemcpyDeviceToHost));
  for (int64_t e = 0; e < take; ++e) {
    const int64_t k = keys[(size_t)e];
    int ii = 0, jj = 0;
    int64_t c = INT64_MAX;
    if (k != MLP_KEY_NONE) {
      c = k >> MLP_KEY_SHIFT;
      move_ij(s.total_nb, 0, s.N, (int)(e / s.total_block), (int)(k & MLP_KEY_MASK), ii, jj);
    }
    if (cost) cost[e] = c;
    if (i) i[e] = ii;
    if (j) j[e] = jj;
  }
  return MLP_OK;
}

int mlp_merge_total(mlp_ctx* ctx, const int32_t* slots, int32_t count) {
  NO_CAPTURE(ctx);
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  for (int q = 0; q < count; ++q) {
    const Slot& s = ctx->slots[slots[q]];
    if (s.total_nb < 0)
      return fail(ctx, MLP_ESTATE, "slot %d has no TOTAL evaluation", slots[q]);
    if (s.total_merged)
      return fail(ctx, MLP_ESTATE, "TOTAL evaluation of slot %d already merged", slots[q]);
  }
  for (int q = 0; q < count; ++q) {
    const int id = slots[q];
    Slot& s = ctx->slots[id];
    const int N = s.N, nb = s.total_nb, B = s.total_block;
    const int64_t M = (int64_t)s.total_rows * B;
    // every |delta| <= N (N + 32) maxd (2-opt: N (Ttot + 6 maxd) + 3 Ttot,
    // or-opt-k, k <= 13: k Ttot + (k + 5) N maxd, Ttot <= (N - 1) maxd)
    const uint64_t bias = (uint64_t)N * (uint64_t)(N + 32) * (uint64_t)s.dev.maxd;
    int end_bit = 1;
    while (end_bit < 64 && (bias >> end_bit) != 0) ++end_bit;
    // tour copies in LDS unless they (with the kernel's static LDS) exceed
    // it, or global arrays are forced
    const int gmem = ctx->merge_force_global ||
                     mlp_total_merge_lds_bytes(N) + mlp_total_static_lds() > ctx->max_total_lds;
    if (M == 0) {
      HCK(mlp_launch_total_merge(ctx->stream, ctx->d_table, id, nb, N, B, 0, bias, nullptr,
                                 nullptr, gmem));
    } else {
      void *ku = nullptr, *ko = nullptr, *vi = nullptr, *vo = nullptr, *tmp = nullptr;
      size_t cku = 0, cko = 0, cvi = 0, cvo = 0, ctmp = 0, tbytes = 0;
      HCK(mlp_sort_pairs_u64(nullptr, &tbytes, nullptr, nullptr, nullptr, nullptr, (int)M, end_bit,
                             ctx->stream));
      st = pool_alloc(ctx, 8 * (size_t)M, &ku, &cku);
      if (!st) st = pool_alloc(ctx, 8 * (size_t)M, &ko, &cko);
      if (!st) st = pool_alloc(ctx, 4 * (size_t)M, &vi, &cvi);
      if (!st) st = pool_alloc(ctx, 4 * (size_t)M, &vo, &cvo);
      if (!st) st = pool_alloc(ctx, std::max<size_t>(tbytes, 1), &tmp, &ctmp);
      hipError_t e = hipSuccess;
      if (!st) {
        e = mlp_launch_total_keys(ctx->stream, s.total, M, B, bias, (uint64_t*)ku,
                                  (int32_t*)vi);
        if (e == hipSuccess)
          e = mlp_sort_pairs_u64(tmp, &tbytes, (const uint64_t*)ku, (uint64_t*)ko,
                                 (const int32_t*)vi, (int32_t*)vo, (int)M, end_bit, ctx->stream);
        if (e == hipSuccess)
          e = mlp_launch_total_merge(ctx->stream, ctx->d_table, id, nb, N, B, M, bias,
                                     (const uint64_t*)ko, (const int32_t*)vo, gmem);
      }
      pool_release(ctx, ku, cku);  // stream-ordered: reused only by later work
      pool_release(ctx, ko, cko);
      pool_release(ctx, vi, cvi);
      pool_release(ctx, vo, cvo);
      pool_release(ctx, tmp, ctmp);
      if (st) return st;
      HCK(e);
    }
    s.total_merged = true;
    s.merged[nb] = true;  // mlp_get_applied reads the kept moves
  }
  return MLP_OK;
}

namespace {

// One DVND round over `slots`: round begin, {eval, merge + apply} for each
// neighborhood in order, round end.  Slots that stopped are skipped by every
// kernel, so the same launches serve all rounds of a chunk.
int enqueue_round(mlp_ctx* ctx, const int32_t* slots, int count, int rowmap, int max_rounds) {
  SlotListInline inl;
  const int32_t* dev = nullptr;
  int st = slot_list(ctx, slots, count, inl, &dev);
  if (st) return st;
  // the persistent grids' work counters are cleared at the round's start;
  // the round's moves are counted from zero since the previous round's end
  if (ctx->dvnd_persistent)
    HCK(mlp_launch_round_begin(ctx->stream, ctx->d_table, inl, dev, count, ctx->d_counter + 1,
                               ctx->dvnd_nbs));
  int last = -1;
  for (int nb = 0; nb < ctx->dvnd_nbs; ++nb)
    for (int q = 0; q < count && last != nb; ++q)
      if (rows_of(nb, rowmap, ctx->slots[slots[q]].N) > 0) last = nb;
  for (int nb = 0; nb <= last; ++nb) {
    bool any = false;
    for (int q = 0; q < count && !any; ++q)
      if (rows_of(nb, rowmap, ctx->slots[slots[q]].N) > 0) any = true;
    if (!any) continue;
    st = eval_impl(ctx, slots, count, 1u << nb, rowmap, ctx->dvnd_persistent, 1,
                   ctx->d_counter + 1 + nb);
    if (st) return st;
    // the round's last merge also ends the round for each slot
    st = merge_impl(ctx, slots, count, nb, 1, nb == last ? max_rounds + 1 : 0);
    if (st) return st;
  }
  if (last < 0) HCK(mlp_launch_round_end(ctx->stream, ctx->d_table, inl, dev, count, max_rounds));
  return MLP_OK;
}

// The fused round kernel applies: ADS row map, the five neighborhoods, every
// slot's merge in LDS, tours of at least 8 positions.
bool fused_ok(const mlp_ctx* ctx, const int32_t* slots, int count, int rowmap, int* pcap,
              int* ncap) {
  if (!ctx->dvnd_fused || rowmap != 0 || ctx->dvnd_nbs != MLP_NB_COUNT || ctx->merge_force_global)
    return false;
  int p = 1, n = 1;
  for (int q = 0; q < count; ++q) {
    const int N = ctx->slots[slots[q]].N;
    if (N < 8) return false;
    p = std::max(p, merge_pcap(N));
    n = std::max(n, N);
  }
  if (mlp_merge_lds_bytes(p, n) > ctx->max_merge_lds) return false;
  *pcap = p;
  *ncap = n;
  return true;
}

// One fused launch of `rounds` DVND rounds over the slots.
int enqueue_fused(mlp_ctx* ctx, const int32_t* slots, int count, int max_rounds, int rounds,
                  int pcap, int ncap) {
  constexpr int R = MLP_FUSED_R, C = MLP_FUSED_C;
  // MLP_FUSED_G tiles per item: a workgroup takes them with one work-counter
  // atomic, one dependency wait and one completion signal
  constexpr int G = MLP_FUSED_G;
  std::vector<int32_t> prefix((size_t)MLP_NB_COUNT * (count + 1)), ncols((size_t)2 * count),
      ntiles((size_t)MLP_NB_COUNT * count);
  FusedPlan P{};
  P.count = count;
  P.group = G;
  P.rounds = rounds;
  P.tail = max_rounds + 1;
  P.pcap = pcap;
  P.ncap = ncap;
  int at = 0;
  for (int nb = 0; nb < MLP_NB_COUNT; ++nb) {
    P.phase[2 * nb] = at;
    int32_t* pre = prefix.data() + (size_t)nb * (count + 1);
    pre[0] = 0;
    for (int q = 0; q < count; ++q) {
      const int N = ctx->slots[slots[q]].N;
      const int nc = (N - 2 + C - 1) / C, nc2 = (N - 3 + C - 1) / C;
      ncols[2 * (size_t)q] = nc;
      ncols[2 * (size_t)q + 1] = nc2;
      const int rows = rows_of(nb, 0, N);
      int tiles = 0;
      if (rows > 0) tiles = nb == 1 ? (rows + MLP_2OPT_ROWS - 1) / MLP_2OPT_ROWS * nc2
                                    : (rows + R - 1) / R * nc;
      ntiles[(size_t)nb * count + q] = tiles;
      pre[q + 1] = pre[q] + (tiles + G - 1) / G;
    }
    at += pre[count];
    P.phase[2 * nb + 1] = at;
    at += count;
  }
  P.phase[10] = at;
  P.per_round = at;
  int st = device_list(ctx, slots, count, &P.slots);
  if (st == MLP_OK) st = device_list(ctx, prefix.data(), (int)prefix.size(), &P.prefix);
  if (st == MLP_OK) st = device_list(ctx, ncols.data(), (int)ncols.size(), &P.ncols);
  if (st == MLP_OK) st = device_list(ctx, ntiles.data(), (int)ntiles.size(), &P.tiles);
  if (st) return st;
  if (count > ctx->sync_cap) {  // counters start at zero; each launch leaves them at zero
    if (ctx->capturing) return fail(ctx, MLP_ESTATE, "fused round counters grown during a capture");
    HCK(hipStreamSynchronize(ctx->stream));
    const int cap = std::max(count, 2 * ctx->sync_cap + 64);
    if (ctx->d_sync) HCK(hipFree(ctx->d_sync));
    ctx->d_sync = nullptr;
    ctx->sync_cap = 0;
    HCK(hipMalloc((void**)&ctx->d_sync, sizeof(int32_t) * (4 + 2 * (size_t)cap)));
    HCK(hipMemset(ctx->d_sync, 0, sizeof(int32_t) * (4 + 2 * (size_t)cap)));
    ctx->sync_cap = cap;
  }
  P.sync = ctx->d_sync;
  const size_t lds = mlp_fused_lds_bytes(pcap, ncap) + 4096;
  const int per_cu = std::max(1, std::min(4, (int)((160 * 1024) / lds)));
  const int grid = std::max(1, std::min(P.per_round * rounds, ctx->cu_count * per_cu));
  HCK(mlp_launch_fused(ctx->stream, ctx->d_table, P, grid));
  for (int q = 0; q < count; ++q) {
    Slot& s = ctx->slots[slots[q]];
    s.parity_stale = true;
    for (int nb = 0; nb < MLP_NB_COUNT; ++nb) {
      s.eval_rowmap[nb] = 0;
      s.merged[nb] = true;
    }
  }
  return MLP_OK;
}

// `rounds` DVND rounds over `slots`, as one captured graph when enabled.  A
// chunk is captured the second time the same (slots, settings) is seen, so
// every device list it needs already exists (no copies inside a capture).
int enqueue_rounds(mlp_ctx* ctx, const int32_t* slots, int count, int rowmap, int max_rounds,
                   int rounds) {
  int pcap = 0, ncap = 0;
  const bool fused = fused_ok(ctx, slots, count, rowmap, &pcap, &ncap);
  for (int r = 0; r < rounds; ++r) {
    const int st = fused ? enqueue_fused(ctx, slots, count, max_rounds, 1, pcap, ncap)
                         : enqueue_round(ctx, slots, count, rowmap, max_rounds);
    if (st) return st;
  }
  return MLP_OK;
}

int run_chunk(mlp_ctx* ctx, const int32_t* slots, int count, int rowmap, int max_rounds,
              int rounds, int gc = 0) {
  if (!ctx->dvnd_graph || ctx->user_capture)
    return enqueue_rounds(ctx, slots, count, rowmap, max_rounds, rounds);
  int tile = ctx->tile_override;  // (the per-neighborhood shapes follow from the slots)
  // the captured launches embed the slot table address, every slot's id and
  // size (segment tables, merge capacities) and the settings
  const int32_t hdr[9] = {rowmap, max_rounds, rounds, ctx->dvnd_persistent, tile, count,
                          ctx->dvnd_nbs, ctx->merge_force_global ? 1 : 0, ctx->dvnd_fused};
  std::string key((const char*)hdr, sizeof hdr);
  key.append((const char*)&ctx->d_table, sizeof ctx->d_table);
  key.append((const char*)slots, 4 * (size_t)count);
  for (int q = 0; q < count; ++q) {
    const int32_t Nn[2] = {ctx->slots[slots[q]].N, ctx->slots[slots[q]].nbs};
    key.append((const char*)Nn, sizeof Nn);
  }
  if (key == ctx->graph_key[gc] && ctx->graph_exec[gc]) {
    HCK(hipGraphLaunch(ctx->graph_exec[gc], ctx->stream));
    // the launches' host bookkeeping (evaluated / merged flags)
    for (int q = 0; q < count; ++q) {
      Slot& s = ctx->slots[slots[q]];
      s.parity_stale = true;
      for (int nb = 0; nb < ctx->dvnd_nbs; ++nb) {
        s.eval_rowmap[nb] = rowmap;
        s.merged[nb] = true;
      }
    }
    return MLP_OK;
  }
  if (key != ctx->graph_key[gc]) {  // first sight: run eagerly, remember the key
    if (ctx->graph_exec[gc]) {
      HCK(hipGraphExecDestroy(ctx->graph_exec[gc]));
      ctx->graph_exec[gc] = nullptr;
    }
    ctx->graph_key[gc] = key;
    return enqueue_rounds(ctx, slots, count, rowmap, max_rounds, rounds);
  }
  hipGraph_t g = nullptr;
  HCK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
  ctx->capturing = true;
  int st = enqueue_rounds(ctx, slots, count, rowmap, max_rounds, rounds);
  ctx->capturing = false;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
  if (st || ec != hipSuccess) {  // nothing captured ran: run the chunk eagerly instead
    if (g) hipGraphDestroy(g);
    ctx->graph_key[gc].clear();
    return enqueue_rounds(ctx, slots, count, rowmap, max_rounds, rounds);
  }
  const hipError_t ei = hipGraphInstantiate(&ctx->graph_exec[gc], g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ei != hipSuccess) {
    ctx->graph_exec[gc] = nullptr;
    return fail(ctx, MLP_EDEVICE, "graph instantiate: %s", hipGetErrorString(ei));
  }
  HCK(hipGraphLaunch(ctx->graph_exec[gc], ctx->stream));
  return MLP_OK;
}

}  // namespace

namespace {

// Slots sorted by N (N^2 is the work of a round) from the largest, cut into
// `groups` groups at 40% of the work and then at even shares of the rest;
// slot order is kept inside each group, and every group has a slot.
int split_groups(mlp_ctx* ctx, const std::vector<int32_t>& active, int groups,
                 std::vector<int32_t>* g) {
  const int c = (int)active.size();
  std::vector<int> order((size_t)c);
  double total = 0;
  for (int q = 0; q < c; ++q) {
    order[(size_t)q] = q;
    const double N = ctx->slots[active[(size_t)q]].N;
    total += N * N;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return ctx->slots[active[(size_t)a]].N > ctx->slots[active[(size_t)b]].N;
  });
  std::vector<int> of((size_t)c, groups - 1);
  double acc = 0;
  int k = 0;
  for (int grp = 0; grp < groups - 1; ++grp) {
    const double cut = (0.4 + 0.6 * grp / (groups - 1)) * total;
    // at least one slot per group, and one left for each later group
    const int kmax = c - (groups - 1 - grp);
    const int k0 = k;
    while (k < kmax && (k == k0 || acc < cut)) {
      const double N = ctx->slots[active[(size_t)order[(size_t)k]]].N;
      acc += N * N;
      of[(size_t)order[(size_t)k]] = grp;
      ++k;
    }
  }
  for (int grp = 0; grp < groups; ++grp) g[grp].clear();
  for (int q = 0; q < c; ++q) g[of[(size_t)q]].push_back(active[(size_t)q]);
  return MLP_OK;
}

// The group streams and events of `groups` slot groups, created once.
int group_streams(mlp_ctx* ctx, int groups) {
  if (!ctx->ev_fork &&
      hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess)
    return fail(ctx, MLP_EDEVICE, "DVND group event: %s", hipGetErrorString(hipGetLastError()));
  for (int g = 0; g < groups; ++g) {
    if (g && !ctx->stream_grp[g] &&
        (hipStreamCreateWithFlags(&ctx->stream_grp[g], hipStreamNonBlocking) != hipSuccess ||
         hipEventCreateWithFlags(&ctx->ev_join[g], hipEventDisableTiming) != hipSuccess))
      return fail(ctx, MLP_EDEVICE, "DVND group stream: %s", hipGetErrorString(hipGetLastError()));
    if (!ctx->ev_chunk[g] &&
        hipEventCreateWithFlags(&ctx->ev_chunk[g], hipEventDisableTiming) != hipSuccess)
      return fail(ctx, MLP_EDEVICE, "DVND group event: %s", hipGetErrorString(hipGetLastError()));
  }
  return MLP_OK;
}

// `rounds` DVND rounds over the slots, split into groups on streams of
// their own when enabled (the fused and persistent forms share device
// counters across slots, so they keep one stream; so does a caller's
// capture).  Work queued after the call on the context stream follows every
// group.
int run_grouped(mlp_ctx* ctx, const int32_t* slots, int count, int rowmap, int max_rounds,
                int rounds) {
  const int groups = std::min(ctx->dvnd_groups, count);
  if (groups < 2 || ctx->dvnd_fused || ctx->dvnd_persistent || ctx->user_capture)
    return run_chunk(ctx, slots, count, rowmap, max_rounds, rounds);
  int st = group_streams(ctx, groups);
  if (st) return st;
  std::vector<int32_t> active(slots, slots + count), part[MLP_DVND_GROUPS];
  st = split_groups(ctx, active, groups, part);
  if (st) return st;
  // the group streams start after the work queued so far on the context stream
  HCK(hipEventRecord(ctx->ev_fork, ctx->stream_main));
  for (int g = 1; g < groups; ++g) HCK(hipStreamWaitEvent(ctx->stream_grp[g], ctx->ev_fork, 0));
  for (int g = 0; g < groups && !st; ++g) {
    ctx->stream = g ? ctx->stream_grp[g] : ctx->stream_main;
    st = run_chunk(ctx, part[g].data(), (int)part[g].size(), rowmap, max_rounds, rounds, g);
  }
  ctx->stream = ctx->stream_main;
  // joined also after a failed enqueue, so later work still follows it
  for (int g = 1; g < groups; ++g) {
    HCK(hipEventRecord(ctx->ev_join[g], ctx->stream_grp[g]));
    HCK(hipStreamWaitEvent(ctx->stream_main, ctx->ev_join[g], 0));
  }
  return st;
}

// mlp_dvnd over slot groups that advance independently: each group's chunk
// of rounds is followed by its own convergence check as soon as that chunk
// is done (the host polls the groups' chunk events), so a group of small
// slots does not wait at every check for the group of large ones.  The
// groups are fixed for the call; the context stream waits for every group
// at the end.
int dvnd_independent(mlp_ctx* ctx, const int32_t* slots, int count, int rowmap,
                     int max_rounds) {
  const int groups = std::min(ctx->dvnd_groups, count);
  int st = group_streams(ctx, groups);
  if (st) return st;
  std::vector<int32_t> active(slots, slots + count), part[MLP_DVND_GROUPS], flag;
  st = split_groups(ctx, active, groups, part);
  if (st) return st;
  HCK(hipEventRecord(ctx->ev_fork, ctx->stream_main));
  for (int g = 1; g < groups; ++g) HCK(hipStreamWaitEvent(ctx->stream_grp[g], ctx->ev_fork, 0));
  bool pending[MLP_DVND_GROUPS] = {};
  for (;;) {
    for (int g = 0; g < groups && !st; ++g)
      if (!pending[g] && !part[g].empty()) {
        ctx->stream = g ? ctx->stream_grp[g] : ctx->stream_main;
        st = run_chunk(ctx, part[g].data(), (int)part[g].size(), rowmap, max_rounds,
                       ctx->dvnd_chunk, g);
        if (!st && hipEventRecord(ctx->ev_chunk[g], ctx->stream) != hipSuccess)
          st = fail(ctx, MLP_EDEVICE, "DVND chunk event");
        pending[g] = !st;
      }
    int g = -1;
    for (int h = 0; h < groups && !st; ++h) g = pending[h] ? h : g;
    if (st || g < 0) break;  // an error, or every group converged
    // the first group whose chunk is done (a spin on the chunks' events that
    // yields the core between polls)
    for (g = -1; g < 0;) {
      for (int h = 0; h < groups && g < 0; ++h)
        if (pending[h]) {
          const hipError_t q = hipEventQuery(ctx->ev_chunk[h]);
          if (q == hipSuccess) g = h;
          else if (q != hipErrorNotReady) {
            ctx->stream = ctx->stream_main;
            return fail(ctx, MLP_EDEVICE, "DVND chunk: %s", hipGetErrorString(q));
          }
        }
      if (g < 0) std::this_thread::yield();
    }
    pending[g] = false;
    ctx->stream = g ? ctx->stream_grp[g] : ctx->stream_main;
    flag.resize(part[g].size());
    st = mlp_dvnd_status(ctx, part[g].data(), (int)part[g].size(), flag.data(), nullptr);
    std::vector<int32_t> next;
    for (size_t q = 0; q < part[g].size() && !st; ++q)
      if (flag[q]) next.push_back(part[g][q]);
    if (!st) part[g].swap(next);
  }
  ctx->stream = ctx->stream_main;
  // joined also after an error, so later work still follows every group
  for (int g = 1; g < groups; ++g) {
    HCK(hipEventRecord(ctx->ev_join[g], ctx->stream_grp[g]));
    HCK(hipStreamWaitEvent(ctx->stream_main, ctx->ev_join[g], 0));
  }
  return st;
}

}  // namespace

int mlp_dvnd_groups(mlp_ctx* ctx, int32_t groups) {
  if (!ctx) return MLP_EINVAL;
  if (groups < 1 || groups > MLP_DVND_GROUPS)
    return fail(ctx, MLP_EINVAL, "groups must be in [1, %d]", MLP_DVND_GROUPS);
  ctx->dvnd_groups = groups;
  return MLP_OK;
}

int mlp_dvnd_start(mlp_ctx* ctx, const int32_t* slots, int32_t count) {
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  for (int q0 = 0; q0 < count; q0 += 65535) {
    const int c = std::min(65535, count - q0);
    SlotListInline inl;
    const int32_t* dev = nullptr;
    st = slot_list(ctx, slots + q0, c, inl, &dev);
    if (st) return st;
    HCK(mlp_launch_dvnd_start(ctx->stream, ctx->d_table, inl, dev, c));
  }
  return MLP_OK;
}

int mlp_dvnd_rounds(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t rowmap,
                    int32_t rounds, int32_t max_rounds) {
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  st = check_mask(ctx, slots, count, (1u << ctx->dvnd_nbs) - 1u);
  if (st) return st;
  if (rowmap != 0 && rowmap != 1) return fail(ctx, MLP_EINVAL, "rowmap must be 0 or 1");
  if (rounds < 0) return fail(ctx, MLP_EINVAL, "rounds must be >= 0");
  if (count == 0 || rounds == 0) return MLP_OK;
  return run_grouped(ctx, slots, count, rowmap, max_rounds, rounds);
}

int mlp_dvnd_status(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t* active,
                    int32_t* rounds) {
  NO_CAPTURE(ctx);
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  if (count == 0) return MLP_OK;
  if (ctx->d_sync) {  // a fused round whose dependency wait gave up
    int32_t flag = 0;
    HCK(hipMemcpyAsync(&flag, ctx->d_sync + 2, 4, hipMemcpyDeviceToHost, ctx->stream));
    HCK(hipStreamSynchronize(ctx->stream));
    if (flag) return fail(ctx, MLP_EDEVICE, "a fused DVND round timed out waiting for its queue");
  }
  std::vector<int64_t> a((size_t)count), r((size_t)count);
  st = gather_info(ctx, slots, count, MLP_INFO_ACTIVE, a.data(), MLP_INFO_ROUNDS, r.data());
  if (st) return st;
  for (int q = 0; q < count; ++q) {
    if (active) active[q] = (int32_t)a[(size_t)q];
    if (rounds) rounds[q] = (int32_t)r[(size_t)q];
  }
  return MLP_OK;
}

int mlp_dvnd_config(mlp_ctx* ctx, int32_t chunk, int32_t use_graph, int32_t persistent) {
  if (!ctx) return MLP_EINVAL;
  if (chunk < 1 || chunk > 64) return fail(ctx, MLP_EINVAL, "chunk must be in [1, 64]");
  ctx->dvnd_chunk = chunk;
  ctx->dvnd_graph = use_graph ? 1 : 0;
  ctx->dvnd_persistent = persistent ? 1 : 0;
  return MLP_OK;
}

int mlp_dvnd_fused(mlp_ctx* ctx, int32_t on) {
  if (!ctx) return MLP_EINVAL;
  ctx->dvnd_fused = on ? 1 : 0;
  return drop_dvnd_graph(ctx);
}

int mlp_dvnd(mlp_ctx* ctx, const int32_t* slots, int32_t count, int32_t rowmap,
             int32_t max_rounds, int32_t* rounds_out) {
  NO_CAPTURE(ctx);
  int st = check_slots(ctx, slots, count);
  if (st) return st;
  if (rowmap != 0 && rowmap != 1) return fail(ctx, MLP_EINVAL, "rowmap must be 0 or 1");
  if (count == 0) return MLP_OK;
  st = check_mask(ctx, slots, count, (1u << ctx->dvnd_nbs) - 1u);
  if (st) return st;
  st = mlp_dvnd_start(ctx, slots, count);
  if (st) return st;
  std::vector<int32_t> active(slots, slots + count), flag((size_t)count);
  if (count >= 2 && ctx->dvnd_groups > 1 && !ctx->dvnd_fused && !ctx->dvnd_persistent) {
    st = dvnd_independent(ctx, slots, count, rowmap, max_rounds);
    if (st) return st;
    active.clear();
  }
  while (!active.empty()) {
    const int c = (int)active.size();
    st = run_grouped(ctx, active.data(), c, rowmap, max_rounds, ctx->dvnd_chunk);
    if (st) return st;
    st = mlp_dvnd_status(ctx, active.data(), c, flag.data(), nullptr);
    if (st) return st;
    std::vector<int32_t> next;
    for (int q = 0; q < c; ++q)
      if (flag[(size_t)q]) next.push_back(active[(size_t)q]);
    active.swap(next);
  }
  if (rounds_out) return mlp_dvnd_status(ctx, slots, count, nullptr, rounds_out);
  return MLP_OK;
}

int mlp_dev_alloc(mlp_ctx* ctx, int64_t bytes, void** dev) {
  NO_CAPTURE(ctx);
  if (!ctx || !dev || bytes < 0) return MLP_EINVAL;
  *dev = nullptr;
  HCK(hipSetDevice(ctx->device));
  void* p = nullptr;
  if (hipMalloc(&p, (size_t)std::max<int64_t>(bytes, 8)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(ctx, MLP_ENOMEM, "device allocation of %lld bytes failed", (long long)bytes);
  }
  ctx->dev_bufs.push_back(p);
  *dev = p;
  return MLP_OK;
}

int mlp_dev_free(mlp_ctx* ctx, void* dev) {
  NO_CAPTURE(ctx);
  if (!ctx) return MLP_EINVAL;
  if (!dev) return MLP_OK;
  auto it = std::find(ctx->dev_bufs.begin(), ctx->dev_bufs.end(), dev);
  if (it == ctx->dev_bufs.end()) return fail(ctx, MLP_EINVAL, "not a buffer of mlp_dev_alloc");
  ctx->dev_bufs.erase(it);
  HCK(hipStreamSynchronize(ctx->stream));
  HCK(hipFree(dev));
  return MLP_OK;
}

int mlp_dev_memset(mlp_ctx* ctx, void* dev, int32_t value, int64_t bytes) {
  NO_CAPTURE(ctx);  // a memset node in a captured graph is not replay-safe here
  if (!ctx || !dev || bytes < 0) return MLP_EINVAL;
  HCK(hipMemsetAsync(dev, value & 0xff, (size_t)bytes, ctx->stream));
  return MLP_OK;
}

int mlp_copy_to_host(mlp_ctx* ctx, void* host, const void* dev, int64_t bytes) {
  NO_CAPTURE(ctx);
  if (!ctx || bytes < 0 || (bytes && (!host || !dev))) return MLP_EINVAL;
  if (bytes == 0) return MLP_OK;
  HCK(hipMemcpyAsync(host, dev, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
  HCK(hipStreamSynchronize(ctx->stream));
  return MLP_OK;
}

int mlp_copy_to_device(mlp_ctx* ctx, void* dev, const void* host, int64_t bytes) {
  NO_CAPTURE(ctx);
  if (!ctx || bytes < 0 || (bytes && (!host || !dev))) return MLP_EINVAL;
  if (bytes == 0) return MLP_OK;
  HCK(hipMemcpyAsync(dev, host, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
  HCK(hipStreamSynchronize(ctx->stream));
  return MLP_OK;
}

int mlp_rows_moves64_device(mlp_ctx* ctx, int32_t slot, int32_t nb, uint64_t* dev, int64_t cap,
                            int32_t* rows) {
  if (!ctx || !dev) return MLP_EINVAL;
  if (!slot_ok(ctx, slot)) return fail(ctx, MLP_EINVAL, "slot %d does not exist", slot);
  if (nb < 0 || nb >= ctx->slots[slot].nbs) return fail(ctx, MLP_EINVAL, "bad neighborhood %d", nb);
  const Slot& s = ctx->slots[slot];
  if (s.N > 16384) return fail(ctx, MLP_ERANGE, "N > 16384 (14-bit i/j of MLMove64)");
  const int rm = s.eval_rowmap[nb];
  if (rm < 0) return fail(ctx, MLP_ESTATE, "neighborhood %d of slot %d not evaluated", nb, slot);
  const int R = std::max(0, rows_of(nb, rm, s.N));
  if (R > cap) return fail(ctx, MLP_EINVAL, "%d rows do not fit %lld entries", R, (long long)cap);
  if (rows) *rows = R;
  if (R > 0) HCK(mlp_launch_rows_moves64(ctx->stream, ctx->d_table, slot, nb, rm, R, dev));
  return MLP_OK;
}

int mlp_total_moves64_device(mlp_ctx* ctx, int32_t slot, uint64_t* dev, int64_t cap,
                             int64_t* entries) {
  if (!ctx || !dev) return MLP_EINVAL;
  if (!slot_ok(ctx, slot)) return fail(ctx, MLP_EINVAL, "slot %d does not exist", slot);
  const Slot& s = ctx->slots[slot];
  if (s.N > 16384) return fail(ctx, MLP_ERANGE, "N > 16384 (14-bit i/j of MLMove64)");
  if (s.total_nb < 0) return fail(ctx, MLP_ESTATE, "slot %d has no TOTAL evaluation", slot);
  const int64_t M = (int64_t)s.total_rows * s.total_block;
  if (M > cap) return fail(ctx, MLP_EINVAL, "%lld entries do not fit %lld", (long long)M, (long long)cap);
  if (entries) *entries = M;
  if (M > 0)
    HCK(mlp_launch_total_moves64(ctx->stream, s.total, M, s.total_block, s.total_nb, s.N, dev));
  return MLP_OK;
}

int mlp_timer_create(mlp_ctx* ctx, int32_t* timer) {
  if (!ctx || !timer) return MLP_EINVAL;
  hipEvent_t a = nullptr, b = nullptr;
  HCK(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    hipEventDestroy(a);
    return fail(ctx, MLP_EDEVICE, "hipEventCreate failed");
  }
  for (size_t q = 0; q < ctx->timers.size(); ++q)
    if (!ctx->timers[q].first) {
      ctx->timers[q] = {a, b};
      *timer = (int32_t)q;
      return MLP_OK;
    }
  ctx->timers.emplace_back(a, b);
  *timer = (int32_t)ctx->timers.size() - 1;
  return MLP_OK;
}

static bool timer_ok(mlp_ctx* ctx, int32_t t) {
  return t >= 0 && t < (int32_t)ctx->timers.size() && ctx->timers[(size_t)t].first;
}

int mlp_timer_start(mlp_ctx* ctx, int32_t timer) {
  if (!ctx || !timer_ok(ctx, timer)) return MLP_EINVAL;
  HCK(hipEventRecord(ctx->timers[(size_t)timer].first, ctx->stream));
  return MLP_OK;
}

int mlp_timer_stop(mlp_ctx* ctx, int32_t timer) {
  if (!ctx || !timer_ok(ctx, timer)) return MLP_EINVAL;
  HCK(hipEventRecord(ctx->timers[(size_t)timer].second, ctx->stream));
  return MLP_OK;
}

int mlp_timer_elapsed(mlp_ctx* ctx, int32_t timer, float* ms) {
  NO_CAPTURE(ctx);
  if (!ctx || !ms || !timer_ok(ctx, timer)) return MLP_EINVAL;
  HCK(hipEventSynchronize(ctx->timers[(size_t)timer].second));
  HCK(hipEventElapsedTime(ms, ctx->timers[(size_t)timer].first,
                          ctx->timers[(size_t)timer].second));
  return MLP_OK;
}

int mlp_timer_destroy(mlp_ctx* ctx, int32_t timer) {
  if (!ctx || !timer_ok(ctx, timer)) return MLP_EINVAL;
  auto& t = ctx->timers[(size_t)timer];
  hipEventDestroy(t.first);
  hipEventDestroy(t.second);
  t = {nullptr, nullptr};
  return MLP_OK;
}

int mlp_eval_static(mlp_ctx* ctx, int32_t inst, const int32_t* tour, uint32_t nb_mask,
                    int32_t rowmap, int64_t* cost, int32_t* is, int32_t* js) {
  if (!ctx) return MLP_EINVAL;
  if (inst < 0 || inst >= (int)ctx->insts.size() || !ctx->insts[inst].alive)
    return fail(ctx, MLP_EINVAL, "instance %d does not exist", inst);
  if (!tour || !cost || !is || !js) return fail(ctx, MLP_EINVAL, "NULL device array");
  if (rowmap != 0 && rowmap != 1) return fail(ctx, MLP_EINVAL, "rowmap must be 0 or 1");
  if (nb_mask == 0 || (nb_mask & ~0x1fu))
    return fail(ctx, MLP_EINVAL, "neighborhood mask must be a non-empty subset of 0x1f");
  if (!ctx->scratch)
    return fail(ctx, MLP_EDEVICE, "no static scratch address on device %d", ctx->device);
  const Instance& in = ctx->insts[inst];
  const int N = in.n + 1;
  if (N > MLP_STATIC_MAX_N)
    return fail(ctx, MLP_ERANGE, "N = %d exceeds MLP_STATIC_MAX_N = %d", N, MLP_STATIC_MAX_N);
  // the scratch holds the slot-table entry, then the slot's block up to its
  // merge arrays (no merge runs here)
  const SlotLayout L = slot_layout(N, MLP_NB_COUNT);
  const size_t head = align_up(sizeof(SlotDev), 256);
  if (head + L.ms > (size_t)MLP_STATIC_SCRATCH_BYTES)
    return fail(ctx, MLP_ERANGE, "N = %d does not fit the static scratch", N);
  char* base = (char*)ctx->scratch;
  SlotDev* table = (SlotDev*)base;
  char* b = base + head;
  SlotDev d{};
  d.N = N;
  d.n = in.n;
  d.round = in.round;
  d.maxd = in.maxd;
  d.rowcap = row_cap(N);
  d.xs = in.xs;
  d.ys = in.ys;
  d.tour = const_cast<int32_t*>(tour);  // read only: prep and the sweeps never write it
  d.xy = (int32_t*)(b + L.xy);
  d.w = (int32_t*)(b + L.w);
  d.t32 = (int32_t*)(b + L.t32);
  d.q32 = (int32_t*)(b + L.q32);
  d.g32 = (int32_t*)(b + L.g32);
  d.h32 = (int32_t*)(b + L.h32);
  d.t64 = (int64_t*)(b + L.t64);
  d.q64 = (int64_t*)(b + L.q64);
  d.info = (int64_t*)(b + L.info);
  d.nbs = MLP_NB_COUNT;
  SlotInitRows rows{};
  rows.nbs = d.nbs;
  for (int nb = 0; nb < d.nbs; ++nb) {
    d.rows[nb] = (int64_t*)(b + L.rows[nb]);
    d.kept_cost[nb] = (int64_t*)(b + L.kc[nb]);
    d.kept_ij[nb] = (int32_t*)(b + L.kij[nb]);
    rows.start[nb] = (int64_t)(L.rows[nb] / 8);
    rows.count[nb] = 2 * (int64_t)d.rowcap;
  }
  d.mscratch = nullptr;
  // the block is cleared by a kernel (info zero, row buffers KEY_NONE) and
  // the table entry written with it; then the ADS of the tour
  HCK(mlp_launch_slot_init(ctx->stream, table, 0, d, (int64_t*)b, (int64_t)(L.ms / 8), rows));
  SlotListInline one{};
  one.id[0] = 0;
  HCK(mlp_launch_prep(ctx->stream, table, one, nullptr, 1));
  // every neighborhood writes buffer 1 (the first tile records it)
  int total_rows = 0;
  for (int nb = 0; nb < MLP_NB_COUNT; ++nb)
    if (nb_mask & (1u << nb)) total_rows += std::max(0, rows_of(nb, rowmap, N));
  {
    const bool odc = rowmap == 1;  // fixed tiles folding into ODC rows
    int tile = ctx->tile_override;
    if (tile < 0) {
      tile = auto_tile_moves(ctx, tour_moves(N, nb_mask), 1, nb_mask);
      if (odc && N >= 512 && (nb_mask & 0x1cu)) tile = MLP_TILE_8x256;  // 16 x 128
    }
    if (odc && tile == MLP_TILE_16x1024) tile = MLP_TILE_8x512;
    int TR, TC;
    mlp_tile_shape(tile | (odc ? MLP_TILE_ODC : 0), &TR, &TC);
    SweepSegsInline segs{};
    int nsegs = 0, total = 0;
    const bool two = !odc && (nb_mask & 3u) == 3u && TR >= 8;
    if (two) mlp_tile_shape(tile | MLP_TILE_TWO, &TR, &TC);
    const uint32_t fixed = (nb_mask & 0x1du) | (two ? 2u : 0u);
    int rmax = 0;
    for (int nb : {0, 2, 3, 4})
      if (fixed & (1u << nb))
        rmax = std::max({rmax, rows_of(nb, 0, N), odc ? rows_of(nb, 1, N) : 0});
    if (two) rmax = std::max({rmax, N - 3, rows_of(1, 0, N)});
    if (fixed && rmax > 0) {
      const int ncols = (N - 2 + TC - 1) / TC;
      segs.s[nsegs++] = SweepSeg{0, 0, (int)fixed, total, ncols, 0x1f};
      total += (rmax + TR - 1) / TR * ncols;
    }
    if (two) tile |= MLP_TILE_TWO;
    if (odc) tile |= MLP_TILE_ODC;
    if ((nb_mask & 2u) && !two && rows_of(1, 0, N) > 0 && N >= 4) {
      const int ncols = (N - 3 + TC - 1) / TC;
      segs.s[nsegs++] = SweepSeg{0, 1, 2, total, ncols, 0x1f};
      total += (rows_of(1, 0, N) + MLP_2OPT_ROWS - 1) / MLP_2OPT_ROWS * ncols;
    }
    if (total > 0)
      HCK(mlp_launch_sweep(ctx->stream, tile, table, segs, nullptr, nsegs, total, nullptr, total, 0));
  }
  if (total_rows > 0)
    HCK(mlp_launch_rows_out(ctx->stream, table, 0, (int)nb_mask, rowmap, total_rows, cost, is,
                            js));
  return MLP_OK;
}

int mlp_ads_materialize(mlp_ctx* ctx, int32_t slot, uint32_t* host_out, int64_t bytes) {
  NO_CAPTURE(ctx);
  if (!ctx || !host_out) return MLP_EINVAL;
  if (!slot_ok(ctx, slot)) return fail(ctx, MLP_EINVAL, "slot %d does not exist", slot);
  const int N = ctx->slots[slot].N;
  const int64_t need = 128 + 4 * (int64_t)N * (N + 3);
  if (bytes < need) return fail(ctx, MLP_EINVAL, "ADS buffer needs %lld bytes", (long long)need);
  uint32_t *d = nullptr, *S = nullptr;
  HCK(hipMalloc(&d, (size_t)need));
  if (hipMalloc(&S, 4 * (size_t)N) != hipSuccess) {
    hipFree(d);
    return fail(ctx, MLP_ENOMEM, "ADS scratch");
  }
  hipError_t e = mlp_launch_ads(ctx->stream, ctx->d_table, slot, N, S, d);
  if (e == hipSuccess)
    e = hipMemcpyAsync(host_out, d, (size_t)need, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(S);
  hipFree(d);
  if (e != hipSuccess) return fail(ctx, MLP_EDEVICE, "ADS: %s", hipGetErrorString(e));
  return MLP_OK;
}

int mlp_dist_table(mlp_ctx* ctx, int32_t s0, int32_t count, int32_t round, int32_t* out) {
  NO_CAPTURE(ctx);
  if (!ctx || !out || count < 0 || s0 < 0) return MLP_EINVAL;
  if ((int64_t)s0 + count > INT32_MAX) return fail(ctx, MLP_ERANGE, "s range");
  int32_t* d = nullptr;
  HCK(hipMalloc(&d, std::max<size_t>(4, 4 * (size_t)count)));
  hipError_t e = mlp_launch_dist_table(ctx->stream, s0, count, round, d);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, d, 4 * (size_t)count, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return fail(ctx, MLP_EDEVICE, "dist table: %s", hipGetErrorString(e));
  return MLP_OK;
}

int mlp_sqrt_selfcheck(mlp_ctx* ctx, int64_t* mismatches) {
  NO_CAPTURE(ctx);
  if (!ctx || !mismatches) return MLP_EINVAL;
  unsigned long long* d = nullptr;
  HCK(hipMalloc(&d, 8));
  hipError_t e = hipMemsetAsync(d, 0, 8, ctx->stream);
  if (e == hipSuccess) e = mlp_launch_sqrt_selfcheck(ctx->stream, 0u, 0x80000000u, d);
  unsigned long long h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return fail(ctx, MLP_EDEVICE, "sqrt selfcheck: %s", hipGetErrorString(e));
  *mismatches = (int64_t)h;
  return MLP_OK;
}

int mlp_dist_list(mlp_ctx* ctx, const int32_t* s, int32_t count, int32_t round, int32_t* out) {
  NO_CAPTURE(ctx);
  if (!ctx || !out || !s || count < 0) return MLP_EINVAL;
  for (int q = 0; q < count; ++q)
    if (s[q] < 0) return fail(ctx, MLP_ERANGE, "negative squared length");
  int32_t *d = nullptr, *ds = nullptr;
  const size_t bytes = std::max<size_t>(4, 4 * (size_t)count);
  HCK(hipMalloc(&d, bytes));
  if (hipMalloc(&ds, bytes) != hipSuccess) {
    hipFree(d);
    return fail(ctx, MLP_ENOMEM, "dist list");
  }
  hipError_t e = hipMemcpyAsync(ds, s, 4 * (size_t)count, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = mlp_launch_dist_list(ctx->stream, ds, count, round, d);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, d, 4 * (size_t)count, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  hipFree(ds);
  if (e != hipSuccess) return fail(ctx, MLP_EDEVICE, "dist list: %s", hipGetErrorString(e));
  return MLP_OK;
}

}  // extern "C"

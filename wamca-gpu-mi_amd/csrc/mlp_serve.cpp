// mlp_serve.cpp -- continuous-batching DVND server (include/mlp_serve.h) over
// the public C ABI of the engine.  Host-side bookkeeping only: admission by
// device footprint, one chunk of DVND rounds per step over every resident job
// of a lane, retirement and result hand-off.
//
// Two lanes: jobs below cfg.big_n share the caller's context; larger jobs run
// on a second context (its own HIP stream) on the same GPU, so the short
// rounds of small jobs do not queue behind the long kernels of large ones.
// The small lane is waited for at each step; the large lane is only polled.
#include "mlp_serve.h"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <deque>
#include <new>
#include <vector>

namespace {

struct Job {
  int64_t id = 0, tag = 0;
  int32_t n = 0, round = 0;
  std::vector<int32_t> xs, ys, tour;
  double t_submit = 0, t_admit = 0, t_finish = 0;
  int32_t inst = -1, slot = -1;
  int64_t bytes = 0, latency = 0;
  int32_t rounds = 0, status = MLP_OK;
  int lane = 0;
  int bypassed = 0;  // later jobs admitted while this one waited for memory
};

// The pool rounds every block up to a power of two of at least 4 KiB.
int64_t pow2_class(int64_t b) {
  int64_t c = 4096;
  while (c < b) c <<= 1;
  return c;
}

struct Lane {
  mlp_ctx* ctx = nullptr;
  bool owned = false;
  std::vector<Job*> active;
  bool chunk_pending = false;
};

}  // namespace

struct mlp_server {
  mlp_server_config cfg{};
  Lane lane[2];
  std::deque<Job*> queued;
  std::vector<Job*> done;    // finished, not yet polled
  std::vector<Job*> handed;  // returned by the last poll
  int64_t used = 0, next_id = 1;
  std::chrono::steady_clock::time_point t0;
};

extern "C" {

double mlp_server_now(const mlp_server* srv) {
  if (!srv) return 0.0;
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - srv->t0).count();
}

int64_t mlp_server_job_bytes(int32_t n) {
  const int64_t slot = mlp_slot_footprint(n);
  if (slot < 0) return -1;
  const int64_t half = (4 * (int64_t)n + 255) / 256 * 256;
  return pow2_class(slot) + pow2_class(2 * half);
}

int mlp_server_create(mlp_ctx* ctx, const mlp_server_config* cfg, mlp_server** out) {
  if (!ctx || !out) return MLP_EINVAL;
  mlp_server* s = new (std::nothrow) mlp_server();
  if (!s) return MLP_ENOMEM;
  if (cfg) s->cfg = *cfg;
  if (s->cfg.hbm_budget <= 0) {
    int64_t fr = 0;
    const int st = mlp_mem_info(ctx, &fr, nullptr);
    if (st) {
      delete s;
      return st;
    }
    s->cfg.hbm_budget = fr / 100 * 85;
  }
  if (s->cfg.max_bypass == 0) s->cfg.max_bypass = 64;
  if (s->cfg.rounds_per_step <= 0) s->cfg.rounds_per_step = 4;
  if (s->cfg.big_n == 0) s->cfg.big_n = 4096;
  if (s->cfg.rowmap != MLP_ROWMAP_ADS && s->cfg.rowmap != MLP_ROWMAP_ODC) {
    delete s;
    return MLP_EINVAL;
  }
  s->lane[0].ctx = ctx;
  s->t0 = std::chrono::steady_clock::now();
  *out = s;
  return MLP_OK;
}

static void release(mlp_server* s, Job* j) {
  mlp_ctx* ctx = s->lane[j->lane].ctx;
  if (j->slot >= 0) mlp_slot_destroy(ctx, j->slot);
  if (j->inst >= 0) mlp_instance_destroy(ctx, j->inst);
  j->slot = j->inst = -1;
}

int mlp_server_destroy(mlp_server* s) {
  if (!s) return MLP_EINVAL;
  for (Lane& l : s->lane) {
    if (!l.ctx) continue;
    mlp_sync(l.ctx);
    for (Job* j : l.active) {
      release(s, j);
      delete j;
    }
  }
  if (s->lane[1].owned) mlp_ctx_destroy(s->lane[1].ctx);
  for (Job* j : s->queued) delete j;
  for (Job* j : s->done) delete j;
  for (Job* j : s->handed) delete j;
  delete s;
  return MLP_OK;
}

int mlp_server_submit(mlp_server* s, const int32_t* xs, const int32_t* ys, int32_t n,
                      int32_t round, const int32_t* tour, int64_t tag, int64_t* job) {
  if (!s || !xs || !ys || !tour || n < 2) return MLP_EINVAL;
  // the checks instance and slot set-up would apply, so that no queued job
  // fails them at admission
  const int chk = mlp_check_job(s->lane[0].ctx, xs, ys, n, round, tour);
  if (chk) return chk;
  const int64_t bytes = mlp_server_job_bytes(n);
  if (bytes < 0) return MLP_ERANGE;
  if (bytes > s->cfg.hbm_budget) return MLP_ENOMEM;
  Job* j = new (std::nothrow) Job();
  if (!j) return MLP_ENOMEM;
  j->id = s->next_id++;
  j->tag = tag;
  j->n = n;
  j->round = round;
  j->xs.assign(xs, xs + n);
  j->ys.assign(ys, ys + n);
  j->tour.assign(tour, tour + n + 1);
  j->bytes = bytes;
  j->lane = (s->cfg.big_n > 0 && n >= s->cfg.big_n) ? 1 : 0;
  j->t_submit = mlp_server_now(s);
  s->queued.push_back(j);
  if (job) *job = j->id;
  return MLP_OK;
}

// Retires the finished jobs of a lane whose last chunk is done.
static int retire(mlp_server* s, Lane& l) {
  if (!l.chunk_pending || l.active.empty()) {
    l.chunk_pending = false;
    return MLP_OK;
  }
  const int c = (int)l.active.size();
  std::vector<int32_t> slots((size_t)c), act((size_t)c), rnd((size_t)c);
  for (int q = 0; q < c; ++q) slots[(size_t)q] = l.active[(size_t)q]->slot;
  int st = mlp_dvnd_status(l.ctx, slots.data(), c, act.data(), rnd.data());
  if (st) return st;
  const double now = mlp_server_now(s);
  std::vector<Job*> still, fin;
  std::vector<int32_t> fslots;
  size_t ftotal = 0;
  for (int q = 0; q < c; ++q) {
    Job* j = l.active[(size_t)q];
    if (act[(size_t)q]) {
      still.push_back(j);
      continue;
    }
    j->rounds = rnd[(size_t)q];
    fin.push_back(j);
    fslots.push_back(j->slot);
    ftotal += (size_t)j->n + 1;
  }
  if (!fin.empty()) {
    // the finished jobs' tours and latencies in one read-back each
    std::vector<int32_t> tours(ftotal);
    std::vector<int64_t> lat(fin.size());
    const int nf = (int)fin.size();
    st = mlp_get_tours(l.ctx, fslots.data(), nf, tours.data());
    if (st == MLP_OK) st = mlp_get_latency(l.ctx, fslots.data(), nf, lat.data());
    if (st) return st;
    size_t off = 0;
    for (size_t q = 0; q < fin.size(); ++q) {
      Job* j = fin[q];
      std::copy(tours.begin() + (long)off, tours.begin() + (long)(off + j->n + 1), j->tour.begin());
      off += (size_t)j->n + 1;
      j->latency = lat[q];
      j->t_finish = now;
      release(s, j);
      s->used -= j->bytes;
      s->done.push_back(j);
    }
  }
  l.active.swap(still);
  l.chunk_pending = false;
  return MLP_OK;
}

int mlp_server_step(mlp_server* s, int32_t* active_out, int32_t* queued_out) {
  if (!s) return MLP_EINVAL;
  int st = MLP_OK;
  // 1. retire: the small lane is waited for, the large lane only polled
  bool ready[2] = {true, true};
  if (s->lane[1].ctx && s->lane[1].chunk_pending) {
    int32_t idle = 0;
    st = mlp_query(s->lane[1].ctx, &idle);
    if (st) return st;
    ready[1] = idle != 0;
  }
  for (int L = 0; L < 2; ++L)
    if (s->lane[L].ctx && ready[L]) {
      st = retire(s, s->lane[L]);
      if (st) return st;
    }
  // 2. admission in arrival order while the footprint fits; when the oldest
  // queued job does not fit yet, later jobs that fit may pass it, at most
  // max_bypass of them in all, so that it is not starved
  std::vector<Job*> admitted[2];
  const int total_active = (int)(s->lane[0].active.size() + s->lane[1].active.size());
  int added = 0;
  const double now0 = mlp_server_now(s);
  auto fail_job = [&](Job* j, int code) {  // retired with an error status
    release(s, j);
    j->status = code;
    j->latency = 0;
    j->t_finish = now0;
    s->done.push_back(j);
  };
  Job* blocked = nullptr;  // oldest queued job whose footprint does not fit
  for (auto it = s->queued.begin(); it != s->queued.end();) {
    Job* j = *it;
    if (s->cfg.max_active > 0 && total_active + added >= s->cfg.max_active) break;
    if (s->used + j->bytes > s->cfg.hbm_budget) {
      if (!blocked) blocked = j;
      if (s->cfg.max_bypass < 0 || blocked->bypassed >= s->cfg.max_bypass) break;
      ++it;
      continue;
    }
    if (!ready[j->lane]) {  // its lane is still running a chunk: later jobs may go first
      ++it;
      continue;
    }
    Lane& l = s->lane[j->lane];
    if (!l.ctx) {  // the large lane is created on first use, on the same device
      int32_t dev = 0;
      st = mlp_ctx_device(s->lane[0].ctx, &dev);
      if (st == MLP_OK) st = mlp_ctx_create(dev, &l.ctx);
      if (st) return st;
      l.owned = true;
      int8_t tune[MLP_TUNE_TABLE_BYTES];  // same tile choices as the first lane
      st = mlp_tune_get(s->lane[0].ctx, tune);
      if (st == MLP_OK) st = mlp_tune_set(l.ctx, tune);
      if (st) return st;
    }
    st = mlp_instance_create_async(l.ctx, j->xs.data(), j->ys.data(), j->n, j->round, &j->inst);
    if (st == MLP_OK) st = mlp_slot_create(l.ctx, j->inst, &j->slot);
    if (st == MLP_ENOMEM) {  // the device is full: the job waits for memory to be freed
      release(s, j);
      break;
    }
    it = s->queued.erase(it);
    if (st) {
      fail_job(j, st);
      continue;
    }
    if (blocked) ++blocked->bypassed;
    s->used += j->bytes;
    admitted[j->lane].push_back(j);
    ++added;
  }
  for (int L = 0; L < 2; ++L) {
    if (admitted[L].empty()) continue;
    Lane& l = s->lane[L];
    std::vector<int32_t> slots, tours;
    for (Job* j : admitted[L]) {
      slots.push_back(j->slot);
      tours.insert(tours.end(), j->tour.begin(), j->tour.end());
    }
    st = mlp_set_tours(l.ctx, slots.data(), (int32_t)slots.size(), tours.data());
    if (st == MLP_OK) st = mlp_dvnd_start(l.ctx, slots.data(), (int32_t)slots.size());
    if (st) {  // none of these jobs started: retire them with the error
      for (Job* j : admitted[L]) {
        s->used -= j->bytes;
        fail_job(j, st);
      }
      continue;
    }
    const double now = mlp_server_now(s);
    for (Job* j : admitted[L]) {
      j->t_admit = now;
      l.active.push_back(j);
    }
  }
  // 3. one chunk of rounds over every resident job of each ready lane
  for (int L = 0; L < 2; ++L) {
    Lane& l = s->lane[L];
    if (!ready[L] || l.active.empty()) continue;
    std::vector<int32_t> slots;
    for (Job* j : l.active) slots.push_back(j->slot);
    st = mlp_dvnd_rounds(l.ctx, slots.data(), (int32_t)slots.size(), s->cfg.rowmap,
                         s->cfg.rounds_per_step, s->cfg.max_rounds);
    if (st) return st;
    l.chunk_pending = true;
  }
  if (active_out) *active_out = (int32_t)(s->lane[0].active.size() + s->lane[1].active.size());
  if (queued_out) *queued_out = (int32_t)s->queued.size();
  return MLP_OK;
}

int mlp_server_poll(mlp_server* s, mlp_job_result* out, int32_t cap, int32_t* count) {
  if (!s || !count || (cap > 0 && !out)) return MLP_EINVAL;
  for (Job* j : s->handed) delete j;
  s->handed.clear();
  const int take = std::min<int>(cap, (int)s->done.size());
  for (int q = 0; q < take; ++q) {
    Job* j = s->done[(size_t)q];
    out[q] = mlp_job_result{j->id, j->tag, j->latency, j->rounds, j->n,
                            j->t_submit, j->t_admit, j->t_finish, j->tour.data(), j->status};
    s->handed.push_back(j);
  }
  s->done.erase(s->done.begin(), s->done.begin() + take);
  *count = take;
  return MLP_OK;
}

}  // extern "C"

/*
 * mlp_serve.h -- continuous-batching DVND server over the MI355X engine
 * (libwamca_mi.so).
 *
 * Jobs (an instance, a start tour) are submitted at any time and queue on the
 * host.  Each mlp_server_step retires the jobs that reached their local
 * optimum (releasing their memory to the pool), admits queued jobs in arrival
 * order while their device footprint (coordinates + one slot: tour, ADS
 * prefix arrays, row buffers, merge scratch) fits the memory budget and the
 * active-job limit (a job that does not fit yet may be passed by a bounded
 * number of later ones), and enqueues a chunk of DVND rounds over every resident
 * job in one set of launches (jobs of any n share the launches; a finished
 * job is skipped on the device).  Jobs of n >= big_n run in a second lane
 * with its own context and HIP stream, so the GPU runs both lanes' kernels
 * side by side; the first lane is waited for at each step, the second only
 * polled, and a queued job whose lane is busy lets later jobs of the other
 * lane go first.  A job's result does not depend on the other
 * jobs it ran beside: every kernel works on one job's slot, and a job's
 * rounds are the same rounds mlp_dvnd runs for it alone.
 *
 * The reference has no serving layer; its unit of work is one DVND run of one
 * tour (WamcaExperiment.hpp:368-581, dvnd.cu:8-83).  Timestamps are host
 * steady-clock seconds since the server was created.
 */
#ifndef MLP_SERVE_H_
#define MLP_SERVE_H_

#include <stdint.h>

#include "mlp_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlp_server mlp_server;

typedef struct {
  int64_t hbm_budget;      /* device bytes resident jobs may hold (<= 0: 85% of the device
                              memory free when the server is created) */
  int32_t max_active;      /* resident jobs at most (<= 0: no limit but the budget) */
  int32_t rowmap;          /* MLP_ROWMAP_ADS or MLP_ROWMAP_ODC */
  int32_t rounds_per_step; /* DVND rounds per step (<= 0: 4) */
  int32_t max_rounds;      /* per job, 0 = to the local optimum */
  int32_t big_n;           /* jobs of n >= big_n run in a second lane (own context and
                              stream, same GPU); 0: 4096, < 0: one lane */
  int32_t max_bypass;      /* later jobs that may be admitted past the oldest queued job
                              while its footprint does not fit the free budget (0: 64,
                              < 0: none -- strict arrival order) */
} mlp_server_config;

typedef struct {
  int64_t job;       /* id returned by mlp_server_submit */
  int64_t tag;       /* caller's tag */
  int64_t latency;   /* final tour latency (int64) */
  int32_t rounds;    /* DVND rounds run */
  int32_t n;         /* clients */
  double t_submit, t_admit, t_finish;
  const int32_t* tour; /* final tour, N = n + 1 entries; valid until the next poll */
  int32_t status;      /* MLP_OK, or the error that stopped the job's admission (its tour is
                          then the start tour and latency 0) */
} mlp_job_result;

int mlp_server_create(mlp_ctx* ctx, const mlp_server_config* cfg, mlp_server** out);
int mlp_server_destroy(mlp_server* srv);
/* Copies the job in.  tour = [0, perm(1..n-1), 0] (n + 1 entries).  The job
 * is checked as mlp_check_job does (coordinates, sizes, tour) and refused
 * with that status (MLP_EINVAL / MLP_ERANGE) before it is queued; MLP_ENOMEM
 * when the job alone exceeds the memory budget. */
int mlp_server_submit(mlp_server* srv, const int32_t* xs, const int32_t* ys, int32_t n,
                      int32_t round, const int32_t* tour, int64_t tag, int64_t* job);
/* Admission, one chunk of rounds, retirement.  Blocks until the chunk of the
 * previous step has finished on the device.  active/queued may be NULL.
 * Queued jobs are admitted in arrival order; when the oldest one does not fit
 * the free budget yet, up to max_bypass later jobs that fit may go first.  A
 * job whose device set-up fails for a reason other than memory is retired
 * with that status (mlp_job_result.status) and the step goes on with the
 * other jobs; a job refused for memory waits in the queue. */
int mlp_server_step(mlp_server* srv, int32_t* active, int32_t* queued);
/* Finished jobs since the last poll (up to cap; the rest stay for the next). */
int mlp_server_poll(mlp_server* srv, mlp_job_result* out, int32_t cap, int32_t* count);
/* Footprint the admission test charges for a job of n clients. */
int64_t mlp_server_job_bytes(int32_t n);
/* Seconds since the server was created (the timestamps' clock). */
double mlp_server_now(const mlp_server* srv);

#ifdef __cplusplus
}
#endif

#endif /* MLP_SERVE_H_ */

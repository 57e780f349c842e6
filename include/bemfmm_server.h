/*
 * bemfmm_server.h — request loop for many FMM-GMRES BVPs on one MI355X (continuous batching).
 *
 * The reference solves one BVP at a time per process: BEM_BVP::solve (bem/BEM_solveBVP.hpp:624),
 * called once per RK stage by the time-stepping driver (bem/main_time_domain.cpp:2095), and
 * solveForPhiPhin_t (bem/BEM_solveBVP.hpp:999) for the phi_t BVP of the same stage. This server
 * runs many such solves at once on one device: a pool of `slots` worker threads, each driving one
 * bemfmm context (include/bemfmm.h) on its own HIP stream, so that the kernels of different requests
 * overlap on the GPU.
 *
 * Scheduling: requests are admitted in arrival order as slots and device memory free up. A request
 * waits while its session has an older request queued or running, or while its estimated device
 * memory does not fit `hbm_budget_bytes` (idle resident sessions are released first, least recently
 * used first). Sessions (session >= 0) keep their context — tree, operators, near field, rigid-mode
 * diagonal, preconditioner — resident in HBM between requests: a request whose geometry equals the
 * session's previous one only refreshes the boundary values (the phi_t BVP of an RK stage); moved nodes
 * re-run the setup in the session's context. session = -1 is a one-shot request.
 *
 * All input pointers are copied at submit. Results are claimed once with bemfmm_server_wait.
 */
#ifndef BEMFMM_SERVER_H
#define BEMFMM_SERVER_H

#include <stdint.h>

#include "bemfmm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bemfmm_server bemfmm_server;

typedef struct bemfmm_server_opts {
  int device;                /* HIP device */
  int slots;                 /* concurrent requests (worker threads / streams), 1..64 */
  double hbm_budget_bytes;   /* device memory the server may hold (estimated, then measured per context) */
  int omp_threads_per_slot;  /* host threads for each request's setup; 0 = omp_get_max_threads() / slots */
} bemfmm_server_opts;

/* One BVP: the arrays of bemfmm_set_geometry / bemfmm_set_bc, the context options and the arguments
 * of bemfmm_solve (x0 may be NULL for a zero initial guess). */
typedef struct bemfmm_problem {
  int n_nodes;
  const double* X;
  int n_faces;
  const int32_t* tri;
  const int32_t* face_dof;
  int n_dof;
  const int32_t* dof_kind;
  const int32_t* dof_point;
  const double* phi;
  const double* phin;
  const double* x0;
  bemfmm_opts opts;
  double tol;
  int max_iter;
  int restart;
} bemfmm_problem;

typedef struct bemfmm_server_stats {
  int64_t completed, failed, queued, evicted;
  int running, resident_sessions;
  double p50_latency_s, p99_latency_s;  /* submit -> result, completed requests */
  double used_bytes, peak_bytes;        /* device memory accounted to contexts */
  int pooled_contexts;                  /* idle contexts kept for reuse by later requests */
} bemfmm_server_stats;

void bemfmm_server_default_opts(bemfmm_server_opts* opts);
/* NULL on invalid options. */
bemfmm_server* bemfmm_server_create(const bemfmm_server_opts* opts);
/* Completes every queued request, then stops the workers and frees all contexts. */
void bemfmm_server_destroy(bemfmm_server* srv);

/* Queues a solve; returns its ticket (> 0) or a negative BEMFMM_E_* code. */
int64_t bemfmm_server_submit(bemfmm_server* srv, const bemfmm_problem* problem, int64_t session);
/* Blocks until the ticket's request has finished. x: n_dof (physical units, as bemfmm_solve); times3 (may be
 * NULL): queueing, service and total latency in seconds. Returns the request's code (non-convergence is
 * reported through result->converged). */
int bemfmm_server_wait(bemfmm_server* srv, int64_t ticket, double* x, bemfmm_result* result, double* times3);
/* Releases a session's context after its queued requests. */
int bemfmm_server_close_session(bemfmm_server* srv, int64_t session);
int bemfmm_server_get_stats(bemfmm_server* srv, bemfmm_server_stats* stats);
const char* bemfmm_server_last_error(bemfmm_server* srv);

#ifdef __cplusplus
}
#endif

#endif /* BEMFMM_SERVER_H */

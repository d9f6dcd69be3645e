/*
 * libphpc — GEMM serving loop: requests of different sizes and accuracy classes arrive
 * at any time. Their operands live in an HBM arena sized to the memory budget; a
 * request is admitted (FIFO) when a worker slot is free and its operands fit. Workers
 * run concurrently on their own HIP streams. Requests too small to fill the GPU alone
 * are batched with waiting requests of the same kind into one persistent grouped
 * launch; large requests run alone.
 *
 * No counterpart in the reference (its driver runs one GEMM per program run,
 * src/main.c:88-107).
 */
#ifndef PHPC_SERVE_H
#define PHPC_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "phpc_gemm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FP64 inputs, FP64-level result: the split-precision CRT GEMM at err_cw <= 1e-14 when
 * min(m, n, k) >= the split threshold (phpc_server_set_split_min, default 4096), else the
 * plain FP64 MFMA kernel, batched with other such requests. */
#define PHPC_ACC_FP64 0
/* FP32 inputs, FP32-level result: the FP32 MFMA grouped GEMM, batched; requests that fill
 * the GPU alone with m*n >= 2048^2 and k >= 256 use the FP32-accurate 3 x BF16 slice
 * planes (phpc_sgemm_device, PHPC_SGEMM_BF16X3). */
#define PHPC_ACC_FP32 1

typedef struct phpc_request {
  int accuracy;      /* PHPC_ACC_FP64 or PHPC_ACC_FP32 */
  int m, n, k;       /* C[m x n] = A[m x k] B[k x n], C starting at zero */
  int dist;          /* input distribution (phpc_fill_device): 0 uniform[-1,1), 1 wide, 2 ramp */
  uint64_t seed;     /* A from stream `seed`, B from stream `seed + 1` */
  void *result_host; /* optional: C is copied here (double or float, row-major m x n) */
} phpc_request;

typedef struct phpc_result {
  int status;            /* PHPC_OK or an error code */
  int batch_size;        /* requests computed by the same launch */
  double latency_s;      /* arrival -> completion */
  double queue_s;        /* arrival -> admission */
  double compute_s;      /* admission -> completion */
  double error_estimate; /* split GEMM a-priori error estimate (FP64 class) */
  double fill_s;         /* device time generating the batch's operands (event-timed) */
  double gemm_s;         /* device time of the batch's GEMM launches (event-timed) */
} phpc_result;

typedef struct phpc_server *phpc_server_t;

/* budget_bytes: HBM the server holds: per-slot workspaces, and the rest as the operand
 * arena (allocated here, once); slots: concurrent workers (each with its own stream); max_dim:
 * workspaces are sized up front for requests up to max_dim x max_dim x max_dim;
 * batch_max: most FP32 requests folded into one grouped launch. */
int phpc_server_create(size_t budget_bytes, int slots, int max_dim, int batch_max, phpc_server_t *server);
/* Enqueue (non-blocking). PHPC_ERR_BUDGET if the request can never fit the budget, or if
 * any of m, n, k exceeds max_dim (its workspaces would have to grow outside the budget). */
int phpc_server_submit(phpc_server_t server, const phpc_request *req, uint64_t *ticket);
/* Block until the request is done; the ticket is released afterwards. */
int phpc_server_wait(phpc_server_t server, uint64_t ticket, phpc_result *result);
/* Block until every submitted request is done. */
int phpc_server_drain(phpc_server_t server);
/* FP64-class requests with min(m, n, k) >= dim use the split GEMM (applies to later
 * submissions). */
int phpc_server_set_split_min(phpc_server_t server, int dim);
#define PHPC_SCHED_FIFO 0 /* oldest request first (default) */
/* Least estimated device time first (estimate from the flop count), less waiting time
 * divided by `aging` so that large requests are not starved; aging = 0: no credit. */
#define PHPC_SCHED_SJF 1
/* Two lanes: while fewer than the big limit of GPU-filling requests run, the waiting one
 * with the least estimated device time (less waiting time / aging) starts first; every
 * other request starts in arrival order beside them. */
#define PHPC_SCHED_LANES 2
int phpc_server_set_policy(phpc_server_t server, int policy, double aging);
/* Requests that fill the GPU alone (output tiles >= compute units) and whose estimated
 * device time is at least min_seconds: at most `limit` of them in flight at once
 * (defaults 1 and 0); every other request may start beside them. Applies to later
 * submissions. */
int phpc_server_set_big_limit(phpc_server_t server, int limit, double min_seconds);
/* Operand bytes currently admitted, and the high-water mark. */
int phpc_server_memory(phpc_server_t server, size_t *in_use, size_t *peak);
void phpc_server_destroy(phpc_server_t server);

#ifdef __cplusplus
}
#endif
#endif

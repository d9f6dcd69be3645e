/*
 * libphpc — grouped GEMM plans: many independent row-major GEMMs (FP64 or FP32) in ONE
 * persistent launch, with autotuned tiles and optional hipGraph capture.
 *
 * No single counterpart in the reference. Its closest relative is the per-rank local
 * GEMM of the SUMMA loop (src/phpc_summa.c:93), which the reference issues as one call
 * per rank and step; a plan covers any set of GEMMs (uneven rank blocks, unrelated
 * shapes) with one launch, so a caller can replay it from a captured graph.
 */
#ifndef PHPC_GROUPED_H
#define PHPC_GROUPED_H

#include "phpc_gemm.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PHPC_F64 0
#define PHPC_F32 1
/* FP32 operands, FP32-accurate result from three BF16 slices per operand on the BF16
 * matrix cores (see phpc_sgemm_device) */
#define PHPC_F32X3 2

/* C[m x n] (ldc) += A[m x k] (lda) * B[k x n] (ldb) on device pointers of the plan's
 * element type. panel > 0 (FP64 only): group k in panels, as phpc_dgemm_device. */
typedef struct phpc_gemm_item {
  const void *A;
  const void *B;
  void *C;
  long lda, ldb, ldc;
  int m, n, k;
  int panel;
} phpc_gemm_item;

typedef struct phpc_plan *phpc_plan_t;

#define PHPC_PLAN_AUTOTUNE 1 /* time every tile configuration once (C is saved and restored) */

/* Builds a plan: copies the item list to the device and chooses a tile configuration
 * (planner rule, or the fastest measured one with PHPC_PLAN_AUTOTUNE; tuned choices are
 * cached per shape list for the life of the process). */
int phpc_plan_create(int dtype, const phpc_gemm_item *items, int count, int flags, void *stream, phpc_plan_t *plan);
/* One persistent launch computing every item, stream-ordered. */
int phpc_plan_launch(phpc_plan_t plan, void *stream);
/* Same launch replayed from a hipGraph captured on the first call (stream must not be
 * NULL: the capture needs a created stream). */
int phpc_plan_graph_launch(phpc_plan_t plan, void *stream);
/* config: tile configuration index; tiles: total tiles; grid: persistent workgroups;
 * tuned_ms: best time measured by the autotuner (0 if not tuned). */
int phpc_plan_info(phpc_plan_t plan, int *config, int *tiles, int *grid, double *tuned_ms);
void phpc_plan_destroy(phpc_plan_t plan);

#ifdef __cplusplus
}
#endif
#endif

/*
 * libphpc — SUMMA k-panel loop on an emulated p x q process grid, one GPU.
 *
 * Drop-in for the reference header src/phpc_summa.h. The reference's first argument,
 * an MPI Cartesian communicator, becomes an opaque grid handle: this build runs every
 * rank of the grid on one GPU, with no MPI and no inter-GPU collectives. All other
 * arguments keep their names, order and meaning. Inputs are full N x N row-major A
 * and B on the host; C (N x N) is accumulated (C += A*B) and returned in full.
 */
#ifndef PHPC_SUMMA_H
#define PHPC_SUMMA_H

#include "phpc_gemm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Emulated process grid (replaces MPI_Comm grid_comm). */
typedef struct phpc_grid *phpc_grid_t;

/* Grid of `nranks` ranks: 1x1 for one rank, else the balanced factorisation of
 * MPI_Dims_create(nranks, 2) (d0 >= d1), as src/main.c:38-45 builds it. */
int phpc_grid_create(int nranks, phpc_grid_t *grid);
/* Grid with explicit dimensions p (rows) x q (columns). */
int phpc_grid_create_dims(int p, int q, phpc_grid_t *grid);
void phpc_grid_destroy(phpc_grid_t grid);
/* dims[0] = p, dims[1] = q. */
int phpc_grid_dims(phpc_grid_t grid, int dims[2]);
/* k-panel width; 0 (default) selects the reference's N / lcm(p, q) (src/phpc_summa.c:36-38). */
int phpc_grid_set_panel(phpc_grid_t grid, int panel);
int phpc_grid_get_panel(phpc_grid_t grid);
/* Execution mode of phpc_gemm_summa_cuda / phpc_summa_device:
 *   PHPC_SUMMA_FUSED    one persistent launch over all ranks, walking the panels with
 *                       each rank's C tile kept in registers (default)
 *   PHPC_SUMMA_STEPWISE per panel: pack the broadcast panels (strided -> packed, as the
 *                       MPI_Type_vector sends do), then one grouped local-update launch
 *   PHPC_SUMMA_VENDOR   the STEPWISE schedule with rocBLAS dgemm as the local GEMM
 *                       (phpc_summa_device only)
 *   PHPC_SUMMA_CHAINED  one persistent launch, one FMA chain per output over all panels
 *                       (C is not rounded between panels): the arithmetic of a single
 *                       local-GEMM call over the full k; the fastest FP64 schedule
 *   PHPC_SUMMA_SPLIT    every rank's update over the full k by the split-precision CRT
 *                       GEMM (include/phpc_split.h): FP64-accurate (err_cw <= 1e-14),
 *                       int8 matrix cores
 * FUSED and STEPWISE produce bit-identical C (the reference's per-panel grouping). */
#define PHPC_SUMMA_FUSED 0
#define PHPC_SUMMA_STEPWISE 1
#define PHPC_SUMMA_VENDOR 2
#define PHPC_SUMMA_CHAINED 3
#define PHPC_SUMMA_SPLIT 4
int phpc_grid_set_mode(phpc_grid_t grid, int mode); /* FUSED, STEPWISE, CHAINED or SPLIT */

/* Least common multiple, src/phpc_summa.c:9-22. */
int find_lcm(int a, int b);

/*
 * Replaces phpc_gemm_summa_cuda, src/phpc_summa.h:6 (definition src/phpc_summa.c:124-126).
 * Rank (i, j) owns C block rows R_i x columns C_j; blocks are N/p (N/q) with the first
 * N % p (N % q) blocks one larger, so N need not divide evenly. For every panel t in
 * order, C_ij += A(R_i, panel t) * B(panel t, C_j), each panel product summed from zero
 * (src/phpc_summa.c:63-95). *compute_time = device time of the SUMMA in seconds.
 */
void phpc_gemm_summa_cuda(phpc_grid_t grid, const double *A, const double *B, double *C, int n, int gpu_count,
                          int grid_width, int grid_height, int block_width, float *compute_time);

/* Replaces phpc_gemm_summa_cublas, src/phpc_summa.h:8 (definition src/phpc_summa.c:128-130):
 * the same panel loop with rocBLAS dgemm as the local GEMM. *compute_time = 0 as in the
 * reference (its cuBLAS local GEMM reports 0, src/phpc_gemm.cu:173). */
void phpc_gemm_summa_cublas(phpc_grid_t grid, const double *A, const double *B, double *C, int n, int gpu_count,
                            float *compute_time);

/* The pluggable local GEMM, src/phpc_summa.c:7. phpc_gemm_cuda and phpc_gemm_cublas
 * have this type. */
typedef void (*gemm_t)(const double *a, int lda, const double *b, int ldb, double *c, int ldc, int m, int k, int n,
                       int gpu_count, int grid_width, int grid_height, int block_width, float *gpu_time);

/*
 * Generic SUMMA loop over host pointers, src/phpc_summa.c:24-122: for each panel t and
 * rank (i, j), calls f on the rank's blocks. The broadcasts become pointer aliases
 * (every emulated rank sees the full A and B, lda = ldb = N) and the gather is a no-op
 * (ranks write their C blocks in place, ldc = N). *gpu_time = sum of f's times.
 * Returns PHPC_OK or an error code.
 */
int phpc_gemm_summa(gemm_t f, phpc_grid_t grid, const double *A, const double *B, double *C, int N, int gpu_count,
                    int grid_width, int grid_height, int block_width, float *gpu_time);

/* ---------------- device-resident extensions ---------------- */

/* SUMMA on device pointers (N x N, row-major, ld = N), stream-ordered, any mode above.
 * panel 0 = N / lcm(p, q). wg > 0 fixes the persistent grid. */
int phpc_summa_device(int p, int q, int panel, int mode, const double *dA, const double *dB, double *dC, int N,
                      int wg, void *stream);

/* Strided block -> packed block copy, the MPI_Type_vector panel and block types of
 * src/phpc_summa.c:53-59 (count = rows, blocklength = cols, stride = ld_src) on device
 * pointers: dst[r * ld_dst + c] = src[r * ld_src + c], bytes copied unchanged. */
int phpc_pack_device(const double *src, long ld_src, double *dst, long ld_dst, int rows, int cols, void *stream);

/* Host-pointer SUMMA through the device path: copies A, B, C in once, runs
 * phpc_summa_device, copies C out. *kernel_seconds = device time of the SUMMA. */
int phpc_summa_host(int p, int q, int panel, int mode, const double *A, const double *B, double *C, int N, int wg,
                    float *kernel_seconds);

#ifdef __cplusplus
}
#endif
#endif

/*
 * libphpc — MI355X (gfx950) FP64 GEMM hot path: local GEMM entry points.
 *
 * Drop-in for the reference header src/phpc_gemm.cuh (Redy1908/HPC-MultiGPU-MatrixMult):
 * the two functions below keep the reference's names, argument order (m, k, n),
 * row-major layout, accumulate semantics (C += A*B) and seconds-as-float timing.
 * Everything after them is an extension with no reference counterpart.
 *
 * Plain C ABI: no HIP or torch types. Streams are passed as `void *` (a hipStream_t,
 * or NULL for the default stream).
 */
#ifndef PHPC_GEMM_H
#define PHPC_GEMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (phpc_last_error and the int-returning entry points). */
#define PHPC_OK 0
#define PHPC_ERR_ARG 1    /* invalid argument (sizes, leading dimensions, tile width) */
#define PHPC_ERR_HIP 2    /* a HIP runtime call failed */
#define PHPC_ERR_NODEV 3  /* no GPU visible */
#define PHPC_ERR_MEM 4    /* device or host allocation failed */
#define PHPC_ERR_VENDOR 5 /* vendor BLAS comparator failed or unavailable */
#define PHPC_ERR_BUDGET 6 /* request exceeds the serving memory budget */

/*
 * C[m x n] (ldc) += A[m x k] (lda) * B[k x n] (ldb) on host row-major pointers.
 * Replaces phpc_gemm_cuda, src/phpc_gemm.cuh:8 (definition src/phpc_gemm.cu:59-156).
 *   gpu_count   : GPUs to split B/C columns over; this build drives one GPU, so any
 *                 value >= 1 is accepted and the work runs on device 0.
 *   grid_width, grid_height : persistent workgroup count hint (the reference's
 *                 grid-stride launch, src/phpc_gemm.cu:21-25). Results do not depend
 *                 on it. Used as the launch grid only under phpc_set_launch_policy(1,..).
 *   block_width : the reference's tile width T (1..32, src/phpc_gemm.cu:83-84); it is
 *                 validated and recorded; the MFMA macro tile is chosen by the planner.
 *   compute_time: kernel time in seconds (device events around the launch, :116-118).
 * Per-element arithmetic equals the reference kernel's: c = 0; c = fma(a_k, b_k, c)
 * for ascending k; C = C + c. On invalid arguments C is left unchanged and
 * phpc_last_error() reports PHPC_ERR_ARG.
 */
void phpc_gemm_cuda(const double *a, int lda, const double *b, int ldb, double *c, int ldc, int m,
                    int k, int n, int gpu_count, int grid_width, int grid_height, int block_width,
                    float *compute_time);

/*
 * Vendor comparator: same contract through rocBLAS dgemm (row-major trick: computes
 * C^T += B^T A^T). Replaces phpc_gemm_cublas, src/phpc_gemm.cuh:14 (definition
 * src/phpc_gemm.cu:158-174). Like the reference, *gpu_time is set to 0.
 */
void phpc_gemm_cublas(const double *a, int lda, const double *b, int ldb, double *c, int ldc,
                      int m, int k, int n, int gpu_count, int grid_width, int grid_height,
                      int block_width, float *gpu_time);

/* ---------------- extensions ---------------- */

/* Status of the last call made on this thread, and its message. */
int phpc_last_error(void);
const char *phpc_last_error_message(void);

/* Initialise device 0 (idempotent). Every entry point calls it on first use. */
int phpc_init(void);

/*
 * Device-resident GEMM: C += A*B on device pointers, stream-ordered, no host sync.
 * panel > 0 groups the k range in panels of that width (C = ((C0 + s_0) + s_1) + ...,
 * each s_t summed from zero), which is what a SUMMA step loop produces; panel <= 0
 * means one panel (the phpc_gemm_cuda arithmetic).
 */
int phpc_dgemm_device(const double *dA, long lda, const double *dB, long ldb, double *dC, long ldc,
                      int m, int k, int n, int panel, void *stream);

/* rocBLAS dgemm on device pointers (comparator), C += A*B row-major. */
int phpc_dgemm_vendor_device(const double *dA, long lda, const double *dB, long ldb, double *dC,
                             long ldc, int m, int k, int n, void *stream);

/*
 * Launch policy for the FP64 kernels:
 *   honor_grid : 1 -> use grid_width*grid_height workgroups where the caller passes them
 *   config_id  : -1 -> planner's choice; else a fixed tile configuration (0..n-1)
 *   grid       : > 0 -> fixed persistent workgroup count; 0 -> planner's choice
 */
int phpc_set_launch_policy(int honor_grid, int config_id, int grid);
int phpc_num_configs(void);
/* Tuning knobs used by the autotuner: key 0 = FP64 config for large panel-grouped
 * problems on device-resident descriptor lists (plans fall back to it), key 1 = tile-row
 * swizzle width (1..64), key 2 = XCD-contiguous tile order (0/1), key 3 = FP64 config for
 * large single-panel problems on device-resident descriptor lists, key 4 / key 5 = FP64
 * config for large single-panel / panel-grouped problems through the entry points that
 * may stage a transposed copy of A (configs 10-12 read such a copy), key 6 = large
 * single FP32 products through phpc_sgemm_device on the transposed-A FP32 kernel (1,
 * default) or on the register-staged one (0); 0 also keeps FP32 plans created without
 * autotuning off configuration 3; key 7 = FP32 configuration of plans created
 * without autotuning (-1 = by work, default: 3 for at least two 128x128 tiles per CU, else
 * 0; 0-2 = register-staged 64x64, 128x64, 128x128; 3 = transposed-A 128x128, with the plan
 * keeping staged operand copies); key 8 = configuration of 3 x BF16 (PHPC_F32X3) plans
 * created without autotuning (-1 = by work, default: 6 for at least one 256x256 tile per
 * CU, else 2; 2 = BF16 slices cut in registers, 128x128 tiles; 4, 5, 6 = BF16 slice planes
 * cut once per operand, then one grouped level launch on 128x128, 256x128 or 256x256 tiles,
 * the plan keeping the planes; configurations 4, 5 and 6 give the same bits). */
int phpc_set_tuning(int key, int value);
/* Describe configuration `id`: macro tile bm x bn, threads per workgroup. */
int phpc_config_info(int id, int *bm, int *bn, int *threads);

/*
 * Seeded synthetic inputs on the device. Element i of stream `seed` uses
 * z = SplitMix64 output mix64(seed + (i+1) * 0x9E3779B97F4A7C15):
 *   dist 0: uniform[-1,1)  2*((z >> 11) * 2^-53) - 1
 *   dist 1: wide exponent  sign(z & 1) * 2^e, e = ((z >> 32) * 41 >> 32) - 20
 *   dist 2: ramp           x[i] = i   (the reference's input, src/main.c:85-86)
 * Fills an m x n block with leading dimension ld, element (r, c) using index r*n + c.
 */
int phpc_fill_device(double *dX, long ld, int m, int n, int dist, uint64_t seed, void *stream);
/* The same values rounded to FP32. */
int phpc_fill_device_f32(float *dX, long ld, int m, int n, int dist, uint64_t seed, void *stream);

/*
 * Sampled accuracy check on the device, for any C += A*B product computed on the GPU:
 * for each (rows[e], cols[e]) computes the double-double value of C0 + sum_k A B and
 * returns max_e |C - ref| / (|C0| + (|A||B|)) in *err_cw and max|C-ref| / max|ref| in
 * *err_nw. C0 may be NULL (zero). rows/cols are host arrays of `count` entries, each
 * checked against m and n before anything runs on the device.
 */
int phpc_check_sampled(const double *dA, long lda, const double *dB, long ldb, const double *dC0,
                       long ldc0, const double *dC, long ldc, int m, int k, int n, const int *rows,
                       const int *cols, int count, double *err_cw, double *err_nw, void *stream);

/* The same check of C = A * B on host matrices (copied to the device): `count` entries
 * drawn from `seed` plus the four corners. Used by the CLI's --check. */
int phpc_check_sampled_host(const double *A, long lda, const double *B, long ldb, const double *C, long ldc, int m,
                            int k, int n, int count, uint64_t seed, double *err_cw, double *err_nw);

/*
 * Device-event timing of the main GEMM kernel launches, for measurement: while on != 0,
 * every FP64 and FP32 tile-kernel launch and every split-precision level-kernel launch
 * (int8 residue / slice GEMMs, BF16 slice GEMMs) is bracketed by a pair of HIP events on the
 * stream it runs on; launches into a stream being captured into a graph are not timed.
 * phpc_kernel_timing() discards what was recorded; phpc_kernel_timing_read() waits for
 * the recorded launches, returns their summed duration (ms) and count, and clears them.
 */
int phpc_kernel_timing(int on);
int phpc_kernel_timing_read(double *ms_total, int *launches);

#ifdef __cplusplus
}
#endif
#endif

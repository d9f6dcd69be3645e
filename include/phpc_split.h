/*
 * libphpc — split-precision GEMM: FP64-accurate products from int8 matrix-core slices.
 *
 * No counterpart in the reference (Redy1908/HPC-MultiGPU-MatrixMult computes only in
 * plain FP64, src/phpc_gemm.cu:6-57). Same contract as phpc_dgemm_device: row-major
 * FP64 operands on the device, C += A*B, stream-ordered.
 *
 * Method (Ozaki scheme with integer slices): rows of A and columns of B are scaled by
 * powers of two and cut exactly into `slices` signed 7-bit digits; the s(s+1)/2 slice
 * products with t + u <= s + 1 are int8 GEMMs accumulated exactly in int32 on the
 * MFMA units, and the levels are recombined in FP64 with the exact power-of-two scales.
 * The result differs from the exact product only by the dropped slice terms (about
 * 2^(-7 s) relative to the row/column scales) and a few FP64 roundings.
 */
#ifndef PHPC_SPLIT_H
#define PHPC_SPLIT_H

#include "phpc_gemm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phpc_split_info {
  int slices;            /* slices (PHPC_SPLIT_SLICES, PHPC_SPLIT_FP8) or moduli (PHPC_SPLIT_CRT) used */
  int products;          /* int8 GEMMs computed: s(s+1)/2 slice products, or one per modulus */
  int launches;          /* kernel launches issued */
  double error_estimate; /* a-priori componentwise error estimate used to pick the size */
  int method;            /* PHPC_SPLIT_SLICES, PHPC_SPLIT_CRT or PHPC_SPLIT_FP8 */
  int int_bits;          /* bits kept per operand entry (7 or 4 * slices, or p of the CRT scheme) */
} phpc_split_info;

/* Methods:
 *   PHPC_SPLIT_SLICES  Ozaki scheme I: `slices` signed 7-bit digits per entry, the
 *                      s(s+1)/2 slice products with t + u <= s + 1
 *   PHPC_SPLIT_CRT     Ozaki scheme II (default): entries scaled to p-bit integers, the
 *                      exact integer product recovered from its residues modulo
 *                      `slices` pairwise coprime moduli <= 256 (one int8 GEMM each) by
 *                      Garner's mixed-radix algorithm; p and the modulus count follow
 *                      from target_err (or p from a fixed modulus count). With the
 *                      automatic count, inputs whose rows (columns) all fit fewer bits
 *                      than target_err asks for (integers, short mantissas, powers of
 *                      two) are scaled at that exact width, with no truncation error
 *   PHPC_SPLIT_FP8     Ozaki scheme I with FP8-E4M3 slices: 4-bit digits (|d| <= 15,
 *                      exact in E4M3), the slice products on the FP8 matrix cores with
 *                      FP32 accumulation, exact in k pieces of at most 74565 products;
 *                      about 3x the products of PHPC_SPLIT_SLICES for the same accuracy */
#define PHPC_SPLIT_SLICES 0
#define PHPC_SPLIT_CRT 1
#define PHPC_SPLIT_FP8 2
int phpc_split_set_method(int method);

/*
 * C[m x n] (ldc) += A[m x k] (lda) * B[k x n] (ldb), FP64, through int8 (or FP8) slices.
 *   slices     : slices 1..12 (PHPC_SPLIT_SLICES), 1..16 (PHPC_SPLIT_FP8) or moduli 2..20
 *                (PHPC_SPLIT_CRT), or 0 to choose the smallest size whose error estimate
 *                is at most target_err
 *   target_err : componentwise target, max |C - C*| / (|A||B|); <= 0 means 1e-14
 *   info       : optional, filled on success
 * Inputs must be finite (PHPC_ERR_ARG otherwise). The call synchronises the stream once
 * (after the exponent pass) to size the slices.
 */
int phpc_dgemm_split_device(const double *dA, long lda, const double *dB, long ldb, double *dC, long ldc, int m,
                            int k, int n, int slices, double target_err, void *stream, phpc_split_info *info);

/*
 * The same product by the CRT method (PHPC_SPLIT_CRT, whatever phpc_split_set_method
 * says), allocating nothing: no hipMalloc, hipFree or other allocating call. No
 * counterpart in the reference; it is for callers that cannot allocate device memory,
 * such as tight serving loops and harnesses that forbid it. Its scratch is a static
 * device array of 64 MiB that each GPU holds until the process ends, which limits the
 * size: the statistics block, m x (L kp) + n x (L kp) residue operands and L x m x n
 * residue planes (kp = k rounded up to 128, L = slices, or 20 when slices = 0) must fit,
 * e.g. m = n = k = 1024 does; larger products fail with PHPC_ERR_ARG and leave C
 * unchanged. Every call on one GPU shares that array: calls that overlap on one GPU
 * (other threads, other streams) race, so the caller serialises them. The call
 * synchronises the stream once, as phpc_dgemm_split_device does.
 */
int phpc_dgemm_split_device_noalloc(const double *dA, long lda, const double *dB, long ldb, double *dC, long ldc,
                                    int m, int k, int n, int slices, double target_err, void *stream,
                                    phpc_split_info *info);

/* The a-priori estimate: (s+1) 2^(-7s) * ratio_a * ratio_b / sqrt(k), where ratio_a is the
 * largest row scale 2^e_i over the row's mean |a| (ratio_b likewise for columns). */
double phpc_split_error_estimate(int slices, double ratio_a, double ratio_b, int k);

/* int8 level-GEMM tile configuration for large problems (autotuner knob; 0 ..
 * phpc_split_num_configs() - 1, default 13: 256x256 tiles, 16 waves, 16x16x64 MFMAs). Every
 * configuration gives the same bits. */
int phpc_split_set_config(int cfg);
int phpc_split_num_configs(void);

/* ---------------- FP32-accurate GEMM from BF16 slices ----------------
 * C[m x n] (ldc) += A[m x k] (lda) * B[k x n] (ldb), FP32 row-major device pointers,
 * stream-ordered.
 *   PHPC_SGEMM_FP32    FP32 matrix cores (v_mfma_f32_32x32x2f32)
 *   PHPC_SGEMM_BF16X3  every operand cut exactly into three BF16 slices (hi + mid + lo,
 *                      round to nearest); the six products hi.hi, hi.mid, mid.hi,
 *                      hi.lo, mid.mid, lo.hi on the BF16 matrix cores
 *                      (v_mfma_f32_32x32x16_bf16) with FP32 accumulation. Error:
 *                      |C - exact| <= ((k + 2) 2^-24 + 2^-23) (|C0| + |A||B|), the FP32
 *                      dot-product bound plus the dropped slice terms. Inputs must be
 *                      finite with |x| < 3.39e38 (the largest BF16).
 * No counterpart in the reference (FP64 only). */
#define PHPC_SGEMM_FP32 0
#define PHPC_SGEMM_BF16X3 1
int phpc_sgemm_device(const float *dA, long lda, const float *dB, long ldb, float *dC, long ldc, int m, int k, int n,
                      int method, void *stream);

#ifdef __cplusplus
}
#endif
#endif

/*
 * CPU oracle for the dense FP64 GEMM hot path (tiled kernel + SUMMA k-panel loop).
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (the HIP library, the C host
 * code, the CLI) may include, link or execute this code. Only the test suite,
 * the smoke check and the benchmark's CPU-baseline leg use it, and only as the
 * checker.
 *
 * Every function restates one piece of the reference's behaviour (file:line into
 * Redy1908/HPC-MultiGPU-MatrixMult) or supplies an independent reference value
 * (double-double / binary128 dot products, closed forms). How each function is
 * pinned is listed beside it.
 *
 * Row-major throughout; leading dimensions are in elements.
 */
#ifndef PHPC_ORACLE_H
#define PHPC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- input generators ----------------
 * Counter-based SplitMix64: element i of stream `seed` is mix64(seed + (i+1)*GAMMA),
 * GAMMA = 0x9E3779B97F4A7C15 (Steele, Lea, Flood 2014). uniform: 2*((z>>11)*2^-53)-1
 * in [-1,1). wide: sign(z&1) * 2^e with e = ((z>>32)*41 >> 32) - 20 in [-20,20].
 * ramp: x[i] = i (src/main.c:85-86, src/iterative.c:30-31).
 * Pinned by: the published SplitMix64 sequence for seed 0 (tests/test_oracle.py),
 * range and moment checks. */
uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i);
void orc_fill_uniform(double *x, size_t n, uint64_t seed);
void orc_fill_wide(double *x, size_t n, uint64_t seed);
void orc_fill_ramp(double *x, size_t n);

/* ---------------- reference arithmetic ----------------
 * orc_iterative: src/iterative.c:8-13. Loop nest i -> j -> k, j is the reduction
 * index, unfused multiply then add straight into C (compiled with
 * -ffp-contract=off). Pinned by: bit-exact match with the reference source built
 * from /root/reference (oracle/_ref), the ramp closed form, brute-force int64. */
void orc_iterative(const double *A, const double *B, double *C, int n);

/* orc_kernel_emul: per-element arithmetic of gemm_kernel, src/phpc_gemm.cu:25-55, as
 * launched by phpc_gemm_cuda (src/phpc_gemm.cu:117) on packed operands with leading
 * dimensions lda/ldb/ldc: c = 0.0; for kk ascending c = fma(A[i,kk], B[kk,j], c)
 * (nvcc contracts `c += a*b` into an FMA by default); then C[i,j] = C[i,j] + c. The
 * zero padding of partial tiles (:38-46) only adds fma(0, 0, c) = c, so it is left
 * out. Pinned by: the ramp closed form, brute-force int64, a hand-computed
 * FMA-vs-unfused known answer, and the double-double bound. */
void orc_kernel_emul(const double *A, int lda, const double *B, int ldb, double *C, int ldc,
                     int m, int k, int n);

/* find_lcm, src/phpc_summa.c:9-22 (Euclid). Pinned by math.lcm. */
int orc_find_lcm(int a, int b);

/* MPI_Dims_create(P, 2) as used by src/main.c:38-45 (1x1 when P == 1): the factor
 * pair (d0 >= d1, d0*d1 = P) with the smallest d0 - d1. Pinned by the MPI standard's
 * examples and brute force over factor pairs. */
void orc_dims_create(int nranks, int dims[2]);

/* Rank block split: part `i` of `n` split into `parts` has size n/parts + (i < n%parts)
 * and starts after the earlier parts (the split rule of src/phpc_gemm.cu:98, applied to
 * rank blocks so that N need not be divisible by the grid). */
int orc_block_size(int n, int parts, int i);
int orc_block_start(int n, int parts, int i);

/* orc_summa_emul: src/phpc_summa.c:36-95 on an emulated p x q grid. The panel width is
 * `panel` (0 selects the reference's N / lcm(p, q), :36-38). Panel t of A (columns
 * [t*w, t*w+w)) and of B (rows [t*w, t*w+w)) are what the owners broadcast at step t
 * (:63-89); every rank then calls the local GEMM (:93), so its C block becomes
 * ((C0 + s_0) + s_1) + ... with each s_t the fresh FMA chain of orc_kernel_emul.
 * Rank blocks follow orc_block_size. Pinned by: equality with orc_kernel_emul for one
 * panel, the ramp closed form, and a hand-computed grouping known answer. */
void orc_summa_emul(const double *A, const double *B, double *C, int N, int p, int q, int panel);
/* orc_summa_entries: the value orc_summa_emul (panel > 0) or orc_kernel_emul (panel <= 0)
 * gives entry (rows[e], cols[e]) of C0 + A B with k = N, computed for the listed entries
 * only: out[e] = ((c0 + s_0) + s_1) + ..., s_t the FMA chain from zero over panel t. C0
 * may be NULL. Pinned by: equality with orc_summa_emul and orc_kernel_emul over every
 * entry of small matrices, for several grids and panel widths (tests/test_oracle.py). */
void orc_summa_entries(const double *A, long lda, const double *B, long ldb, const double *C0, long ldc0,
                       const int *rows, const int *cols, long count, int k, int panel, double *out);

/* ---------------- independent reference values ----------------
 * orc_gemm_dd: C_ref[i,j] = fl(C0[i,j] + sum_k A[i,k] B[k,j]) computed in
 * double-double (TwoProduct by FMA, TwoSum; Ogita, Rump, Oishi 2005 "Dot2"), rounded
 * once to double. C0 may be NULL (treated as 0). Pinned by binary128 (__float128)
 * sums, exact integer results and cancellation known answers. */
void orc_gemm_dd(const double *A, int lda, const double *B, int ldb, const double *C0, int ldc0,
                 double *Cref, int ldcr, int m, int k, int n);
/* Same value for a list of (row, col) entries only (for large N). */
void orc_gemm_dd_entries(const double *A, int lda, const double *B, int ldb, const double *C0,
                         int ldc0, const int *rows, const int *cols, double *out, size_t count,
                         int k);
/* binary128 dot product (independent check of the double-double path). */
double orc_dot_q(const double *a, int inca, const double *b, int incb, int k);

/* |A||B| entry (i,j) in double, used to normalise errors (only the magnitude matters). */
double orc_absdot(const double *A, int lda, const double *B, int ldb, int i, int j, int k);

/* Componentwise error  max_ij |C - Cref| / (|C0| + |A||B|)_ij and normwise
 * max|C - Cref| / max|Cref| over a full m x n matrix. C0 may be NULL. */
void orc_err_full(const double *C, int ldc, const double *Cref, int ldcr, const double *A, int lda,
                  const double *B, int ldb, const double *C0, int ldc0, int m, int k, int n,
                  double *err_cw, double *err_nw);

/* Ramp closed form: with A[r,j] = r*N + j and B[j,c] = j*N + c (src/main.c:85-86),
 * C[r,c] = r*N^2*S1 + r*c*N^2 + N*S2 + c*S1, S1 = N(N-1)/2, S2 = (N-1)N(2N-1)/6.
 * Exact in unsigned 128-bit integers; returned as double (exact while < 2^53, i.e.
 * N <= 1536). Pinned by brute-force int64 for small N. */
double orc_ramp_closed(int N, int r, int c);

/* Freivalds residual: max_i |(C x)_i - (A (B x))_i| / ((|A| (|B| |x|))_i) with
 * x[j] = uniform(seed), all products in double-double. O(N^2). */
double orc_freivalds(const double *A, const double *B, const double *C, int n, uint64_t seed);

/* ---------------- split-precision (Ozaki scheme, int8 slices) ----------------
 * Restatement of the split GEMM the library runs (no reference counterpart; the method
 * is Ozaki, Ogita, Oishi, Rump 2012 with integer slices as in Ootomo, Ozaki, Yokota
 * 2024): row/column power-of-two exponents, exact 7-bit digits by repeated
 * x <- 2^7 x, d = trunc(x), x <- x - d; level sums S_g = sum_{t+u=g} a_t b_u in exact
 * integers, split in k pieces of at most 131072 products exactly as the device splits
 * them; FP64 recombination from level s+1 down to 2; C += 2^(e_i + f_j) * acc.
 * Pinned by: exactness on inputs that fit the slices (small integers, powers of two),
 * the double-double product within the truncation bound, and error decreasing with s. */
void orc_split_gemm(const double *A, int lda, const double *B, int ldb, double *C, int ldc, int m, int k, int n,
                    int s);
/* The same with `bits`-bit digits (x <- 2^bits x) and k pieces of at most `max_terms`
 * products: bits = 7, max_terms = 131072 is orc_split_gemm; bits = 4, max_terms = 74565
 * restates the FP8-E4M3 slices (digits |d| <= 15, exact FP32 sums below 2^24). */
void orc_split_gemm_w(const double *A, int lda, const double *B, int ldb, double *C, int ldc, int m, int k, int n,
                      int s, int bits, long max_terms);
/* Digits of one value: d[0..s-1] and exponent e such that x = 2^e sum d_t 2^{-7t} + r. */
void orc_split_digits(double x, int e, int s, int *d);
/* The same with `bits`-bit digits: x = 2^e sum d_t 2^{-bits t} + r. */
void orc_split_digits_w(double x, int e, int s, int bits, int *d);
int orc_exp_of_max(double mx);

/* ---------------- split-precision, CRT form (Ozaki scheme II) ----------------
 * Restatement of the library's CRT split GEMM: A' = rint(A[i,l] 2^(p - e_i)),
 * B' = rint(B[l,j] 2^(p - f_j)); the first L moduli of
 * {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191,
 *  181, 179, 173}; symmetric residues; exact residue products; Garner mixed-radix digits;
 * double-double Horner; C += 2^(e_i + f_j - 2p) C'. orc_crt_exact returns the exact
 * integer C' = A' B' (as two 64-bit halves) so tests can pin the reconstruction. */
void orc_crt_gemm(const double *A, int lda, const double *B, int ldb, double *C, int ldc, int m, int k, int n,
                  int p, int L);
void orc_crt_digits(const int *residues, int L, int *digits);

#ifdef __cplusplus
}
#endif
#endif

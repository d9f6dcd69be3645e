// Split-precision FP64 GEMM (Ozaki scheme, integer slices) on gfx950 int8 matrix cores.
//
// C[m x n] += A[m x k] * B[k x n], FP64 in and out, computed from exact int8 slices:
//
//   A[i,l] = 2^e_i * sum_{t=1..s} a_t[i,l] 2^{-7t} + (truncation),  a_t in [-127, 127]
//   B[l,j] = 2^f_j * sum_{u=1..s} b_u[l,j] 2^{-7u} + (truncation)
//
// with e_i (f_j) the binary exponent of row i of A (column j of B): |A[i,:]| < 2^e_i.
// The digits come from repeated exact steps x <- 2^7 x, d = trunc(x), x <- x - d.
// Keeping the products with t + u <= s + 1 (triangular truncation, s(s+1)/2 of them),
//
//   C[i,j] += 2^(e_i + f_j) * sum_{g=2..s+1} 2^{-7g} S_g[i,j],   S_g = sum_{t+u=g} a_t b_u.
//
// Each S_g is ONE int8 GEMM: A's slices are stored side by side along k (row i of
// A8 = [a_1 | a_2 | ... | a_s], each slice padded to kp bytes) and B's slices
// transposed (row j of B8 = column j of B) in reverse order ([b_s | ... | b_1]), so
// S_g = A8[:, 0 : (g-1)kp] . B8[:, (s-g+1)kp : s kp]^T, a contiguous k range in both.
// int8 products accumulate exactly in int32 as long as a launch sums at most 133143
// products per output (127^2 * 133143 < 2^31); longer ranges are split into pieces.
// The FP64 epilogue adds the levels smallest first into an FP64 accumulator and the
// last one applies the exact scale 2^(e_i + f_j) and adds into C.
#include "phpc_internal.h"
#include "phpc_split.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

namespace {

typedef int i4v __attribute__((ext_vector_type(4)));
typedef int i16v __attribute__((ext_vector_type(16)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef int i8v __attribute__((ext_vector_type(8)));
typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef __attribute__((address_space(3))) void lds_void;
typedef double __attribute__((address_space(1))) gdouble;

constexpr int kMaxSlices = 12;
constexpr int kMaxSlicesF8 = 16;  // FP8-E4M3 slices carry 4 bits each
// FP32 sums of FP8-digit products (|d| <= 15) stay exact integers below 2^24 over at most
// 74565 products: 225 * 74565 < 2^24
constexpr long kMaxExactTermsF8 = 74565;
constexpr long kMaxExactTerms = 131072;  // products per int32 accumulation (127^2 * 2^17 < 2^31)
// residues reach -128: 128^2 * 130944 < 2^31 (130944 = 2^17 - 128, a whole number of
// 128-byte slabs)
constexpr long kMaxExactTermsCrt = 130944;
// Widest integer operand of the CRT scheme: residues_of cuts |x| < 2^72 into nine bytes,
// and the scaled operands satisfy |x| <= 2^p.
constexpr int kMaxIntBits = 71;

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ long uni(long v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (long)(((uint64_t)hi << 32) | lo);
}
template <typename T>
__device__ __forceinline__ T *uni(T *p) {
  return reinterpret_cast<T *>(uni((long)reinterpret_cast<uintptr_t>(p)));
}
__device__ __forceinline__ rsrc_t make_rsrc(const void *base, long bytes) {
  const uint32_t nb = bytes > 0x7fffffffL ? 0x7fffffffu : (uint32_t)(bytes < 0 ? 0 : bytes);
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)nb, 0x00020000);
}
template <int N>
__device__ __forceinline__ void wait_vm_c() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt(0x0F70 | (N & 0xF) | (((N >> 4) & 3) << 14));
}
// Workgroup barrier of the slab rings. __syncthreads() brackets s_barrier with
// workgroup-scope fences, and the release fence makes the compiler wait for every
// outstanding vector-memory operation (vmcnt(0)) first: that drains the direct-to-LDS
// loads of the slabs still in flight and leaves a ring one slab deep whatever its number
// of stages. The rings wait for the loads they need themselves (wait_vm_c) and write no
// LDS with ordinary stores, so the barrier only has to keep each wave's LDS reads of a
// slab ahead of the loads that refill its buffer; the memory clobber keeps the compiler
// from moving memory operations across it.
__device__ __forceinline__ void slab_barrier() { asm volatile("s_barrier" ::: "memory"); }

// ---------------------------------------------------------------------------
// slicing
// ---------------------------------------------------------------------------

// Binary exponent E with |x| < 2^E for every x of the row/column (max = m * 2^E,
// m in [0.5, 1)); 0 for an all-zero vector.
__device__ __forceinline__ int exp_of_max(double mx) {
  int e = 0;
  if (mx > 0.0) (void)frexp(mx, &e);
  return e;
}

// Low-bit code of a finite x: 1075 - q for x = (odd integer) * 2^q, in [104, 2149]; 0 for
// x = 0. The largest code over a row or column gives its lowest set bit q_min, and with
// its exponent E (every |x| < 2^E) the width W = E - q_min = E + code - 1075: scaled by
// 2^(p - E) with p >= W, every entry of that row or column is an integer, exactly.
__device__ __forceinline__ int low_bit_code(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x) & 0x7FFFFFFFFFFFFFFFull;
  if (u == 0) return 0;
  const int be = (int)(u >> 52);
  const uint64_t man = u & 0xFFFFFFFFFFFFFull;
  const int q = be ? be - 1075 + __builtin_ctzll(man | (1ull << 52)) : -1074 + __builtin_ctzll(man);
  return 1075 - q;
}

// One workgroup per row of A: exponent e_i, 2^e_i / mean|a_i.|, non-finite flag, and the
// largest exact width over the rows (low_bit_code) into *width.
__global__ __launch_bounds__(256) void row_stats_kernel(const double *__restrict__ A, long lda, int k,
                                                        int *__restrict__ eA, double *__restrict__ ratioA,
                                                        int *__restrict__ bad, int *__restrict__ width) {
  __shared__ double red_max[256];
  __shared__ double red_sum[256];
  __shared__ int red_code[256];
  const int i = blockIdx.x;
  const double *row = A + (long)i * lda;
  double mx = 0.0, sm = 0.0;
  int code = 0;
  bool nonfinite = false;
  // four loads in flight per thread; the zeros past k change neither the maximum, the
  // sum (+0.0 added to a sum of magnitudes) nor the code, so the order of the additions,
  // and the bits, are those of one element at a time
  for (int l = threadIdx.x; l < k; l += 1024) {
    double xs[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) xs[q] = load_or_zero(row, l + 256 * q, l + 256 * q < k);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double x = xs[q];
      const double v = fabs(x);
      nonfinite |= !(v <= 1.79769313486231570815e308);  // inf or nan
      mx = fmax(mx, v);
      sm += v;
      code = max(code, low_bit_code(x));
    }
  }
  if (nonfinite) *bad = 1;
  red_max[threadIdx.x] = mx;
  red_sum[threadIdx.x] = sm;
  red_code[threadIdx.x] = code;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      red_max[threadIdx.x] = fmax(red_max[threadIdx.x], red_max[threadIdx.x + off]);
      red_sum[threadIdx.x] += red_sum[threadIdx.x + off];
      red_code[threadIdx.x] = max(red_code[threadIdx.x], red_code[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int e = exp_of_max(red_max[0]);
    eA[i] = e;
    ratioA[i] = red_sum[0] > 0.0 ? ldexp((double)k, e) / red_sum[0] : 0.0;
    if (red_code[0] > 0) atomicMax(width, e + red_code[0] - 1075);
  }
}

// Slice digits: BITS = 7 gives int8 digits |d| <= 127; BITS = 4 gives |d| <= 15, stored as
// FP8-E4M3 (OCP: sign, 4 exponent bits of bias 7, 3 mantissa bits), which holds every
// integer up to 16 exactly: d = 1.mmm * 2^q with q = floor(log2 |d|) <= 3.
__device__ __forceinline__ int8_t e4m3_of(int d) {
  if (d == 0) return 0;
  const int a = d < 0 ? -d : d;
  const int q = 31 - __builtin_clz((unsigned)a);
  const int mant = ((a << 3) >> q) & 7;
  return (int8_t)((d < 0 ? 0x80 : 0) | ((q + 7) << 3) | mant);
}
template <int BITS>
__device__ __forceinline__ int8_t slice_digit(double d) {
  if constexpr (BITS == 4)
    return e4m3_of((int)d);
  else
    return (int8_t)(int)d;
}

// A8 row i: slice t (1-based) at bytes [(t-1)*kp, (t-1)*kp + k), zero up to kp.
template <int BITS>
__global__ __launch_bounds__(256) void slice_rows_kernel(const double *__restrict__ A, long lda, int k, int s,
                                                         long kp, const int *__restrict__ eA,
                                                         int8_t *__restrict__ A8, long ld8) {
  const int i = blockIdx.x;
  const double *row = A + (long)i * lda;
  const int e = eA[i];
  int8_t *out = A8 + (long)i * ld8;
  for (long l = threadIdx.x; l < kp; l += 256) {
    double x = l < k ? ldexp(row[l], -e) : 0.0;
    for (int t = 0; t < s; ++t) {
      x = ldexp(x, BITS);
      const double d = trunc(x);
      x -= d;
      out[(long)t * kp + l] = slice_digit<BITS>(d);
    }
  }
}

__global__ void zero_words_kernel(uint32_t *p, long nwords) {
  for (long w = (long)blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += (long)gridDim.x * blockDim.x) p[w] = 0u;
}

// What the modulus-count rule needs from the row and column ratios (all >= 0 for finite
// inputs): their maxima, their smallest positive values (+inf when none is positive),
// the non-finite flag, and the largest exact widths of A's rows and B's columns (0 when
// all entries are zero). max and min are exact, so the result does not depend on the
// reduction order.
struct RatioSummary {
  double max_a, max_b, minpos_a, minpos_b;
  int bad;
  int width_a, width_b;
  int pad;
};

__global__ __launch_bounds__(256) void ratio_summary_kernel(const double *__restrict__ ra, int m,
                                                            const double *__restrict__ rb, int n,
                                                            const int *__restrict__ bad,
                                                            RatioSummary *__restrict__ out) {
  __shared__ double red[4][256];
  double v[4] = {0.0, 0.0, INFINITY, INFINITY};
  for (int i = threadIdx.x; i < m; i += 256) {
    const double x = ra[i];
    if (x > v[0]) v[0] = x;
    if (x > 0.0 && x < v[2]) v[2] = x;
  }
  for (int j = threadIdx.x; j < n; j += 256) {
    const double x = rb[j];
    if (x > v[1]) v[1] = x;
    if (x > 0.0 && x < v[3]) v[3] = x;
  }
  for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + h]);
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + h]);
      red[2][threadIdx.x] = fmin(red[2][threadIdx.x], red[2][threadIdx.x + h]);
      red[3][threadIdx.x] = fmin(red[3][threadIdx.x], red[3][threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out->max_a = red[0][0];
    out->max_b = red[1][0];
    out->minpos_a = red[2][0];
    out->minpos_b = red[3][0];
    out->bad = bad[0];
    out->width_a = bad[1];
    out->width_b = bad[2];
    out->pad = 0;
  }
}

// Column maxima / sums of B over a chunk of rows; combined with atomics.
__global__ __launch_bounds__(256) void col_stats_kernel(const double *__restrict__ B, long ldb, int k, int n,
                                                        int rows_per_block, unsigned long long *__restrict__ cmax,
                                                        double *__restrict__ csum, int *__restrict__ ccode,
                                                        int *__restrict__ bad) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const int r0 = blockIdx.y * rows_per_block;
  const int r1 = min(k, r0 + rows_per_block);
  double mx = 0.0, sm = 0.0;
  int code = 0;
  bool nonfinite = false;
  // four loads in flight per thread, the zeros past r1 neutral (as in row_stats_kernel)
  for (int r = r0; r < r1; r += 4) {
    double xs[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) xs[q] = load_or_zero(B, (long)(r + q) * ldb + j, r + q < r1);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double x = xs[q];
      const double v = fabs(x);
      // per entry: fmax drops a NaN, so the maximum alone would not show one
      nonfinite |= !(v <= 1.79769313486231570815e308);
      mx = fmax(mx, v);
      sm += v;
      code = max(code, low_bit_code(x));
    }
  }
  if (nonfinite) *bad = 1;
  // non-negative doubles order like their bit patterns
  atomicMax(&cmax[j], (unsigned long long)__double_as_longlong(mx));
  atomicAdd(&csum[j], sm);
  if (code > 0) atomicMax(&ccode[j], code);
}

__global__ void col_exps_kernel(const unsigned long long *__restrict__ cmax, const double *__restrict__ csum,
                                const int *__restrict__ ccode, int k, int n, int *__restrict__ eB,
                                double *__restrict__ ratioB, int *__restrict__ width) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const double mx = __longlong_as_double((long long)cmax[j]);
  const int e = exp_of_max(mx);
  eB[j] = e;
  ratioB[j] = csum[j] > 0.0 ? ldexp((double)k, e) / csum[j] : 0.0;
  if (ccode[j] > 0) atomicMax(width, e + ccode[j] - 1075);
}

// B slices, transposed: B8 row j, slice u at bytes [(s-u)*kp, (s-u)*kp + k), zero to kp.
// 64 (k) x 64 (j) tiles go through LDS so both the read of B and the write of B8 are
// coalesced.
template <int BITS>
__global__ __launch_bounds__(256) void slice_cols_kernel(const double *__restrict__ B, long ldb, int k, int n, int s,
                                                         long kp, const int *__restrict__ eB, int8_t *__restrict__ B8,
                                                         long ld8) {
  __shared__ double tile[64][65];
  const int k0 = blockIdx.x * 64;
  const int j0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4) {
    const int kk = k0 + r, jj = j0 + tx;
    tile[r][tx] = (kk < k && jj < n) ? B[(long)kk * ldb + jj] : 0.0;
  }
  __syncthreads();
  // each thread: column j0 + (row of this loop), 64 consecutive k
  for (int c = ty; c < 64; c += 4) {
    const int jj = j0 + c;
    if (jj >= n) continue;
    const int kk = k0 + tx;
    if (kk >= kp) continue;
    double x = ldexp(tile[tx][c], -eB[jj]);
    int8_t *out = B8 + (long)jj * ld8;
    for (int u = 1; u <= s; ++u) {
      x = ldexp(x, BITS);
      const double d = trunc(x);
      x -= d;
      out[(long)(s - u) * kp + kk] = slice_digit<BITS>(d);
    }
  }
}

// ---------------------------------------------------------------------------
// int8 level GEMM with FP64 epilogue
// ---------------------------------------------------------------------------
// S[i,j] = sum_{l < kl} A8[i, a0 + l] * B8[j, b0 + l]   (exact int32)
// epilogue: v = S * 2^scale_exp;  first ? acc = v : acc += v   (acc: FP64 m x n, ld n)
//           last: C[i,j] += 2^(eA[i]+eB[j]) * acc   (acc not written)
struct LevelArgs {
  const int8_t *A8;
  const int8_t *B8;
  long lda8, ldb8;  // bytes
  long a0, b0;      // k offsets (bytes) of this level piece
  int m, n, kl;
  int scale_exp;
  double *acc;
  long ldacc;
  double *C;
  long ldc;
  const int *eA;
  const int *eB;
  int first, last;
  // residue mode (CRT scheme): res[i,j] = smod(S + (first ? 0 : res[i,j]), mod_m)
  int mod_m;  // 0: level mode (FP64 epilogue above)
  int8_t *res;
  long ldres;
  // BF16 mode (operands are BF16 slices): C32[i,j] += the FP32 sum
  float *C32;
  long ldc32;
  // BF16 mode: the k range may be up to three segments of one sum. Segment s starts at
  // byte seg[s] of the kl range (seg[0] = 0; seg1 = 0: one segment) and reads A from
  // row byte (k - seg[s]) and B from row byte (k - seg[s]) + segb[s].
  int seg1, seg2;
  int segb1, segb2;
};

// v mod m in the symmetric range for |v| < 2^31 and a wave-uniform m: m = 256 gives
// [-128, 127] (v's low byte), an odd m <= 255 gives |r| <= (m-1)/2, in int32 and FP32
// arithmetic with rmf = 1/m rounded to FP32, on a = |v| (the residue of -a is minus that
// of a for odd m). First round: float(a) is within 64 of a and the rounded product within
// about 2 of a/m (< 2^24), so q1 = rint(a * rmf) leaves an exact remainder a - q1 m (a
// 24-bit unsigned multiply) with |r| < 3m < 2^11. Second round: r is exact in FP32, r * rmf
// is within 2^-20 of r/m, which for odd m lies at least 1/(2m) from a half-integer, so
// q2 = rint(r * rmf) is the nearest quotient and r - q2 m the symmetric residue.
__device__ __forceinline__ int smod_i32(int v, int m, float rmf) {
  if (m == 256) return (int)(int8_t)(v & 0xff);
  const uint32_t a = (uint32_t)(v < 0 ? -v : v);
  const uint32_t q1 = (uint32_t)rintf((float)a * rmf);
  const int r = (int)(a - __umul24(q1, (uint32_t)m));
  const int q2 = (int)rintf((float)r * rmf);
  const int s = r - __mul24(q2, m);
  return v < 0 ? -s : s;
}

// LDS layout per stage: A then B, BM (BN) rows of BK k-bytes each, in 16-byte chunks.
// Chunk c of row r sits at slot c ^ sw(r), sw(r) = (r * CPR / 16) mod CPR (CPR chunks per
// row), so the 16-byte operand reads of 16 consecutive rows cover 16 distinct 16-byte
// bank groups. Level ranges are whole multiples of 64 bytes, so every slab is full.
template <int BM, int BN, int WM, int WN, int BK_, int STAGES>
struct I8Cfg {
  static constexpr int BK = BK_;        // bytes of k per slab
  static constexpr int CPR = BK_ / 16;  // chunks per row
  static constexpr int KSTEPS = BK_ / 32;
  static constexpr int NW = WM * WN;
  static constexpr int WTM = BM / WM, WTN = BN / WN;
  static constexpr int TM = WTM / 32, TN = WTN / 32;
  static constexpr int A_BLOCKS = BM * CPR / 64, B_BLOCKS = BN * CPR / 64;
  static constexpr int A_BYTES = A_BLOCKS * 1024, B_BYTES = B_BLOCKS * 1024;
  static constexpr int STAGE_BYTES = A_BYTES + B_BYTES;
  static constexpr int BLOCKS = A_BLOCKS + B_BLOCKS;
  static constexpr int PER = BLOCKS / NW;
  static constexpr int ROWS_PER_BLOCK = 64 / CPR;
  static_assert(BK_ == 64 || BK_ == 128, "slab depth");
  static_assert(BLOCKS % NW == 0 && (BM * CPR) % 64 == 0 && (BN * CPR) % 64 == 0, "blocks must split evenly");
  static_assert(STAGES * STAGE_BYTES <= 160 * 1024, "LDS budget");
  static_assert(STAGES >= 2 && STAGES <= 5 && 3 * PER < 64, "pipeline depth");
  static_assert(TM >= 1 && TN >= 1, "wave tile");
};

// LDS row swizzle: chunk c of row r is stored at slot c ^ i8_swz(r). SD = 16 spreads the
// 16-byte chunks of 16 consecutive rows over 64 banks; SD = 8 over 32 banks.
template <int CPR, int SD = 16>
__device__ __forceinline__ int i8_swz(int r) {
  return (r * CPR / SD) & (CPR - 1);
}

// BF = false: int8 operands, v_mfma_i32_32x32x32_i8, int32 sums (level or residue
// epilogue). BF = true: BF16 operands (8 per 16-byte chunk), v_mfma_f32_32x32x16_bf16,
// FP32 sums added to C32. F8 = true: FP8-E4M3 operands (16 per chunk); the chunks of two
// k steps make one lane's 32 bytes of v_mfma_scale_f32_32x32x64_f8f6f4 at unit scales
// (A and B take the same bytes, so the sum runs over the same 64 k), FP32 sums that stay
// exact integers below 2^24, FP64 level epilogue. The byte layout and the pipeline are
// the same.
// The operands of one level product as the tile code reads them (wave-uniform values):
// read once per product by the plain walk, once per tile by the grouped one.
struct LevelOps {
  const int8_t *A8, *B8;  // at the product's k offset
  long lda8, ldb8;
  int m, n, nslab;
  int sg1, sg2, sb1, sb2;  // BF16 segments
  float *C32;              // BF16 epilogue
  long ldc32;
};
template <int BK, bool BF>
__device__ __forceinline__ LevelOps level_ops(const LevelArgs &args) {
  LevelOps o;
  o.m = uni(args.m);
  o.n = uni(args.n);
  o.nslab = uni(args.kl) / BK;
  o.lda8 = uni(args.lda8);
  o.ldb8 = uni(args.ldb8);
  o.A8 = uni(args.A8) + uni(args.a0);
  o.B8 = uni(args.B8) + uni(args.b0);
  o.sg1 = BF ? uni(args.seg1) : 0;
  o.sg2 = BF ? uni(args.seg2) : 0;
  o.sb1 = BF ? uni(args.segb1) : 0;
  o.sb2 = BF ? uni(args.segb2) : 0;
  o.C32 = BF ? uni(args.C32) : nullptr;
  o.ldc32 = BF ? uni(args.ldc32) : 0;
  return o;
}

// One BM x BN output tile at (row0, col0) of the product ops describes; args carries the
// int8 and FP8 epilogues' fields (the tile walk is the caller's). MI: the MFMA block,
// 32 x 32 (32 bytes of k per int8 product, 16 per BF16 one) or 16 x 16 (v_mfma_i32_16x16x64_i8
// and v_mfma_f32_16x16x32_bf16: 64 bytes of k); a lane's 16-byte operand is row l % MI,
// chunk (64 / MI) kk + l / MI of the k step. Not for FP8 at 16.
template <int BM, int BN, int WM, int WN, int BK_, int STAGES, int SD, bool BF, bool F8, int MI = 32>
__device__ __forceinline__ void level_tile(const LevelArgs &args, const LevelOps &ops, int row0, int col0) {
  using K = I8Cfg<BM, BN, WM, WN, BK_, STAGES>;
  static_assert(MI == 32 || (MI == 16 && !F8), "MFMA block");
  constexpr int TM = K::WTM / MI, TN = K::WTN / MI;
  constexpr int KSTEPS = MI == 32 ? K::KSTEPS : K::BK / 64;
  static_assert(TM >= 1 && TN >= 1 && KSTEPS >= 1, "wave tile");
  extern __shared__ __attribute__((aligned(1024))) char smem[];
  const int tid = threadIdx.x;
  const int wave = uni(tid >> 6);
  const int lane = tid & 63;
  const int wm = wave / WN, wn = wave % WN;

  const int m = ops.m, n = ops.n;
  const long lda8 = ops.lda8, ldb8 = ops.ldb8;
  const int8_t *A8 = ops.A8;
  const int8_t *B8 = ops.B8;
  const int nslab = ops.nslab;
  const int sg1 = ops.sg1, sg2 = ops.sg2;
  const int sb1 = ops.sb1, sb2 = ops.sb2;

  // LDS operand offsets per k step: lane reads 16 bytes of row (l % 32), chunk
  // 2*kk + l/32, at its swizzled slot. The rows a lane reads are l % 32 plus multiples of
  // 32, and the swizzle of a multiple of 32 is 0, so the slot depends on l and kk alone:
  // fragment i of A sits at a_row + i*32*BK + cofs[kk] (B likewise), the constant part
  // folding into the LDS instruction's offset.
  // (MI = 16: rows l % 16 plus multiples of 16, whose swizzle is 0 too)
  const int sw = i8_swz<K::CPR, SD>(lane & (MI - 1));
  const int a_row = (wm * K::WTM + (lane & (MI - 1))) * K::BK;
  const int b_row = K::A_BYTES + (wn * K::WTN + (lane & (MI - 1))) * K::BK;
  int cofs[KSTEPS];
#pragma unroll
  for (int kk = 0; kk < KSTEPS; ++kk) cofs[kk] = (((64 / MI) * kk + lane / MI) ^ sw) << 4;

  {
    const int m_rem = m - row0, n_rem = n - col0;
    const rsrc_t rA = make_rsrc(A8 + (long)row0 * lda8, (long)min(m_rem, BM) * lda8);
    const rsrc_t rB = make_rsrc(B8 + (long)col0 * ldb8, (long)min(n_rem, BN) * ldb8);

    // this wave's blocks: wave, wave + NW, ...; lane p of a block fills slot p % CPR of
    // row p / CPR, i.e. chunk (p % CPR) ^ sw(row); offsets for k = 0 (kb in soffset).
    // Block u of A (u < PA) covers rows (wave + u NW) RPB + lane / CPR: the row advances
    // by NW RPB per u, a multiple of 32, so the swizzled slot is the same for every u and
    // the offset is a0 + u * (NW RPB lda8) (B likewise, its blocks numbered after A's).
    // Rows past the tile's last one (m_rem, n_rem) fall at or beyond the buffer's
    // num_records (rows x ld), so they read zero.
    constexpr int PA = K::A_BLOCKS / K::NW;
    constexpr int RSTEP = K::NW * K::ROWS_PER_BLOCK;
    static_assert(K::A_BLOCKS % K::NW == 0 && RSTEP % 32 == 0, "blocks split evenly between A and B");
    const int lrow = lane / K::CPR;
    const uint32_t ga0 =
        (uint32_t)((long)(wave * K::ROWS_PER_BLOCK + lrow) * lda8 +
                   (((lane % K::CPR) ^ i8_swz<K::CPR, SD>(wave * K::ROWS_PER_BLOCK + lrow)) << 4));
    const uint32_t gb0 =
        (uint32_t)((long)(wave * K::ROWS_PER_BLOCK + lrow) * ldb8 +
                   (((lane % K::CPR) ^ i8_swz<K::CPR, SD>(wave * K::ROWS_PER_BLOCK + lrow)) << 4));
    const uint32_t gstep_a = (uint32_t)(RSTEP * lda8), gstep_b = (uint32_t)(RSTEP * ldb8);

    // k byte offsets of slab sidx in A's and B's rows
    auto slab_k = [&](int sidx, int &ka, int &kbb) {
      const int kb = sidx * K::BK;
      ka = kb;
      kbb = kb;
      if constexpr (BF) {
        // segments start on slab boundaries, so a slab lies in one segment
        if (sg1 > 0 && kb >= sg1) {
          if (sg2 > 0 && kb >= sg2) {
            ka = kb - sg2;
            kbb = ka + sb2;
          } else {
            ka = kb - sg1;
            kbb = ka + sb1;
          }
        }
      }
    };
    auto issue = [&](int sidx) {
      int ka, kbb;
      slab_k(sidx, ka, kbb);
      char *stage = smem + (sidx % STAGES) * K::STAGE_BYTES;
#pragma unroll
      for (int u = 0; u < K::PER; ++u) {
        const int blk = wave + u * K::NW;
        const bool isA = u < PA;
        // the range check applies to the lane offset (past num_records reads zero); k
        // rides in soffset
        const uint32_t go = isA ? ga0 + (uint32_t)u * gstep_a : gb0 + (uint32_t)(u - PA) * gstep_b;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(isA ? rA : rB, (lds_void *)(stage + blk * 1024), 16, go,
                                                  isA ? ka : kbb, 0, 0);
      }
    };

    using acc32_t = typename std::conditional<BF || F8, f16v, i16v>::type;
    using acc16_t = typename std::conditional<BF, f4v, i4v>::type;
    using acc_t = typename std::conditional<MI == 32, acc32_t, acc16_t>::type;
    acc_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{};

    slab_barrier();
#pragma unroll
    for (int s0 = 0; s0 < STAGES - 1; ++s0)
      if (s0 < nslab) issue(s0);
    for (int sidx = 0; sidx < nslab; ++sidx) {
      // slabs issued after this one that may stay in flight
      const int later = min(STAGES - 2, nslab - 1 - sidx);
      if (later <= 0)
        wait_vm_c<0>();
      else if (later == 1)
        wait_vm_c<K::PER>();
      else if (later == 2)
        wait_vm_c<2 * K::PER>();
      else
        wait_vm_c<3 * K::PER>();
      slab_barrier();
      if (sidx + STAGES - 1 < nslab) issue(sidx + STAGES - 1);

      const char *stage = smem + (sidx % STAGES) * K::STAGE_BYTES;
      if constexpr (F8) {
        // two k steps' 16-byte chunks make one lane's 32 bytes of a 64-deep FP8 product
        static_assert(K::KSTEPS % 2 == 0, "FP8 products take k steps in pairs");
#pragma unroll
        for (int kk = 0; kk < K::KSTEPS; kk += 2) {
          i8v a[K::TM], b[K::TN];
#pragma unroll
          for (int i = 0; i < K::TM; ++i) {
            const i4v lo = *reinterpret_cast<const i4v *>(stage + a_row + cofs[kk] + i * 32 * K::BK);
            const i4v hi = *reinterpret_cast<const i4v *>(stage + a_row + cofs[kk + 1] + i * 32 * K::BK);
            a[i] = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
          }
#pragma unroll
          for (int j = 0; j < K::TN; ++j) {
            const i4v lo = *reinterpret_cast<const i4v *>(stage + b_row + cofs[kk] + j * 32 * K::BK);
            const i4v hi = *reinterpret_cast<const i4v *>(stage + b_row + cofs[kk + 1] + j * 32 * K::BK);
            b[j] = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
          }
          // formats 0/0: FP8-E4M3 operands; scales 127: 2^0 (E8M0)
#pragma unroll
          for (int i = 0; i < K::TM; ++i)
#pragma unroll
            for (int j = 0; j < K::TN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[i], b[j], acc[i][j], 0, 0, 0, 127, 0, 127);
        }
      } else {
#pragma unroll
        for (int kk = 0; kk < KSTEPS; ++kk) {
          i4v a[TM], b[TN];
#pragma unroll
          for (int i = 0; i < TM; ++i)
            a[i] = *reinterpret_cast<const i4v *>(stage + a_row + cofs[kk] + i * MI * K::BK);
#pragma unroll
          for (int j = 0; j < TN; ++j)
            b[j] = *reinterpret_cast<const i4v *>(stage + b_row + cofs[kk] + j * MI * K::BK);
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              if constexpr (BF && MI == 32)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8v, a[i]),
                                                                   __builtin_bit_cast(bf8v, b[j]), acc[i][j], 0, 0, 0);
              else if constexpr (BF)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a[i]),
                                                                   __builtin_bit_cast(bf8v, b[j]), acc[i][j], 0, 0, 0);
              else if constexpr (MI == 32)
                acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[i], b[j], acc[i][j], 0, 0, 0);
              else
                acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b[j], acc[i][j], 0, 0, 0);
        }
      }
    }

    // epilogue: 32x32 block, lane l / reg r -> row 8(r/4) + 4(l/32) + r%4, column l%32;
    // 16x16 block: row 4(l/16) + r, column l%16
    constexpr int NR = MI == 32 ? 16 : 4;
    auto erow = [&](int i, int r) {
      return MI == 32 ? i * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4) : i * 16 + 4 * (lane >> 4) + r;
    };
    if constexpr (BF) {
      float __attribute__((address_space(1))) *Cg = (float __attribute__((address_space(1))) *)ops.C32;
      const long ldc32 = ops.ldc32;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int col = col0 + wn * K::WTN + j * MI + (lane & (MI - 1));
          if (col >= n) continue;
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const int row = row0 + wm * K::WTM + erow(i, r);
            if (row >= m) continue;
            auto *pc = Cg + (long)row * ldc32 + col;
            *pc = *pc + acc[i][j][r];
          }
        }
      return;
    }
    const int mod_m = uni(args.mod_m);
    if (mod_m > 0) {
      const int firstr = uni(args.first);
      const float rmf = 1.0f / (float)mod_m;
      int8_t *resg = uni(args.res);
      const long ldres = uni(args.ldres);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int col = col0 + wn * K::WTN + j * MI + (lane & (MI - 1));
          if (col >= n) continue;
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const int row = row0 + wm * K::WTM + erow(i, r);
            if (row >= m) continue;
            int8_t *pr = resg + (long)row * ldres + col;
            // |sum| <= 128^2 x 130944 < 2^31 - 128 (kMaxExactTermsCrt), so adding the
            // previous piece's residue stays in int32
            int v = (int)acc[i][j][r];
            if (!firstr) v += *pr;
            *pr = (int8_t)smod_i32(v, mod_m, rmf);
          }
        }
      return;
    }
    const int sc = uni(args.scale_exp);
    const int first = uni(args.first), last = uni(args.last);
    gdouble *accg = (gdouble *)uni(args.acc);
    gdouble *Cg = (gdouble *)uni(args.C);
    const long ldacc = uni(args.ldacc), ldc = uni(args.ldc);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int col = col0 + wn * K::WTN + j * MI + (lane & (MI - 1));
        if (col >= n) continue;
        const int ej = last ? args.eB[col] : 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const int row = row0 + wm * K::WTM + erow(i, r);
          if (row >= m) continue;
          double v = ldexp((double)acc[i][j][r], sc);
          gdouble *pa = accg + (long)row * ldacc + col;
          if (!first) v = *pa + v;
          if (last) {
            gdouble *pc = Cg + (long)row * ldc + col;
            *pc = *pc + ldexp(v, args.eA[row] + ej);
          } else {
            *pa = v;
          }
        }
      }
  }
}

template <int BM, int BN, int WM, int WN, int BK_, int STAGES, int SD = 16, bool BF = false, bool F8 = false,
          int MI = 32>
__device__ __forceinline__ void i8_level_body(const LevelArgs &args) {
  const LevelOps ops = level_ops<BK_, BF>(args);
  const int m = ops.m, n = ops.n;
  const int tiles_n = (n + BN - 1) / BN;
  const int tiles_m = (m + BM - 1) / BM;
  const int total = tiles_m * tiles_n;
  const int G = gridDim.x;
  const int bid = blockIdx.x;
  const int slot = uni((G % 8) == 0 ? (bid % 8) * (G / 8) + bid / 8 : bid);
  for (int round = 0;; ++round) {
    const int t = round * G + slot;
    if (t >= total) break;
    // group 8 tile-rows for L2 reuse
    const int gm = 8;
    const int grp = t / (gm * tiles_n);
    const int first_m = grp * gm;
    const int gsz = min(tiles_m - first_m, gm);
    const int in_grp = t - grp * gm * tiles_n;
    const int row0 = uni((first_m + in_grp % gsz) * BM);
    const int col0 = uni((in_grp / gsz) * BN);
    level_tile<BM, BN, WM, WN, BK_, STAGES, SD, BF, F8, MI>(args, ops, row0, col0);
  }
}

__global__ __launch_bounds__(512) void i8_level_256x256_w8(LevelArgs a) { i8_level_body<256, 256, 2, 4, 64, 3>(a); }
__global__ __launch_bounds__(256) void i8_level_128x128_w4(LevelArgs a) { i8_level_body<128, 128, 2, 2, 64, 4>(a); }
__global__ __launch_bounds__(512) void i8_level_256x256_w8_k128(LevelArgs a) {
  i8_level_body<256, 256, 2, 4, 128, 2>(a);
}
__global__ __launch_bounds__(512) void i8_level_256x128_w8(LevelArgs a) { i8_level_body<256, 128, 4, 2, 64, 4>(a); }
__global__ __launch_bounds__(512) void i8_level_256x256_w8_s4(LevelArgs a) { i8_level_body<256, 256, 2, 4, 64, 4>(a); }
__global__ __launch_bounds__(512) void i8_level_256x256_w8_s5(LevelArgs a) { i8_level_body<256, 256, 2, 4, 64, 5>(a); }
__global__ __launch_bounds__(512) void i8_level_256x256_w8_k128_sd8(LevelArgs a) {
  i8_level_body<256, 256, 2, 4, 128, 2, 8>(a);
}
__global__ __launch_bounds__(512) void i8_level_256x256_w8_sd8(LevelArgs a) { i8_level_body<256, 256, 2, 4, 64, 3, 8>(a); }
// 16 waves (4 per SIMD) with 64x64 wave tiles: more waves to cover the slab loads
__global__ __launch_bounds__(1024) void i8_level_256x256_w16_k128(LevelArgs a) {
  i8_level_body<256, 256, 4, 4, 128, 2>(a);
}
// deeper rings: a 256x128 tile with three 128-byte stages (48 KiB each), and the
// 16-wave 256x256 tile with four or five 64-byte stages (32 KiB each)
__global__ __launch_bounds__(512) void i8_level_256x128_w8_k128_s3(LevelArgs a) {
  i8_level_body<256, 128, 4, 2, 128, 3>(a);
}
__global__ __launch_bounds__(1024) void i8_level_256x256_w16_k64_s4(LevelArgs a) {
  i8_level_body<256, 256, 4, 4, 64, 4>(a);
}
__global__ __launch_bounds__(1024) void i8_level_256x256_w16_k64_s5(LevelArgs a) {
  i8_level_body<256, 256, 4, 4, 64, 5>(a);
}
__global__ __launch_bounds__(1024) void i8_level_256x128_w16_k128_s3(LevelArgs a) {
  i8_level_body<256, 128, 8, 2, 128, 3>(a);
}
// the default tile on 16 x 16 MFMA blocks (v_mfma_i32_16x16x64_i8)
__global__ __launch_bounds__(1024) void i8_level_256x256_w16_k128_mi16(LevelArgs a) {
  i8_level_body<256, 256, 4, 4, 128, 2, 16, false, false, 16>(a);
}
__global__ __launch_bounds__(512) void i8_level_256x256_w8_k128_mi16(LevelArgs a) {
  i8_level_body<256, 256, 2, 4, 128, 2, 16, false, false, 16>(a);
}
__global__ __launch_bounds__(256) void i8_level_128x128_w4_mi16(LevelArgs a) {
  i8_level_body<128, 128, 2, 2, 64, 4, 16, false, false, 16>(a);
}
// FP8-E4M3 digit slices on the same pipeline (FP64-accurate, scheme I with 4-bit digits)
__global__ __launch_bounds__(1024) void f8_level_256x256_w16_k128(LevelArgs a) {
  i8_level_body<256, 256, 4, 4, 128, 2, 16, false, true>(a);
}
__global__ __launch_bounds__(256) void f8_level_128x128_w4(LevelArgs a) {
  i8_level_body<128, 128, 2, 2, 64, 4, 16, false, true>(a);
}
// BF16 slices on the same pipeline (FP32-accurate 3 x BF16 GEMM). Every BF16 level kernel,
// single-product and grouped, uses the same MFMA block, so that they sum each output in the
// same order.
constexpr int kBfMI = 16;
__global__ __launch_bounds__(512) void bf16_level_256x256_w8_k128(LevelArgs a) {
  i8_level_body<256, 256, 2, 4, 128, 2, 16, true, false, kBfMI>(a);
}
__global__ __launch_bounds__(1024) void bf16_level_256x256_w16_k128(LevelArgs a) {
  i8_level_body<256, 256, 4, 4, 128, 2, 16, true, false, kBfMI>(a);
}
__global__ __launch_bounds__(256) void bf16_level_128x128_w4(LevelArgs a) {
  i8_level_body<128, 128, 2, 2, 64, 4, 16, true, false, kBfMI>(a);
}

typedef void (*LevelFn)(LevelArgs);
struct I8Config {
  int bm, bn, bk, threads;
  size_t lds;
  LevelFn fn;
};
#define I8CFG(BM_, BN_, WM_, WN_, BK_, ST_, FN_) \
  I8Config { BM_, BN_, BK_, WM_ *WN_ * 64, (size_t)(ST_)*I8Cfg<BM_, BN_, WM_, WN_, BK_, ST_>::STAGE_BYTES, FN_ }
const I8Config kI8Configs[] = {
    I8CFG(256, 256, 2, 4, 64, 3, i8_level_256x256_w8),            // 0: 8 waves (wave tile 128x64), 3 stages
    I8CFG(128, 128, 2, 2, 64, 4, i8_level_128x128_w4),            // 1: 4 waves, 64x64 wave tiles (small problems)
    I8CFG(256, 256, 2, 4, 128, 2, i8_level_256x256_w8_k128),      // 2: 8 waves, 128-byte slabs
    I8CFG(256, 128, 4, 2, 64, 4, i8_level_256x128_w8),            // 3: 8 waves, wave tile 64x64
    I8CFG(256, 256, 2, 4, 64, 4, i8_level_256x256_w8_s4),         // 4: 8 waves, 4-stage pipeline
    I8CFG(256, 256, 2, 4, 64, 5, i8_level_256x256_w8_s5),         // 5: 8 waves, 5-stage pipeline
    I8CFG(256, 256, 2, 4, 128, 2, i8_level_256x256_w8_k128_sd8),  // 6: as 2, 32-bank swizzle
    I8CFG(256, 256, 2, 4, 64, 3, i8_level_256x256_w8_sd8),        // 7: as 0, 32-bank swizzle
    I8CFG(256, 256, 4, 4, 128, 2, i8_level_256x256_w16_k128),     // 8: 16 waves, 64x64 wave tiles
    I8CFG(256, 128, 4, 2, 128, 3, i8_level_256x128_w8_k128_s3),   // 9: 8 waves, 3 stages of 128-byte slabs
    I8CFG(256, 256, 4, 4, 64, 4, i8_level_256x256_w16_k64_s4),    // 10: 16 waves, 4 stages of 64-byte slabs
    I8CFG(256, 256, 4, 4, 64, 5, i8_level_256x256_w16_k64_s5),    // 11: 16 waves, 5 stages of 64-byte slabs
    I8CFG(256, 128, 8, 2, 128, 3, i8_level_256x128_w16_k128_s3),  // 12: 16 waves (32x64 wave tiles), 3 stages
    I8CFG(256, 256, 4, 4, 128, 2, i8_level_256x256_w16_k128_mi16),  // 13: as 8 on 16x16 MFMA blocks (default)
    I8CFG(256, 256, 2, 4, 128, 2, i8_level_256x256_w8_k128_mi16),   // 14: as 2 on 16x16 MFMA blocks
    I8CFG(128, 128, 2, 2, 64, 4, i8_level_128x128_w4_mi16),         // 15: as 1 on 16x16 MFMA blocks
};
constexpr int kNumI8 = sizeof(kI8Configs) / sizeof(kI8Configs[0]);
int g_i8_big = 13;  // kI8Configs index for large problems

// BF16-slice level kernels (FP32-accurate 3 x BF16 GEMM), kept apart from the int8
// table so that phpc_split_set_config can only pick int8 kernels
const I8Config kBfConfigs[] = {
    I8CFG(256, 256, 4, 4, 128, 2, bf16_level_256x256_w16_k128),  // 0: large problems, 16 waves
    I8CFG(128, 128, 2, 2, 64, 4, bf16_level_128x128_w4),         // 1: small problems
    I8CFG(256, 256, 2, 4, 128, 2, bf16_level_256x256_w8_k128),   // 2: 8 waves
};
constexpr int kNumBf = sizeof(kBfConfigs) / sizeof(kBfConfigs[0]);

// FP8-digit level kernels (scheme I with FP8-E4M3 slices)
const I8Config kF8Configs[] = {
    I8CFG(256, 256, 4, 4, 128, 2, f8_level_256x256_w16_k128),  // 0: large problems, 16 waves
    I8CFG(128, 128, 2, 2, 64, 4, f8_level_128x128_w4),         // 1: small problems
};
constexpr int kNumF8 = sizeof(kF8Configs) / sizeof(kF8Configs[0]);

std::mutex g_mu;
bool g_ready = false;
int g_occ[kNumI8];
int g_occ_bf[kNumBf];
int g_occ_f8[kNumF8];

int occupancy_of(const I8Config &c) {
  const void *fn = reinterpret_cast<const void *>(c.fn);
  (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.lds);
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, c.threads, c.lds) != hipSuccess || n < 1) n = 1;
  return n;
}

void init_locked() {
  if (g_ready) return;
  for (int c = 0; c < kNumI8; ++c) g_occ[c] = occupancy_of(kI8Configs[c]);
  for (int c = 0; c < kNumBf; ++c) g_occ_bf[c] = occupancy_of(kBfConfigs[c]);
  for (int c = 0; c < kNumF8; ++c) g_occ_f8[c] = occupancy_of(kF8Configs[c]);
  g_ready = true;
}

int launch_with(const I8Config &c, int occ, const LevelArgs &a, hipStream_t s) {
  // the kernels walk k in whole slabs of c.bk bytes: a range that is not a whole number
  // of slabs, or starts inside one, would lose its tail
  if (a.kl <= 0 || a.kl % c.bk != 0 || a.a0 % 16 != 0 || a.b0 % 16 != 0 ||
      (a.seg1 | a.seg2 | a.segb1 | a.segb2) % c.bk != 0)
    return phpc_set_error(PHPC_ERR_ARG, "level range of %d bytes is not whole %d-byte slabs", a.kl, c.bk);
  // buffer offsets are 32-bit: a tile's rows of A and B must span less than 2 GiB
  if ((long)c.bm * a.lda8 >= 0x7fffffffL || (long)c.bn * a.ldb8 >= 0x7fffffffL)
    return phpc_set_error(PHPC_ERR_ARG, "operand rows of %ld / %ld bytes too long for %d x %d tiles", a.lda8,
                          a.ldb8, c.bm, c.bn);
  const long tiles = (long)((a.m + c.bm - 1) / c.bm) * ((a.n + c.bn - 1) / c.bn);
  long G = (long)occ * phpc_device().cus;
  if (G > tiles) G = tiles;
  if (G < 1) G = 1;
  const int kt = phpc_ktime_begin(s);
  hipLaunchKernelGGL(c.fn, dim3((unsigned)G), dim3(c.threads), c.lds, s, a);
  phpc_ktime_end(kt, s);
  PHPC_HIP(hipGetLastError());
  return PHPC_OK;
}

int launch_level(const LevelArgs &a, int cfg, hipStream_t s) {
  int occ;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    init_locked();
    occ = g_occ[cfg];
  }
  return launch_with(kI8Configs[cfg], occ, a, s);
}

int launch_f8_level(const LevelArgs &a, int cfg, hipStream_t s) {
  int occ;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    init_locked();
    occ = g_occ_f8[cfg];
  }
  return launch_with(kF8Configs[cfg], occ, a, s);
}

int launch_bf_level(const LevelArgs &a, int cfg, hipStream_t s) {
  int occ;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    init_locked();
    occ = g_occ_bf[cfg];
  }
  return launch_with(kBfConfigs[cfg], occ, a, s);
}

// ---------------------------------------------------------------------------
// FP32 -> three BF16 slice planes for the level kernels: x = hi + mid + lo exactly,
// each rounded to nearest BF16 (the same cut as the FP32-accurate grouped kernel)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ void split3(float x, uint32_t &h, uint32_t &mi, uint32_t &lo) {
  h = bf16_rne(x);
  const float r1 = x - __uint_as_float(h << 16);
  mi = bf16_rne(r1);
  const float r2 = r1 - __uint_as_float(mi << 16);
  lo = bf16_rne(r2);
}
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

// A3 row i = [hi | mid | lo], each plane kp BF16 wide (zero beyond k); 8 k per thread
__global__ __launch_bounds__(256) void bf16x3_rows_kernel(const float *__restrict__ A, long lda, int m, int k,
                                                          long kp, uint16_t *__restrict__ A3) {
  const int i = blockIdx.y;
  const float *row = A + (long)i * lda;
  uint16_t *out = A3 + (long)i * 3 * kp;
  for (long c = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 8; c < kp; c += (long)gridDim.x * blockDim.x * 8) {
    uint32_t h[8], mi[8], lo[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) split3(load_or_zero(row, c + e, c + e < k), h[e], mi[e], lo[e]);
    u4v vh, vm, vl;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      vh[q] = h[2 * q] | (h[2 * q + 1] << 16);
      vm[q] = mi[2 * q] | (mi[2 * q + 1] << 16);
      vl[q] = lo[2 * q] | (lo[2 * q + 1] << 16);
    }
    *reinterpret_cast<u4v *>(out + c) = vh;
    *reinterpret_cast<u4v *>(out + kp + c) = vm;
    *reinterpret_cast<u4v *>(out + 2 * kp + c) = vl;
  }
}

// B3 row j (column j of B) = [lo | mid | hi], through a 64 x 64 LDS transpose
__global__ __launch_bounds__(256) void bf16x3_cols_kernel(const float *__restrict__ B, long ldb, int k, int n,
                                                          long kp, uint16_t *__restrict__ B3) {
  __shared__ float tile[64][65];
  const int k0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  float v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int kk = k0 + ty + 4 * i, jj = j0 + tx;
    v[i] = load_or_zero(B, (long)kk * ldb + jj, kk < k && jj < n);
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) tile[ty + 4 * i][tx] = v[i];
  __syncthreads();
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int item = threadIdx.x + it * 256;  // 64 columns x 8 groups of 8 k
    const int c = item >> 3, kq = (item & 7) * 8;
    const int jj = j0 + c;
    if (jj >= n) continue;
    uint32_t h[8], mi[8], lo[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) split3(tile[kq + e][c], h[e], mi[e], lo[e]);
    u4v vh, vm, vl;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      vh[q] = h[2 * q] | (h[2 * q + 1] << 16);
      vm[q] = mi[2 * q] | (mi[2 * q + 1] << 16);
      vl[q] = lo[2 * q] | (lo[2 * q + 1] << 16);
    }
    uint16_t *out = B3 + (long)jj * 3 * kp + k0 + kq;
    *reinterpret_cast<u4v *>(out) = vl;
    *reinterpret_cast<u4v *>(out + kp) = vm;
    *reinterpret_cast<u4v *>(out + 2 * kp) = vh;
  }
}

// Every cut of a grouped 3 x BF16 plan in one launch: one workgroup per 64 plane rows x 64
// k of a cut, its cut read from a per-workgroup table. Each thread cuts 8 consecutive k of
// one plane row (two per thread) and writes 16 bytes per plane. A's rows are its plane
// rows, so a thread reads its 8 k straight from the row (16-byte loads where rows are
// aligned); B's tile passes through LDS as tile[plane row][k], read along B's rows
// (coalesced). Every load of a thread is issued before the first wait. The cut is split3,
// as in bf16x3_rows_kernel / bf16x3_cols_kernel.
__global__ __launch_bounds__(256) void bf16x3_cut_kernel(const Bf3Cut *__restrict__ cuts,
                                                         const int *__restrict__ block_cut) {
  __shared__ float tile[64][65];
  const int b = blockIdx.x;
  const Bf3Cut c = cuts[block_cut[b]];
  const int lb = b - c.block_start;
  const int r0 = (lb / c.tiles_k) * 64;  // first plane row
  const int k0 = (lb % c.tiles_k) * 64;
  const int nrows = c.is_b ? c.cols : c.rows;  // plane rows
  const int kk = c.is_b ? c.rows : c.cols;     // k
  float x2[2][8];
  if (c.is_b) {
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    float v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      // B row k0 + r, columns r0 + tx: plane row tx, k r
      const int r = ty + 4 * i;
      const int sk = k0 + r, sj = r0 + tx;
      v[i] = load_or_zero(c.src, (long)sk * c.ld + sj, sk < kk && sj < nrows);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) tile[tx][ty + 4 * i] = v[i];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const int item = threadIdx.x + it * 256;
#pragma unroll
      for (int e = 0; e < 8; ++e) x2[it][e] = tile[item >> 3][(item & 7) * 8 + e];
    }
  } else {
    // both items' loads go out before the first wait
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const int item = threadIdx.x + it * 256;
      const int pr = min(item >> 3, nrows - 1 - r0), kq = (item & 7) * 8;  // rows past the edge: unused
      const long base = (long)(r0 + pr) * c.ld + k0 + kq;
      const int krem = kk - (k0 + kq);
      if (c.vec && krem >= 8) {
        const float4 v0 = *reinterpret_cast<const float4 *>(c.src + base);
        const float4 v1 = *reinterpret_cast<const float4 *>(c.src + base + 4);
        x2[it][0] = v0.x, x2[it][1] = v0.y, x2[it][2] = v0.z, x2[it][3] = v0.w;
        x2[it][4] = v1.x, x2[it][5] = v1.y, x2[it][6] = v1.z, x2[it][7] = v1.w;
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) x2[it][e] = load_or_zero(c.src, base + e, e < krem);
      }
    }
  }
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int item = threadIdx.x + it * 256;  // 64 plane rows x 8 groups of 8 k
    const int pr = item >> 3, kq = (item & 7) * 8;
    if (r0 + pr >= nrows) continue;
    const float *x = x2[it];
    uint32_t h[8], mi[8], lo[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) split3(x[e], h[e], mi[e], lo[e]);
    u4v vh, vm, vl;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      vh[q] = h[2 * q] | (h[2 * q + 1] << 16);
      vm[q] = mi[2 * q] | (mi[2 * q + 1] << 16);
      vl[q] = lo[2 * q] | (lo[2 * q + 1] << 16);
    }
    uint16_t *out = c.dst + (long)(r0 + pr) * 3 * c.kp + k0 + kq;
    *reinterpret_cast<u4v *>(out) = c.is_b ? vl : vh;
    *reinterpret_cast<u4v *>(out + c.kp) = vm;
    *reinterpret_cast<u4v *>(out + 2 * c.kp) = c.is_b ? vh : vl;
  }
}

// One grouped launch of the BF16 level kernel over a 3 x BF16 plan's flat tile list
// (items in the plan's order, each item's tiles in groups of 8 tile-rows): the persistent
// workgroups take tiles from XCD-contiguous slots, find the tile's item by binary search
// on tile_start, and run the level tile on its planes: the three k segments of
// phpc_sgemm_bf16x3_planes, so the same sum per output.
struct Bf3Group {
  const Bf3Item *items;
  int count;
  int total_tiles;
};
template <int BM, int BN, int WM, int WN, int BK_, int STAGES>
__device__ __forceinline__ void bf16_group_body(const Bf3Group &g) {
  const int G = gridDim.x;
  const int bid = blockIdx.x;
  const int slot = uni((G % 8) == 0 ? (bid % 8) * (G / 8) + bid / 8 : bid);
  const int total = uni(g.total_tiles);
  const int count = uni(g.count);
  const Bf3Item *items = uni(g.items);
  const LevelArgs none{};  // the BF16 epilogue reads ops only
  for (int round = 0;; ++round) {
    const int t = round * G + slot;
    if (t >= total) break;
    int lo = 0, hi = count - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (uni(items[mid].tile_start) <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    const Bf3Item *it = items + lo;
    const int kp2 = uni(it->kp2);
    LevelOps o;
    o.A8 = (const int8_t *)uni(it->A3);
    o.B8 = (const int8_t *)uni(it->B3);
    o.lda8 = 3L * kp2;
    o.ldb8 = 3L * kp2;
    o.m = uni(it->m);
    o.n = uni(it->n);
    o.nslab = 6 * kp2 / BK_;
    o.sg1 = 3 * kp2;
    o.sg2 = 5 * kp2;
    o.sb1 = kp2;
    o.sb2 = 2 * kp2;
    o.C32 = uni(it->C);
    o.ldc32 = uni(it->ldc);
    const int lt = t - uni(it->tile_start);
    const int tiles_m = uni(it->tiles_m), tiles_n = uni(it->tiles_n);
    const int gm = 8;
    const int grp = lt / (gm * tiles_n);
    const int first_m = grp * gm;
    const int gsz = min(tiles_m - first_m, gm);
    const int in_grp = lt - grp * gm * tiles_n;
    const int row0 = uni((first_m + in_grp % gsz) * BM);
    const int col0 = uni((in_grp / gsz) * BN);
    level_tile<BM, BN, WM, WN, BK_, STAGES, 16, true, false, kBfMI>(none, o, row0, col0);
  }
}
__global__ __launch_bounds__(256) void bf16_group_128x128_w4(Bf3Group g) { bf16_group_body<128, 128, 2, 2, 64, 4>(g); }
__global__ __launch_bounds__(512) void bf16_group_256x128_w8_k128(Bf3Group g) {
  bf16_group_body<256, 128, 4, 2, 128, 2>(g);
}
__global__ __launch_bounds__(1024) void bf16_group_256x256_w16_k128(Bf3Group g) {
  bf16_group_body<256, 256, 4, 4, 128, 2>(g);
}

struct Bf3GroupConfig {
  int bm, bn, bk, threads;
  size_t lds;
  void (*fn)(Bf3Group);
};
#define BF3CFG(BM_, BN_, WM_, WN_, BK_, ST_, FN_) \
  Bf3GroupConfig { BM_, BN_, BK_, WM_ *WN_ * 64, (size_t)(ST_)*I8Cfg<BM_, BN_, WM_, WN_, BK_, ST_>::STAGE_BYTES, FN_ }
const Bf3GroupConfig kBf3Group[] = {
    BF3CFG(128, 128, 2, 2, 64, 4, bf16_group_128x128_w4),         // 0: 4 waves, 64x64 wave tiles
    BF3CFG(256, 128, 4, 2, 128, 2, bf16_group_256x128_w8_k128),   // 1: 8 waves, 64x64 wave tiles
    BF3CFG(256, 256, 4, 4, 128, 2, bf16_group_256x256_w16_k128),  // 2: 16 waves, 64x64 wave tiles
};
constexpr int kNumBf3Group = sizeof(kBf3Group) / sizeof(kBf3Group[0]);
int g_occ_bf3[kNumBf3Group];
bool g_bf3_ready = false;

int bf3_occupancy(int g) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_bf3_ready) {
    for (int c = 0; c < kNumBf3Group; ++c) {
      const void *fn = reinterpret_cast<const void *>(kBf3Group[c].fn);
      (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kBf3Group[c].lds);
      int n = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kBf3Group[c].threads, kBf3Group[c].lds) != hipSuccess ||
          n < 1)
        n = 1;
      g_occ_bf3[c] = n;
    }
    g_bf3_ready = true;
  }
  return g_occ_bf3[g];
}

// Error estimate used to pick the slice count: the truncated terms of one product are
// below (s+1) 2^{-7s} 2^{e_i+f_j}; they add up like a random walk over k, while
// (|A||B|)_ij ~ k * mean|a_i.| * mean|b_.j| = k 2^{e_i+f_j} / (ratio_i ratio_j).
double split_error_estimate(int s, double ratio_a, double ratio_b, int k, int bits = 7) {
  return (s + 1) * std::ldexp(1.0, -bits * s) * ratio_a * ratio_b / std::sqrt((double)std::max(k, 1));
}

// ---------------------------------------------------------------------------
// CRT scheme (Ozaki scheme II): integer operands, int8 residues, Garner reconstruction
// ---------------------------------------------------------------------------
// A' = rint(A[i,l] 2^(p - e_i)) and B' = rint(B[l,j] 2^(p - f_j)) are integers of at most
// p+1 bits; C' = A' B' is computed exactly from its residues modulo L pairwise coprime
// moduli m_t <= 256 (residues fit int8), one int8 GEMM per modulus, and
// C += 2^(e_i + f_j - 2p) C'. The moduli product M must exceed 2 max|C'|: with the
// caller's modulus count, sum log2 m_t >= 2p + ceil(log2 k) + 2 (any inputs); with the
// automatic count, the bound from the row and column ratios (crt_needed_bits).
constexpr int kMaxMods = 20;
constexpr int kModuli[kMaxMods] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                   223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

// Garner's recurrence for digit t: v_t = r_t inv_t + sum_{s<t} v_s q[s][t] (mod m_t),
// inv_t = (m_1 ... m_{t-1})^-1 and q[s][t] = -(m_1 ... m_{s-1}) inv_t mod m_t. Both are
// stored in the symmetric range, which fits int8 for every modulus <= 256, and q is
// packed four to a word by s (qp[t][w] byte b = q[4w+b][t], zero for 4w+b >= t) so the
// sum runs as int8 dot products against the packed digits.
struct CrtConsts {
  int L;
  int m[kMaxMods];
  int invs[kMaxMods];   // inv_t, symmetric
  int qp[kMaxMods][5];  // packed symmetric q[s][t]
  float rmf[kMaxMods];  // 1 / m, rounded to FP32
  uint32_t c8[kMaxMods][3];  // byte i of word w: 2^(8 (4w + i)) mod m, for the 9 bytes of |x| < 2^72
  // |x| <= 2^46: the six bytes of x's 48-bit two's complement as c8's first six, then
  // byte 2 of word 1: (-2^48) mod m, the weight of the sign
  uint32_t c8n[kMaxMods][2];
};
static_assert(sizeof(CrtConsts) <= 2560, "CRT constants travel as a kernel argument");

// Lowest digit index S whose Horner prefix (digits S..L-1, |value| <= prod m_S..m_{L-1} / 2)
// stays below 2^53: those steps are exact in plain FP64.
constexpr int crt_plain_start(int L) {
  double prod = kModuli[L - 1];
  int s = L - 1;
  while (s > 0 && prod * kModuli[s - 1] <= 9007199254740992.0) {
    prod *= kModuli[s - 1];
    --s;
  }
  return s;
}

// Lowest digit index F whose Horner value (digits F..L-1, |value| <= prod m_F..m_{L-1} / 2)
// stays below 2^104: double-double steps there are exact integer arithmetic.
constexpr int crt_exact_floor(int L) {
  double prod = 1.0;
  int f = L;
  while (f > 0 && prod * kModuli[f - 1] <= 0x1p105) {
    prod *= kModuli[f - 1];
    --f;
  }
  return f;
}

// Symmetric residues of one integer-valued double x (|x| < 2^72) for the first L
// moduli. |x| is cut into nine bytes (three words); for each odd modulus the sum
// sum_i byte_i (2^(8 i) mod m) <= 9 * 255^2 < 2^20 is two int8 dot products, exact, and
// its symmetric residue is sum - m rint(sum / m): |sum / m| < 3400 keeps the FP32
// quotient's rounding (< 2^-11) below the 1 / (2m) distance of sum / m from a
// half-integer, so rint is exact. m_0 = 256 takes the low byte.
template <int L>
__device__ __forceinline__ void residues_of(double x, const CrtConsts &cc, int (&r)[L]) {
  const double ax = fabs(x);
  const double h = trunc(ldexp(ax, -48));              // < 2^24
  const uint64_t lo = (uint64_t)(ax - ldexp(h, 48));   // exact, < 2^48
  const uint32_t hi = (uint32_t)h;
  const uint32_t w0 = (uint32_t)lo;                          // bytes 0-3
  const uint32_t w1 = (uint32_t)(lo >> 32) | (hi << 16);     // bytes 4-5 of lo, 0-1 of hi
  const uint32_t w2 = hi >> 16;                              // byte 2 of hi
  const bool neg = x < 0.0;
  {
    int v = (int)(int8_t)(w0 & 0xffu);
    if (neg) v = -v;
    if (v == 128) v = -128;  // keep [-128, 127]
    r[0] = v;
  }
#pragma unroll
  for (int t = 1; t < L; ++t) {
    uint32_t sum = __builtin_amdgcn_udot4(w0, cc.c8[t][0], 0u, false);
    sum = __builtin_amdgcn_udot4(w1, cc.c8[t][1], sum, false);
    sum = __builtin_amdgcn_udot4(w2, cc.c8[t][2], sum, false);
    const float sf = (float)sum;
    const float q = rintf(sf * cc.rmf[t]);
    const int v = (int)fmaf(-q, (float)cc.m[t], sf);
    r[t] = neg ? -v : v;
  }
}

// The same residues for |x| <= 2^46 (p <= kNarrowBits), from x's 48-bit two's complement
// u = x mod 2^48 and its sign bit s: x = u - s 2^48, so x = sum_i byte_i(u) 2^(8 i) +
// s (-2^48) (mod m), two int8 dot products with s riding in word 1's byte 2 (sum < 2^19, the
// same rint argument), and no sign fix-up. u's low byte is x's residue mod 256. hi =
// floor(x 2^-24) and lo = x - hi 2^24 are exact (|hi| <= 2^22, 0 <= lo < 2^24); bits 8..24
// of hi are u's bytes 4-5 and, in bit 16, the sign.
constexpr int kNarrowBits = 46;
template <int L>
__device__ __forceinline__ void residues_narrow(double x, const CrtConsts &cc, int (&r)[L]) {
  const double hd = floor(x * 0x1p-24);
  const int hi = (int)hd;
  const uint32_t lo = (uint32_t)(x - hd * 0x1p24);
  const uint32_t w0 = lo | ((uint32_t)hi << 24);
  const uint32_t w1 = __builtin_amdgcn_ubfe((uint32_t)hi, 8, 17);
  r[0] = (int)(int8_t)(w0 & 0xffu);
#pragma unroll
  for (int t = 1; t < L; ++t) {
    uint32_t sum = __builtin_amdgcn_udot4(w0, cc.c8n[t][0], 0u, false);
    sum = __builtin_amdgcn_udot4(w1, cc.c8n[t][1], sum, false);
    const float sf = (float)sum;
    const float q = rintf(sf * cc.rmf[t]);
    r[t] = (int)fmaf(-q, (float)cc.m[t], sf);
  }
}

// residues of x for the first L moduli, by the narrow form where the caller's p allows it
template <int L, bool NARROW>
__device__ __forceinline__ void residues_any(double x, const CrtConsts &cc, int (&r)[L]) {
  if constexpr (NARROW)
    residues_narrow<L>(x, cc, r);
  else
    residues_of<L>(x, cc, r);
}

__device__ __forceinline__ uint32_t pack4(int a, int b, int c, int d) {
  return (uint32_t)(a & 0xff) | ((uint32_t)(b & 0xff) << 8) | ((uint32_t)(c & 0xff) << 16) | ((uint32_t)(d & 0xff) << 24);
}

// residue_cols tile: k rows (a multiple of 4 dividing 128, the plane padding) by columns
constexpr int kResColsTK = 128, kResColsTJ = 16;
static_assert(128 % kResColsTK == 0 && kResColsTK == 4 * 32 && 256 % kResColsTJ == 0, "residue_cols tile");

// A8 row i: residue plane t at bytes [t*kp, t*kp + k), zero to kp. Each thread handles
// four consecutive k and stores one 32-bit word per plane; the next four k are loaded
// before these are cut.
template <int L, bool NARROW>
__global__ __launch_bounds__(256) void residue_rows_kernel(const double *__restrict__ A, long lda, int k, int p,
                                                           long kp, const int *__restrict__ eA, CrtConsts cc,
                                                           int8_t *__restrict__ A8, long ld8) {
  const int i = blockIdx.x;
  const double *row = A + (long)i * lda;
  const int sh = p - eA[i];
  int8_t *out = A8 + (long)i * ld8;
  double xs[4];
  long l0 = (long)threadIdx.x * 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) xs[q] = load_or_zero(row, l0 + q, l0 + q < k);
  for (; l0 < kp; l0 += 1024) {
    double nx[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) nx[q] = load_or_zero(row, l0 + 1024 + q, l0 + 1024 + q < k);
    int r[4][L];
#pragma unroll
    for (int q = 0; q < 4; ++q) residues_any<L, NARROW>(rint(ldexp(xs[q], sh)), cc, r[q]);  // 0 beyond k
#pragma unroll
    for (int q = 0; q < 4; ++q) xs[q] = nx[q];
#pragma unroll
    for (int t = 0; t < L; ++t)
      *reinterpret_cast<uint32_t *>(out + (long)t * kp + l0) = pack4(r[0][t], r[1][t], r[2][t], r[3][t]);
  }
}

// B8 row j (column j of B): residue plane t at bytes [t*kp, t*kp + k). 128 (k) x 16 (j)
// tiles pass through LDS (17 KiB: registers, not LDS, bound the occupancy); each thread
// then takes four consecutive k of two columns, so 32 lanes write one column's 128 bytes
// of a plane.
template <int L, bool NARROW>
__global__ __launch_bounds__(256) void residue_cols_kernel(const double *__restrict__ B, long ldb, int k, int n, int p,
                                                           long kp, const int *__restrict__ eB, CrtConsts cc,
                                                           int8_t *__restrict__ B8, long ld8) {
  constexpr int TK = kResColsTK, TJ = kResColsTJ;
  __shared__ double tile[TK][TJ + 1];
  const int k0 = blockIdx.x * TK;
  const int j0 = blockIdx.y * TJ;
  const int tx = threadIdx.x % TJ, ty = threadIdx.x / TJ;
  constexpr int RS = 256 / TJ, NL = TK / RS;  // row step, loads per thread
  double v[NL];
#pragma unroll
  for (int u = 0; u < NL; ++u) {
    const int kk = k0 + ty + u * RS, jj = j0 + tx;
    v[u] = load_or_zero(B, (long)kk * ldb + jj, kk < k && jj < n);
  }
#pragma unroll
  for (int u = 0; u < NL; ++u) tile[ty + u * RS][tx] = v[u];
  __syncthreads();
  const int kq = (threadIdx.x & 31) * 4;  // four consecutive k
  for (int c = threadIdx.x >> 5; c < TJ; c += 8) {
    const int jj = j0 + c;
    if (jj >= n) continue;
    const int sh = p - eB[jj];
    int r[4][L];
#pragma unroll
    for (int q = 0; q < 4; ++q) residues_any<L, NARROW>(rint(ldexp(tile[kq + q][c], sh)), cc, r[q]);
    int8_t *out = B8 + (long)jj * ld8 + k0 + kq;
#pragma unroll
    for (int t = 0; t < L; ++t)
      *reinterpret_cast<uint32_t *>(out + (long)t * kp) = pack4(r[0][t], r[1][t], r[2][t], r[3][t]);
  }
}

#pragma clang fp contract(off)
// C[i,j] += 2^(e_i + f_j - 2p) * C'[i,j], C' from its residues (planes of m x n bytes):
// Garner's mixed-radix digits v_t in symmetric ranges (exact small-integer arithmetic),
// then C' = v_1 + m_1 (v_2 + m_2 (v_3 + ...)) evaluated in double-double.
template <int L>
__device__ __forceinline__ double crt_value(const int (&r)[L], const CrtConsts &cc) {
  int v[L];
  uint32_t vp[(L + 3) / 4] = {};  // digits packed four to a word
  // m_0 = 256 and inv_0 = 1: the first digit is the first residue
  v[0] = r[0];
  vp[0] = (uint32_t)(r[0] & 0xff);
#pragma unroll
  for (int t = 1; t < L; ++t) {
    // |r_t inv_t| and each |v_s q[s][t]| are at most 2^14, so |y| < 2^18; m_t is odd, so
    // y / m_t is at least 1 / (2 m_t) from a half-integer, far beyond the FP32 rounding of
    // y * (1 / m_t) (< 2^-12): rint gives the exact nearest quotient and the symmetric
    // residue y - q m_t is exact.
    int y = __mul24(r[t], cc.invs[t]);  // |r|, |inv| <= 128: a full-rate 24-bit multiply
#pragma unroll
    for (int w = 0; w < (t + 3) / 4; ++w) y = __builtin_amdgcn_sdot4((int)vp[w], cc.qp[t][w], y, false);
    const float yf = (float)y;
    const float q = rintf(yf * cc.rmf[t]);
    v[t] = (int)fmaf(-q, (float)cc.m[t], yf);
    vp[t / 4] |= (uint32_t)(v[t] & 0xff) << (8 * (t % 4));
  }
  // Horner from the top digit: plain FP64 while the prefix stays below 2^53, then
  // double-double. While the value stays below 2^104 (digits F..L-1) every step is exact
  // integer arithmetic, so two digits go in a step there (multiplier m_t m_{t-1} < 2^16,
  // digit v_t m_{t-1} + v_{t-1}, |.| < 2^15) and leave the same (hi, lo) as two steps of
  // one digit; the digits below F, where only moduli counts above 13 reach, go one at a
  // time, so the result is the same bits as a one-digit Horner in every case.
  constexpr int S = crt_plain_start(L);
  constexpr int F = crt_exact_floor(L);
  double hi = (double)v[L - 1], lo = 0.0;
  int t = L - 2;
#pragma unroll
  for (; t - 1 >= S; t -= 2)
    hi = fma(hi, (double)(cc.m[t] * cc.m[t - 1]), (double)(__mul24(v[t], cc.m[t - 1]) + v[t - 1]));
  if (t >= S) {
    hi = fma(hi, (double)cc.m[t], (double)v[t]);
    --t;
  }
  auto dd_step = [&](double mt, double vt) {
    // (hi, lo) = (hi, lo) * mt + vt
    const double ph = hi * mt;
    const double pe = fma(hi, mt, -ph);
    const double pl = lo * mt + pe;
    const double sh = ph + vt;
    const double bb = sh - ph;
    const double se = (ph - (sh - bb)) + (vt - bb);
    const double tl = pl + se;
    hi = sh + tl;
    lo = tl - (hi - sh);
  };
#pragma unroll
  for (; t - 1 >= F; t -= 2) dd_step((double)(cc.m[t] * cc.m[t - 1]), (double)(__mul24(v[t], cc.m[t - 1]) + v[t - 1]));
#pragma unroll
  for (; t >= 0; --t) dd_step((double)cc.m[t], (double)v[t]);
  return hi + lo;
}

template <int L>
__global__ __launch_bounds__(256) void crt_reconstruct_kernel(const int8_t *__restrict__ res, long plane, int m, int n,
                                                              CrtConsts cc, int p, const int *__restrict__ eA,
                                                              const int *__restrict__ eB, double *__restrict__ C,
                                                              long ldc) {
  // four consecutive outputs of one row per thread (n is a multiple of 4 on this path)
  const long total4 = (long)m * n / 4;
  for (long e4 = (long)blockIdx.x * blockDim.x + threadIdx.x; e4 < total4; e4 += (long)gridDim.x * blockDim.x) {
    const long e = e4 * 4;
    uint32_t w[L];
#pragma unroll
    for (int t = 0; t < L; ++t) w[t] = *reinterpret_cast<const uint32_t *>(res + (long)t * plane + e);
    const long i = e / n, j = e % n;
    double *pc = C + i * ldc + j;
    const int ei = eA[i] - 2 * p;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int r[L];
#pragma unroll
      for (int t = 0; t < L; ++t) r[t] = (int)(int8_t)(w[t] >> (8 * q));
      const double cp = crt_value<L>(r, cc);
      pc[q] = pc[q] + ldexp(cp, ei + eB[j + q]);
    }
  }
}

// general shapes: one output per thread
template <int L>
__global__ __launch_bounds__(256) void crt_reconstruct1_kernel(const int8_t *__restrict__ res, long plane, int m, int n,
                                                               CrtConsts cc, int p, const int *__restrict__ eA,
                                                               const int *__restrict__ eB, double *__restrict__ C,
                                                               long ldc) {
  const long total = (long)m * n;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    int r[L];
#pragma unroll
    for (int t = 0; t < L; ++t) r[t] = res[(long)t * plane + e];
    const double cp = crt_value<L>(r, cc);
    const long i = e / n, j = e % n;
    double *pc = C + i * ldc + j;
    *pc = *pc + ldexp(cp, eA[i] + eB[j] - 2 * p);
  }
}
#pragma clang fp contract(on)

// launches the kernels above for a runtime modulus count
struct CrtLaunch {
  template <int L>
  static int run(hipStream_t s, const double *dA, long lda, const double *dB, long ldb, int m, int k, int n, int p,
                 long kp, const int *eA, const int *eB, const CrtConsts &cc, int8_t *A8, int8_t *B8, long ld8,
                 int stage, const int8_t *R8, double *dC, long ldc) {
    if (stage == 0) {
      // |A'|, |B'| <= 2^p: the narrow residue form takes p <= kNarrowBits
      const bool narrow = p <= kNarrowBits;
      auto rows_fn = narrow ? residue_rows_kernel<L, true> : residue_rows_kernel<L, false>;
      auto cols_fn = narrow ? residue_cols_kernel<L, true> : residue_cols_kernel<L, false>;
      hipLaunchKernelGGL(rows_fn, dim3(m), dim3(256), 0, s, dA, lda, k, p, kp, eA, cc, A8, ld8);
      PHPC_HIP(hipGetLastError());
      hipLaunchKernelGGL(cols_fn,
                         dim3((unsigned)((kp + kResColsTK - 1) / kResColsTK),
                              (unsigned)((n + kResColsTJ - 1) / kResColsTJ)),
                         dim3(256), 0, s, dB, ldb, k, n, p, kp, eB, cc, B8, ld8);
      PHPC_HIP(hipGetLastError());
    } else {
      const long total = (long)m * n;
      if (n % 4 == 0) {
        long blocks = std::min<long>((total / 4 + 255) / 256, 65536);
        hipLaunchKernelGGL(crt_reconstruct_kernel<L>, dim3((unsigned)blocks), dim3(256), 0, s, R8, total, m, n, cc, p,
                           eA, eB, dC, ldc);
      } else {
        long blocks = std::min<long>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(crt_reconstruct1_kernel<L>, dim3((unsigned)blocks), dim3(256), 0, s, R8, total, m, n, cc,
                           p, eA, eB, dC, ldc);
      }
      PHPC_HIP(hipGetLastError());
    }
    return PHPC_OK;
  }
};

template <int L>
int crt_dispatch_t(int Lrt, hipStream_t s, const double *dA, long lda, const double *dB, long ldb, int m, int k, int n,
                   int p, long kp, const int *eA, const int *eB, const CrtConsts &cc, int8_t *A8, int8_t *B8, long ld8,
                   int stage, const int8_t *R8, double *dC, long ldc) {
  if (Lrt == L) return CrtLaunch::run<L>(s, dA, lda, dB, ldb, m, k, n, p, kp, eA, eB, cc, A8, B8, ld8, stage, R8, dC, ldc);
  if constexpr (L < 20) {
    return crt_dispatch_t<L + 1>(Lrt, s, dA, lda, dB, ldb, m, k, n, p, kp, eA, eB, cc, A8, B8, ld8, stage, R8, dC, ldc);
  } else {
    return phpc_set_error(PHPC_ERR_ARG, "modulus count %d out of range", Lrt);
  }
}

int modinv(int a, int m) {
  // extended Euclid, a and m coprime
  long t = 0, nt = 1, r = m, nr = ((a % m) + m) % m;
  while (nr) {
    const long q = r / nr;
    long tmp = t - q * nt;
    t = nt;
    nt = tmp;
    tmp = r - q * nr;
    r = nr;
    nr = tmp;
  }
  return (int)((t % m + m) % m);
}

CrtConsts make_crt(int L) {
  CrtConsts c{};
  c.L = L;
  for (int t = 0; t < L; ++t) {
    const int m = kModuli[t];
    c.m[t] = m;
    long pmod[kMaxMods];  // (m_1..m_{s-1}) mod m_t with s = s2 + 1
    long prod = 1;
    for (int s2 = 0; s2 < t; ++s2) {
      pmod[s2] = prod;
      prod = prod * kModuli[s2] % m;
    }
    const int inv = t == 0 ? 1 : modinv((int)prod, m);
    auto sym = [m](long x) { return (int)(x > (m - 1) / 2 ? x - m : x); };  // [0, m) -> symmetric
    c.invs[t] = sym(inv);
    for (int s2 = 0; s2 < t; ++s2) {
      const int q = sym((m - pmod[s2] * inv % m) % m);
      c.qp[t][s2 / 4] |= (int)((uint32_t)(q & 0xff) << (8 * (s2 % 4)));
    }
    c.rmf[t] = 1.0f / (float)m;
    long w = 1;  // 2^(8 i) mod m
    for (int i = 0; i < 9; ++i) {
      c.c8[t][i / 4] |= (uint32_t)w << (8 * (i % 4));
      if (i < 6) c.c8n[t][i / 4] |= (uint32_t)w << (8 * (i % 4));
      if (i == 6) c.c8n[t][1] |= (uint32_t)((m - w) % m) << 16;  // w = 2^48 mod m
      w = w * 256 % m;
    }
  }
  return c;
}

// Bytes of the statistics block (slot 3): exponents (m + n ints), ratios (m + n
// doubles, 16-byte aligned), column maxima and sums (2 n words), the non-finite flag and
// the ratio summary.
size_t split_meta_bytes(int m, int n) {
  return (size_t)(m + n) * (sizeof(int) + sizeof(double)) + (size_t)n * 20 + 64 + sizeof(RatioSummary) + 16;
}

double crt_bits(int L) {
  double b = 0;
  for (int t = 0; t < L; ++t) b += std::log2((double)kModuli[t]);
  return b;
}

// Bits the moduli product needs for C' = A' B' at integer width p. |A'_il| <= |a_il|
// 2^(p - e_i) + 1/2, so sum_l |A'_il| <= 2^p k / ratio_i + k / 2, and with |B'_lj| <= 2^p,
// |C'_ij| <= 2^p k (2^p / ratio_i + 1/2); the same holds with the column ratio of B. R is
// the larger of min_i ratio_i and min_j ratio_j over nonzero rows and columns (zero ones
// give C' = 0), R >= 1 since 2^e_i exceeds every |a_il|; it is shrunk by 1e-6 to cover the
// rounding of the FP64 row sums (k u < 2.4e-7 for k < 2^31). M > 2 max|C'| then holds with
// one bit to spare when sum log2 m_t >= the returned value. For R = 1 this is the
// worst-case 2p + log2 k + 2.
double crt_needed_bits(int p, double R, int k) {
  const double r = std::max(R * (1.0 - 1e-6), 1.0);
  return p + std::log2((double)std::max(k, 1)) + std::log2(std::ldexp(1.0, p) / r + 0.5) + 2.0;
}


// a-priori estimate for integer operands of p bits: each entry moves by at most
// 2^(e - p - 1); the errors add like a random walk over k. The factor 2 covers the tail
// of the maximum over all m*n outputs (the random-walk figure matches the measured
// maximum of a few thousand sampled outputs).
double crt_error_estimate(int p, double ratio_a, double ratio_b, int k) {
  return 2.0 * std::ldexp(1.0, -p) * (ratio_a + ratio_b) / std::sqrt((double)std::max(k, 1));
}

int g_method = PHPC_SPLIT_CRT;

}  // namespace

extern "C" int phpc_split_set_config(int cfg) {
  if (cfg < 0 || cfg >= kNumI8) return phpc_set_error(PHPC_ERR_ARG, "split config %d out of range", cfg);
  g_i8_big = cfg;
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_split_num_configs(void) { return kNumI8; }

extern "C" int phpc_split_set_method(int method) {
  if (method != PHPC_SPLIT_SLICES && method != PHPC_SPLIT_CRT && method != PHPC_SPLIT_FP8)
    return phpc_set_error(PHPC_ERR_ARG, "unknown split method %d", method);
  g_method = method;
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" double phpc_split_error_estimate(int slices, double ratio_a, double ratio_b, int k) {
  return split_error_estimate(slices, ratio_a, ratio_b, k);
}

namespace {
SharedWorkspace g_default_ws;
}  // namespace

extern "C" int phpc_dgemm_split_device(const double *dA, long lda, const double *dB, long ldb, double *dC, long ldc,
                                       int m, int k, int n, int slices, double target_err, void *stream_,
                                       phpc_split_info *info) {
  // the process-wide workspace serves one call at a time, each on its stream after the
  // previous call's use
  std::lock_guard<std::mutex> lk(g_default_ws.mu);
  hipStream_t s = (hipStream_t)stream_;
  g_default_ws.after_previous(s);
  const int rc =
      phpc_dgemm_split_ws(dA, lda, dB, ldb, dC, ldc, m, k, n, slices, target_err, g_method, s, &g_default_ws.ws, info);
  g_default_ws.mark_used(s);
  return rc;
}

int phpc_dgemm_split_ws(const double *dA, long lda, const double *dB, long ldb, double *dC, long ldc, int m, int k,
                        int n, int slices, double target_err, int method, hipStream_t stream_, Workspace *ws,
                        phpc_split_info *info) {
  const bool use_crt = method == PHPC_SPLIT_CRT;
  const bool use_f8 = method == PHPC_SPLIT_FP8;
  if (method != PHPC_SPLIT_SLICES && !use_crt && !use_f8)
    return phpc_set_error(PHPC_ERR_ARG, "unknown split method %d", method);
  const int bits = use_f8 ? 4 : 7;                          // digit width of the slices
  const int max_slices = use_f8 ? kMaxSlicesF8 : kMaxSlices;
  if (info) *info = phpc_split_info{};
  if (phpc_init()) return phpc_last_error();
  if (m < 0 || k < 0 || n < 0) return phpc_set_error(PHPC_ERR_ARG, "negative size");
  if ((m > 0 && k > 0 && (!dA || lda < k)) || (k > 0 && n > 0 && (!dB || ldb < n)) || (m > 0 && n > 0 && (!dC || ldc < n)))
    return phpc_set_error(PHPC_ERR_ARG, "bad pointers or leading dimensions");
  if (!use_crt && (slices < 0 || slices > max_slices))
    return phpc_set_error(PHPC_ERR_ARG, "slices must be 0 (auto) or 1..%d", max_slices);
  if (use_crt && (slices < 0 || slices == 1 || slices > kMaxMods))
    return phpc_set_error(PHPC_ERR_ARG, "moduli must be 0 (auto) or 2..%d", kMaxMods);
  if (m == 0 || n == 0 || k == 0) return phpc_set_error(PHPC_OK, nullptr);
  hipStream_t s = (hipStream_t)stream_;
  const long kp = phpc_split_kp(k);

  // scratch: exponents, ratios, column stats (slices and the FP64 accumulator below)
  const size_t small = split_meta_bytes(m, n);
  char *meta = (char *)ws->get(small, 3, s);
  if (!meta) return phpc_set_error(PHPC_ERR_MEM, "split scratch allocation failed");
  int *eA = (int *)meta;
  int *eB = eA + m;
  double *ratioA = (double *)(((uintptr_t)(eB + n) + 15) & ~(uintptr_t)15);
  double *ratioB = ratioA + m;
  unsigned long long *cmax = (unsigned long long *)(ratioB + n);
  double *csum = (double *)(cmax + n);
  int *bad = (int *)(csum + n);  // [0] non-finite flag, [1] width of A's rows, [2] of B's columns
  int *ccode = bad + 4;          // per column of B: largest low-bit code
  RatioSummary *summary = (RatioSummary *)(((uintptr_t)(ccode + n) + 15) & ~(uintptr_t)15);

  {
    const long words = ((long)n * 20 + 16) / 4;  // cmax, csum, the flag and widths, ccode
    hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)std::min<long>((words + 255) / 256, 4096)), dim3(256), 0, s,
                       (uint32_t *)cmax, words);
    PHPC_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(row_stats_kernel, dim3(m), dim3(256), 0, s, dA, lda, k, eA, ratioA, bad, bad + 1);
  PHPC_HIP(hipGetLastError());
  const int rows_per_block = 256;
  hipLaunchKernelGGL(col_stats_kernel, dim3((n + 255) / 256, (k + rows_per_block - 1) / rows_per_block), dim3(256), 0,
                     s, dB, ldb, k, n, rows_per_block, cmax, csum, ccode, bad);
  PHPC_HIP(hipGetLastError());
  hipLaunchKernelGGL(col_exps_kernel, dim3((n + 255) / 256), dim3(256), 0, s, cmax, csum, ccode, k, n, eB, ratioB,
                     bad + 2);
  PHPC_HIP(hipGetLastError());

  if (use_crt) {
    // CRT scheme: choose the integer width p, then the modulus count
    RatioSummary hs{};
    hipLaunchKernelGGL(ratio_summary_kernel, dim3(1), dim3(256), 0, s, ratioA, m, ratioB, n, bad, summary);
    PHPC_HIP(hipGetLastError());
    PHPC_HIP(hipMemcpyAsync(&hs, summary, sizeof(hs), hipMemcpyDeviceToHost, s));
    PHPC_HIP(hipStreamSynchronize(s));
    if (hs.bad) return phpc_set_error(PHPC_ERR_ARG, "split-precision GEMM needs finite inputs");
    const double mra = hs.max_a, mrb = hs.max_b;
    const int logk = (int)std::ceil(std::log2((double)std::max(k, 2)));
    int pbits = 0, L = 0;
    bool exact_scaling = false;
    if (slices > 0) {
      // the caller fixes the modulus count; use the widest integers it can hold, but
      // no wider than residues_of takes: |A'| <= 2^p (|a| < 2^e), and p <= kMaxIntBits
      // keeps every integer below 2^72, its nine-byte range
      L = std::min(slices, kMaxMods);
      pbits = std::min(kMaxIntBits, (int)std::floor((crt_bits(L) - logk - 2) / 2));
    } else {
      const double target = target_err > 0 ? target_err : 1e-14;
      // the smallest positive ratio of each side, 1 when there is none
      const double R = std::max(std::isinf(hs.minpos_a) ? 1.0 : hs.minpos_a,
                                std::isinf(hs.minpos_b) ? 1.0 : hs.minpos_b);
      for (pbits = 8; pbits < 70; ++pbits)
        if (crt_error_estimate(pbits, mra, mrb, k) <= target) break;
      // entries that all fit a narrower width (integers, short mantissas, powers of two)
      // scale to integers exactly at that width: no truncation error at all, and fewer
      // bits for the moduli to cover
      const int wexact = std::max(hs.width_a, hs.width_b);
      if (wexact >= 1 && wexact < pbits) {
        pbits = wexact;
        exact_scaling = true;
      }
      for (L = 2; L < kMaxMods; ++L)
        if (crt_bits(L) >= crt_needed_bits(pbits, R, k)) break;
      while (pbits > 0 && crt_bits(L) < crt_needed_bits(pbits, R, k)) --pbits;
      if (pbits < wexact) exact_scaling = false;
    }
    if (pbits < 1) return phpc_set_error(PHPC_ERR_ARG, "k = %d too large for the CRT scheme", k);
    const CrtConsts cc = make_crt(L);
    const long ld8 = kp * L;
    int8_t *A8 = (int8_t *)ws->get((size_t)m * ld8, 0, s);
    int8_t *B8 = (int8_t *)ws->get((size_t)n * ld8, 1, s);
    int8_t *R8 = (int8_t *)ws->get((size_t)L * m * n, 4, s);
    if (!A8 || !B8 || !R8) return phpc_set_error(PHPC_ERR_MEM, "CRT scratch allocation failed");
    {
      int rc = crt_dispatch_t<2>(L, s, dA, lda, dB, ldb, m, k, n, pbits, kp, eA, eB, cc, A8, B8, ld8, 0, nullptr,
                                 nullptr, 0);
      if (rc) return rc;
    }
    long tiles256 = (long)((m + 255) / 256) * ((n + 255) / 256);
    const int cfg = tiles256 >= (long)phpc_device().cus / 2 ? g_i8_big : 1;
    const long piece = std::min<long>(kp, kMaxExactTermsCrt);
    int launches = 0;
    for (int t = 0; t < L; ++t) {
      for (long off = 0; off < kp; off += piece) {
        LevelArgs a{};
        a.A8 = A8;
        a.B8 = B8;
        a.lda8 = ld8;
        a.ldb8 = ld8;
        a.a0 = (long)t * kp + off;
        a.b0 = (long)t * kp + off;
        a.m = m;
        a.n = n;
        a.kl = (int)std::min(piece, kp - off);
        a.first = off == 0 ? 1 : 0;
        a.mod_m = cc.m[t];
        a.res = R8 + (size_t)t * m * n;
        a.ldres = n;
        int rc = launch_level(a, cfg, s);
        if (rc) return rc;
        ++launches;
      }
    }
    {
      int rc = crt_dispatch_t<2>(L, s, dA, lda, dB, ldb, m, k, n, pbits, kp, eA, eB, cc, A8, B8, ld8, 1, R8, dC, ldc);
      if (rc) return rc;
    }
    if (info) {
      info->slices = L;
      info->products = L;
      info->launches = launches + 7;
      // exact scaling leaves the two roundings of C += 2^(e_i + f_j - 2p) C'
      info->error_estimate = exact_scaling ? std::ldexp(1.0, -52) : crt_error_estimate(pbits, mra, mrb, k);
      info->method = 1;
      info->int_bits = pbits;
    }
    return phpc_set_error(PHPC_OK, nullptr);
  }

  int s_use = slices;
  double est = 0.0;
  {
    // exponents and ratios decide the slice count (and reveal non-finite input)
    RatioSummary hs{};
    hipLaunchKernelGGL(ratio_summary_kernel, dim3(1), dim3(256), 0, s, ratioA, m, ratioB, n, bad, summary);
    PHPC_HIP(hipGetLastError());
    PHPC_HIP(hipMemcpyAsync(&hs, summary, sizeof(hs), hipMemcpyDeviceToHost, s));
    PHPC_HIP(hipStreamSynchronize(s));
    if (hs.bad) return phpc_set_error(PHPC_ERR_ARG, "split-precision GEMM needs finite inputs");
    const double mra = hs.max_a, mrb = hs.max_b;
    if (s_use == 0) {
      const double target = target_err > 0 ? target_err : 1e-14;
      s_use = max_slices;
      for (int c = 2; c <= max_slices; ++c)
        if (split_error_estimate(c, mra, mrb, k, bits) <= target) {
          s_use = c;
          break;
        }
    }
    est = split_error_estimate(s_use, mra, mrb, k, bits);
  }
  const long ld8a = kp * s_use, ld8b = kp * s_use;
  int8_t *A8 = (int8_t *)ws->get((size_t)m * ld8a, 0, s);
  int8_t *B8 = (int8_t *)ws->get((size_t)n * ld8b, 1, s);
  double *acc = (double *)ws->get((size_t)m * n * 8, 2, s);
  if (!A8 || !B8 || !acc) return phpc_set_error(PHPC_ERR_MEM, "split scratch allocation failed");
  hipLaunchKernelGGL(use_f8 ? slice_rows_kernel<4> : slice_rows_kernel<7>, dim3(m), dim3(256), 0, s, dA, lda, k, s_use,
                     kp, eA, A8, ld8a);
  PHPC_HIP(hipGetLastError());
  // B slices in reverse order: slice u at block (s_use - u)
  {
    dim3 grid((unsigned)((kp + 63) / 64), (unsigned)((n + 63) / 64));
    hipLaunchKernelGGL(use_f8 ? slice_cols_kernel<4> : slice_cols_kernel<7>, grid, dim3(256), 0, s, dB, ldb, k, n,
                       s_use, kp, eB, B8, ld8b);
    PHPC_HIP(hipGetLastError());
  }

  // level GEMMs, smallest level first
  long tiles256 = (long)((m + 255) / 256) * ((n + 255) / 256);
  const bool big = tiles256 >= (long)phpc_device().cus / 2;
  const int cfg = use_f8 ? (big ? 0 : 1) : (big ? g_i8_big : 1);
  // whole slices per piece keep offsets aligned; else whole 128-byte slabs
  const long max_terms = use_f8 ? kMaxExactTermsF8 : kMaxExactTerms;
  const long piece = max_terms / kp * kp;
  const long max_piece = piece > 0 ? piece : (max_terms / 128) * 128;
  bool first = true;
  int launches = 0;
  for (int g = s_use + 1; g >= 2; --g) {
    const long klen = (long)(g - 1) * kp;
    for (long off = 0; off < klen; off += max_piece) {
      const long len = std::min(max_piece, klen - off);
      LevelArgs a{};
      a.A8 = A8;
      a.B8 = B8;
      a.lda8 = ld8a;
      a.ldb8 = ld8b;
      a.a0 = off;
      a.b0 = (long)(s_use - g + 1) * kp + off;
      a.m = m;
      a.n = n;
      a.kl = (int)len;
      a.scale_exp = -bits * g;
      a.acc = acc;
      a.ldacc = n;
      a.C = dC;
      a.ldc = ldc;
      a.eA = eA;
      a.eB = eB;
      a.first = first ? 1 : 0;
      a.last = (g == 2 && off + len >= klen) ? 1 : 0;
      int rc = use_f8 ? launch_f8_level(a, cfg, s) : launch_level(a, cfg, s);
      if (rc) return rc;
      first = false;
      ++launches;
    }
  }
  if (info) {
    info->slices = s_use;
    info->products = s_use * (s_use + 1) / 2;
    info->launches = launches + 6;
    info->error_estimate = est;
    info->method = method;
    info->int_bits = bits * s_use;
  }
  return phpc_set_error(PHPC_OK, nullptr);
}

// FP32-accurate GEMM from BF16 slice planes: C += A*B (FP32, row-major, device pointers).
// A3 rows hold [hi | mid | lo] and B3 rows (columns of B) [lo | mid | hi], so the six
// products with slice-index sum <= 2 form three contiguous k ranges, summed smallest first
// into one FP32 accumulator by one launch (three k segments): A3[:, 0:3kp] . B3[:, 0:3kp]
// = hi.lo + mid.mid + lo.hi, then A3[:, 0:2kp] . B3[:, kp:3kp] = hi.mid + mid.hi, then
// A3[:, 0:kp] . B3[:, 2kp:3kp] = hi.hi; the sum is added to C once.
int phpc_sgemm_bf16x3_planes(const float *dA, long lda, const float *dB, long ldb, float *dC, long ldc, int m, int k,
                             int n, hipStream_t s, Workspace *ws) {
  const long kp = ((long)k + 63) / 64 * 64;
  if (3 * kp * 2 >= 0x7fffffffL / 256) return phpc_set_error(PHPC_ERR_ARG, "k = %d too large for the slice planes", k);
  uint16_t *A3 = (uint16_t *)ws->get((size_t)m * 3 * kp * 2, 0, s);
  uint16_t *B3 = (uint16_t *)ws->get((size_t)n * 3 * kp * 2, 1, s);
  if (!A3 || !B3) return phpc_set_error(PHPC_ERR_MEM, "slice plane allocation failed");
  {
    const long per_row = (kp / 8 + 255) / 256;
    hipLaunchKernelGGL(bf16x3_rows_kernel, dim3((unsigned)std::min<long>(per_row, 64), (unsigned)m), dim3(256), 0, s,
                       dA, lda, m, k, kp, A3);
    PHPC_HIP(hipGetLastError());
    hipLaunchKernelGGL(bf16x3_cols_kernel, dim3((unsigned)(kp / 64), (unsigned)((n + 63) / 64)), dim3(256), 0, s, dB,
                       ldb, k, n, kp, B3);
    PHPC_HIP(hipGetLastError());
  }
  const long tiles256 = (long)((m + 255) / 256) * ((n + 255) / 256);
  const int cfg = tiles256 >= (long)phpc_device().cus / 2 ? 0 : 1;
  const int kp2 = (int)(kp * 2);  // bytes of one plane row
  LevelArgs a{};
  a.A8 = (const int8_t *)A3;
  a.B8 = (const int8_t *)B3;
  a.lda8 = 3L * kp2;
  a.ldb8 = 3L * kp2;
  a.m = m;
  a.n = n;
  a.kl = 6 * kp2;
  a.seg1 = 3 * kp2;
  a.seg2 = 5 * kp2;
  a.segb1 = kp2;
  a.segb2 = 2 * kp2;
  a.C32 = dC;
  a.ldc32 = ldc;
  return launch_bf_level(a, cfg, s);
}

// ---------------------------------------------------------------------------
// Grouped 3 x BF16 plans on slice planes (declared in phpc_internal.h)
// ---------------------------------------------------------------------------
int phpc_bf3_num_configs() { return kNumBf3Group; }

namespace {
size_t bf3_align(size_t b) { return (b + 255) / 256 * 256; }
int g_f32x3_plan_cfg = -1;
}  // namespace

int phpc_f32x3_set_plan_config(int id) {
  if (id != -1 && id != 2 && (id < kF32X3PlanesConfig || id >= kF32X3PlanesConfig + kNumBf3Group))
    return phpc_set_error(PHPC_ERR_ARG, "3 x BF16 plan config %d out of range", id);
  g_f32x3_plan_cfg = id;
  return phpc_set_error(PHPC_OK, nullptr);
}

int phpc_f32x3_pick_plan_config(const F32Desc *descs, int ndesc) {
  if (g_f32x3_plan_cfg >= 0) return g_f32x3_plan_cfg;
  // slice planes on 256 x 256 tiles (the fastest tile on the 512-GEMM mix) once the items
  // give every CU at least one such tile
  long work = 0;
  for (int i = 0; i < ndesc; ++i) work += (long)((descs[i].m + 255) / 256) * ((descs[i].n + 255) / 256);
  return work >= (long)phpc_device().cus ? kF32X3PlanesConfig + 2 : 2;
}

int Bf3Planes::build(const F32Desc *descs, int ndesc, int gcfg) {
  release();
  cuts.clear();
  off.clear();
  block_cut.clear();
  items.clear();
  if (gcfg < 0 || gcfg >= kNumBf3Group) return phpc_set_error(PHPC_ERR_ARG, "slice-plane tile %d out of range", gcfg);
  g = gcfg;
  const Bf3GroupConfig &cf = kBf3Group[g];
  blocks = 0;
  total_tiles = 0;
  ws_bytes = 0;
  // one cut per distinct operand (an A or B shared by several items is cut once)
  auto find_or_add = [&](const float *src, long ld, int rows, int cols, int is_b, long kp) {
    for (size_t j = 0; j < cuts.size(); ++j) {
      const Bf3Cut &c = cuts[j];
      if (c.src == src && c.ld == ld && c.rows == rows && c.cols == cols && c.is_b == is_b) return (int)j;
    }
    Bf3Cut c{};
    c.src = src;
    c.ld = ld;
    c.rows = rows;
    c.cols = cols;
    c.kp = kp;
    c.is_b = is_b;
    c.vec = !is_b && (reinterpret_cast<uintptr_t>(src) & 15) == 0 && ld % 4 == 0;
    c.tiles_k = (int)(kp / 64);
    c.block_start = blocks;
    const int prow = is_b ? cols : rows;
    const int nt = c.tiles_k * ((prow + 63) / 64);
    blocks += nt;
    block_cut.insert(block_cut.end(), nt, (int)cuts.size());
    cuts.push_back(c);
    off.push_back(ws_bytes);
    ws_bytes += bf3_align((size_t)prow * 3 * kp * 2);
    return (int)cuts.size() - 1;
  };
  std::vector<int> ca, cb;
  for (int i = 0; i < ndesc; ++i) {
    const F32Desc &d = descs[i];
    if (d.m <= 0 || d.n <= 0 || d.k <= 0) continue;
    const long kp = ((long)d.k + 63) / 64 * 64;
    // a tile's plane rows are read through 32-bit buffer offsets, and so is the k offset
    if ((long)std::max(cf.bm, cf.bn) * 3 * kp * 2 >= 0x7fffffffL)
      return phpc_set_error(PHPC_ERR_ARG, "k = %d too large for the slice planes", d.k);
    ca.push_back(find_or_add(d.A, d.lda, d.m, d.k, 0, kp));
    cb.push_back(find_or_add(d.B, d.ldb, d.k, d.n, 1, kp));
    Bf3Item it{};
    it.C = d.C;
    it.ldc = d.ldc;
    it.m = d.m;
    it.n = d.n;
    it.kp2 = (int)(kp * 2);
    it.tiles_m = (d.m + cf.bm - 1) / cf.bm;
    it.tiles_n = (d.n + cf.bn - 1) / cf.bn;
    it.tile_start = total_tiles;
    total_tiles += it.tiles_m * it.tiles_n;
    items.push_back(it);
  }
  if (items.empty()) return PHPC_OK;
  if (hipMalloc(&ws, ws_bytes) != hipSuccess) {
    ws = nullptr;
    return phpc_set_error(PHPC_ERR_MEM, "slice-plane workspace allocation failed (%zu bytes)", ws_bytes);
  }
  char *base = static_cast<char *>(ws);
  for (size_t j = 0; j < cuts.size(); ++j) cuts[j].dst = reinterpret_cast<uint16_t *>(base + off[j]);
  for (size_t i = 0; i < items.size(); ++i) {
    items[i].A3 = cuts[ca[i]].dst;
    items[i].B3 = cuts[cb[i]].dst;
  }
  const size_t cuts_b = bf3_align(cuts.size() * sizeof(Bf3Cut));
  const size_t bc_b = bf3_align(block_cut.size() * sizeof(int));
  const size_t it_b = items.size() * sizeof(Bf3Item);
  if (hipMalloc(&list, cuts_b + bc_b + it_b) != hipSuccess) {
    list = nullptr;
    release();
    return phpc_set_error(PHPC_ERR_MEM, "slice-plane list allocation failed");
  }
  char *l = static_cast<char *>(list);
  PHPC_HIP(hipMemcpy(l, cuts.data(), cuts.size() * sizeof(Bf3Cut), hipMemcpyHostToDevice));
  PHPC_HIP(hipMemcpy(l + cuts_b, block_cut.data(), block_cut.size() * sizeof(int), hipMemcpyHostToDevice));
  PHPC_HIP(hipMemcpy(l + cuts_b + bc_b, items.data(), it_b, hipMemcpyHostToDevice));
  return PHPC_OK;
}

int Bf3Planes::run(hipStream_t s, int grid) const {
  if (items.empty() || total_tiles <= 0) return PHPC_OK;
  const char *l = static_cast<const char *>(list);
  const size_t cuts_b = bf3_align(cuts.size() * sizeof(Bf3Cut));
  const size_t bc_b = bf3_align(block_cut.size() * sizeof(int));
  hipLaunchKernelGGL(bf16x3_cut_kernel, dim3((unsigned)blocks), dim3(256), 0, s, reinterpret_cast<const Bf3Cut *>(l),
                     reinterpret_cast<const int *>(l + cuts_b));
  PHPC_HIP(hipGetLastError());
  const Bf3GroupConfig &cf = kBf3Group[g];
  const long G = std::min<long>(grid > 0 ? grid : (long)bf3_occupancy(g) * phpc_device().cus, total_tiles);
  Bf3Group a{};
  a.items = reinterpret_cast<const Bf3Item *>(l + cuts_b + bc_b);
  a.count = (int)items.size();
  a.total_tiles = total_tiles;
  const int kt = phpc_ktime_begin(s);
  hipLaunchKernelGGL(cf.fn, dim3((unsigned)G), dim3(cf.threads), cf.lds, s, a);
  phpc_ktime_end(kt, s);
  PHPC_HIP(hipGetLastError());
  return PHPC_OK;
}

void Bf3Planes::release() {
  if (list) (void)hipFree(list);
  if (ws) (void)hipFree(ws);
  list = nullptr;
  ws = nullptr;
}

// ---------------------------------------------------------------------------
// Allocation-free split GEMM: the CRT path above with its scratch (the statistics block,
// the int8 residue operands A8 and B8, the residue planes R8) carved out of one static
// device array, `scratch`, instead of a workspace the library allocates. Each GPU the
// library runs on holds its own copy of the array until the process ends; every call on
// that GPU reuses it (the path writes everything it later reads: the statistics are
// cleared by a kernel, A8, B8 and R8 are written whole), so calls that overlap on one GPU,
// from other host threads or on other streams, race.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kNoAllocMaxDevices] = {};  // this device's address of `scratch`

// Bytes the CRT path carves for an m x k x n product at L moduli, in its order, each at a
// 256-byte aligned offset: the statistics block, A8 (m x L kp), B8 (n x L kp), R8
// (L x m x n). In double precision, so that no size overflows.
double crt_scratch_need(int m, int k, int n, int L) {
  auto up = [](double x) { return std::ceil(x / 256.0) * 256.0; };
  const double kp = (double)phpc_split_kp(k);
  return up((double)split_meta_bytes(m, n)) + up((double)m * L * kp) + up((double)n * L * kp) + (double)L * m * n;
}
}  // namespace

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

extern "C" int phpc_dgemm_split_device_noalloc(const double *dA, long lda, const double *dB, long ldb, double *dC,
                                               long ldc, int m, int k, int n, int slices, double target_err,
                                               void *stream, phpc_split_info *info) {
  if (info) *info = phpc_split_info{};
  if (phpc_init()) return phpc_last_error();
  if (m < 0 || k < 0 || n < 0) return phpc_set_error(PHPC_ERR_ARG, "negative size");
  if (slices < 0 || slices == 1 || slices > kMaxMods)
    return phpc_set_error(PHPC_ERR_ARG, "moduli must be 0 (auto) or 2..%d", kMaxMods);
  // sized for the most moduli the call may pick, so it cannot run out part-way
  const int L = slices > 0 ? slices : kMaxMods;
  const double need = crt_scratch_need(m, k, n, L);
  if (need > (double)kNoAllocScratchBytes)
    return phpc_set_error(PHPC_ERR_ARG,
                          "%d x %d x %d needs %.0f bytes of scratch at %d moduli; the allocation-free split GEMM "
                          "holds %zu",
                          m, k, n, need, L, kNoAllocScratchBytes);
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices)
    return phpc_set_error(PHPC_ERR_HIP, "no usable current device (%d)", dev);
  void *ptr = nullptr;
  {
    // once per device, with that device current: the address of its copy of the array
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
      void *p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return phpc_set_error(PHPC_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      g_noalloc_ptr[dev] = p;
    }
    ptr = g_noalloc_ptr[dev];
  }
  Workspace ws;
  ws.use_fixed(ptr, kNoAllocScratchBytes);
  return phpc_dgemm_split_ws(dA, lda, dB, ldb, dC, ldc, m, k, n, slices, target_err, PHPC_SPLIT_CRT,
                             (hipStream_t)stream, &ws, info);
}

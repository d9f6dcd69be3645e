// FP32 grouped GEMM on gfx950 matrix cores (v_mfma_f32_32x32x2f32): C += A*B for a list
// of independent row-major FP32 GEMMs of any shapes, one persistent launch.
//
// Same structure as the FP64 kernel: a BM x BN macro tile per workgroup, 4 waves in a
// 2 x 2 arrangement, operands staged HBM -> registers (raw buffer loads, out-of-range
// lanes read zero) -> LDS (A transposed, At[k][m]; B as Bs[k][n]) in 16-deep slabs with
// double buffering, and a persistent workgroup loop over the flat tile list of all GEMMs.
// MFMA 32x32x2: lane l holds A[row l%32][k l/32] and B[k l/32][col l%32]; the 16
// accumulators of lane l are rows 8(r/4) + 4(l/32) + r%4, column l%32.
#include "phpc_internal.h"
#include "phpc_split.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

namespace {

typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef float __attribute__((address_space(1))) gfloat;

constexpr uint32_t kOOB = 0x80000000u;
constexpr int BK = 16;

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

struct F32Batch {
  const F32Desc *descs;
  int ndesc;
  int total_tiles;
  F32Desc d0;
};

__device__ __forceinline__ F32Desc uniform_desc(const F32Batch &b, int t) {
  const F32Desc *d = &b.d0;
  if (b.ndesc > 1) {
    int lo = 0, hi = b.ndesc - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (b.descs[mid].tile_start <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    d = &b.descs[lo];
  }
  F32Desc u;
  u.A = uni(d->A);
  u.B = uni(d->B);
  u.C = uni(d->C);
  u.lda = uni(d->lda);
  u.ldb = uni(d->ldb);
  u.ldc = uni(d->ldc);
  u.m = uni(d->m);
  u.n = uni(d->n);
  u.k = uni(d->k);
  u.tiles_n = uni(d->tiles_n);
  u.tile_start = uni(d->tile_start);
  return u;
}

template <int BM, int BN, int VEC>
__global__ __launch_bounds__(256) void f32_gemm_kernel(F32Batch batch) {
  constexpr int WTM = BM / 2, WTN = BN / 2;
  constexpr int TM = WTM / 32, TN = WTN / 32;
  constexpr int APITCH = BM + 4, BPITCH = BN + 4;
  constexpr int STAGE = BK * APITCH + BK * BPITCH;
  constexpr int CA = BM * BK / VEC / 256, CB = BN * BK / VEC / 256;
  static_assert(CA * VEC * 256 == BM * BK && CB * VEC * 256 == BN * BK, "slab split");
  static_assert(TM >= 1 && TN >= 1, "wave tile");
  __shared__ float lds[2 * STAGE];

  const int tid = threadIdx.x;
  const int wave = uni(tid >> 6);
  const int lane = tid & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const int G = gridDim.x, bid = blockIdx.x;
  const int slot = uni((G % 8) == 0 ? (bid % 8) * (G / 8) + bid / 8 : bid);

  for (int round = 0;; ++round) {
    const int t = round * G + slot;
    if (t >= batch.total_tiles) break;
    const F32Desc d = uniform_desc(batch, uni(t));
    const int lt = t - d.tile_start;
    const int row0 = uni((lt / d.tiles_n) * BM);
    const int col0 = uni((lt % d.tiles_n) * BN);
    const int m_rem = d.m - row0, n_rem = d.n - col0;
    const long lda = d.lda, ldb = d.ldb, ldc = d.ldc;
    const float *Ag = d.A + (long)row0 * lda;
    const float *Bg = d.B + col0;
    gfloat *Cg = (gfloat *)(d.C + (long)row0 * ldc + col0);
    const rsrc_t rA = make_rsrc(Ag, (long)min(m_rem, BM) * lda * 4);

    f16v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = f16v{};

    float ra[CA][VEC], rb[CB][VEC];
    auto load = [&](int kb) {
      const int krem = d.k - kb;
#pragma unroll
      for (int i = 0; i < CA; ++i) {
        const int q = tid + i * 256;
        const int row = q / (BK / VEC), kc = (q % (BK / VEC)) * VEC;
        const uint32_t off = (row < m_rem && kc < krem) ? (uint32_t)(((long)row * lda + kb + kc) * 4) : kOOB;
        if constexpr (VEC == 4) {
          // bit-cast the whole 128-bit result (element-wise extraction from the builtin's
          // integer vector is lowered to a single-dword load)
          const f4v v = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(rA, off, 0, 0));
#pragma unroll
          for (int e = 0; e < 4; ++e) ra[i][e] = v[e];
        } else {
          ra[i][0] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rA, off, 0, 0));
        }
      }
      const rsrc_t rB = make_rsrc(Bg + (long)kb * ldb, (long)krem * ldb * 4);
#pragma unroll
      for (int i = 0; i < CB; ++i) {
        const int q = tid + i * 256;
        const int kr = q / (BN / VEC), cc = (q % (BN / VEC)) * VEC;
        const uint32_t off = (kr < krem && cc < n_rem) ? (uint32_t)(((long)kr * ldb + cc) * 4) : kOOB;
        if constexpr (VEC == 4) {
          const f4v v = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(rB, off, 0, 0));
#pragma unroll
          for (int e = 0; e < 4; ++e) rb[i][e] = v[e];
        } else {
          rb[i][0] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rB, off, 0, 0));
        }
      }
    };
    auto store = [&](int buf) {
      float *At = lds + buf * STAGE;
      float *Bs = At + BK * APITCH;
#pragma unroll
      for (int i = 0; i < CA; ++i) {
        const int q = tid + i * 256;
        const int row = q / (BK / VEC), kc = (q % (BK / VEC)) * VEC;
#pragma unroll
        for (int e = 0; e < VEC; ++e) At[(kc + e) * APITCH + row] = ra[i][e];
      }
#pragma unroll
      for (int i = 0; i < CB; ++i) {
        const int q = tid + i * 256;
        const int kr = q / (BN / VEC), cc = (q % (BN / VEC)) * VEC;
#pragma unroll
        for (int e = 0; e < VEC; ++e) Bs[kr * BPITCH + cc + e] = rb[i][e];
      }
    };

    __syncthreads();
    const int nslab = (d.k + BK - 1) / BK;
    load(0);
    store(0);
    __syncthreads();
    for (int sidx = 0; sidx < nslab; ++sidx) {
      const int buf = sidx & 1;
      if (sidx + 1 < nslab) load((sidx + 1) * BK);
      const float *At = lds + buf * STAGE;
      const float *Bs = At + BK * APITCH;
#pragma unroll
      for (int kk = 0; kk < BK / 2; ++kk) {
        const int k = kk * 2 + (lane >> 5);
        float a[TM], b[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) a[i] = At[k * APITCH + wm * WTM + i * 32 + (lane & 31)];
#pragma unroll
        for (int j = 0; j < TN; ++j) b[j] = Bs[k * BPITCH + wn * WTN + j * 32 + (lane & 31)];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
      }
      if (sidx + 1 < nslab) {
        store(buf ^ 1);
        __syncthreads();
      }
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int cc = wn * WTN + j * 32 + (lane & 31);
        if (cc >= n_rem) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int rr = wm * WTM + i * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
          if (rr < m_rem) {
            gfloat *pc = Cg + (long)rr * ldc + cc;
            *pc = *pc + acc[i][j][r];
          }
        }
      }
  }
}


// ---------------------------------------------------------------------------
// FP32-accurate GEMM from three BF16 slices (v_mfma_f32_32x32x16_bf16)
// ---------------------------------------------------------------------------
// Every FP32 operand is cut exactly into x = hi + mid + lo, each a BF16 value: hi is x
// rounded to nearest BF16, mid the rounded remainder, lo the rest (|mid| <= 2^-8|x|,
// |lo| <= 2^-16|x|; with round-to-nearest the remainder after two slices has at most 8
// significant bits, so lo is exact for finite inputs whose slices stay normal). The six
// products with slice indices summing to at most 2 (hi.hi, hi.mid, mid.hi, hi.lo,
// mid.mid, lo.hi) run on the BF16 matrix cores with FP32 accumulation; the dropped
// mid.lo, lo.mid and lo.lo terms are below 2^-23 |a||b|. The result carries the error
// of an FP32 dot product: |C - exact| <= ((k + 2) 2^-24 + 2^-23) (|C0| + |A||B|).
//
// Operands are loaded as FP32 from HBM (raw buffer loads, zero outside the matrix),
// split in registers and stored to LDS as three BF16 planes with k contiguous: A as
// [plane][row][k], B transposed as [plane][col][k], so each MFMA operand is one 16-byte
// LDS read (lane l: row / column l % 32, k = 8 (l / 32) .. + 7 of the 16-deep step).
// The next slab's FP32 values are loaded into registers while the current one computes.
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef unsigned short us8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint32_t bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_val(uint32_t h) { return __uint_as_float(h << 16); }

template <int BM, int BN, int VEC>
__global__ __launch_bounds__(256) void f32x3_gemm_kernel(F32Batch batch) {
  constexpr int BK = 32;
  constexpr int KP = BK + 8;  // bf16 per LDS row (80 bytes: 16-byte aligned, staggered banks)
  constexpr int WTM = BM / 2, WTN = BN / 2;
  constexpr int TM = WTM / 32, TN = WTN / 32;
  constexpr int AE = BM * BK / 256;  // A floats per thread
  constexpr int BE = BN * BK / 256;  // B floats per thread
  static_assert(AE == 16 && BE == 16, "thread mapping assumes 128 x 128 tiles");
  __shared__ __attribute__((aligned(16))) unsigned short lds[3 * BM * KP + 3 * BN * KP];
  unsigned short *Al = lds;
  unsigned short *Bl = lds + 3 * BM * KP;

  const int tid = threadIdx.x;
  const int wave = uni(tid >> 6);
  const int lane = tid & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const int G = gridDim.x, bid = blockIdx.x;
  const int slot = uni((G % 8) == 0 ? (bid % 8) * (G / 8) + bid / 8 : bid);
  // A: thread -> row tid / 2, k offset (tid % 2) * 16; B: column tid % 128, k half tid / 128
  const int a_row = tid >> 1, a_k = (tid & 1) * 16;
  const int b_col = tid & 127, b_k = (tid >> 7) * 16;

  for (int round = 0;; ++round) {
    const int t = round * G + slot;
    if (t >= batch.total_tiles) break;
    const F32Desc d = uniform_desc(batch, uni(t));
    const int lt = t - d.tile_start;
    const int row0 = uni((lt / d.tiles_n) * BM);
    const int col0 = uni((lt % d.tiles_n) * BN);
    const int m_rem = d.m - row0, n_rem = d.n - col0;
    const long lda = d.lda, ldb = d.ldb, ldc = d.ldc;
    const float *Ag = d.A + (long)row0 * lda;
    const float *Bg = d.B + col0;
    gfloat *Cg = (gfloat *)(d.C + (long)row0 * ldc + col0);
    const rsrc_t rA = make_rsrc(Ag, (long)min(m_rem, BM) * lda * 4);

    f16v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = f16v{};

    float ra[16], rb[16];
    auto load = [&](int kb) {
      const int krem = d.k - kb;
      if constexpr (VEC == 4) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int kc = a_k + 4 * q;
          const uint32_t off = (a_row < m_rem && kc < krem) ? (uint32_t)(((long)a_row * lda + kb + kc) * 4) : kOOB;
          const f4v v = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(rA, off, 0, 0));
#pragma unroll
          for (int e = 0; e < 4; ++e) ra[4 * q + e] = v[e];
        }
      } else {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int kc = a_k + e;
          const uint32_t off = (a_row < m_rem && kc < krem) ? (uint32_t)(((long)a_row * lda + kb + kc) * 4) : kOOB;
          ra[e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rA, off, 0, 0));
        }
      }
      const rsrc_t rB = make_rsrc(Bg + (long)kb * ldb, (long)krem * ldb * 4);
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int kr = b_k + e;
        const uint32_t off = (kr < krem && b_col < n_rem) ? (uint32_t)(((long)kr * ldb + b_col) * 4) : kOOB;
        rb[e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rB, off, 0, 0));
      }
    };
    // split 16 consecutive-k values into three 16-element BF16 runs at dst + plane * pstride
    auto split_store = [&](const float *x, unsigned short *dst, int pstride) {
      us8v h[2], m[2], l[2];
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const uint32_t hb = bf16_rne(x[e]);
        const float r1 = x[e] - bf16_val(hb);
        const uint32_t mb = bf16_rne(r1);
        const float r2 = r1 - bf16_val(mb);
        h[e >> 3][e & 7] = (unsigned short)hb;
        m[e >> 3][e & 7] = (unsigned short)mb;
        l[e >> 3][e & 7] = (unsigned short)bf16_rne(r2);
      }
#pragma unroll
      for (int hlf = 0; hlf < 2; ++hlf) {
        *(us8v *)(dst + 8 * hlf) = h[hlf];
        *(us8v *)(dst + pstride + 8 * hlf) = m[hlf];
        *(us8v *)(dst + 2 * pstride + 8 * hlf) = l[hlf];
      }
    };

    const int nslab = (d.k + BK - 1) / BK;
    load(0);
    for (int sidx = 0; sidx < nslab; ++sidx) {
      __syncthreads();  // the previous slab's MFMA reads are done
      split_store(ra, Al + a_row * KP + a_k, BM * KP);
      split_store(rb, Bl + b_col * KP + b_k, BN * KP);
      __syncthreads();
      if (sidx + 1 < nslab) load((sidx + 1) * BK);
#pragma unroll
      for (int kk = 0; kk < BK / 16; ++kk) {
        const int ko = kk * 16 + 8 * (lane >> 5);
        bf8v a[TM][3], b[TN][3];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int p = 0; p < 3; ++p)
            a[i][p] = __builtin_bit_cast(
                bf8v, *(const us8v *)(Al + p * BM * KP + (wm * WTM + i * 32 + (lane & 31)) * KP + ko));
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int p = 0; p < 3; ++p)
            b[j][p] = __builtin_bit_cast(
                bf8v, *(const us8v *)(Bl + p * BN * KP + (wn * WTN + j * 32 + (lane & 31)) * KP + ko));
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            // small terms first
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][2], b[j][0], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][1], b[j][1], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][0], b[j][2], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][1], b[j][0], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][0], b[j][1], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][0], b[j][0], acc[i][j], 0, 0, 0);
          }
      }
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int cc = wn * WTN + j * 32 + (lane & 31);
        if (cc >= n_rem) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int rr = wm * WTM + i * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
          if (rr < m_rem) {
            gfloat *pc = Cg + (long)rr * ldc + cc;
            *pc = *pc + acc[i][j][r];
          }
        }
      }
    __syncthreads();  // LDS is rewritten by the next tile
  }
}

// ---------------------------------------------------------------------------
// Transposed-A direct-to-LDS FP32 GEMM for large products: the caller stores A
// transposed (At[k][m], row pitch = the descriptor's lda, a multiple of 4 floats, pad
// columns zero) and B with 16-byte aligned rows, so both 16-row slabs are k-row-major and
// reach LDS by 16-byte buffer loads with no staging registers and no LDS write
// instructions. A 1 KiB load instruction fills two 128-float k rows; within a row the
// 16-byte chunks of odd rows are XOR-permuted by 8, so the two half-waves of an MFMA
// operand read (rows k and k+1) hit opposite halves of the banks. Two stages.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void lds_void;

template <int N>
__device__ __forceinline__ void f32_wait_vm() {
  __builtin_amdgcn_s_waitcnt(0x0F70 | (N & 0xF) | (((N >> 4) & 3) << 14));
}

// B4: B rows need no alignment. Each B k row then arrives as two 4-byte-per-lane loads of 64
// columns (lane i of half h at position 64 h + i holds column (64 h + i) ^ 32 on odd rows,
// the same permutation as the 16-byte chunks' XOR by 8), exact to the last column of the
// row; A still arrives in 1 KiB blocks.
template <bool B4>
__device__ __forceinline__ void f32_at_body(const F32Batch &batch) {
  constexpr int BM = 128, BN = 128, WTM = 64, WTN = 64, TM = 2, TN = 2;
  constexpr int ROW_BYTES = 512;                  // one 128-float k row
  constexpr int A_BYTES = BK * ROW_BYTES;         // 16 rows
  constexpr int STAGE_BYTES = 2 * A_BYTES;        // A then B
  constexpr int A_BLOCKS = A_BYTES / 1024, BLOCKS = 2 * A_BLOCKS, PER = BLOCKS / 4;
  constexpr int B4_PER = 2 * BK / 4;              // 4-byte B loads per wave and slab
  extern __shared__ __attribute__((aligned(1024))) char smem[];
  const int tid = threadIdx.x;
  const int wave = uni(tid >> 6);
  const int lane = tid & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const int G = gridDim.x, bid = blockIdx.x;
  const int slot = uni((G % 8) == 0 ? (bid % 8) * (G / 8) + bid / 8 : bid);
  // this lane's place in any 1 KiB block: row pair offset (lane / 32) and chunk
  const int lrow = lane >> 5;
  const int lchunk = (lane & 31) ^ (lrow << 3);

  for (int round = 0;; ++round) {
    const int t = round * G + slot;
    if (t >= batch.total_tiles) break;
    const F32Desc d = uniform_desc(batch, uni(t));
    const int lt = t - d.tile_start;
    const int row0 = uni((lt / d.tiles_n) * BM);
    const int col0 = uni((lt % d.tiles_n) * BN);
    const int m_rem = d.m - row0, n_rem = d.n - col0;
    const long ldat = d.lda, ldb = d.ldb, ldc = d.ldc;
    const float *At_g = d.A + row0;
    const float *Bg = d.B + col0;
    gfloat *Cg = (gfloat *)(d.C + (long)row0 * ldc + col0);
    const int Kdim = d.k;
    const int nslab = (Kdim + BK - 1) / BK;

    uint32_t goff[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int blk = wave + u * 4;
      const bool isA = blk < A_BLOCKS;
      const int r = 2 * (isA ? blk : blk - A_BLOCKS) + lrow;  // k row within the slab
      const int c = lchunk * 4;                               // column of the chunk
      const int rem = isA ? m_rem : n_rem;
      goff[u] = c < rem ? (uint32_t)(((long)r * (isA ? ldat : ldb) + c) * 4) : kOOB;
    }
    uint32_t goff_b4[B4 ? B4_PER : 1];
    if constexpr (B4) {
#pragma unroll
      for (int v = 0; v < B4_PER; ++v) {
        const int j = wave + v * 4;  // row j / 2, half j % 2
        const int r = j >> 1, h = j & 1;
        const int c = (64 * h + lane) ^ ((r & 1) << 5);
        goff_b4[v] = c < n_rem ? (uint32_t)(((long)r * ldb + c) * 4) : kOOB;
      }
    }
    const int full_a = (int)(BK * ldat * 4), full_b = (int)(BK * ldb * 4);
    auto issue = [&](int sidx) {
      const int kb = sidx * BK;
      const int krem = min(BK, Kdim - kb);
      char *stage = smem + (sidx & 1) * STAGE_BYTES;
      const int na = krem == BK ? full_a : krem * (int)ldat * 4;
      const int nb = krem == BK ? full_b : krem * (int)ldb * 4;
      const rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(At_g + (long)kb * ldat), (short)0, na,
                                                          0x00020000);
      const rsrc_t rB =
          __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(Bg + (long)kb * ldb), (short)0, nb, 0x00020000);
      if constexpr (B4) {
#pragma unroll
        for (int u = 0; u < A_BLOCKS / 4; ++u) {
          const int blk = wave + u * 4;
          __builtin_amdgcn_raw_ptr_buffer_load_lds(rA, (lds_void *)(stage + blk * 1024), 16, goff[u], 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < B4_PER; ++v) {
          const int j = wave + v * 4;
          __builtin_amdgcn_raw_ptr_buffer_load_lds(rB, (lds_void *)(stage + A_BYTES + (j >> 1) * ROW_BYTES + (j & 1) * 256),
                                                    4, goff_b4[v], 0, 0, 0);
        }
      } else {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
          const int blk = wave + u * 4;
          __builtin_amdgcn_raw_ptr_buffer_load_lds(blk < A_BLOCKS ? rA : rB, (lds_void *)(stage + blk * 1024), 16,
                                                    goff[u], 0, 0, 0);
        }
      }
    };

    f16v acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = f16v{};

    // LDS read offsets of this lane's operands for k step kk: row k = 2 kk + lane / 32
    int a_off[TM], b_off[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int col = wm * WTM + i * 32 + (lane & 31);
      a_off[i] = lrow * ROW_BYTES + (((col >> 2) ^ (lrow << 3)) << 4) + (col & 3) * 4;
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int col = wn * WTN + j * 32 + (lane & 31);
      b_off[j] = A_BYTES + lrow * ROW_BYTES + (((col >> 2) ^ (lrow << 3)) << 4) + (col & 3) * 4;
    }

    __syncthreads();
    if (nslab > 0) issue(0);
    for (int sidx = 0; sidx < nslab; ++sidx) {
      f32_wait_vm<0>();
      __syncthreads();
      const char *stage = smem + (sidx & 1) * STAGE_BYTES;
      auto kstep = [&](int kk) {
        float a[TM], b[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) a[i] = *reinterpret_cast<const float *>(stage + a_off[i] + kk * 2 * ROW_BYTES);
#pragma unroll
        for (int j = 0; j < TN; ++j) b[j] = *reinterpret_cast<const float *>(stage + b_off[j] + kk * 2 * ROW_BYTES);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
      };
      kstep(0);
      if (sidx + 1 < nslab) issue(sidx + 1);
#pragma unroll
      for (int kk = 1; kk < BK / 2; ++kk) kstep(kk);
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int cc = wn * WTN + j * 32 + (lane & 31);
        if (cc >= n_rem) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int rr = wm * WTM + i * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
          if (rr < m_rem) {
            gfloat *pc = Cg + (long)rr * ldc + cc;
            *pc = *pc + acc[i][j][r];
          }
        }
      }
  }
}

__global__ __launch_bounds__(256) void f32_at_128x128(F32Batch batch) { f32_at_body<false>(batch); }
__global__ __launch_bounds__(256) void f32_at_128x128_b4(F32Batch batch) { f32_at_body<true>(batch); }

// At[k][m] = A[m][k] for m < M, k < K; 0 for M <= m < ldat
// 64 x 64 tiles, 16 entries per thread
constexpr int kF32TransTile = 64;
__global__ __launch_bounds__(256) void f32_transpose_kernel(const float *__restrict__ A, long lda, int M, int Kd,
                                                            float *__restrict__ At, long ldat) {
  constexpr int T = kF32TransTile;
  __shared__ float tile[T][T + 1];
  const int k0 = blockIdx.x * T, m0 = blockIdx.y * T;
  const int tx = threadIdx.x % T, ty = threadIdx.x / T;
  constexpr int RS = 256 / T;  // row step
  float v[T / RS];
#pragma unroll
  for (int u = 0; u < T / RS; ++u) {
    const int m = m0 + ty + u * RS, k = k0 + tx;
    v[u] = load_or_zero(A, (long)m * lda + k, m < M && k < Kd);
  }
#pragma unroll
  for (int u = 0; u < T / RS; ++u) tile[ty + u * RS][tx] = v[u];
  __syncthreads();
  // 16-byte stores: a thread writes 4 consecutive m of 4 rows k (ldat and m are multiples
  // of 4, so a group lies wholly below or wholly past ldat)
  static_assert(T == 64, "store groups assume 64 x 64 tiles");
  const int q = threadIdx.x & 15;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int r = (threadIdx.x >> 4) + 16 * u;
    const int k = k0 + r, m = m0 + 4 * q;
    if (k < Kd && m < ldat)
      *reinterpret_cast<float4 *>(At + (long)k * ldat + m) =
          make_float4(tile[4 * q][r], tile[4 * q + 1][r], tile[4 * q + 2][r], tile[4 * q + 3][r]);
  }
}

// All transposed A copies of a grouped transposed-A launch in one kernel: one workgroup per
// 64 x 64 tile of a copy's destination, its copy read from a per-tile table.
__global__ __launch_bounds__(256) void f32_stage_kernel(const F32StageCopy *__restrict__ copies,
                                                        const int *__restrict__ block_copy) {
  __shared__ float tile[64][65];
  const int b = blockIdx.x;
  const F32StageCopy c = copies[block_copy[b]];
  const int lb = b - c.block_start;
  const int d0 = (lb / c.tiles_x) * 64;  // first destination row (k) of the tile
  const int e0 = (lb % c.tiles_x) * 64;  // first destination column (m)
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  float v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int sr = e0 + ty + 4 * i, sc = d0 + tx;
    v[i] = load_or_zero(c.src, (long)sr * c.ld + sc, sr < c.rows && sc < c.cols);
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) tile[ty + 4 * i][tx] = v[i];
  __syncthreads();
  // 16-byte stores of 4 consecutive destination columns (pitch and dc are multiples of 4,
  // copies 256-byte aligned)
  const int q = threadIdx.x & 15;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = (threadIdx.x >> 4) + 16 * i;
    const int dr = d0 + r, dc = e0 + 4 * q;
    if (dr < c.cols && dc < c.pitch)
      *reinterpret_cast<float4 *>(c.dst + (long)dr * c.pitch + dc) =
          make_float4(tile[4 * q][r], tile[4 * q + 1][r], tile[4 * q + 2][r], tile[4 * q + 3][r]);
  }
}

typedef void (*F32Fn)(F32Batch);
struct F32Cfg {
  int bm, bn;
  F32Fn fn[2];  // [vec4]
};
const F32Cfg kF32[] = {
    {64, 64, {f32_gemm_kernel<64, 64, 1>, f32_gemm_kernel<64, 64, 4>}},
    {128, 64, {f32_gemm_kernel<128, 64, 1>, f32_gemm_kernel<128, 64, 4>}},
    {128, 128, {f32_gemm_kernel<128, 128, 1>, f32_gemm_kernel<128, 128, 4>}},
};
constexpr int kNumF32 = sizeof(kF32) / sizeof(kF32[0]);

std::mutex g_mu;
bool g_ready = false;
int g_occ[kNumF32][2];

int occ(int c, int v) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ready) {
    for (int i = 0; i < kNumF32; ++i)
      for (int j = 0; j < 2; ++j) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(kF32[i].fn[j]), 256, 0) !=
                hipSuccess ||
            n < 1)
          n = 1;
        g_occ[i][j] = n;
      }
    g_ready = true;
  }
  return g_occ[c][v];
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

int g_f32_at = 1;  // large single FP32 products through the transposed-A kernel
int g_f32_plan_cfg = -1;  // FP32 plans' configuration without autotuning (-1: by work)
SharedWorkspace g_f32_at_ws;
int g_f32_at_occ[2] = {0, 0};

constexpr size_t kF32AtLds = 2 * 2 * 16 * 512;  // two stages of A and B slabs

// b4: some B rows are not 16-byte aligned (or their pitch is not a multiple of 4 floats);
// the kernel then loads B 4 bytes per lane
int f32_at_run(const F32Batch &b, bool b4, int grid, hipStream_t stream) {
  const void *fn = b4 ? reinterpret_cast<const void *>(f32_at_128x128_b4) : reinterpret_cast<const void *>(f32_at_128x128);
  int &occ_v = g_f32_at_occ[b4 ? 1 : 0];
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!occ_v) {
      (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kF32AtLds);
      int n = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 256, kF32AtLds) != hipSuccess || n < 1) n = 1;
      occ_v = n;
    }
  }
  int G = grid > 0 ? grid : occ_v * phpc_device().cus;
  if (G > b.total_tiles) G = b.total_tiles;
  if (G < 1) G = 1;
  const int kt = phpc_ktime_begin(stream);
  if (b4)
    hipLaunchKernelGGL(f32_at_128x128_b4, dim3(G), dim3(256), kF32AtLds, stream, b);
  else
    hipLaunchKernelGGL(f32_at_128x128, dim3(G), dim3(256), kF32AtLds, stream, b);
  phpc_ktime_end(kt, stream);
  PHPC_HIP(hipGetLastError());
  return PHPC_OK;
}

bool b16_ok(const F32Desc &d) { return aligned16(d.B) && d.ldb % 4 == 0; }

// One large FP32 GEMM: A staged transposed, then the transposed-A kernel over 128 x 128
// tiles, B read in place (16-byte loads when its rows are 16-byte aligned with a pitch of 4
// floats, else 4-byte loads). With 16-byte loads, the chunk holding a row's last column
// starts on a 16-byte boundary and so lies in the same page as that column; its extra
// columns only feed output columns that are never stored.
int f32_launch_transposed(F32Desc d, hipStream_t stream) {
  const long ldat = ((long)d.m + 3) / 4 * 4;
  const size_t total = ((size_t)d.k * ldat + 63) / 64 * 64;
  if (16L * ldat * 4 >= 0x7fffffffL || 16L * d.ldb * 4 >= 0x7fffffffL)
    return phpc_set_error(PHPC_ERR_ARG, "leading dimension too large");
  std::lock_guard<std::mutex> lk(g_f32_at_ws.mu);
  g_f32_at_ws.after_previous(stream);
  float *ws = static_cast<float *>(g_f32_at_ws.ws.get(total * sizeof(float), 0, stream));
  if (!ws) return phpc_set_error(PHPC_ERR_MEM, "FP32 staging workspace allocation failed");
  hipLaunchKernelGGL(f32_transpose_kernel,
                     dim3((unsigned)((d.k + kF32TransTile - 1) / kF32TransTile),
                          (unsigned)((ldat + kF32TransTile - 1) / kF32TransTile)),
                     dim3(256), 0, stream, d.A, d.lda, d.m, d.k, ws, ldat);
  PHPC_HIP(hipGetLastError());
  d.A = ws;
  d.lda = ldat;
  d.tiles_n = (d.n + 127) / 128;
  d.tile_start = 0;
  const int total_tiles = ((d.m + 127) / 128) * d.tiles_n;
  F32Batch b{};
  b.ndesc = 1;
  b.total_tiles = total_tiles;
  b.d0 = d;
  if (int rc = f32_at_run(b, !b16_ok(d), 0, stream)) return rc;
  g_f32_at_ws.mark_used(stream);
  return PHPC_OK;
}

}  // namespace

static_assert(kF32AtConfig == kNumF32, "the transposed-A configuration follows the register-staged ones");

int phpc_f32_num_configs() { return kNumF32 + 1; }

void phpc_f32_config_tile(int id, int *bm, int *bn) {
  *bm = id == kF32AtConfig ? 128 : kF32[id].bm;
  *bn = id == kF32AtConfig ? 128 : kF32[id].bn;
}

// Fills tiles_n / tile_start of descs for config `id`; returns the total tile count.
int phpc_f32_plan_tiles(F32Desc *descs, int ndesc, int id) {
  int bm, bn;
  phpc_f32_config_tile(id, &bm, &bn);
  int start = 0;
  for (int i = 0; i < ndesc; ++i) {
    F32Desc &d = descs[i];
    d.tiles_n = (d.n + bn - 1) / bn;
    d.tile_start = start;
    start += ((d.m + bm - 1) / bm) * d.tiles_n;
  }
  return start;
}

bool phpc_f32_at_ok(const F32Desc *descs, int ndesc) {
  for (int i = 0; i < ndesc; ++i) {
    const F32Desc &d = descs[i];
    const long ldat = ((long)d.m + 3) / 4 * 4;
    if (16L * ldat * 4 >= 0x7fffffffL || 16L * d.ldb * 4 >= 0x7fffffffL) return false;
  }
  return true;
}

void F32Staging::plan(const F32Desc *descs, int ndesc) {
  copies.clear();
  off.clear();
  block_copy.clear();
  a_copy.assign(ndesc, -1);
  total = 0;
  blocks = 0;
  auto find_or_add = [&](const float *src, long ld, int rows, int cols) {
    for (size_t j = 0; j < copies.size(); ++j) {
      const F32StageCopy &c = copies[j];
      if (c.src == src && c.ld == ld && c.rows == rows && c.cols == cols) return (int)j;
    }
    F32StageCopy c{};
    c.src = src;
    c.ld = ld;
    c.rows = rows;
    c.cols = cols;
    c.pitch = ((long)rows + 3) / 4 * 4;
    c.tiles_x = (int)((c.pitch + 63) / 64);
    c.block_start = blocks;
    const int nt = c.tiles_x * ((cols + 63) / 64);
    blocks += nt;
    block_copy.insert(block_copy.end(), nt, (int)copies.size());
    copies.push_back(c);
    off.push_back(total);
    const size_t n = (size_t)cols * c.pitch;
    total += (n + 63) / 64 * 64;  // 256-byte aligned copies
    return (int)copies.size() - 1;
  };
  for (int i = 0; i < ndesc; ++i) {
    const F32Desc &d = descs[i];
    if (d.m <= 0 || d.n <= 0 || d.k <= 0) continue;
    a_copy[i] = find_or_add(d.A, d.lda, d.m, d.k);
  }
}

void F32Staging::apply(float *ws, F32Desc *descs, int ndesc) {
  for (size_t j = 0; j < copies.size(); ++j) copies[j].dst = ws + off[j];
  for (int i = 0; i < ndesc; ++i) {
    if (a_copy[i] >= 0) {
      descs[i].A = copies[a_copy[i]].dst;
      descs[i].lda = copies[a_copy[i]].pitch;
    }
  }
}

static size_t copies_bytes(size_t n) { return (n * sizeof(F32StageCopy) + 255) / 256 * 256; }

size_t F32Staging::list_bytes() const { return copies_bytes(copies.size()) + block_copy.size() * sizeof(int); }

int F32Staging::upload(void *list_dev) const {
  if (copies.empty()) return PHPC_OK;
  char *p = static_cast<char *>(list_dev);
  PHPC_HIP(hipMemcpy(p, copies.data(), copies.size() * sizeof(F32StageCopy), hipMemcpyHostToDevice));
  PHPC_HIP(hipMemcpy(p + copies_bytes(copies.size()), block_copy.data(), block_copy.size() * sizeof(int),
                     hipMemcpyHostToDevice));
  return PHPC_OK;
}

int F32Staging::run(const void *list_dev, hipStream_t stream) const {
  if (copies.empty() || blocks <= 0) return PHPC_OK;
  const char *p = static_cast<const char *>(list_dev);
  hipLaunchKernelGGL(f32_stage_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                     reinterpret_cast<const F32StageCopy *>(p), reinterpret_cast<const int *>(p + copies_bytes(copies.size())));
  PHPC_HIP(hipGetLastError());
  return PHPC_OK;
}

int phpc_f32_launch_at(const F32Desc *descs_dev, const F32Desc *descs_host, int ndesc, int total_tiles, int grid,
                       hipStream_t stream) {
  if (ndesc <= 0 || total_tiles <= 0) return PHPC_OK;
  bool b4 = false;
  for (int i = 0; i < ndesc; ++i) {
    const F32Desc &d = descs_host[i];
    if (d.lda % 4 || !aligned16(d.A) || d.lda < ((long)d.m + 3) / 4 * 4)
      return phpc_set_error(PHPC_ERR_ARG, "transposed-A FP32 launch needs a staged A");
    if (16L * d.lda * 4 >= 0x7fffffffL || 16L * d.ldb * 4 >= 0x7fffffffL)
      return phpc_set_error(PHPC_ERR_ARG, "leading dimension too large");
    if (!b16_ok(d)) b4 = true;
  }
  F32Batch b{};
  b.descs = descs_dev;
  b.ndesc = ndesc;
  b.total_tiles = total_tiles;
  b.d0 = descs_host[0];
  return f32_at_run(b, b4, grid, stream);
}

bool phpc_f32_vec4_ok(const F32Desc *descs, int ndesc) {
  for (int i = 0; i < ndesc; ++i) {
    const F32Desc &d = descs[i];
    if (!aligned16(d.A) || !aligned16(d.B) || d.lda % 4 || d.ldb % 4 || d.k % 4 || d.n % 4) return false;
  }
  return true;
}

// FP32-accurate 3 x BF16 kernel: 128 x 128 tiles (plan with config 2's tile size).
int phpc_f32x3_launch(const F32Desc *descs_dev, const F32Desc *descs_host, int ndesc, int total_tiles, bool vec4,
                      int grid, hipStream_t stream) {
  if (ndesc <= 0 || total_tiles <= 0) return PHPC_OK;
  for (int i = 0; i < ndesc; ++i) {
    const F32Desc &d = descs_host[i];
    if (128L * d.lda * 4 >= 0x7fffffffL || 32L * d.ldb * 4 >= 0x7fffffffL)
      return phpc_set_error(PHPC_ERR_ARG, "leading dimension too large");
  }
  int occ;
  {
    static int occ_cache[2] = {0, 0};
    std::lock_guard<std::mutex> lk(g_mu);
    const void *fn = vec4 ? reinterpret_cast<const void *>(f32x3_gemm_kernel<128, 128, 4>)
                          : reinterpret_cast<const void *>(f32x3_gemm_kernel<128, 128, 1>);
    if (occ_cache[vec4 ? 1 : 0] == 0) {
      int n = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 256, 0) != hipSuccess || n < 1) n = 1;
      occ_cache[vec4 ? 1 : 0] = n;
    }
    occ = occ_cache[vec4 ? 1 : 0];
  }
  F32Batch b{};
  b.descs = descs_dev;
  b.ndesc = ndesc;
  b.total_tiles = total_tiles;
  b.d0 = descs_host[0];
  int G = grid > 0 ? grid : occ * phpc_device().cus;
  if (G > total_tiles) G = total_tiles;
  if (G < 1) G = 1;
  if (vec4)
    hipLaunchKernelGGL((f32x3_gemm_kernel<128, 128, 4>), dim3(G), dim3(256), 0, stream, b);
  else
    hipLaunchKernelGGL((f32x3_gemm_kernel<128, 128, 1>), dim3(G), dim3(256), 0, stream, b);
  PHPC_HIP(hipGetLastError());
  return PHPC_OK;
}

// Launch the grouped kernel over device-resident descriptors (planned for config id).
int phpc_f32_launch(const F32Desc *descs_dev, const F32Desc *descs_host, int ndesc, int total_tiles, int id, bool vec4,
                    int grid, hipStream_t stream) {
  if (ndesc <= 0 || total_tiles <= 0) return PHPC_OK;
  if (id < 0 || id >= kNumF32) return phpc_set_error(PHPC_ERR_ARG, "FP32 config %d is not register-staged", id);
  for (int i = 0; i < ndesc; ++i) {
    const F32Desc &d = descs_host[i];
    if ((long)kF32[id].bm * d.lda * 4 >= 0x7fffffffL || 16L * d.ldb * 4 >= 0x7fffffffL)
      return phpc_set_error(PHPC_ERR_ARG, "leading dimension too large");
  }
  F32Batch b{};
  b.descs = descs_dev;
  b.ndesc = ndesc;
  b.total_tiles = total_tiles;
  b.d0 = descs_host[0];
  int G = grid > 0 ? grid : occ(id, vec4) * phpc_device().cus;
  if (G > total_tiles) G = total_tiles;
  if (G < 1) G = 1;
  const int kt = phpc_ktime_begin(stream);
  hipLaunchKernelGGL(kF32[id].fn[vec4 ? 1 : 0], dim3(G), dim3(256), 0, stream, b);
  phpc_ktime_end(kt, stream);
  PHPC_HIP(hipGetLastError());
  return PHPC_OK;
}

// Single FP32 GEMM on device pointers: FP32 matrix cores or the 3 x BF16 kernel.
extern "C" int phpc_sgemm_device(const float *dA, long lda, const float *dB, long ldb, float *dC, long ldc, int m, int k,
                                 int n, int method, void *stream) {
  if (phpc_init()) return phpc_last_error();
  if (method != PHPC_SGEMM_FP32 && method != PHPC_SGEMM_BF16X3) return phpc_set_error(PHPC_ERR_ARG, "bad method");
  if (m < 0 || k < 0 || n < 0) return phpc_set_error(PHPC_ERR_ARG, "negative size");
  if ((m > 0 && k > 0 && (!dA || lda < k)) || (k > 0 && n > 0 && (!dB || ldb < n)) || (m > 0 && n > 0 && (!dC || ldc < n)))
    return phpc_set_error(PHPC_ERR_ARG, "bad pointers or leading dimensions");
  if (m == 0 || n == 0 || k == 0) return phpc_set_error(PHPC_OK, nullptr);
  F32Desc d{};
  d.A = dA;
  d.B = dB;
  d.C = dC;
  d.lda = lda;
  d.ldb = ldb;
  d.ldc = ldc;
  d.m = m;
  d.n = n;
  d.k = k;
  const bool vec4 = phpc_f32_vec4_ok(&d, 1);
  int rc;
  if (method == PHPC_SGEMM_BF16X3 && (long)m * n >= 2048L * 2048 && k >= 256) {
    // large products: slice planes cut once, then the three slice levels on the
    // direct-to-LDS pipeline
    static SharedWorkspace ws;
    std::lock_guard<std::mutex> lk(ws.mu);
    ws.after_previous((hipStream_t)stream);
    rc = phpc_sgemm_bf16x3_planes(dA, lda, dB, ldb, dC, ldc, m, k, n, (hipStream_t)stream, &ws.ws);
    ws.mark_used((hipStream_t)stream);
  } else if (method == PHPC_SGEMM_BF16X3) {
    const int tiles = phpc_f32_plan_tiles(&d, 1, 2);
    rc = phpc_f32x3_launch(nullptr, &d, 1, tiles, vec4, 0, (hipStream_t)stream);
  } else {
    const long work = (long)((m + 127) / 128) * ((n + 127) / 128);
    const int cfg = work >= 2L * phpc_device().cus ? 2 : 0;
    // the transpose's grid.y is the staged pitch / 32
    if (cfg == 2 && g_f32_at && (m + 3L) / 4 * 4 / 32 < 65535)
      rc = f32_launch_transposed(d, (hipStream_t)stream);
    else {
      const int tiles = phpc_f32_plan_tiles(&d, 1, cfg);
      rc = phpc_f32_launch(nullptr, &d, 1, tiles, cfg, vec4, 0, (hipStream_t)stream);
    }
  }
  return rc ? rc : phpc_set_error(PHPC_OK, nullptr);
}

void phpc_f32_set_transposed_a(int on) { g_f32_at = on ? 1 : 0; }

int phpc_f32_set_plan_config(int id) {
  if (id < -1 || id > kF32AtConfig) return phpc_set_error(PHPC_ERR_ARG, "FP32 config %d out of range", id);
  g_f32_plan_cfg = id;
  return phpc_set_error(PHPC_OK, nullptr);
}

int phpc_f32_pick_plan_config(const F32Desc *descs, int ndesc) {
  if (g_f32_plan_cfg >= 0 && (g_f32_plan_cfg != kF32AtConfig || phpc_f32_at_ok(descs, ndesc))) return g_f32_plan_cfg;
  long work = 0;
  for (int i = 0; i < ndesc; ++i) work += (long)((descs[i].m + 127) / 128) * ((descs[i].n + 127) / 128);
  if (work < 2L * phpc_device().cus) return 0;
  return g_f32_at && phpc_f32_at_ok(descs, ndesc) ? kF32AtConfig : 2;
}

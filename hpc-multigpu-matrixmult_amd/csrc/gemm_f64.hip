// FP64 GEMM on gfx950 matrix cores: C += A*B, grouped and persistent.
//
// Replaces the reference's gemm_kernel (src/phpc_gemm.cu:6-57). The reference gives
// each thread one output and walks the K dimension in T-wide shared-memory phases.
// Here a workgroup owns a BM x BN macro tile, its waves own 16x16 MFMA sub-tiles
// (v_mfma_f64_16x16x4f64), operands are staged through LDS in BK=16 slabs with
// double buffering, and the workgroups form a persistent grid that walks a flat list
// of tiles covering any number of GEMMs (one launch for all SUMMA ranks, or for a
// batch of unrelated GEMMs).
//
// Arithmetic. v_mfma_f64_16x16x4f64 computes, per output, the FMA chain
// d = fma(a3,b3, fma(a2,b2, fma(a1,b1, fma(a0,b0,c)))) in k order (measured on the
// device against 10^6 random cases). With the accumulator started at zero for each
// panel, k visited in ascending order and the panel sum added to C afterwards, every
// output equals the reference kernel's `c = 0; c += a*b (contracted to FMA); C += c`
// (src/phpc_gemm.cu:32,50-51,55). Zero padding of partial slabs adds fma(0,0,c) = c,
// as the reference's own tile padding does (:38-46).
//
// Layout. Row-major, arbitrary leading dimensions. LDS holds A transposed
// (At[k][m]) and B as is (Bs[k][n]) so each MFMA operand fetch is 16 consecutive
// doubles per lane group.
#include "phpc_internal.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

typedef double d4 __attribute__((ext_vector_type(4)));

namespace {

struct GemmBatch {
  const GemmDesc *descs;  // device array (ndesc > 1)
  int ndesc;
  int total_tiles;
  int group_m;            // tile-row swizzle width
  int xcd_remap;          // 1: contiguous tile runs per XCD
  GemmDesc d0;            // inline copy when ndesc == 1
};

typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef double __attribute__((address_space(1))) gdouble;

// Out-of-range offset: raw buffer loads past num_records return zero, which is the
// zero padding of partial tiles (rows >= m, columns >= n, k beyond the panel).
constexpr uint32_t kOOB = 0x80000000u;

__device__ __forceinline__ rsrc_t make_rsrc(const double *base, long bytes) {
  const uint32_t nb = bytes > 0x7fffffffL ? 0x7fffffffu : (uint32_t)(bytes < 0 ? 0 : bytes);
  // 0x00020000: gfx9-family buffer descriptor word 3 (32-bit data format, raw access)
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), (short)0, (int)nb, 0x00020000);
}

template <int VEC>
struct Chunk;
template <>
struct Chunk<1> {
  double x;
};
template <>
struct Chunk<2> {
  double x, y;
};

template <int VEC>
__device__ __forceinline__ Chunk<VEC> buf_load(rsrc_t r, uint32_t off) {
  Chunk<VEC> c;
  if constexpr (VEC == 2) {
    const double2 v = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    c.x = v.x;
    c.y = v.y;
  } else {
    c.x = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  }
  return c;
}

template <int BM, int BN, int WM, int WN, int BK, bool MP, int VEC>
struct F64Gemm {
  static constexpr int NT = WM * WN * 64;
  static constexpr int WTM = BM / WM;
  static constexpr int WTN = BN / WN;
  static constexpr int TM = WTM / 16;
  static constexpr int TN = WTN / 16;
  static constexpr int APITCH = BM + 4;
  static constexpr int BPITCH = BN + 4;
  static constexpr int STAGE = BK * APITCH + BK * BPITCH;  // doubles per LDS stage
  // global->register chunks per thread (VEC doubles per chunk)
  static constexpr int CA = BM * BK / VEC / NT;
  static constexpr int CB = BN * BK / VEC / NT;
  static_assert(CA * VEC * NT == BM * BK, "A slab must split evenly over threads");
  static_assert(CB * VEC * NT == BN * BK, "B slab must split evenly over threads");
  static_assert(TM >= 1 && TN >= 1, "wave tile must hold at least one 16x16 block");
};

// Values that are the same for every lane of the workgroup (tile descriptors), moved to
// scalar registers so that buffer descriptors built from them need no per-lane loop.
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

// Decode global tile id -> (descriptor, tile row, tile col).
__device__ __forceinline__ const GemmDesc *find_desc(const GemmBatch &b, int t) {
  if (b.ndesc == 1) return &b.d0;
  int lo = 0, hi = b.ndesc - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (b.descs[mid].tile_start <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return &b.descs[lo];
}

// Descriptor of tile t with every field moved to scalar registers (the tile id and
// hence the descriptor are the same for the whole workgroup).
__device__ __forceinline__ GemmDesc uniform_desc(const GemmBatch &b, int t) {
  const GemmDesc &d = *find_desc(b, t);
  GemmDesc u;
  u.A = uni(d.A);
  u.B = uni(d.B);
  u.C = uni(d.C);
  u.lda = uni(d.lda);
  u.ldb = uni(d.ldb);
  u.ldc = uni(d.ldc);
  u.m = uni(d.m);
  u.n = uni(d.n);
  u.k = uni(d.k);
  u.panel = uni(d.panel);
  u.tiles_n = uni(d.tiles_n);
  u.tile_start = uni(d.tile_start);
  return u;
}

template <int BM, int BN, int WM, int WN, int BK, bool MP, int VEC>
// two waves per SIMD at least: a 4-wave workgroup then fits twice on a CU
__global__ __launch_bounds__(WM *WN * 64) __attribute__((amdgpu_waves_per_eu(2))) void f64_gemm_kernel(
    GemmBatch batch) {
  using K = F64Gemm<BM, BN, WM, WN, BK, MP, VEC>;
  __shared__ double lds[2 * K::STAGE];

  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int g = lane >> 4;   // k index inside an MFMA (0..3)
  const int li = lane & 15;  // row (A) / column (B) inside a 16-block
  const int wm = wave / WN;
  const int wn = wave % WN;

  const int G = gridDim.x;
  const int bid = blockIdx.x;
  // XCD-aware order: workgroup bid runs on XCD bid % 8; give each XCD a contiguous
  // run of the tiles processed in the same round so neighbouring tiles share its L2.
  const bool xcd_remap = batch.xcd_remap && (G % 8) == 0;
  const int slot = uni(xcd_remap ? (bid % 8) * (G / 8) + bid / 8 : bid);

  for (int round = 0;; ++round) {
    const int t = round * G + slot;
    if (t >= batch.total_tiles) break;
    const GemmDesc d = uniform_desc(batch, uni(t));
    const int lt = t - d.tile_start;
    const int tiles_n = d.tiles_n;
    const int tiles_m = (d.m + BM - 1) / BM;
    // group-M swizzle
    const int gm = batch.group_m;
    const int grp = lt / (gm * tiles_n);
    const int first_m = grp * gm;
    const int gsz = min(tiles_m - first_m, gm);
    const int in_grp = lt - grp * gm * tiles_n;
    const int tm_idx = first_m + in_grp % gsz;
    const int tn_idx = in_grp / gsz;
    const int row0 = tm_idx * BM;
    const int col0 = tn_idx * BN;
    const int m_rem = uni(d.m - row0);
    const int n_rem = uni(d.n - col0);
    const long lda = uni(d.lda), ldb = uni(d.ldb), ldc = uni(d.ldc);
    const double *Ag = uni(d.A) + (long)row0 * lda;
    const double *Bg = uni(d.B) + col0;
    gdouble *Cg = (gdouble *)(uni(d.C) + (long)row0 * ldc + col0);
    const rsrc_t rA = make_rsrc(Ag, (long)min(m_rem, BM) * lda * 8);
    const int Kdim = uni(d.k);
    const int dpanel = uni(d.panel);
    const int w = dpanel > 0 && dpanel < Kdim ? dpanel : Kdim;

    d4 acc[K::TM][K::TN];
    d4 creg[MP ? K::TM : 1][MP ? K::TN : 1];
#pragma unroll
    for (int i = 0; i < K::TM; ++i)
#pragma unroll
      for (int j = 0; j < K::TN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    if constexpr (MP) {
#pragma unroll
      for (int i = 0; i < K::TM; ++i)
#pragma unroll
        for (int j = 0; j < K::TN; ++j) {
          const int cc = wn * K::WTN + j * 16 + li;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int rr = wm * K::WTM + i * 16 + 4 * r + g;
            creg[i][j][r] = (rr < m_rem && cc < n_rem) ? Cg[(long)rr * ldc + cc] : 0.0;
          }
        }
    }

    Chunk<VEC> ra[K::CA];
    Chunk<VEC> rb[K::CB];

    auto load_slab = [&](int kb, int ke) {
      const int krem = ke - kb;
#pragma unroll
      for (int i = 0; i < K::CA; ++i) {
        const int q = tid + i * K::NT;
        const int row = q / (BK / VEC);
        const int kc = (q % (BK / VEC)) * VEC;
        const uint32_t off = (row < m_rem && kc < krem) ? (uint32_t)(((long)row * lda + kb + kc) * 8) : kOOB;
        ra[i] = buf_load<VEC>(rA, off);
      }
      const rsrc_t rB = make_rsrc(Bg + (long)kb * ldb, (long)krem * ldb * 8);
#pragma unroll
      for (int i = 0; i < K::CB; ++i) {
        const int q = tid + i * K::NT;
        const int kr = q / (BN / VEC);
        const int cc = (q % (BN / VEC)) * VEC;
        const uint32_t off = (kr < krem && cc < n_rem) ? (uint32_t)(((long)kr * ldb + cc) * 8) : kOOB;
        rb[i] = buf_load<VEC>(rB, off);
      }
    };

    auto store_slab = [&](int buf) {
      double *At = lds + buf * K::STAGE;
      double *Bs = At + BK * K::APITCH;
#pragma unroll
      for (int i = 0; i < K::CA; ++i) {
        const int q = tid + i * K::NT;
        const int row = q / (BK / VEC);
        const int kc = (q % (BK / VEC)) * VEC;
        At[kc * K::APITCH + row] = ra[i].x;
        if constexpr (VEC == 2) At[(kc + 1) * K::APITCH + row] = ra[i].y;
      }
#pragma unroll
      for (int i = 0; i < K::CB; ++i) {
        const int q = tid + i * K::NT;
        const int kr = q / (BN / VEC);
        const int cc = (q % (BN / VEC)) * VEC;
        if constexpr (VEC == 2) {
          *reinterpret_cast<double2 *>(&Bs[kr * K::BPITCH + cc]) = double2{rb[i].x, rb[i].y};
        } else {
          Bs[kr * K::BPITCH + cc] = rb[i].x;
        }
      }
    };

    auto mfma_step = [&](const double *At, const double *Bs, int kk) {
      double a[K::TM], b[K::TN];
#pragma unroll
      for (int i = 0; i < K::TM; ++i) a[i] = At[(kk * 4 + g) * K::APITCH + wm * K::WTM + i * 16 + li];
#pragma unroll
      for (int j = 0; j < K::TN; ++j) b[j] = Bs[(kk * 4 + g) * K::BPITCH + wn * K::WTN + j * 16 + li];
#pragma unroll
      for (int i = 0; i < K::TM; ++i)
#pragma unroll
        for (int j = 0; j < K::TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    };

    auto compute = [&](int buf, int klen) {
      const double *At = lds + buf * K::STAGE;
      const double *Bs = At + BK * K::APITCH;
      // k beyond the slab was loaded as zero, so a partial slab adds fma(0, 0, c) = c
      // terms only (the reference kernel pads its own T-wide phases the same way)
      (void)klen;
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) mfma_step(At, Bs, kk);
    };

    // All waves must be done with the previous tile's last slab before the prologue
    // overwrites LDS.
    __syncthreads();
    int kb = 0;
    int pe = min(w, Kdim);
    load_slab(kb, min(kb + BK, pe));
    store_slab(0);
    __syncthreads();
    int buf = 0;
    while (true) {
      const int ke = min(kb + BK, pe);
      const bool panel_end = (ke == pe);
      const int nkb = panel_end ? pe : kb + BK;
      const int npe = panel_end ? min(pe + w, Kdim) : pe;
      const bool has_next = nkb < Kdim;
      if (has_next) load_slab(nkb, min(nkb + BK, npe));
      compute(buf, ke - kb);
      if constexpr (MP) {
        if (panel_end) {
#pragma unroll
          for (int i = 0; i < K::TM; ++i)
#pragma unroll
            for (int j = 0; j < K::TN; ++j) {
              creg[i][j] = creg[i][j] + acc[i][j];
              acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
            }
        }
      }
      if (!has_next) break;
      store_slab(buf ^ 1);
      __syncthreads();
      buf ^= 1;
      kb = nkb;
      pe = npe;
    }

    // epilogue: C = C + (panel sum) ; or store the carried C
#pragma unroll
    for (int i = 0; i < K::TM; ++i)
#pragma unroll
      for (int j = 0; j < K::TN; ++j) {
        const int cc = wn * K::WTN + j * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rr = wm * K::WTM + i * 16 + 4 * r + g;
          if (rr < m_rem && cc < n_rem) {
            gdouble *pc = Cg + (long)rr * ldc + cc;
            if constexpr (MP)
              *pc = creg[i][j][r];
            else
              *pc = *pc + acc[i][j][r];
          }
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Direct-to-LDS variant (16-byte aligned operands, even k and n).
//
// Slabs are copied HBM -> LDS by buffer_load_dwordx4 ... lds: each wave instruction
// fills one contiguous 1 KiB LDS block, lane p writing bytes [16p, 16p+16). Which
// global 16-byte chunk each lane fetches is free, so the lane -> address map sets
// the LDS layout. No registers stage the data, so STAGES slabs can be in flight.
//
// LDS layout of one stage, in 16-byte chunks:
//   A: BM rows x 16 doubles. Rows go in pairs of 17 chunks (8 + 8 + 1 pad), row r at
//      chunk (r/2)*17 + (r%2)*8. A 16-lane MFMA operand read (16 rows, one k) then
//      touches 16 distinct 8-byte bank slots, and k steps are plain address offsets.
//   B: 16 rows x BN doubles, row k at chunk k*BN/2, chunk c at slot c ^ ((k&1)<<3), so
//      the two 16-lane groups of a half-wave (rows k, k+1) use opposite bank halves.
// Blocks of a stage are dealt to waves round-robin; pad chunks are fetched with an
// out-of-range offset (zero) and never read.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void lds_void;

// s_waitcnt vmcnt(N) leaving expcnt and lgkmcnt unconstrained (gfx9 encoding: vmcnt
// bits [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8])
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

template <int BM, int BN_, int WM, int WN, int STAGES, bool MP>
struct F64Lds {
  static constexpr int BK = 16;
  static constexpr int BN = BN_;
  static constexpr int NT = WM * WN * 64;
  static constexpr int NW = WM * WN;
  static constexpr int WTM = BM / WM;
  static constexpr int WTN = BN_ / WN;
  static constexpr int TM = WTM / 16;
  static constexpr int TN = WTN / 16;
  static constexpr int A_CHUNKS = (BM / 2) * 17;
  static constexpr int A_BLOCKS = (A_CHUNKS + 63) / 64;
  static constexpr int A_BYTES = A_BLOCKS * 1024;
  static constexpr int B_BLOCKS = BK * BN_ * 8 / 1024;
  static constexpr int B_BYTES = B_BLOCKS * 1024;
  static constexpr int STAGE_BYTES = A_BYTES + B_BYTES;
  static constexpr int BLOCKS = A_BLOCKS + B_BLOCKS;
  static constexpr int PER_LO = BLOCKS / NW;                  // blocks per wave (low)
  static constexpr int PER_HI = (BLOCKS + NW - 1) / NW;       // blocks per wave (high)
  static_assert(BN_ % 32 == 0 && BM % 32 == 0, "tile shape");
  static_assert(STAGES >= 2 && STAGES <= 4, "pipeline depth");
  static_assert(STAGES * STAGE_BYTES <= 160 * 1024, "LDS budget");
};

// wait until at most `n` vector-memory operations of this thread are outstanding,
// n = later * per with per in {PER_LO, PER_HI} and later <= STAGES-2
template <int LO, int HI>
__device__ __forceinline__ void wait_outstanding(int later, int per) {
  const int n = later * per;
  if (n <= 0) { wait_vm_c<0>(); return; }
  if (n == LO) { wait_vm_c<LO>(); return; }
  if (n == HI) { wait_vm_c<HI>(); return; }
  if (n == 2 * LO) { wait_vm_c<2 * LO>(); return; }
  if (n == 2 * HI) { wait_vm_c<2 * HI>(); return; }
  wait_vm_c<0>();
}

template <int BM, int BN, int WM, int WN, int STAGES, bool MP>
__global__ __launch_bounds__(WM *WN * 64) __attribute__((amdgpu_waves_per_eu(2))) void f64_lds_kernel(
    GemmBatch batch) {
  using K = F64Lds<BM, BN, WM, WN, STAGES, MP>;
  extern __shared__ __attribute__((aligned(1024))) char smem[];

  const int tid = threadIdx.x;
  const int wave = uni(tid >> 6);  // wave-uniform: keeps LDS block bases scalar
  const int lane = tid & 63;
  const int g = lane >> 4;
  const int li = lane & 15;
  const int wm = wave / WN;
  const int wn = wave % WN;
  // blocks owned by this wave: wave, wave + NW, ... (uniform per wave)
  const int my_blocks = (K::BLOCKS - wave + K::NW - 1) / K::NW;

  // per-lane LDS read offsets (bytes, stage-relative); k steps add constants
  int a_off[K::TM], b_off[K::TN];
#pragma unroll
  for (int i = 0; i < K::TM; ++i) {
    const int r = wm * K::WTM + i * 16 + li;
    a_off[i] = ((r >> 1) * 17 + (r & 1) * 8) * 16 + g * 8;
  }
#pragma unroll
  for (int j = 0; j < K::TN; ++j) {
    const int col = wn * K::WTN + j * 16 + li;
    b_off[j] = K::A_BYTES + g * K::BN * 8 + ((((col >> 1) ^ ((g & 1) << 3))) << 4) + (col & 1) * 8;
  }

  const int G = gridDim.x;
  const int bid = blockIdx.x;
  const bool xcd_remap = batch.xcd_remap && (G % 8) == 0;
  const int slot = uni(xcd_remap ? (bid % 8) * (G / 8) + bid / 8 : bid);

  for (int round = 0;; ++round) {
    const int t = round * G + slot;
    if (t >= batch.total_tiles) break;
    const GemmDesc d = uniform_desc(batch, uni(t));
    const int lt = t - d.tile_start;
    const int tiles_n = d.tiles_n;
    const int tiles_m = (d.m + BM - 1) / BM;
    const int gm = batch.group_m;
    const int grp = lt / (gm * tiles_n);
    const int first_m = grp * gm;
    const int gsz = min(tiles_m - first_m, gm);
    const int in_grp = lt - grp * gm * tiles_n;
    const int row0 = (first_m + in_grp % gsz) * BM;
    const int col0 = (in_grp / gsz) * BN;
    const int m_rem = uni(d.m - row0);
    const int n_rem = uni(d.n - col0);
    const long lda = uni(d.lda), ldb = uni(d.ldb), ldc = uni(d.ldc);
    const double *Ag = uni(d.A) + (long)row0 * lda;
    const double *Bg = uni(d.B) + col0;
    gdouble *Cg = (gdouble *)(uni(d.C) + (long)row0 * ldc + col0);
    const int Kdim = uni(d.k);
    const int dpanel = uni(d.panel);
    const int w = dpanel > 0 && dpanel < Kdim ? dpanel : Kdim;
    const rsrc_t rA = make_rsrc(Ag, (long)min(m_rem, BM) * lda * 8);

    // Per-lane global offsets of this wave's blocks for a full slab starting at k = 0
    // (A: add kb*8 through soffset; B: the descriptor base moves to row kb). For a
    // partial slab the k bound is applied on top (kcol < krem).
    uint32_t goff[K::PER_HI];
    int gk[K::PER_HI];  // k column (A) or k row (B) of the chunk; -1 for pad chunks
#pragma unroll
    for (int u = 0; u < K::PER_HI; ++u) {
      const int blk = wave + u * K::NW;
      goff[u] = kOOB;
      gk[u] = 1 << 30;
      if (blk < K::A_BLOCKS) {
        const int q = blk * 64 + lane;  // chunk slot in the A region
        const int pr = q / 17, pc = q % 17;
        const int row = 2 * pr + (pc >= 8 ? 1 : 0);
        const int c = pc >= 8 ? pc - 8 : pc;
        if (pc < 16 && q < K::A_CHUNKS && row < m_rem) {
          goff[u] = (uint32_t)(((long)row * lda + 2 * c) * 8);
          gk[u] = 2 * c;
        }
      } else if (blk < K::BLOCKS) {
        const int q = (blk - K::A_BLOCKS) * 64 + lane;
        const int kr = q / (BN / 2);
        const int c = (q % (BN / 2)) ^ ((kr & 1) << 3);
        if (2 * c < n_rem) {
          goff[u] = (uint32_t)(((long)kr * ldb + 2 * c) * 8);
          gk[u] = kr;
        }
      }
    }

    d4 acc[K::TM][K::TN];
    d4 creg[MP ? K::TM : 1][MP ? K::TN : 1];
#pragma unroll
    for (int i = 0; i < K::TM; ++i)
#pragma unroll
      for (int j = 0; j < K::TN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    if constexpr (MP) {
#pragma unroll
      for (int i = 0; i < K::TM; ++i)
#pragma unroll
        for (int j = 0; j < K::TN; ++j) {
          const int cc = wn * K::WTN + j * 16 + li;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int rr = wm * K::WTM + i * 16 + 4 * r + g;
            creg[i][j][r] = (rr < m_rem && cc < n_rem) ? Cg[(long)rr * ldc + cc] : 0.0;
          }
        }
    }

    const int spp = (w + K::BK - 1) / K::BK;  // slabs per full panel
    const int full_panels = Kdim / w;
    const int tail = Kdim - full_panels * w;
    const int nslab = full_panels * spp + (tail + K::BK - 1) / K::BK;

    auto issue = [&](int sidx) {
      const int p = sidx / spp;
      const int kb = p * w + (sidx - p * spp) * K::BK;
      const int krem = min(min(kb + K::BK, p * w + w), Kdim) - kb;
      char *stage = smem + (sidx % STAGES) * K::STAGE_BYTES;
      const rsrc_t rB = make_rsrc(Bg + (long)kb * ldb, (long)krem * ldb * 8);
#pragma unroll
      for (int u = 0; u < K::PER_HI; ++u) {
        const int blk = wave + u * K::NW;
        if (u < K::PER_LO || blk < K::BLOCKS) {
          const uint32_t off = gk[u] < krem ? goff[u] : kOOB;
          if (blk < K::A_BLOCKS)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rA, (lds_void *)(stage + blk * 1024), 16, off, kb * 8, 0, 0);
          else
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rB, (lds_void *)(stage + blk * 1024), 16, off, 0, 0, 0);
        }
      }
    };

    auto mfma_step = [&](const char *stage, int kk) {
      double a[K::TM], b[K::TN];
#pragma unroll
      for (int i = 0; i < K::TM; ++i) a[i] = *reinterpret_cast<const double *>(stage + a_off[i] + kk * 32);
#pragma unroll
      for (int j = 0; j < K::TN; ++j)
        b[j] = *reinterpret_cast<const double *>(stage + b_off[j] + kk * 4 * K::BN * 8);
#pragma unroll
      for (int i = 0; i < K::TM; ++i)
#pragma unroll
        for (int j = 0; j < K::TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    };

    // every wave must be done reading the previous tile's stages before refilling them
    slab_barrier();
#pragma unroll
    for (int s0 = 0; s0 < STAGES - 1; ++s0)
      if (s0 < nslab) issue(s0);

    for (int sidx = 0; sidx < nslab; ++sidx) {
      // loads of slab sidx are complete once at most the later slabs' are in flight
      wait_outstanding<K::PER_LO, K::PER_HI>(min(STAGES - 2, nslab - 1 - sidx), my_blocks);
      slab_barrier();
      if (sidx + STAGES - 1 < nslab) issue(sidx + STAGES - 1);
      const int p = sidx / spp;
      const int j = sidx - p * spp;
      const int kb = p * w + j * K::BK;
      const int pe = min(p * w + w, Kdim);
      const int klen = min(kb + K::BK, pe) - kb;
      const char *stage = smem + (sidx % STAGES) * K::STAGE_BYTES;
      // k beyond the slab reads as zero (out-of-range loads), so a partial slab simply
      // adds fma(0, 0, c) = c terms, like the reference kernel's own tile padding
#pragma unroll
      for (int kk = 0; kk < K::BK / 4; ++kk) mfma_step(stage, kk);
      if constexpr (MP) {
        if (kb + klen == pe) {
#pragma unroll
          for (int i = 0; i < K::TM; ++i)
#pragma unroll
            for (int jj = 0; jj < K::TN; ++jj) {
              creg[i][jj] = creg[i][jj] + acc[i][jj];
              acc[i][jj] = d4{0.0, 0.0, 0.0, 0.0};
            }
        }
      }
    }

#pragma unroll
    for (int i = 0; i < K::TM; ++i)
#pragma unroll
      for (int jj = 0; jj < K::TN; ++jj) {
        const int cc = wn * K::WTN + jj * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rr = wm * K::WTM + i * 16 + 4 * r + g;
          if (rr < m_rem && cc < n_rem) {
            gdouble *pc = Cg + (long)rr * ldc + cc;
            if constexpr (MP)
              *pc = creg[i][jj][r];
            else
              *pc = *pc + acc[i][jj][r];
          }
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Transposed-A direct-to-LDS variant. The caller stores A transposed (At[k][m], row
// pitch = the descriptor's lda, 16-byte aligned rows, pad columns zero), so both
// operand slabs are k-row-major and are copied HBM -> LDS by 16-byte buffer loads
// with no register staging and no LDS write instructions: each wave instruction fills
// one contiguous 1 KiB piece of a k row. LDS rows are padded by 64 bytes. The MFMA
// operand reads are the register-staged kernel's (16 consecutive doubles per lane
// group), so the arithmetic is identical. STAGES slabs of BK k rows ring through LDS:
// slab s + STAGES - 1 is issued while slab s is computed (two stages: a slab's loads land
// while the previous slab is computed). MP: k is walked in panels, each panel's sum starts
// from zero and is added to the C value carried in registers (as the register-staged MP
// kernels).
// ---------------------------------------------------------------------------
template <int BM, int BN, int WM, int WN, int STAGES_ = 2, int BK_ = 16>
struct F64At {
  static constexpr int BK = BK_;
  static constexpr int STAGES = STAGES_;
  static constexpr int NW = WM * WN;
  static constexpr int WTM = BM / WM, WTN = BN / WN;
  static constexpr int TM = WTM / 16, TN = WTN / 16;
  static constexpr int AP = BM + 8, BP = BN + 8;  // LDS row pitches (doubles)
  static constexpr int A_BYTES = BK * AP * 8;
  static constexpr int STAGE_BYTES = A_BYTES + BK * BP * 8;
  static constexpr int A_PIECES = BM / 128, B_PIECES = BN / 128;  // 1 KiB pieces per k row
  static constexpr int A_BLOCKS = BK * A_PIECES, BLOCKS = A_BLOCKS + BK * B_PIECES;
  static constexpr int PER = BLOCKS / NW;
  static_assert(BM % 128 == 0 && BN % 128 == 0 && BLOCKS % NW == 0, "tile shape");
  static_assert(STAGES * STAGE_BYTES <= 160 * 1024, "LDS budget");
  static_assert(STAGES >= 2 && (STAGES - 2) * PER < 64 && BK % 4 == 0, "pipeline depth");
};

template <int BM, int BN, int WM, int WN, bool MP, int STAGES = 2, int BK = 16>
__device__ __forceinline__ void f64_at_body(const GemmBatch &batch) {
  using K = F64At<BM, BN, WM, WN, STAGES, BK>;
  extern __shared__ __attribute__((aligned(1024))) char smem[];
  const int tid = threadIdx.x;
  const int wave = uni(tid >> 6);
  const int lane = tid & 63;
  const int g = lane >> 4;
  const int li = lane & 15;
  const int wm = wave / WN;
  const int wn = wave % WN;

  const int G = gridDim.x;
  const int bid = blockIdx.x;
  const bool xcd_remap = batch.xcd_remap && (G % 8) == 0;
  const int slot = uni(xcd_remap ? (bid % 8) * (G / 8) + bid / 8 : bid);

  for (int round = 0;; ++round) {
    const int t = round * G + slot;
    if (t >= batch.total_tiles) break;
    const GemmDesc d = uniform_desc(batch, uni(t));
    const int lt = t - d.tile_start;
    const int tiles_n = d.tiles_n;
    const int tiles_m = (d.m + BM - 1) / BM;
    const int gm = batch.group_m;
    const int grp = lt / (gm * tiles_n);
    const int first_m = grp * gm;
    const int gsz = min(tiles_m - first_m, gm);
    const int in_grp = lt - grp * gm * tiles_n;
    const int row0 = uni((first_m + in_grp % gsz) * BM);
    const int col0 = uni((in_grp / gsz) * BN);
    const int m_rem = uni(d.m - row0);
    const int n_rem = uni(d.n - col0);
    // single-panel kernels skip the MFMAs of waves with no output rows or columns; the
    // panel kernels run them (skipping measured slower there: it costs the carried-C loop
    // more than the few edge tiles give back)
    const bool wave_live = MP || (wm * K::WTM < m_rem && wn * K::WTN < n_rem);
    const long ldat = uni(d.lda), ldb = uni(d.ldb), ldc = uni(d.ldc);
    const double *At_g = uni(d.A);
    const double *Bg = uni(d.B) + col0;
    gdouble *Cg = (gdouble *)(uni(d.C) + (long)row0 * ldc + col0);
    const int Kdim = uni(d.k);
    const int dpanel = uni(d.panel);
    const int w = MP && dpanel > 0 && dpanel < Kdim ? dpanel : Kdim;
    const int spp = (w + K::BK - 1) / K::BK;  // slabs per full panel
    const int full_panels = Kdim / w;
    const int tail = Kdim - full_panels * w;
    const int nslab = full_panels * spp + (tail + K::BK - 1) / K::BK;

    // lane offsets of this wave's pieces within a slab (k row kr, columns c, c+1);
    // columns past the tile's edge read zero
    uint32_t goff[K::PER];
    int loff[K::PER];
#pragma unroll
    for (int u = 0; u < K::PER; ++u) {
      const int blk = wave + u * K::NW;
      if (blk < K::A_BLOCKS) {
        const int kr = blk / K::A_PIECES;
        const int c = (blk % K::A_PIECES) * 128 + 2 * lane;
        goff[u] = c < m_rem ? (uint32_t)(((long)kr * ldat + row0 + c) * 8) : kOOB;
        loff[u] = kr * K::AP * 8 + (blk % K::A_PIECES) * 1024;
      } else {
        const int q = blk - K::A_BLOCKS;
        const int kr = q / K::B_PIECES;
        const int c = (q % K::B_PIECES) * 128 + 2 * lane;
        goff[u] = c < n_rem ? (uint32_t)(((long)kr * ldb + c) * 8) : kOOB;
        loff[u] = K::A_BYTES + kr * K::BP * 8 + (q % K::B_PIECES) * 1024;
      }
    }

    // slab bases advance by BK rows; a slab's byte range is BK rows (fewer for the last
    // slab, whose missing k rows then read zero). The launcher keeps BK rows < 2 GiB.
    const int full_a = (int)(K::BK * ldat * 8), full_b = (int)(K::BK * ldb * 8);
    // slab sidx: k rows [kb, kb + krem) of panel p (slabs never cross a panel's end)
    auto slab_rows = [&](int sidx, int &kb, int &krem, int &pe) {
      const int p = sidx / spp;
      kb = p * w + (sidx - p * spp) * K::BK;
      pe = min(p * w + w, Kdim);
      krem = min(K::BK, pe - kb);
    };
    auto issue = [&](int sidx) {
      int kb, krem, pe;
      slab_rows(sidx, kb, krem, pe);
      char *stage = smem + (sidx % K::STAGES) * K::STAGE_BYTES;
      const int na = krem == K::BK ? full_a : krem * (int)ldat * 8;
      const int nb = krem == K::BK ? full_b : krem * (int)ldb * 8;
      const rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(At_g + (long)kb * ldat), (short)0, na,
                                                          0x00020000);
      const rsrc_t rB =
          __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(Bg + (long)kb * ldb), (short)0, nb, 0x00020000);
#pragma unroll
      for (int u = 0; u < K::PER; ++u) {
        const int blk = wave + u * K::NW;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(blk < K::A_BLOCKS ? rA : rB, (lds_void *)(stage + loff[u]), 16, goff[u],
                                                  0, 0, 0);
      }
    };

    d4 acc[K::TM][K::TN];
    d4 creg[MP ? K::TM : 1][MP ? K::TN : 1];
#pragma unroll
    for (int i = 0; i < K::TM; ++i)
#pragma unroll
      for (int j = 0; j < K::TN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    if constexpr (MP) {
#pragma unroll
      for (int i = 0; i < K::TM; ++i)
#pragma unroll
        for (int j = 0; j < K::TN; ++j) {
          const int cc = wn * K::WTN + j * 16 + li;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int rr = wm * K::WTM + i * 16 + 4 * r + g;
            creg[i][j][r] = (rr < m_rem && cc < n_rem) ? Cg[(long)rr * ldc + cc] : 0.0;
          }
        }
    }

    // every wave must be done reading the previous tile's stages before refilling them;
    // with more than two stages nothing older than the ring's loads may be outstanding,
    // so that counting them (in issue order) tells which slabs have landed
    if constexpr (K::STAGES > 2) wait_vm_c<0>();
    slab_barrier();
#pragma unroll
    for (int s0 = 0; s0 < K::STAGES - 1; ++s0)
      if (s0 < nslab) issue(s0);
    for (int sidx = 0; sidx < nslab; ++sidx) {
      // slabs issued after this one that may stay in flight
      const int later = min(K::STAGES - 2, nslab - 1 - sidx);
      if constexpr (K::STAGES == 2) {
        wait_vm_c<0>();
      } else {
        if (later <= 0)
          wait_vm_c<0>();
        else if (later == 1)
          wait_vm_c<K::PER>();
        else
          wait_vm_c<(K::STAGES - 2) * K::PER>();
      }
      slab_barrier();
      const char *stage = smem + (sidx % K::STAGES) * K::STAGE_BYTES;
      const double *At = reinterpret_cast<const double *>(stage);
      const double *Bs = reinterpret_cast<const double *>(stage + K::A_BYTES);
      // k beyond the last slab's rows reads as zero: fma(0, 0, c) = c
      auto kstep = [&](int kk) {
        double a[K::TM], b[K::TN];
#pragma unroll
        for (int i = 0; i < K::TM; ++i) a[i] = At[(kk * 4 + g) * K::AP + wm * K::WTM + i * 16 + li];
#pragma unroll
        for (int j = 0; j < K::TN; ++j) b[j] = Bs[(kk * 4 + g) * K::BP + wn * K::WTN + j * 16 + li];
#pragma unroll
        for (int i = 0; i < K::TM; ++i)
#pragma unroll
          for (int j = 0; j < K::TN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
      };
      // the first k step's LDS reads go out right after the barrier; the next slab's
      // loads are issued behind its MFMAs. A wave whose rows or columns all lie past the
      // tile's edge has no output: it only loads its share of the slabs and keeps the
      // barriers.
      if (wave_live) kstep(0);
      if (sidx + K::STAGES - 1 < nslab) issue(sidx + K::STAGES - 1);
      if (wave_live) {
#pragma unroll
        for (int kk = 1; kk < K::BK / 4; ++kk) kstep(kk);
      }
      if constexpr (MP) {
        int kb, krem, pe;
        slab_rows(sidx, kb, krem, pe);
        if (kb + krem == pe) {
#pragma unroll
          for (int i = 0; i < K::TM; ++i)
#pragma unroll
            for (int j = 0; j < K::TN; ++j) {
              creg[i][j] = creg[i][j] + acc[i][j];
              acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
            }
        }
      }
    }

#pragma unroll
    for (int i = 0; i < K::TM; ++i)
#pragma unroll
      for (int jj = 0; jj < K::TN; ++jj) {
        const int cc = wn * K::WTN + jj * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rr = wm * K::WTM + i * 16 + 4 * r + g;
          if (rr < m_rem && cc < n_rem) {
            gdouble *pc = Cg + (long)rr * ldc + cc;
            if constexpr (MP)
              *pc = creg[i][jj][r];
            else
              *pc = *pc + acc[i][jj][r];
          }
        }
      }
  }
}

__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void f64_at_256x128(GemmBatch batch) {
  f64_at_body<256, 128, 4, 2, false>(batch);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void f64_at_128x128(GemmBatch batch) {
  f64_at_body<128, 128, 2, 2, false>(batch);
}
// 8 waves with 64x32 wave tiles, so the carried C fits beside the accumulators
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void f64_at_128x128_w8(GemmBatch batch) {
  f64_at_body<128, 128, 2, 4, false>(batch);
}
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void f64_at_128x128_w8_mp(GemmBatch batch) {
  f64_at_body<128, 128, 2, 4, true>(batch);
}
// One workgroup per CU leaves LDS for a deeper ring: three or four 16-deep stages, or two
// 32-deep ones (half the barriers per k)
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void f64_at_128x128_w8_s3(GemmBatch batch) {
  f64_at_body<128, 128, 2, 4, false, 3>(batch);
}
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void f64_at_128x128_w8_mp_s3(
    GemmBatch batch) {
  f64_at_body<128, 128, 2, 4, true, 3>(batch);
}
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void f64_at_128x128_w8_s4(GemmBatch batch) {
  f64_at_body<128, 128, 2, 4, false, 4>(batch);
}
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void f64_at_128x128_w8_mp_s4(
    GemmBatch batch) {
  f64_at_body<128, 128, 2, 4, true, 4>(batch);
}
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void f64_at_128x128_w8_k32(GemmBatch batch) {
  f64_at_body<128, 128, 2, 4, false, 2, 32>(batch);
}
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void f64_at_128x128_w8_mp_k32(
    GemmBatch batch) {
  f64_at_body<128, 128, 2, 4, true, 2, 32>(batch);
}
// 16 waves with 32x32 wave tiles: the carried C and the accumulators fit 4 waves per SIMD,
// so each barrier has four waves per SIMD to hide behind (LDS reads: one per MFMA)
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4))) void f64_at_128x128_w16_k32(
    GemmBatch batch) {
  f64_at_body<128, 128, 4, 4, false, 2, 32>(batch);
}
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4))) void f64_at_128x128_w16_mp_k32(
    GemmBatch batch) {
  f64_at_body<128, 128, 4, 4, true, 2, 32>(batch);
}

// At[k][m] = A[m][k] for m < M, k < K; At[k][m] = 0 for M <= m < ldat (row pitch ldat)
__global__ __launch_bounds__(256) void f64_transpose_kernel(const double *__restrict__ A, long lda, int M, int Kd,
                                                            double *__restrict__ At, long ldat) {
  __shared__ double tile[32][33];
  const int k0 = blockIdx.x * 32, m0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  double v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int m = m0 + ty + 8 * u, k = k0 + tx;
    v[u] = load_or_zero(A, (long)m * lda + k, m < M && k < Kd);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) tile[ty + 8 * u][tx] = v[u];
  __syncthreads();
#pragma unroll
  for (int r = ty; r < 32; r += 8) {
    const int k = k0 + r, m = m0 + tx;
    if (k < Kd && m < ldat) At[(long)k * ldat + m] = tile[tx][r];
  }
}

// Bc[k][j] = B[k][j] for j < N, 0 for N <= j < ldbc (row pitch ldbc)
__global__ __launch_bounds__(256) void f64_copy_pad_kernel(const double *__restrict__ B, long ldb, int Kd, int N,
                                                           double *__restrict__ Bc, long ldbc) {
  const int k = blockIdx.y;
  for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < ldbc; j += (long)gridDim.x * blockDim.x)
    Bc[(long)k * ldbc + j] = j < N ? B[(long)k * ldb + j] : 0.0;
}

typedef void (*KernelFn)(GemmBatch);

// One tile configuration: kernels for [panel grouping][16-byte loads]. Direct-to-LDS
// configs use f64_lds_kernel when 16-byte loads are legal and fall back to the
// register-staged kernel with the same macro tile otherwise (same tiles, same
// arithmetic, so the planner's tile list is valid for both).
struct ConfigEntry {
  int bm, bn, threads, bk;
  size_t lds_bytes[2];  // dynamic LDS for [vec1][vec2] kernels
  bool panels;          // panel-grouped variants fit without spills
  KernelFn fn[2][2];
  bool trans_a = false;  // the 16-byte kernel reads A transposed (descriptor A = At, lda = its pitch)
};

#define PHPC_V1(BM_, BN_, WM_, WN_, BK_, MP_, VEC_) f64_gemm_kernel<BM_, BN_, WM_, WN_, BK_, MP_, VEC_>
#define PHPC_LDS(BM_, BN_, WM_, WN_, ST_, MP_) f64_lds_kernel<BM_, BN_, WM_, WN_, ST_, MP_>
#define LDS_BYTES(BM_, BN_, ST_) ((size_t)(ST_) * (size_t)F64Lds<BM_, BN_, 2, 2, ST_, false>::STAGE_BYTES)

const ConfigEntry kConfigs[] = {
    // 0: register-staged, 8 waves (wave tile 64x32), 16-deep slabs
    {128, 128, 512, 16, {0, 0}, true,
     {{PHPC_V1(128, 128, 2, 4, 16, false, 1), PHPC_V1(128, 128, 2, 4, 16, false, 2)},
      {PHPC_V1(128, 128, 2, 4, 16, true, 1), PHPC_V1(128, 128, 2, 4, 16, true, 2)}}},
    // 1: register-staged, 8 waves, 32-deep slabs
    {128, 128, 512, 32, {0, 0}, true,
     {{PHPC_V1(128, 128, 2, 4, 32, false, 1), PHPC_V1(128, 128, 2, 4, 32, false, 2)},
      {PHPC_V1(128, 128, 2, 4, 32, true, 1), PHPC_V1(128, 128, 2, 4, 32, true, 2)}}},
    // 2: register-staged, 4 waves (wave tile 32x32), 64x64 tiles for small problems
    {64, 64, 256, 16, {0, 0}, true,
     {{PHPC_V1(64, 64, 2, 2, 16, false, 1), PHPC_V1(64, 64, 2, 2, 16, false, 2)},
      {PHPC_V1(64, 64, 2, 2, 16, true, 1), PHPC_V1(64, 64, 2, 2, 16, true, 2)}}},
    // 3: direct-to-LDS, 8 waves (wave tile 64x32), 4-stage pipeline
    {128, 128, 512, 16, {0, LDS_BYTES(128, 128, 4)}, true,
     {{PHPC_V1(128, 128, 2, 4, 16, false, 1), PHPC_LDS(128, 128, 2, 4, 4, false)},
      {PHPC_V1(128, 128, 2, 4, 16, true, 1), PHPC_LDS(128, 128, 2, 4, 4, true)}}},
    // 4: direct-to-LDS, 4 waves (wave tile 64x64), 4-stage pipeline; single panel only
    {128, 128, 256, 16, {0, LDS_BYTES(128, 128, 4)}, false,
     {{PHPC_V1(128, 128, 2, 2, 16, false, 1), PHPC_LDS(128, 128, 2, 2, 4, false)},
      {PHPC_V1(128, 128, 2, 2, 16, false, 1), PHPC_LDS(128, 128, 2, 2, 4, false)}}},
    // 5: direct-to-LDS, 4 waves (wave tile 32x32), 64x64 tiles, 4-stage pipeline
    {64, 64, 256, 16, {0, LDS_BYTES(64, 64, 4)}, true,
     {{PHPC_V1(64, 64, 2, 2, 16, false, 1), PHPC_LDS(64, 64, 2, 2, 4, false)},
      {PHPC_V1(64, 64, 2, 2, 16, true, 1), PHPC_LDS(64, 64, 2, 2, 4, true)}}},
    // 6: register-staged, 8 waves (wave tile 64x64), 256x128 tiles; single panel only
    {256, 128, 512, 16, {0, 0}, false,
     {{PHPC_V1(256, 128, 4, 2, 16, false, 1), PHPC_V1(256, 128, 4, 2, 16, false, 2)},
      {PHPC_V1(256, 128, 4, 2, 16, false, 1), PHPC_V1(256, 128, 4, 2, 16, false, 2)}}},
    // 7: direct-to-LDS, 8 waves (wave tile 64x64), 256x128 tiles, 3 stages; single panel only
    {256, 128, 512, 16, {0, LDS_BYTES(256, 128, 3)}, false,
     {{PHPC_V1(256, 128, 4, 2, 16, false, 1), PHPC_LDS(256, 128, 4, 2, 3, false)},
      {PHPC_V1(256, 128, 4, 2, 16, false, 1), PHPC_LDS(256, 128, 4, 2, 3, false)}}},
    // 8: register-staged, 4 waves (wave tile 64x64), 128x128 tiles: 68 KiB of LDS, two
    //    workgroups per CU so one's barrier stalls are covered by the other; single panel only
    {128, 128, 256, 16, {0, 0}, false,
     {{PHPC_V1(128, 128, 2, 2, 16, false, 1), PHPC_V1(128, 128, 2, 2, 16, false, 2)},
      {PHPC_V1(128, 128, 2, 2, 16, false, 1), PHPC_V1(128, 128, 2, 2, 16, false, 2)}}},
    // 9: direct-to-LDS, 4 waves (wave tile 64x64), 128x128 tiles, 2 stages (66 KiB, two
    //    workgroups per CU); single panel only
    {128, 128, 256, 16, {0, LDS_BYTES(128, 128, 2)}, false,
     {{PHPC_V1(128, 128, 2, 2, 16, false, 1), PHPC_LDS(128, 128, 2, 2, 2, false)},
      {PHPC_V1(128, 128, 2, 2, 16, false, 1), PHPC_LDS(128, 128, 2, 2, 2, false)}}},
    // 10: transposed-A direct-to-LDS, 8 waves (wave tile 64x64), 256x128 tiles, 2 stages;
    //     single panel only. The 8-byte slots are config 6's kernels: when B rules out
    //     16-byte loads the launcher runs config 6 on the untransposed A instead.
    {256, 128, 512, 16, {0, 2 * (size_t)F64At<256, 128, 4, 2>::STAGE_BYTES}, false,
     {{PHPC_V1(256, 128, 4, 2, 16, false, 1), f64_at_256x128},
      {PHPC_V1(256, 128, 4, 2, 16, false, 1), f64_at_256x128}},
     true},
    // 11: transposed-A direct-to-LDS, 4 waves (wave tile 64x64), 128x128 tiles, 2 stages
    //     (68 KiB, two workgroups per CU); single panel only. 8-byte slots: config 8's kernel.
    {128, 128, 256, 16, {0, 2 * (size_t)F64At<128, 128, 2, 2>::STAGE_BYTES}, false,
     {{PHPC_V1(128, 128, 2, 2, 16, false, 1), f64_at_128x128},
      {PHPC_V1(128, 128, 2, 2, 16, false, 1), f64_at_128x128}},
     true},
    // 12: transposed-A direct-to-LDS, 8 waves (wave tile 64x32), 128x128 tiles, 2 stages,
    //     panel-grouped too (C carried in registers). 8-byte slots: config 0's kernels.
    {128, 128, 512, 16, {0, 2 * (size_t)F64At<128, 128, 2, 4>::STAGE_BYTES}, true,
     {{PHPC_V1(128, 128, 2, 4, 16, false, 1), f64_at_128x128_w8},
      {PHPC_V1(128, 128, 2, 4, 16, true, 1), f64_at_128x128_w8_mp}},
     true},
    // 13: as 12 with three 16-deep stages (102 KiB)
    {128, 128, 512, 16, {0, 3 * (size_t)F64At<128, 128, 2, 4, 3>::STAGE_BYTES}, true,
     {{PHPC_V1(128, 128, 2, 4, 16, false, 1), f64_at_128x128_w8_s3},
      {PHPC_V1(128, 128, 2, 4, 16, true, 1), f64_at_128x128_w8_mp_s3}},
     true},
    // 14: as 12 with four 16-deep stages (136 KiB)
    {128, 128, 512, 16, {0, 4 * (size_t)F64At<128, 128, 2, 4, 4>::STAGE_BYTES}, true,
     {{PHPC_V1(128, 128, 2, 4, 16, false, 1), f64_at_128x128_w8_s4},
      {PHPC_V1(128, 128, 2, 4, 16, true, 1), f64_at_128x128_w8_mp_s4}},
     true},
    // 15: as 12 with two 32-deep stages (136 KiB)
    {128, 128, 512, 32, {0, 2 * (size_t)F64At<128, 128, 2, 4, 2, 32>::STAGE_BYTES}, true,
     {{PHPC_V1(128, 128, 2, 4, 16, false, 1), f64_at_128x128_w8_k32},
      {PHPC_V1(128, 128, 2, 4, 16, true, 1), f64_at_128x128_w8_mp_k32}},
     true},
    // 16: as 15 with 16 waves (wave tile 32x32). There is no 1024-thread register-staged
    //     kernel for the 8-byte slots; they are never launched (a transposed-A launch
    //     needs 16-byte B rows, and phpc_f64_grouped switches to config 0 when A cannot be
    //     staged), so they hold the same kernels, keeping every slot's shape 1024 threads.
    {128, 128, 1024, 32, {2 * (size_t)F64At<128, 128, 4, 4, 2, 32>::STAGE_BYTES,
                          2 * (size_t)F64At<128, 128, 4, 4, 2, 32>::STAGE_BYTES}, true,
     {{f64_at_128x128_w16_k32, f64_at_128x128_w16_k32},
      {f64_at_128x128_w16_mp_k32, f64_at_128x128_w16_mp_k32}},
     true},
};
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);

std::mutex g_mu;
bool g_occ_ready = false;
int g_occ[kNumConfigs][2][2];
GemmDesc *g_desc_dev = nullptr;
size_t g_desc_cap = 0;
GemmDesc *g_desc_pinned = nullptr;
size_t g_pinned_cap = 0;
std::mutex g_desc_mu;  // guards the descriptor buffers below (never held with g_mu)
hipEvent_t g_desc_done = nullptr;  // recorded after the last launch that read g_desc_dev

// Threads per workgroup of a configuration's kernels. Single-panel configurations fill
// their panel-grouped slots with the single-panel kernels, which are never launched with
// panels (phpc_f64_grouped_dev rejects that), so every slot has the configuration's shape.
int kernel_threads(int cfg, int mp, int vec2) {
  (void)mp;
  (void)vec2;
  return kConfigs[cfg].threads;
}

void init_kernels_locked() {
  if (g_occ_ready) return;
  for (int c = 0; c < kNumConfigs; ++c)
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        const void *fn = reinterpret_cast<const void *>(kConfigs[c].fn[a][b]);
        const size_t lds = kConfigs[c].lds_bytes[b];
        if (lds > 0) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kernel_threads(c, a, b), lds) != hipSuccess || n < 1)
          n = 1;
        g_occ[c][a][b] = n;
      }
  g_occ_ready = true;
}

int occupancy(int cfg, int mp, int vec2) {
  std::lock_guard<std::mutex> lk(g_mu);
  init_kernels_locked();
  return g_occ[cfg][mp][vec2];
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace

int phpc_f64_num_configs() { return kNumConfigs; }

F64Config phpc_f64_config(int id) {
  F64Config c{};
  if (id < 0 || id >= kNumConfigs) return c;
  c.id = id;
  c.bm = kConfigs[id].bm;
  c.bn = kConfigs[id].bn;
  c.threads = kConfigs[id].threads;
  c.wg_per_cu = occupancy(id, 0, 1);
  return c;
}

// Planner. Single-panel work on enough 256x128 tiles to cover the CUs takes the
// 256x128 tile (16 independent MFMA chains per wave); panel-grouped work takes the
// 128x128 tile, whose waves also carry C in registers; small problems take 64x64.
int g_big_config = 0;       // large, panel-grouped
int g_big_single = 6;       // large, one panel
int g_group_m = 8;
int g_xcd_remap = 1;

int phpc_f64_pick_config(const GemmDesc *descs, int ndesc) {
  long tiles128 = 0, tiles256 = 0;
  bool mp = false;
  for (int i = 0; i < ndesc; ++i) {
    tiles128 += (long)((descs[i].m + 127) / 128) * ((descs[i].n + 127) / 128);
    tiles256 += (long)((descs[i].m + 255) / 256) * ((descs[i].n + 127) / 128);
    if (descs[i].panel > 0 && descs[i].panel < descs[i].k) mp = true;
  }
  const int cus = phpc_device().cus;
  if (!mp && tiles256 >= (long)cus) return g_big_single;
  if (tiles128 >= (long)cus) return g_big_config;
  return 5;
}

namespace {

// Configuration for large single-panel work reached through host descriptors
// (phpc_f64_grouped), which can stage a transposed copy of A first.
int g_big_single_host = 11;
int g_big_config_host = 16;  // large panel-grouped work through host descriptors

bool b_vec2_ok(const GemmDesc &d) { return aligned16(d.B) && (d.ldb % 2 == 0) && (d.n % 2 == 0); }
// the transposed-A kernels load B in 16-byte column pairs; with an even ldb and an aligned
// base, the pair holding an odd n's last column ends inside the row, and the extra column
// only feeds an output column that is never stored
bool b_at_ok(const GemmDesc &d) { return aligned16(d.B) && (d.ldb % 2 == 0); }

int launch_grouped(const GemmDesc *descs_dev, const GemmDesc *descs_host, int ndesc, int grid, int config_id,
                   hipStream_t stream, bool a_transposed) {
  if (ndesc <= 0) return PHPC_OK;
  if (config_id < 0 || config_id >= kNumConfigs)
    return phpc_set_error(PHPC_ERR_ARG, "bad FP64 config id %d", config_id);
  const ConfigEntry &ce = kConfigs[config_id];
  if (ce.trans_a != a_transposed)
    return phpc_set_error(PHPC_ERR_ARG, "FP64 config %d %s a transposed A", config_id,
                          ce.trans_a ? "needs" : "does not take");
  bool mp = false, vec2 = true;
  int total = 0;
  for (int i = 0; i < ndesc; ++i) {
    const GemmDesc &d = descs_host[i];
    if (d.panel > 0 && d.panel < d.k) mp = true;
    // 32-bit buffer offsets: one A tile (bm rows) and one B slab (bk rows) must span < 2 GiB
    if ((long)ce.bm * d.lda * 8 >= 0x7fffffffL || (long)ce.bk * d.ldb * 8 >= 0x7fffffffL)
      return phpc_set_error(PHPC_ERR_ARG, "leading dimension too large (lda=%ld, ldb=%ld)", d.lda, d.ldb);
    // 16-byte loads need 16-byte aligned rows and even k / n boundaries (a transposed A
    // is stored that way by construction)
    const bool a_ok = a_transposed || (aligned16(d.A) && (d.lda % 2 == 0) && (d.k % 2 == 0) &&
                                       (d.panel <= 0 || d.panel >= d.k || d.panel % 2 == 0));
    if (!a_ok || !(a_transposed ? b_at_ok(d) : b_vec2_ok(d))) vec2 = false;
    total = d.tile_start + ((d.m + ce.bm - 1) / ce.bm) * d.tiles_n;
  }
  if (a_transposed && !vec2) return phpc_set_error(PHPC_ERR_ARG, "transposed-A kernels need 16-byte B rows");
  if (mp && !ce.panels)
    return phpc_set_error(PHPC_ERR_ARG, "FP64 config %d does not support panel grouping", config_id);
  if (total <= 0) return PHPC_OK;
  GemmBatch b{};
  b.descs = descs_dev;
  b.ndesc = ndesc;
  b.total_tiles = total;
  b.group_m = g_group_m;
  b.xcd_remap = g_xcd_remap;
  b.d0 = descs_host[0];
  int G = grid;
  if (G <= 0) {
    G = occupancy(config_id, mp, vec2) * phpc_device().cus;
    if (G > total) G = total;
  }
  if (G < 1) G = 1;
  const int kt = phpc_ktime_begin(stream);
  hipLaunchKernelGGL(ce.fn[mp][vec2], dim3(G), dim3(kernel_threads(config_id, mp, vec2)), ce.lds_bytes[vec2],
                     stream, b);
  phpc_ktime_end(kt, stream);
  PHPC_HIP(hipGetLastError());
  return PHPC_OK;
}

// Transposed copies of A for the transposed-A configurations: one per distinct
// (A, lda, m, k), in a process-wide workspace used in stream order.
SharedWorkspace g_at_ws;

int launch_live(std::vector<GemmDesc> &live, int grid, int config_id, hipStream_t stream, bool a_transposed);

int launch_transposed(std::vector<GemmDesc> &live, int grid, int config_id, hipStream_t stream) {
  F64Staging st;
  st.plan(live.data(), (int)live.size());
  std::lock_guard<std::mutex> lk(g_at_ws.mu);
  g_at_ws.after_previous(stream);
  double *ws = static_cast<double *>(g_at_ws.ws.get(st.total * sizeof(double), 0, stream));
  if (!ws) return phpc_set_error(PHPC_ERR_MEM, "transposed-A workspace allocation failed (%zu bytes)", st.total * 8);
  if (int rc = st.run(ws, stream)) return rc;
  st.apply(ws, live.data(), (int)live.size());
  const int rc = launch_live(live, grid, config_id, stream, true);
  g_at_ws.mark_used(stream);
  return rc;
}

}  // namespace

int phpc_f64_grouped_dev(const GemmDesc *descs_dev, const GemmDesc *descs_host, int ndesc, int grid,
                         int config_id, hipStream_t stream) {
  return launch_grouped(descs_dev, descs_host, ndesc, grid, config_id, stream, false);
}

bool phpc_f64_config_trans_a(int id) { return id >= 0 && id < kNumConfigs && kConfigs[id].trans_a; }

long phpc_f64_at_pitch(int m) { return ((long)m + 1) / 2 * 2; }

bool phpc_f64_trans_a_ok(const GemmDesc *descs, int ndesc) {
  for (int i = 0; i < ndesc; ++i) {
    const GemmDesc &d = descs[i];
    if (d.m <= 0 || d.n <= 0 || d.k <= 0) continue;
    // grid limits of the staging kernels
    if (phpc_f64_at_pitch(d.m) / 32 >= 65535 || (!b_at_ok(d) && d.k >= 65535)) return false;
  }
  return true;
}

void F64Staging::plan(const GemmDesc *descs, int ndesc) {
  copies.clear();
  a_copy.assign(ndesc, -1);
  b_copy.assign(ndesc, -1);
  total = 0;
  auto find_or_add = [&](const double *src, long ld, int rows, int cols, bool tr) {
    for (size_t j = 0; j < copies.size(); ++j) {
      const Copy &c = copies[j];
      if (c.src == src && c.ld == ld && c.rows == rows && c.cols == cols && c.transpose == tr) return (int)j;
    }
    const long pitch = phpc_f64_at_pitch(tr ? rows : cols);
    const size_t n = (size_t)(tr ? cols : rows) * pitch;
    copies.push_back({src, ld, rows, cols, pitch, total, tr});
    total += (n + 31) / 32 * 32;  // 256-byte aligned copies
    return (int)copies.size() - 1;
  };
  for (int i = 0; i < ndesc; ++i) {
    const GemmDesc &d = descs[i];
    if (d.m <= 0 || d.n <= 0 || d.k <= 0) continue;
    a_copy[i] = find_or_add(d.A, d.lda, d.m, d.k, true);
    if (!b_at_ok(d)) b_copy[i] = find_or_add(d.B, d.ldb, d.k, d.n, false);
  }
}

int F64Staging::run(double *ws, hipStream_t stream) const {
  for (const Copy &c : copies) {
    if (c.transpose) {
      if (int rc = phpc_f64_transpose(c.src, c.ld, c.rows, c.cols, ws + c.off, c.pitch, stream)) return rc;
    } else {
      const dim3 grid((unsigned)std::min<long>((c.pitch + 255) / 256, 64), (unsigned)c.rows);
      hipLaunchKernelGGL(f64_copy_pad_kernel, grid, dim3(256), 0, stream, c.src, c.ld, c.rows, c.cols, ws + c.off,
                         c.pitch);
      PHPC_HIP(hipGetLastError());
    }
  }
  return PHPC_OK;
}

void F64Staging::apply(double *ws, GemmDesc *descs, int ndesc) const {
  for (int i = 0; i < ndesc; ++i) {
    if (a_copy[i] >= 0) {
      descs[i].A = ws + copies[a_copy[i]].off;
      descs[i].lda = copies[a_copy[i]].pitch;
    }
    if (b_copy[i] >= 0) {
      descs[i].B = ws + copies[b_copy[i]].off;
      descs[i].ldb = copies[b_copy[i]].pitch;
    }
  }
}

int phpc_f64_reserve_staging(const GemmDesc *descs, int ndesc, hipStream_t stream) {
  const int cfg = phpc_f64_pick_config_at(descs, ndesc);
  if (!kConfigs[cfg].trans_a) return PHPC_OK;
  F64Staging st;
  st.plan(descs, ndesc);
  std::lock_guard<std::mutex> lk(g_at_ws.mu);
  g_at_ws.after_previous(stream);
  if (!g_at_ws.ws.get(st.total * sizeof(double), 0, stream))
    return phpc_set_error(PHPC_ERR_MEM, "transposed-A workspace allocation failed (%zu bytes)", st.total * 8);
  g_at_ws.mark_used(stream);
  return PHPC_OK;
}

int phpc_f64_transpose(const double *A, long lda, int m, int k, double *At, long ldat, hipStream_t stream) {
  if (m <= 0 || k <= 0) return PHPC_OK;
  const dim3 tg((unsigned)((k + 31) / 32), (unsigned)((ldat + 31) / 32));
  hipLaunchKernelGGL(f64_transpose_kernel, tg, dim3(256), 0, stream, A, lda, m, k, At, ldat);
  PHPC_HIP(hipGetLastError());
  return PHPC_OK;
}

int phpc_f64_grouped_dev_at(const GemmDesc *descs_dev, const GemmDesc *descs_host, int ndesc, int grid,
                            int config_id, hipStream_t stream) {
  return launch_grouped(descs_dev, descs_host, ndesc, grid, config_id, stream, true);
}

int phpc_f64_pick_config_at(const GemmDesc *descs, int ndesc) {
  const int cfg = phpc_f64_pick_config(descs, ndesc);
  if (!phpc_f64_trans_a_ok(descs, ndesc)) return cfg;
  if (cfg == g_big_single) return g_big_single_host;
  if (cfg == g_big_config) return g_big_config_host;
  return cfg;
}

int phpc_f64_grouped(const GemmDesc *descs_in, int ndesc, int grid, int config_id,
                     hipStream_t stream) {
  if (ndesc <= 0) return PHPC_OK;
  for (int i = 0; i < ndesc; ++i) {
    const GemmDesc &d = descs_in[i];
    if (d.m < 0 || d.n < 0 || d.k < 0) return phpc_set_error(PHPC_ERR_ARG, "negative GEMM size");
  }
  std::vector<GemmDesc> descs(descs_in, descs_in + ndesc);
  if (config_id < 0) config_id = phpc_f64_pick_config_at(descs.data(), ndesc);
  if (config_id < 0 || config_id >= kNumConfigs)
    return phpc_set_error(PHPC_ERR_ARG, "bad FP64 config id %d", config_id);
  // the transposed copy needs 16-byte B rows and a grid-sized m; otherwise the same tile
  // runs register-staged on A as it is
  if (kConfigs[config_id].trans_a && !phpc_f64_trans_a_ok(descs.data(), ndesc))
    config_id = kConfigs[config_id].bm == 256 ? 6 : kConfigs[config_id].panels ? 0 : 8;
  const ConfigEntry &ce = kConfigs[config_id];
  // drop empty GEMMs; k == 0 leaves C unchanged (C += 0 adds +0.0 only to tiles we skip)
  std::vector<GemmDesc> live;
  live.reserve(ndesc);
  int start = 0;
  for (auto d : descs) {
    if (d.m == 0 || d.n == 0 || d.k == 0) continue;
    d.tiles_n = (d.n + ce.bn - 1) / ce.bn;
    d.tile_start = start;
    start += ((d.m + ce.bm - 1) / ce.bm) * d.tiles_n;
    live.push_back(d);
  }
  if (live.empty()) return PHPC_OK;
  if (ce.trans_a) return launch_transposed(live, grid, config_id, stream);
  return launch_live(live, grid, config_id, stream, false);
}

namespace {

// One grouped launch of prepared descriptors (host copies in `live`).
int launch_live(std::vector<GemmDesc> &live, int grid, int config_id, hipStream_t stream, bool a_transposed) {
  if (live.size() == 1) return launch_grouped(nullptr, live.data(), 1, grid, config_id, stream, a_transposed);
  // Several GEMMs: descriptors go through one process-wide device buffer. Callers on
  // other streams may still have a launch queued that reads it, so the buffer is
  // rewritten only after the previous user's launch has finished (g_desc_done).
  std::lock_guard<std::mutex> lk(g_desc_mu);
  const size_t bytes = live.size() * sizeof(GemmDesc);
  if (g_desc_done) PHPC_HIP(hipEventSynchronize(g_desc_done));
  if (bytes > g_desc_cap) {
    if (g_desc_dev) PHPC_HIP(hipFree(g_desc_dev));
    g_desc_dev = nullptr;
    g_desc_cap = 0;
    const size_t cap = std::max(bytes, (size_t)64 * sizeof(GemmDesc));
    PHPC_HIP(hipMalloc(&g_desc_dev, cap));
    g_desc_cap = cap;
  }
  if (bytes > g_pinned_cap) {
    if (g_desc_pinned) PHPC_HIP(hipHostFree(g_desc_pinned));
    g_desc_pinned = nullptr;
    g_pinned_cap = 0;
    const size_t cap = std::max(bytes, (size_t)64 * sizeof(GemmDesc));
    PHPC_HIP(hipHostMalloc(&g_desc_pinned, cap, hipHostMallocDefault));
    g_pinned_cap = cap;
  }
  if (!g_desc_done) PHPC_HIP(hipEventCreateWithFlags(&g_desc_done, hipEventDisableTiming));
  memcpy(g_desc_pinned, live.data(), bytes);
  PHPC_HIP(hipMemcpyAsync(g_desc_dev, g_desc_pinned, bytes, hipMemcpyHostToDevice, stream));
  const int rc = launch_grouped(g_desc_dev, live.data(), (int)live.size(), grid, config_id, stream, a_transposed);
  PHPC_HIP(hipEventRecord(g_desc_done, stream));
  return rc;
}

}  // namespace

// Tuning knobs for the autotuner: 0 = config for large panel-grouped problems, 1 =
// tile-row swizzle width, 2 = XCD-contiguous tile order on/off, 3 = config for large
// single-panel problems on device descriptors, 4 = the same through host descriptors
// (may be a transposed-A configuration), 5 = large panel-grouped problems through host
// descriptors.
extern "C" int phpc_set_tuning(int key, int value) {
  switch (key) {
    case 0:
      if (value < 0 || value >= kNumConfigs || kConfigs[value].trans_a)
        return phpc_set_error(PHPC_ERR_ARG, "config %d out of range", value);
      g_big_config = value;
      break;
    case 1:
      if (value < 1 || value > 64) return phpc_set_error(PHPC_ERR_ARG, "group width %d out of range", value);
      g_group_m = value;
      break;
    case 2:
      g_xcd_remap = value ? 1 : 0;
      break;
    case 3:
      if (value < 0 || value >= kNumConfigs || kConfigs[value].trans_a)
        return phpc_set_error(PHPC_ERR_ARG, "config %d out of range", value);
      g_big_single = value;
      break;
    case 4:
      if (value < 0 || value >= kNumConfigs) return phpc_set_error(PHPC_ERR_ARG, "config %d out of range", value);
      g_big_single_host = value;
      break;
    case 5:
      if (value < 0 || value >= kNumConfigs || !kConfigs[value].panels)
        return phpc_set_error(PHPC_ERR_ARG, "config %d out of range", value);
      g_big_config_host = value;
      break;
    case 6:
      phpc_f32_set_transposed_a(value);
      break;
    case 7:
      return phpc_f32_set_plan_config(value);
    case 8:
      return phpc_f32x3_set_plan_config(value);

    default:
      return phpc_set_error(PHPC_ERR_ARG, "unknown tuning key %d", key);
  }
  return phpc_set_error(PHPC_OK, nullptr);
}

int phpc_f64_plan_tiles(GemmDesc *descs, int ndesc, int id) {
  const ConfigEntry &ce = kConfigs[id];
  int start = 0;
  for (int i = 0; i < ndesc; ++i) {
    GemmDesc &d = descs[i];
    d.tiles_n = (d.n + ce.bn - 1) / ce.bn;
    d.tile_start = start;
    start += ((d.m + ce.bm - 1) / ce.bm) * d.tiles_n;
  }
  return start;
}

bool phpc_f64_config_panels(int id) { return id >= 0 && id < kNumConfigs && kConfigs[id].panels; }

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
  GemmDesc d0;            // inline copy when ndesc == 1
};

constexpr int BK = 16;

template <int BM, int BN, int WM, int WN, bool MP, int VEC>
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

template <int BM, int BN, int WM, int WN, bool MP, int VEC>
__global__ __launch_bounds__(WM *WN * 64) void f64_gemm_kernel(GemmBatch batch) {
  using K = F64Gemm<BM, BN, WM, WN, MP, VEC>;
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
  const bool xcd_remap = (G % 8) == 0;
  const int slot = xcd_remap ? (bid % 8) * (G / 8) + bid / 8 : bid;

  for (int round = 0;; ++round) {
    const int t = round * G + slot;
    if (t >= batch.total_tiles) break;
    const GemmDesc &d = *find_desc(batch, t);
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
    const int m_rem = d.m - row0;
    const int n_rem = d.n - col0;
    const double *__restrict__ Ag = d.A + (long)row0 * d.lda;
    const double *__restrict__ Bg = d.B + col0;
    const long lda = d.lda, ldb = d.ldb;
    const int Kdim = d.k;
    const int w = d.panel > 0 && d.panel < Kdim ? d.panel : Kdim;

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
            creg[i][j][r] = (rr < m_rem && cc < n_rem) ? d.C[(long)(row0 + rr) * d.ldc + col0 + cc] : 0.0;
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
        const double *p = Ag + (long)row * lda + kb + kc;
        if constexpr (VEC == 2) {
          if (row < m_rem && kc + 1 < krem) {
            const double2 v = *reinterpret_cast<const double2 *>(p);
            ra[i].x = v.x;
            ra[i].y = v.y;
          } else {
            ra[i].x = (row < m_rem && kc < krem) ? p[0] : 0.0;
            ra[i].y = 0.0;
          }
        } else {
          ra[i].x = (row < m_rem && kc < krem) ? p[0] : 0.0;
        }
      }
#pragma unroll
      for (int i = 0; i < K::CB; ++i) {
        const int q = tid + i * K::NT;
        const int kr = q / (BN / VEC);
        const int cc = (q % (BN / VEC)) * VEC;
        const double *p = Bg + (long)(kb + kr) * ldb + cc;
        if constexpr (VEC == 2) {
          if (kr < krem && cc + 1 < n_rem) {
            const double2 v = *reinterpret_cast<const double2 *>(p);
            rb[i].x = v.x;
            rb[i].y = v.y;
          } else {
            rb[i].x = (kr < krem && cc < n_rem) ? p[0] : 0.0;
            rb[i].y = 0.0;
          }
        } else {
          rb[i].x = (kr < krem && cc < n_rem) ? p[0] : 0.0;
        }
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

    auto compute = [&](int buf, int klen) {
      const double *At = lds + buf * K::STAGE;
      const double *Bs = At + BK * K::APITCH;
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) {
        if (kk * 4 < klen) {
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
        }
      }
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
            double *pc = d.C + (long)(row0 + rr) * d.ldc + col0 + cc;
            if constexpr (MP)
              *pc = creg[i][j][r];
            else
              *pc = *pc + acc[i][j][r];
          }
        }
      }
  }
}

typedef void (*KernelFn)(GemmBatch);

struct ConfigEntry {
  int bm, bn, threads;
  KernelFn fn[2][2];  // [multi-panel][vec2]
};

#define PHPC_F64_CFG(BM_, BN_, WM_, WN_)                                                     \
  ConfigEntry {                                                                            \
    BM_, BN_, WM_ *WN_ * 64, {                                                             \
      {f64_gemm_kernel<BM_, BN_, WM_, WN_, false, 1>, f64_gemm_kernel<BM_, BN_, WM_, WN_, false, 2>}, \
      {f64_gemm_kernel<BM_, BN_, WM_, WN_, true, 1>, f64_gemm_kernel<BM_, BN_, WM_, WN_, true, 2>}    \
    }                                                                                      \
  }

const ConfigEntry kConfigs[] = {
    PHPC_F64_CFG(128, 128, 2, 4),  // 0: 8 waves, wave tile 64x32
    PHPC_F64_CFG(128, 128, 2, 2),  // 1: 4 waves, wave tile 64x64
    PHPC_F64_CFG(128, 64, 2, 2),   // 2: 4 waves, wave tile 64x32
    PHPC_F64_CFG(64, 64, 2, 2),    // 3: 4 waves, wave tile 32x32
    PHPC_F64_CFG(256, 128, 4, 2),  // 4: 8 waves, wave tile 64x64
};
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);

struct OccKey {
  int occ[kNumConfigs][2][2];
  bool ready = false;
};

std::mutex g_mu;
OccKey g_occ;
GemmDesc *g_desc_dev = nullptr;
size_t g_desc_cap = 0;
GemmDesc *g_desc_pinned = nullptr;
size_t g_pinned_cap = 0;

int occupancy(int cfg, int mp, int vec2) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_occ.ready) {
    for (int c = 0; c < kNumConfigs; ++c)
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          int n = 0;
          if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(kConfigs[c].fn[a][b]),
                                                           kConfigs[c].threads, 0) != hipSuccess ||
              n < 1)
            n = 1;
          g_occ.occ[c][a][b] = n;
        }
    g_occ.ready = true;
  }
  return g_occ.occ[cfg][mp][vec2];
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

// Planner: large problems take the 128x128 8-wave tile; small ones a tile that still
// yields enough workgroups to cover the 256 CUs.
int phpc_f64_pick_config(const GemmDesc *descs, int ndesc) {
  long work_tiles128 = 0;
  for (int i = 0; i < ndesc; ++i)
    work_tiles128 += (long)((descs[i].m + 127) / 128) * ((descs[i].n + 127) / 128);
  const int cus = phpc_device().cus;
  if (work_tiles128 >= 2L * cus) return 0;
  long work_tiles64 = 0;
  for (int i = 0; i < ndesc; ++i)
    work_tiles64 += (long)((descs[i].m + 127) / 128) * ((descs[i].n + 63) / 64);
  if (work_tiles64 >= 2L * cus) return 2;
  return 3;
}

int phpc_f64_grouped_dev(const GemmDesc *descs_dev, const GemmDesc *descs_host, int ndesc, int grid,
                         int config_id, hipStream_t stream) {
  if (ndesc <= 0) return PHPC_OK;
  if (config_id < 0 || config_id >= kNumConfigs)
    return phpc_set_error(PHPC_ERR_ARG, "bad FP64 config id %d", config_id);
  const ConfigEntry &ce = kConfigs[config_id];
  bool mp = false, vec2 = true;
  int total = 0;
  for (int i = 0; i < ndesc; ++i) {
    const GemmDesc &d = descs_host[i];
    if (d.panel > 0 && d.panel < d.k) mp = true;
    const bool a_ok = aligned16(d.A) && (d.lda % 2 == 0) &&
                      (d.panel <= 0 || d.panel >= d.k || d.panel % 2 == 0);
    const bool b_ok = aligned16(d.B) && (d.ldb % 2 == 0);
    if (!a_ok || !b_ok) vec2 = false;
    total = d.tile_start + ((d.m + ce.bm - 1) / ce.bm) * d.tiles_n;
  }
  if (total <= 0) return PHPC_OK;
  GemmBatch b{};
  b.descs = descs_dev;
  b.ndesc = ndesc;
  b.total_tiles = total;
  b.group_m = 8;
  b.d0 = descs_host[0];
  int G = grid;
  if (G <= 0) {
    G = occupancy(config_id, mp, vec2) * phpc_device().cus;
    if (G > total) G = total;
  }
  if (G < 1) G = 1;
  hipLaunchKernelGGL(ce.fn[mp][vec2], dim3(G), dim3(ce.threads), 0, stream, b);
  PHPC_HIP(hipGetLastError());
  return PHPC_OK;
}

int phpc_f64_grouped(const GemmDesc *descs_in, int ndesc, int grid, int config_id,
                     hipStream_t stream) {
  if (ndesc <= 0) return PHPC_OK;
  for (int i = 0; i < ndesc; ++i) {
    const GemmDesc &d = descs_in[i];
    if (d.m < 0 || d.n < 0 || d.k < 0) return phpc_set_error(PHPC_ERR_ARG, "negative GEMM size");
  }
  std::vector<GemmDesc> descs(descs_in, descs_in + ndesc);
  if (config_id < 0) config_id = phpc_f64_pick_config(descs.data(), ndesc);
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
  const GemmDesc *dev = nullptr;
  if (live.size() > 1) {
    std::lock_guard<std::mutex> lk(g_mu);
    const size_t bytes = live.size() * sizeof(GemmDesc);
    if (bytes > g_desc_cap) {
      if (g_desc_dev) {
        PHPC_HIP(hipStreamSynchronize(stream));
        PHPC_HIP(hipFree(g_desc_dev));
      }
      g_desc_cap = std::max(bytes, (size_t)64 * sizeof(GemmDesc));
      PHPC_HIP(hipMalloc(&g_desc_dev, g_desc_cap));
    }
    if (bytes > g_pinned_cap) {
      if (g_desc_pinned) PHPC_HIP(hipHostFree(g_desc_pinned));
      g_pinned_cap = std::max(bytes, (size_t)64 * sizeof(GemmDesc));
      PHPC_HIP(hipHostMalloc(&g_desc_pinned, g_pinned_cap, hipHostMallocDefault));
    }
    // The pinned staging copy may still be read by an earlier async copy.
    PHPC_HIP(hipStreamSynchronize(stream));
    memcpy(g_desc_pinned, live.data(), bytes);
    PHPC_HIP(hipMemcpyAsync(g_desc_dev, g_desc_pinned, bytes, hipMemcpyHostToDevice, stream));
    dev = g_desc_dev;
  }
  return phpc_f64_grouped_dev(dev, live.data(), (int)live.size(), grid, config_id, stream);
}

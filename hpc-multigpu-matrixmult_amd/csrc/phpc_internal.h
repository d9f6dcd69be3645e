// Internal declarations shared by the HIP sources of libphpc (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>
#include <stdint.h>
#include <stddef.h>

#include "phpc_gemm.h"

// An event destroyed when its scope ends, error returns included.
struct ScopedEvent {
  hipEvent_t e = nullptr;
  explicit ScopedEvent(unsigned flags = hipEventDefault) {
    if (hipEventCreateWithFlags(&e, flags) != hipSuccess) e = nullptr;
  }
  ~ScopedEvent() {
    if (e) (void)hipEventDestroy(e);
  }
  ScopedEvent(const ScopedEvent &) = delete;
  ScopedEvent &operator=(const ScopedEvent &) = delete;
  ScopedEvent(ScopedEvent &&o) noexcept : e(o.e) { o.e = nullptr; }
};

// p[idx] where ok, else 0, through a load that always happens (from p[0] where !ok; p[0]
// must be readable). A guarded load compiles to a branch with a wait for the load inside
// it, so a thread's guarded loads go out one at a time; these all go out before the first
// wait.
template <class T>
__device__ __forceinline__ T load_or_zero(const T *p, long idx, bool ok) {
  const T v = p[ok ? idx : 0];
  return ok ? v : T(0);
}

// ---------------------------------------------------------------------------
// Error handling: every entry point records its status in a thread-local slot
// (phpc_last_error) and returns it where the signature allows.
// ---------------------------------------------------------------------------
extern "C" int phpc_set_error(int code, const char *fmt, ...);
#define PHPC_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return phpc_set_error(PHPC_ERR_HIP, "%s failed: %s (%s:%d)", #call,                \
                            hipGetErrorString(e_), __FILE__, __LINE__);                  \
  } while (0)

// ---------------------------------------------------------------------------
// One GEMM of a grouped launch: C[m x n] (ldc) += A[m x k] (lda) * B[k x n] (ldb),
// row-major FP64. The k range is walked in panels of `panel` columns; each panel's
// product is summed from zero and then added to the C value carried from the
// previous panel, i.e. C = ((C0 + s_0) + s_1) + ... (the grouping produced by one
// local-GEMM call per SUMMA step, src/phpc_summa.c:63-95 with src/phpc_gemm.cu:32,55).
// ---------------------------------------------------------------------------
struct GemmDesc {
  const double *A;
  const double *B;
  double *C;
  long lda, ldb, ldc;
  int m, n, k;
  int panel;       // >= k: one panel
  int tiles_n;     // tiles along n for this GEMM (filled by the planner)
  int tile_start;  // first global tile id of this GEMM (prefix sum)
};

// One FP32 GEMM of a grouped launch (row-major, C += A*B).
struct F32Desc {
  const float *A;
  const float *B;
  float *C;
  long lda, ldb, ldc;
  int m, n, k;
  int tiles_n;
  int tile_start;
};
int phpc_f32_num_configs();
void phpc_f32_config_tile(int id, int *bm, int *bn);
int phpc_f32_plan_tiles(F32Desc *descs, int ndesc, int id);
void phpc_f32_set_transposed_a(int on);
// configuration of FP32 plans created without autotuning: -1 = by work (at least two
// 128x128 tiles per CU: the transposed-A configuration 3, or 2 where its staged copies would
// not fit its 32-bit offsets or tuning key 6 is off; else 0, 64x64); 0-3 = that
// configuration (3 where it applies)
int phpc_f32_set_plan_config(int id);
int phpc_f32_pick_plan_config(const struct F32Desc *descs, int ndesc);
// FP32 configuration ids: 0-2 register-staged (64x64, 128x64, 128x128); kF32AtConfig is the
// transposed-A direct-to-LDS kernel (128x128), which reads staged operand copies.
constexpr int kF32AtConfig = 3;
// One transposed A copy of a grouped FP32 transposed-A launch, filled by one batched kernel:
// dst[c][r] = src[r][c] for r < rows, c < cols, zero for rows <= r < pitch. 64 x 64 tiles of
// dst, tiles_x across a dst row; block_start is the first tile's index in the batched grid.
// B is read in place (the kernel loads unaligned B rows 4 bytes per lane).
struct F32StageCopy {
  const float *src;
  long ld;
  int rows, cols;
  float *dst;
  long pitch;
  int tiles_x;
  int block_start;
};
struct F32Staging {
  std::vector<F32StageCopy> copies;  // dst set by apply
  std::vector<size_t> off;           // floats into the workspace, per copy
  std::vector<int> a_copy;           // per descriptor
  std::vector<int> block_copy;       // the copy of each tile of the batched kernel
  size_t total = 0;                  // floats
  int blocks = 0;                    // tiles of the batched copy kernel
  void plan(const F32Desc *descs, int ndesc);
  void apply(float *ws, F32Desc *descs, int ndesc);
  // the device list the kernel reads: the copies (dst set by apply), then block_copy
  size_t list_bytes() const;
  int upload(void *list_dev) const;
  int run(const void *list_dev, hipStream_t stream) const;
};
// the descriptors fit the transposed-A kernel's 32-bit buffer offsets once staged
bool phpc_f32_at_ok(const F32Desc *descs, int ndesc);
// transposed-A kernel over descriptors already pointing at staged copies (tiles planned
// with config kF32AtConfig); grid 0 = persistent default
int phpc_f32_launch_at(const F32Desc *descs_dev, const F32Desc *descs_host, int ndesc, int total_tiles, int grid,
                       hipStream_t stream);
bool phpc_f32_vec4_ok(const F32Desc *descs, int ndesc);

// FP32-accurate grouped plans on BF16 slice planes (3 x BF16). One batched kernel cuts
// every distinct A into rows [hi | mid | lo] and every distinct B into columns
// [lo | mid | hi] (planes kp = k rounded up to 64 wide, zero past k); one grouped launch
// of the BF16 level kernel then sums each item's three k segments smallest first into one
// FP32 accumulator per output and adds it to C: the planes and sums of the single-product
// slice-plane path (phpc_sgemm_bf16x3_planes), so every item equals that path bit for bit.
// Plan configurations kF32X3PlanesConfig + g, g < phpc_bf3_num_configs(), select the
// level kernel's tile.
constexpr int kF32X3PlanesConfig = 4;
int phpc_bf3_num_configs();
// configuration of 3 x BF16 plans created without autotuning: -1 = by work (slice planes on
// 256 x 256 tiles for at least one such tile per CU; else 2); 2 or kF32X3PlanesConfig + g
int phpc_f32x3_set_plan_config(int id);

int phpc_f32x3_pick_plan_config(const F32Desc *descs, int ndesc);
struct Bf3Cut {  // one operand's planes, cut by 64 x 64 tiles of the source
  const float *src;
  long ld;
  int rows, cols;  // source shape: A is m x k, B is k x n
  uint16_t *dst;   // plane row r (row r of A, column r of B) at dst + r * 3 kp
  long kp;
  int is_b;        // B: the source is read across its rows (a transposed cut)
  int tiles_k;     // 64-wide k tiles across one plane row
  int block_start; // first workgroup of the batched kernel on this cut
  int vec;         // A: rows 16-byte aligned (16-byte loads)
};
struct Bf3Item {  // one GEMM of the grouped level launch
  const uint16_t *A3;
  const uint16_t *B3;
  float *C;
  long ldc;
  int m, n;
  int kp2;  // bytes of one plane row (2 kp)
  int tiles_m, tiles_n, tile_start;
};
struct Bf3Planes {
  std::vector<Bf3Cut> cuts;
  std::vector<size_t> off;      // bytes into the workspace, per cut
  std::vector<int> block_cut;   // the cut of each workgroup of the batched kernel
  std::vector<Bf3Item> items;
  int g = 0;                    // level kernel tile (0 .. phpc_bf3_num_configs() - 1)
  int blocks = 0, total_tiles = 0;
  size_t ws_bytes = 0;
  void *ws = nullptr;   // the planes
  void *list = nullptr; // cuts, block_cut and items on the device
  // plans the cuts and tiles of descs for tile g, allocates the planes and uploads the
  // lists (synchronously); PHPC_ERR_ARG when a plane row is too long for the kernel's
  // 32-bit offsets
  int build(const F32Desc *descs, int ndesc, int g);
  // the cut kernel, then the level kernel; grid 0 = persistent default
  int run(hipStream_t stream, int grid) const;
  void release();
  Bf3Planes() = default;
  Bf3Planes(const Bf3Planes &) = delete;
  Bf3Planes &operator=(const Bf3Planes &) = delete;
  ~Bf3Planes() { release(); }
};

int phpc_f32x3_launch(const F32Desc *descs_dev, const F32Desc *descs_host, int ndesc, int total_tiles, bool vec4,
                      int grid, hipStream_t stream);
int phpc_f32_launch(const F32Desc *descs_dev, const F32Desc *descs_host, int ndesc, int total_tiles, int id, bool vec4,
                    int grid, hipStream_t stream);

// FP64: fill tiles_n / tile_start for config id; returns the total tile count.
int phpc_f64_plan_tiles(GemmDesc *descs, int ndesc, int id);
bool phpc_f64_config_panels(int id);

// Kernel-level tile configuration chosen by the planner / autotuner.
struct F64Config {
  int id;         // index into the instantiated kernel table
  int bm, bn;     // macro tile
  int threads;    // workgroup size
  int wg_per_cu;  // resident workgroups per CU the config is built for
};

// Launch policy knobs (see phpc_set_launch_policy in phpc_gemm.h).
struct LaunchPolicy {
  int honor_grid;     // 1: persistent grid = caller's grid_width * grid_height
  int config_id;      // -1: autotuned
  int grid_override;  // >0: persistent grid size
};
LaunchPolicy &phpc_policy();

// Device properties cached per process (one GPU).
struct DeviceInfo {
  int device;
  int cus;
  int xcds;
  size_t lds_per_cu;
};
const DeviceInfo &phpc_device();

// Grouped FP64 launch. `descs_host` are copied into a library-owned device buffer
// (stream-ordered). `grid` <= 0 selects the autotuned persistent grid. Returns
// PHPC_OK or an error code.
int phpc_f64_grouped(const GemmDesc *descs_host, int ndesc, int grid, int config_id,
                     hipStream_t stream);
// Same, but with the descriptor array already resident on the device (graph-friendly).
int phpc_f64_grouped_dev(const GemmDesc *descs_dev, const GemmDesc *descs_host, int ndesc,
                         int grid, int config_id, hipStream_t stream);
int phpc_f64_num_configs();
F64Config phpc_f64_config(int id);
int phpc_f64_pick_config(const GemmDesc *descs, int ndesc);
// Configurations whose kernels read a transposed copy of A (descriptor A = At[k][m] with
// lda = phpc_f64_at_pitch(m), pad column zero). phpc_f64_grouped stages the copy itself;
// callers with device descriptors stage it with phpc_f64_transpose and launch through
// phpc_f64_grouped_dev_at (phpc_f64_grouped_dev rejects these configurations).
bool phpc_f64_config_trans_a(int id);
bool phpc_f64_trans_a_ok(const GemmDesc *descs, int ndesc);
long phpc_f64_at_pitch(int m);
int phpc_f64_transpose(const double *A, long lda, int m, int k, double *At, long ldat, hipStream_t stream);

// Operand copies that the transposed-A kernels read: At[k][m] for every distinct A block,
// and, for a B block whose rows are not 16-byte aligned (base or odd ldb), a copy with an
// even pitch. `plan` lists them and sizes the workspace; `run` fills a workspace; `apply`
// points descriptors at the copies.
struct F64Staging {
  struct Copy {
    const double *src;
    long ld;
    int rows, cols;  // of the source block
    long pitch;      // of the copy
    size_t off;      // doubles into the workspace
    bool transpose;
  };
  std::vector<Copy> copies;
  std::vector<int> a_copy, b_copy;  // per descriptor; b_copy -1: B is read in place
  size_t total = 0;                 // doubles
  void plan(const GemmDesc *descs, int ndesc);
  int run(double *ws, hipStream_t stream) const;
  void apply(double *ws, GemmDesc *descs, int ndesc) const;
};
int phpc_f64_grouped_dev_at(const GemmDesc *descs_dev, const GemmDesc *descs_host, int ndesc, int grid,
                            int config_id, hipStream_t stream);
// the planner's choice, with the transposed-A configuration for large single-panel work
// whose B allows it
int phpc_f64_pick_config_at(const GemmDesc *descs, int ndesc);
// Size the process-wide staging workspace for these descriptors now, so that a pipeline
// of host-pointer launches does not allocate between its copies.
int phpc_f64_reserve_staging(const GemmDesc *descs, int ndesc, hipStream_t stream);

// Strided block pack: dst[r*ldd + c] = src[r*lds + c] for r < rows, c < cols.
// Host <-> device copies through pinned staging buffers (rows x width bytes, strided).
// h2d returns once the source may be reused (transfers complete in stream order); d2h
// returns once the destination holds the data.
int phpc_stage_h2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t rows,
                   hipStream_t stream);
int phpc_stage_d2h(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t rows,
                   hipStream_t stream);
// Process-wide copy and compute streams of the host-pointer entry points.
int phpc_host_streams(hipStream_t *copy, hipStream_t *compute);
int phpc_pack_launch(const double *src, long lds, double *dst, long ldd, int rows, int cols,
                     hipStream_t stream);

// Scratch device buffer owned by the library (grows; stream-ordered use only).
void *phpc_scratch(size_t bytes, int slot);
// Kernel event timing (phpc_kernel_timing): begin returns a token (-1 when timing is off
// or the stream is capturing); end records the closing event for that token.
int phpc_ktime_begin(hipStream_t stream);
void phpc_ktime_end(int token, hipStream_t stream);

// A set of growable device buffers for one caller at a time (the split path needs four).
// Growing synchronises `stream` before the old buffer is freed.
struct Workspace {
  static constexpr int kSlots = 8;
  void *p[kSlots] = {};
  size_t cap[kSlots] = {};
  // Fixed region (set by use_fixed): get() allocates nothing and instead carves each
  // slot, in call order, out of [fixed, fixed + fixed_size) at 256-byte aligned offsets,
  // returning nullptr once the region is exhausted; release() frees nothing.
  char *fixed = nullptr;
  size_t fixed_size = 0, fixed_used = 0;
  void use_fixed(void *base, size_t size);
  void *get(size_t bytes, int slot, hipStream_t stream);
  size_t total() const;
  void release();
  ~Workspace() { release(); }
};

// A workspace shared by calls that may come on different streams: the holder of the
// lock orders its stream after the previous holder's last use (device-side wait), and
// marks its own last use before unlocking.
struct SharedWorkspace {
  std::mutex mu;
  Workspace ws;
  hipEvent_t last = nullptr;
  void after_previous(hipStream_t stream) {
    if (last) (void)hipStreamWaitEvent(stream, last, 0);
  }
  void mark_used(hipStream_t stream) {
    if (!last) (void)hipEventCreateWithFlags(&last, hipEventDisableTiming);
    if (last) (void)hipEventRecord(last, stream);
  }
};

// Bytes of k per residue / slice plane row of the split GEMM: k rounded up to 128, so
// that every level range is a whole number of the level kernels' k slabs (64 or 128
// bytes) and the 16-byte direct-to-LDS loads never straddle a plane.
inline long phpc_split_kp(long k) { return (k + 127) / 128 * 128; }

#include "phpc_split.h"
// Split-precision GEMM with an explicit workspace and method (the public entry point
// uses a process-wide workspace and the method set by phpc_split_set_method).
// FP32-accurate 3 x BF16 GEMM through BF16 slice planes (workspace slots 0 and 1).
int phpc_sgemm_bf16x3_planes(const float *dA, long lda, const float *dB, long ldb, float *dC, long ldc, int m, int k,
                             int n, hipStream_t s, Workspace *ws);
int phpc_dgemm_split_ws(const double *dA, long lda, const double *dB, long ldb, double *dC, long ldc, int m, int k,
                        int n, int slices, double target_err, int method, hipStream_t stream, Workspace *ws,
                        phpc_split_info *info);

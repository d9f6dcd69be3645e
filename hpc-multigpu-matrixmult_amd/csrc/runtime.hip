// libphpc runtime: errors, device state, the drop-in host-pointer entry points
// (phpc_gemm_cuda / phpc_gemm_cublas), the device-resident API, seeded generators,
// the strided block pack and the on-device sampled accuracy check.
#include "phpc_internal.h"

#include <rocblas/rocblas.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
namespace {
thread_local int t_err = PHPC_OK;
thread_local char t_msg[512] = "";
}  // namespace

extern "C" int phpc_set_error(int code, const char *fmt, ...) {
  t_err = code;
  if (fmt) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_msg, sizeof(t_msg), fmt, ap);
    va_end(ap);
  } else {
    t_msg[0] = 0;
  }
  return code;
}

extern "C" int phpc_last_error(void) { return t_err; }
extern "C" const char *phpc_last_error_message(void) { return t_msg; }

// ---------------------------------------------------------------------------
// device state
// ---------------------------------------------------------------------------
namespace {
std::once_flag g_init_once;
int g_init_status = PHPC_OK;
DeviceInfo g_dev{};
LaunchPolicy g_policy{0, -1, 0};
std::mutex g_scratch_mu;
struct ScratchSlot {
  void *p = nullptr;
  size_t cap = 0;
};
ScratchSlot g_scratch[16];

void do_init() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    g_init_status = PHPC_ERR_NODEV;
    return;
  }
  if (hipSetDevice(0) != hipSuccess) {
    g_init_status = PHPC_ERR_HIP;
    return;
  }
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, 0) != hipSuccess) {
    g_init_status = PHPC_ERR_HIP;
    return;
  }
  g_dev.device = 0;
  g_dev.cus = p.multiProcessorCount;
  g_dev.xcds = 8;
  g_dev.lds_per_cu = p.sharedMemPerMultiprocessor;
}
}  // namespace

int phpc_host_streams(hipStream_t *copy, hipStream_t *compute) {
  static std::once_flag once;
  static hipStream_t sx = nullptr, sc = nullptr;
  static int status = PHPC_OK;
  std::call_once(once, [] {
    if (hipStreamCreateWithFlags(&sx, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&sc, hipStreamNonBlocking) != hipSuccess)
      status = PHPC_ERR_HIP;
  });
  if (status) return phpc_set_error(status, "stream creation failed");
  *copy = sx;
  *compute = sc;
  return PHPC_OK;
}

extern "C" int phpc_init(void) {
  std::call_once(g_init_once, do_init);
  if (g_init_status != PHPC_OK)
    return phpc_set_error(g_init_status, g_init_status == PHPC_ERR_NODEV ? "no GPU visible" : "HIP device initialisation failed");
  return PHPC_OK;
}

const DeviceInfo &phpc_device() { return g_dev; }
LaunchPolicy &phpc_policy() { return g_policy; }

// ---------------------------------------------------------------------------
// Event timing of the main GEMM kernel launches (measurement only). While enabled,
// each FP64 tile launch and each split-precision level launch is bracketed by a pair
// of HIP events on its own stream; reading syncs on them and returns the summed
// duration. Launches on a stream that is being captured into a graph are not timed.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_kt_mu;
bool g_kt_on = false;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_kt_pairs;  // recorded since the last read
std::vector<hipEvent_t> g_kt_pool;                          // idle events, reused
hipEvent_t kt_event() {
  if (!g_kt_pool.empty()) {
    hipEvent_t e = g_kt_pool.back();
    g_kt_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}
}  // namespace

int phpc_ktime_begin(hipStream_t stream) {
  std::lock_guard<std::mutex> lk(g_kt_mu);
  if (!g_kt_on) return -1;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return -1;
  hipEvent_t a = kt_event(), b = kt_event();
  if (!a || !b || hipEventRecord(a, stream) != hipSuccess) {
    if (a) g_kt_pool.push_back(a);
    if (b) g_kt_pool.push_back(b);
    return -1;
  }
  g_kt_pairs.emplace_back(a, b);
  return (int)g_kt_pairs.size() - 1;
}

void phpc_ktime_end(int token, hipStream_t stream) {
  if (token < 0) return;
  std::lock_guard<std::mutex> lk(g_kt_mu);
  if (token < (int)g_kt_pairs.size()) (void)hipEventRecord(g_kt_pairs[token].second, stream);
}

extern "C" int phpc_kernel_timing(int on) {
  std::lock_guard<std::mutex> lk(g_kt_mu);
  for (auto &pr : g_kt_pairs) {
    (void)hipEventSynchronize(pr.second);
    g_kt_pool.push_back(pr.first);
    g_kt_pool.push_back(pr.second);
  }
  g_kt_pairs.clear();
  g_kt_on = on != 0;
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_kernel_timing_read(double *ms_total, int *launches) {
  std::lock_guard<std::mutex> lk(g_kt_mu);
  double total = 0.0;
  int count = 0, rc = PHPC_OK;
  for (auto &pr : g_kt_pairs) {
    float ms = 0.0f;
    if (hipEventSynchronize(pr.second) != hipSuccess || hipEventElapsedTime(&ms, pr.first, pr.second) != hipSuccess)
      rc = PHPC_ERR_HIP;
    total += ms;
    ++count;
    g_kt_pool.push_back(pr.first);
    g_kt_pool.push_back(pr.second);
  }
  g_kt_pairs.clear();
  if (ms_total) *ms_total = total;
  if (launches) *launches = count;
  return phpc_set_error(rc, rc ? "kernel timing events failed" : nullptr);
}

void *phpc_scratch(size_t bytes, int slot) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  ScratchSlot &s = g_scratch[slot];
  if (bytes > s.cap) {
    if (s.p) {
      (void)hipDeviceSynchronize();
      (void)hipFree(s.p);
      s.p = nullptr;
      s.cap = 0;
    }
    if (hipMalloc(&s.p, bytes) != hipSuccess) {
      s.p = nullptr;
      return nullptr;
    }
    s.cap = bytes;
  }
  return s.p;
}

void Workspace::use_fixed(void *base, size_t size) {
  release();
  fixed = static_cast<char *>(base);
  fixed_size = size;
  fixed_used = 0;
}

void *Workspace::get(size_t bytes, int slot, hipStream_t stream) {
  if (slot < 0 || slot >= kSlots) return nullptr;
  if (bytes == 0) bytes = 1;
  if (fixed) {
    if (bytes <= cap[slot]) return p[slot];
    const size_t off = (fixed_used + 255) / 256 * 256;
    if (off > fixed_size || bytes > fixed_size - off) return nullptr;
    p[slot] = fixed + off;
    cap[slot] = bytes;
    fixed_used = off + bytes;
    return p[slot];
  }
  if (bytes > cap[slot]) {
    if (p[slot]) {
      (void)hipStreamSynchronize(stream);
      (void)hipFree(p[slot]);
      p[slot] = nullptr;
      cap[slot] = 0;
    }
    if (hipMalloc(&p[slot], bytes) != hipSuccess) {
      p[slot] = nullptr;
      return nullptr;
    }
    cap[slot] = bytes;
  }
  return p[slot];
}

size_t Workspace::total() const {
  size_t t = 0;
  for (int i = 0; i < kSlots; ++i) t += cap[i];
  return t;
}

void Workspace::release() {
  for (int i = 0; i < kSlots; ++i) {
    if (p[i] && !fixed) (void)hipFree(p[i]);
    p[i] = nullptr;
    cap[i] = 0;
  }
  fixed_used = 0;
}

extern "C" int phpc_set_launch_policy(int honor_grid, int config_id, int grid) {
  if (config_id >= phpc_f64_num_configs()) return phpc_set_error(PHPC_ERR_ARG, "config id %d out of range", config_id);
  g_policy.honor_grid = honor_grid ? 1 : 0;
  g_policy.config_id = config_id < 0 ? -1 : config_id;
  g_policy.grid_override = grid > 0 ? grid : 0;
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_num_configs(void) { return phpc_f64_num_configs(); }

extern "C" int phpc_config_info(int id, int *bm, int *bn, int *threads) {
  if (id < 0 || id >= phpc_f64_num_configs()) return phpc_set_error(PHPC_ERR_ARG, "config id %d out of range", id);
  if (phpc_init()) return phpc_last_error();
  F64Config c = phpc_f64_config(id);
  if (bm) *bm = c.bm;
  if (bn) *bn = c.bn;
  if (threads) *threads = c.threads;
  return PHPC_OK;
}

// ---------------------------------------------------------------------------
// device-resident FP64 GEMM
// ---------------------------------------------------------------------------
static int check_gemm_args(const void *A, long lda, const void *B, long ldb, const void *C, long ldc, int m,
                           int k, int n) {
  if (m < 0 || k < 0 || n < 0) return phpc_set_error(PHPC_ERR_ARG, "negative size (m=%d k=%d n=%d)", m, k, n);
  if (m > 0 && k > 0 && (A == nullptr || lda < k)) return phpc_set_error(PHPC_ERR_ARG, "bad A / lda (%ld < k=%d)", lda, k);
  if (k > 0 && n > 0 && (B == nullptr || ldb < n)) return phpc_set_error(PHPC_ERR_ARG, "bad B / ldb (%ld < n=%d)", ldb, n);
  if (m > 0 && n > 0 && (C == nullptr || ldc < n)) return phpc_set_error(PHPC_ERR_ARG, "bad C / ldc (%ld < n=%d)", ldc, n);
  return PHPC_OK;
}

int phpc_dgemm_launch(const double *dA, long lda, const double *dB, long ldb, double *dC, long ldc, int m, int k,
                      int n, int panel, int grid, hipStream_t stream) {
  GemmDesc d{};
  d.A = dA;
  d.B = dB;
  d.C = dC;
  d.lda = lda;
  d.ldb = ldb;
  d.ldc = ldc;
  d.m = m;
  d.n = n;
  d.k = k;
  d.panel = panel > 0 ? panel : 0;
  const LaunchPolicy &pol = phpc_policy();
  int G = grid > 0 ? grid : pol.grid_override;
  return phpc_f64_grouped(&d, 1, G, pol.config_id, stream);
}

extern "C" int phpc_dgemm_device(const double *dA, long lda, const double *dB, long ldb, double *dC, long ldc, int m,
                                 int k, int n, int panel, void *stream) {
  if (phpc_init()) return phpc_last_error();
  if (check_gemm_args(dA, lda, dB, ldb, dC, ldc, m, k, n)) return phpc_last_error();
  int rc = phpc_dgemm_launch(dA, lda, dB, ldb, dC, ldc, m, k, n, panel, 0, (hipStream_t)stream);
  if (rc) return rc;
  return phpc_set_error(PHPC_OK, nullptr);
}

// ---------------------------------------------------------------------------
// drop-in host-pointer entry point (src/phpc_gemm.cu:59-156)
// ---------------------------------------------------------------------------
static int gemm_cuda_impl(const double *a, int lda, const double *b, int ldb, double *c, int ldc, int m, int k, int n,
                          int gpu_count, int grid_width, int grid_height, int block_width, float *compute_time,
                          bool vendor) {
  if (compute_time) *compute_time = 0.0f;
  if (phpc_init()) return phpc_last_error();
  if (gpu_count < 1) return phpc_set_error(PHPC_ERR_ARG, "gpu_count must be >= 1 (got %d)", gpu_count);
  if (!vendor) {
    // the reference launches block (T, T) with 2*T*T doubles of shared memory
    // (src/phpc_gemm.cu:82-84): T must be 1..32 and the grid non-empty
    if (block_width < 1 || block_width > 32)
      return phpc_set_error(PHPC_ERR_ARG, "block_width must be in 1..32 (got %d)", block_width);
    if (grid_width < 1 || grid_height < 1)
      return phpc_set_error(PHPC_ERR_ARG, "grid must be at least 1x1 (got %dx%d)", grid_width, grid_height);
  }
  if (check_gemm_args(a, lda, b, ldb, c, ldc, m, k, n)) return phpc_last_error();
  if (m == 0 || n == 0) return phpc_set_error(PHPC_OK, nullptr);
  const size_t sa = (size_t)m * k, sb = (size_t)k * n, sc = (size_t)m * n;
  double *dA = (double *)phpc_scratch((sa ? sa : 1) * 8, 0);
  double *dB = (double *)phpc_scratch((sb ? sb : 1) * 8, 1);
  double *dC = (double *)phpc_scratch(sc * 8, 2);
  if (!dA || !dB || !dC) return phpc_set_error(PHPC_ERR_MEM, "device allocation failed");
  hipStream_t s, sc_unused;
  if (int rc = phpc_host_streams(&s, &sc_unused)) return rc;
  if (k > 0) {
    if (int rc = phpc_stage_h2d(dA, (size_t)k * 8, a, (size_t)lda * 8, (size_t)k * 8, m, s)) return rc;
    if (int rc = phpc_stage_h2d(dB, (size_t)n * 8, b, (size_t)ldb * 8, (size_t)n * 8, k, s)) return rc;
  }
  if (int rc = phpc_stage_h2d(dC, (size_t)n * 8, c, (size_t)ldc * 8, (size_t)n * 8, m, s)) return rc;
  ScopedEvent e0, e1;
  if (!e0.e || !e1.e) return phpc_set_error(PHPC_ERR_HIP, "event creation failed");
  PHPC_HIP(hipEventRecord(e0.e, s));
  int rc = PHPC_OK;
  if (k > 0) {
    if (vendor) {
      rc = phpc_dgemm_vendor_device(dA, k, dB, n, dC, n, m, k, n, s);
    } else {
      const LaunchPolicy &pol = phpc_policy();
      int G = pol.honor_grid ? grid_width * grid_height : 0;
      rc = phpc_dgemm_launch(dA, k, dB, n, dC, n, m, k, n, 0, G, s);
    }
  }
  PHPC_HIP(hipEventRecord(e1.e, s));
  if (rc == PHPC_OK) rc = phpc_stage_d2h(c, (size_t)ldc * 8, dC, (size_t)n * 8, (size_t)n * 8, m, s);
  PHPC_HIP(hipStreamSynchronize(s));
  if (rc) return rc;
  float ms = 0;
  PHPC_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
  if (compute_time) *compute_time = vendor ? 0.0f : ms / 1000.0f;
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" void phpc_gemm_cuda(const double *a, int lda, const double *b, int ldb, double *c, int ldc, int m, int k,
                               int n, int gpu_count, int grid_width, int grid_height, int block_width,
                               float *compute_time) {
  gemm_cuda_impl(a, lda, b, ldb, c, ldc, m, k, n, gpu_count, grid_width, grid_height, block_width, compute_time, false);
}

extern "C" void phpc_gemm_cublas(const double *a, int lda, const double *b, int ldb, double *c, int ldc, int m, int k,
                                 int n, int gpu_count, int grid_width, int grid_height, int block_width,
                                 float *gpu_time) {
  gemm_cuda_impl(a, lda, b, ldb, c, ldc, m, k, n, gpu_count, grid_width, grid_height, block_width, gpu_time, true);
}

// ---------------------------------------------------------------------------
// vendor comparator (rocBLAS): row-major C += A*B as column-major C^T += B^T A^T
// (the same trick as src/phpc_gemm.cu:169-170)
// ---------------------------------------------------------------------------
namespace {
std::mutex g_blas_mu;
rocblas_handle g_blas = nullptr;
}  // namespace

extern "C" int phpc_dgemm_vendor_device(const double *dA, long lda, const double *dB, long ldb, double *dC, long ldc,
                                        int m, int k, int n, void *stream) {
  if (phpc_init()) return phpc_last_error();
  if (check_gemm_args(dA, lda, dB, ldb, dC, ldc, m, k, n)) return phpc_last_error();
  if (m == 0 || n == 0 || k == 0) return phpc_set_error(PHPC_OK, nullptr);
  std::lock_guard<std::mutex> lk(g_blas_mu);
  if (!g_blas && rocblas_create_handle(&g_blas) != rocblas_status_success) {
    g_blas = nullptr;
    return phpc_set_error(PHPC_ERR_VENDOR, "rocblas_create_handle failed");
  }
  if (rocblas_set_stream(g_blas, (hipStream_t)stream) != rocblas_status_success)
    return phpc_set_error(PHPC_ERR_VENDOR, "rocblas_set_stream failed");
  const double one = 1.0;
  rocblas_status st = rocblas_dgemm(g_blas, rocblas_operation_none, rocblas_operation_none, n, m, k, &one, dB,
                                    (rocblas_int)ldb, dA, (rocblas_int)lda, &one, dC, (rocblas_int)ldc);
  if (st != rocblas_status_success) return phpc_set_error(PHPC_ERR_VENDOR, "rocblas_dgemm failed (%d)", (int)st);
  return phpc_set_error(PHPC_OK, nullptr);
}

// ---------------------------------------------------------------------------
// seeded generators
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ void fill_kernel(double *x, long ld, int m, int n, int dist, uint64_t seed) {
  const long total = (long)m * n;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long r = e / n, c = e % n;
    double v;
    if (dist == 2) {
      v = (double)e;
    } else {
      const uint64_t z = mix64(seed + (uint64_t)(e + 1) * 0x9E3779B97F4A7C15ULL);
      if (dist == 0) {
        v = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
      } else {
        const int ex = (int)(((z >> 32) * 41ULL) >> 32) - 20;
        v = ldexp((z & 1) ? -1.0 : 1.0, ex);
      }
    }
    x[r * ld + c] = v;
  }
}

extern "C" int phpc_fill_device(double *dX, long ld, int m, int n, int dist, uint64_t seed, void *stream) {
  if (phpc_init()) return phpc_last_error();
  if (m < 0 || n < 0 || ld < n || (m > 0 && n > 0 && !dX) || dist < 0 || dist > 2)
    return phpc_set_error(PHPC_ERR_ARG, "bad fill arguments");
  if ((long)m * n == 0) return phpc_set_error(PHPC_OK, nullptr);
  const long total = (long)m * n;
  long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, dX, ld, m, n, dist, seed);
  PHPC_HIP(hipGetLastError());
  return phpc_set_error(PHPC_OK, nullptr);
}

__global__ void fill_f32_kernel(float *x, long ld, int m, int n, int dist, uint64_t seed) {
  const long total = (long)m * n;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long r = e / n, c = e % n;
    double v;
    if (dist == 2) {
      v = (double)e;
    } else {
      const uint64_t z = mix64(seed + (uint64_t)(e + 1) * 0x9E3779B97F4A7C15ULL);
      if (dist == 0) {
        v = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
      } else {
        const int ex = (int)(((z >> 32) * 41ULL) >> 32) - 20;
        v = ldexp((z & 1) ? -1.0 : 1.0, ex);
      }
    }
    x[r * ld + c] = (float)v;
  }
}

extern "C" int phpc_fill_device_f32(float *dX, long ld, int m, int n, int dist, uint64_t seed, void *stream) {
  if (phpc_init()) return phpc_last_error();
  if (m < 0 || n < 0 || ld < n || (m > 0 && n > 0 && !dX) || dist < 0 || dist > 2)
    return phpc_set_error(PHPC_ERR_ARG, "bad fill arguments");
  if ((long)m * n == 0) return phpc_set_error(PHPC_OK, nullptr);
  const long total = (long)m * n;
  long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, dX, ld, m, n, dist,
                     seed);
  PHPC_HIP(hipGetLastError());
  return phpc_set_error(PHPC_OK, nullptr);
}

// ---------------------------------------------------------------------------
// strided block pack (the MPI_Type_vector panels of src/phpc_summa.c:53-56)
// ---------------------------------------------------------------------------
__global__ void pack_kernel(const double *__restrict__ src, long lds, double *__restrict__ dst, long ldd, int rows,
                            int cols) {
  for (int r = blockIdx.y; r < rows; r += gridDim.y) {
    const double *s = src + (long)r * lds;
    double *d = dst + (long)r * ldd;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += gridDim.x * blockDim.x) d[c] = s[c];
  }
}

int phpc_pack_launch(const double *src, long lds, double *dst, long ldd, int rows, int cols, hipStream_t stream) {
  if (rows <= 0 || cols <= 0) return PHPC_OK;
  int bx = (cols + 255) / 256;
  if (bx > 64) bx = 64;
  int by = rows < 65535 ? rows : 65535;
  hipLaunchKernelGGL(pack_kernel, dim3(bx, by), dim3(256), 0, stream, src, lds, dst, ldd, rows, cols);
  PHPC_HIP(hipGetLastError());
  return PHPC_OK;
}

extern "C" int phpc_pack_device(const double *src, long ld_src, double *dst, long ld_dst, int rows, int cols,
                                void *stream) {
  if (phpc_init()) return phpc_last_error();
  if (rows < 0 || cols < 0 || ld_src < cols || ld_dst < cols || (rows > 0 && cols > 0 && (!src || !dst)))
    return phpc_set_error(PHPC_ERR_ARG, "bad block descriptor");
  int rc = phpc_pack_launch(src, ld_src, dst, ld_dst, rows, cols, (hipStream_t)stream);
  return rc ? rc : phpc_set_error(PHPC_OK, nullptr);
}

// ---------------------------------------------------------------------------
// sampled double-double accuracy check (one 64-lane workgroup per sampled entry)
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)
__device__ __forceinline__ void dd_two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  const double z = s - a;
  e = (a - (s - z)) + (b - z);
}
__device__ __forceinline__ void dd_add(double &hi, double &lo, double h2, double l2) {
  double s, e;
  dd_two_sum(hi, h2, s, e);
  e = e + (lo + l2);
  hi = s + e;
  lo = e - (hi - s);
}

__global__ __launch_bounds__(64) void check_kernel(const double *A, long lda, const double *B, long ldb,
                                                   const double *C0, long ldc0, const double *C, long ldc, int k,
                                                   const int *rows, const int *cols, double *out_cw,
                                                   double *out_absd, double *out_absr) {
  const int e = blockIdx.x;
  const int i = rows[e], j = cols[e];
  const int lane = threadIdx.x;
  double hi = 0.0, lo = 0.0, ab = 0.0;
  for (int kk = lane; kk < k; kk += 64) {
    const double a = A[(long)i * lda + kk];
    const double b = B[(long)kk * ldb + j];
    const double p = a * b;
    const double pe = __builtin_fma(a, b, -p);
    double s, se;
    dd_two_sum(hi, p, s, se);
    hi = s;
    lo = lo + (se + pe);
    ab = ab + fabs(a) * fabs(b);
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double h2 = __shfl_down(hi, off, 64);
    const double l2 = __shfl_down(lo, off, 64);
    const double a2 = __shfl_down(ab, off, 64);
    dd_add(hi, lo, h2, l2);
    ab = ab + a2;
  }
  if (lane == 0) {
    double c0 = C0 ? C0[(long)i * ldc0 + j] : 0.0;
    dd_add(hi, lo, c0, 0.0);
    const double ref = hi + lo;
    const double got = C[(long)i * ldc + j];
    const double d = fabs(got - ref);
    const double den = ab + fabs(c0);
    out_cw[e] = d == 0.0 ? 0.0 : (den > 0.0 ? d / den : INFINITY);
    out_absd[e] = d;
    out_absr[e] = fabs(ref);
  }
}
#pragma clang fp contract(on)

extern "C" int phpc_check_sampled(const double *dA, long lda, const double *dB, long ldb, const double *dC0, long ldc0,
                                  const double *dC, long ldc, int m, int k, int n, const int *rows, const int *cols,
                                  int count, double *err_cw, double *err_nw, void *stream) {
  if (phpc_init()) return phpc_last_error();
  if (count <= 0 || !rows || !cols || !dA || !dB || !dC || m < 1 || n < 1 || k < 0 || lda < k || ldb < n ||
      ldc < n || (dC0 && ldc0 < n))
    return phpc_set_error(PHPC_ERR_ARG, "bad check arguments");
  for (int e = 0; e < count; ++e)
    if (rows[e] < 0 || rows[e] >= m || cols[e] < 0 || cols[e] >= n)
      return phpc_set_error(PHPC_ERR_ARG, "sample %d = (%d, %d) outside %d x %d", e, rows[e], cols[e], m, n);
  hipStream_t s = (hipStream_t)stream;
  int *didx = nullptr;
  double *dout = nullptr;
  std::vector<double> h(3 * (size_t)count);
  // every exit after the first allocation frees both buffers (the stream is drained
  // first, so no launched kernel still reads or writes them)
  auto fail = [&](hipError_t e, const char *what) {
    (void)hipStreamSynchronize(s);
    if (didx) (void)hipFree(didx);
    if (dout) (void)hipFree(dout);
    return phpc_set_error(PHPC_ERR_HIP, "check_sampled: %s failed: %s", what, hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMalloc(&didx, sizeof(int) * 2 * count)) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&dout, sizeof(double) * 3 * count)) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMemcpyAsync(didx, rows, sizeof(int) * count, hipMemcpyHostToDevice, s)) != hipSuccess ||
      (e = hipMemcpyAsync(didx + count, cols, sizeof(int) * count, hipMemcpyHostToDevice, s)) != hipSuccess)
    return fail(e, "sample copy");
  hipLaunchKernelGGL(check_kernel, dim3(count), dim3(64), 0, s, dA, lda, dB, ldb, dC0, ldc0, dC, ldc, k, didx,
                     didx + count, dout, dout + count, dout + 2 * count);
  if ((e = hipGetLastError()) != hipSuccess) return fail(e, "check_kernel launch");
  if ((e = hipMemcpyAsync(h.data(), dout, sizeof(double) * 3 * count, hipMemcpyDeviceToHost, s)) != hipSuccess)
    return fail(e, "result copy");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e, "hipStreamSynchronize");
  (void)hipFree(didx);
  (void)hipFree(dout);
  double cw = 0, md = 0, mr = 0;
  for (int e = 0; e < count; ++e) {
    if (h[e] > cw || std::isnan(h[e])) cw = std::isnan(h[e]) ? INFINITY : h[e];
    if (h[count + e] > md) md = h[count + e];
    if (h[2 * count + e] > mr) mr = h[2 * count + e];
  }
  if (err_cw) *err_cw = cw;
  if (err_nw) *err_nw = mr > 0 ? md / mr : md;
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_check_sampled_host(const double *A, long lda, const double *B, long ldb, const double *C, long ldc,
                                       int m, int k, int n, int count, uint64_t seed, double *err_cw, double *err_nw) {
  if (phpc_init()) return phpc_last_error();
  if (!A || !B || !C || m < 1 || n < 1 || k < 1 || lda < k || ldb < n || ldc < n || count < 1)
    return phpc_set_error(PHPC_ERR_ARG, "bad check arguments");
  // sampled entries from the seed (SplitMix64 stream), plus the four corners
  std::vector<int> rows, cols;
  const int corners[4][2] = {{0, 0}, {0, n - 1}, {m - 1, 0}, {m - 1, n - 1}};
  for (auto &c : corners) {
    rows.push_back(c[0]);
    cols.push_back(c[1]);
  }
  for (int e = 0; e < count; ++e) {
    uint64_t z = seed + (uint64_t)(e + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    rows.push_back((int)((z >> 32) % (uint64_t)m));
    cols.push_back((int)((z & 0xffffffffu) % (uint64_t)n));
  }
  double *dA = nullptr, *dB = nullptr, *dC = nullptr;
  int rc = PHPC_OK;
  if (hipMalloc(&dA, sizeof(double) * (size_t)m * k) != hipSuccess ||
      hipMalloc(&dB, sizeof(double) * (size_t)k * n) != hipSuccess ||
      hipMalloc(&dC, sizeof(double) * (size_t)m * n) != hipSuccess) {
    rc = phpc_set_error(PHPC_ERR_MEM, "check buffers");
  }
  if (!rc && (hipMemcpy2D(dA, sizeof(double) * k, A, sizeof(double) * lda, sizeof(double) * k, m,
                          hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy2D(dB, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n, k,
                          hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy2D(dC, sizeof(double) * n, C, sizeof(double) * ldc, sizeof(double) * n, m,
                          hipMemcpyHostToDevice) != hipSuccess))
    rc = phpc_set_error(PHPC_ERR_HIP, "check copies");
  if (!rc)
    rc = phpc_check_sampled(dA, k, dB, n, nullptr, 0, dC, n, m, k, n, rows.data(), cols.data(), (int)rows.size(),
                            err_cw, err_nw, nullptr);
  if (dA) (void)hipFree(dA);
  if (dB) (void)hipFree(dB);
  if (dC) (void)hipFree(dC);
  return rc;
}

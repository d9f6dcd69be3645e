// Host <-> device staging through pinned memory.
//
// The drop-in entry points take pageable host pointers, as the reference's do
// (src/phpc_gemm.cu:59-156). A pageable hipMemcpy goes through the runtime's own
// staging at a fraction of the link rate. Here a small pool of host threads copies
// row chunks between the caller's memory and a ring of pinned buffers while the DMA
// engine moves the previous chunk, so host copies, PCIe transfers and (in the callers)
// kernels overlap.
#include "phpc_internal.h"

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace {

// Parallel for over [0, n) with a fixed set of worker threads; the caller participates.
class CopyPool {
 public:
  explicit CopyPool(int nthreads) {
    for (int i = 0; i < nthreads; ++i) th_.emplace_back([this] { work(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  void run(int n, const std::function<void(int)> &fn) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      fn_ = &fn;
      n_ = n;
      next_ = 0;
      left_ = n;
      ++gen_;
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return left_ == 0; });
    fn_ = nullptr;
  }

 private:
  void drain() {
    for (;;) {
      int i;
      const std::function<void(int)> *fn;
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (!fn_ || next_ >= n_) return;
        i = next_++;
        fn = fn_;
      }
      (*fn)(i);
      std::lock_guard<std::mutex> lk(mu_);
      if (--left_ == 0) done_.notify_all();
    }
  }
  void work() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      drain();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)> *fn_ = nullptr;
  int n_ = 0, next_ = 0, left_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

constexpr int kSlots = 4;
constexpr size_t kSlotBytes = 32u << 20;

struct Stager {
  std::mutex mu;  // one staged copy at a time
  bool ready = false;
  char *slot[kSlots] = {};
  hipEvent_t done[kSlots] = {};
  CopyPool *pool = nullptr;

  int init() {
    if (ready) return PHPC_OK;
    for (int i = 0; i < kSlots; ++i) {
      if (hipHostMalloc((void **)&slot[i], kSlotBytes, hipHostMallocDefault) != hipSuccess)
        return phpc_set_error(PHPC_ERR_MEM, "pinned staging allocation failed");
      if (hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess)
        return phpc_set_error(PHPC_ERR_HIP, "staging event creation failed");
    }
    const unsigned hw = std::thread::hardware_concurrency();
    pool = new CopyPool((int)std::min(7u, hw > 1 ? hw - 1 : 1u));
    ready = true;
    return PHPC_OK;
  }
};
Stager g_stage;

// copy rows [r0, r1) of width bytes between a strided region and a packed slot, in
// parallel pieces of about 1 MiB
void copy_rows(char *dst, size_t dpitch, const char *src, size_t spitch, size_t width, size_t rows) {
  const size_t per = std::max<size_t>(1, (1u << 20) / std::max<size_t>(width, 1));
  const int pieces = (int)((rows + per - 1) / per);
  auto fn = [&](int p) {
    const size_t a = (size_t)p * per, b = std::min(rows, a + per);
    if (dpitch == width && spitch == width) {
      memcpy(dst + a * width, src + a * width, (b - a) * width);
      return;
    }
    for (size_t r = a; r < b; ++r) memcpy(dst + r * dpitch, src + r * spitch, width);
  };
  if (pieces <= 1) {
    fn(0);
    return;
  }
  g_stage.pool->run(pieces, fn);
}

}  // namespace

// dst (device, dpitch bytes) <- src (host, spitch bytes), rows x width bytes, on stream.
// Returns once src may be reused; the transfers complete in stream order.
int phpc_stage_h2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t rows,
                   hipStream_t stream) {
  if (rows == 0 || width == 0) return PHPC_OK;
  std::lock_guard<std::mutex> lk(g_stage.mu);
  if (int rc = g_stage.init()) return rc;
  if (width > kSlotBytes) {  // a row longer than a slot: let the runtime stage it
    PHPC_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyHostToDevice, stream));
    PHPC_HIP(hipStreamSynchronize(stream));
    return PHPC_OK;
  }
  const size_t per = kSlotBytes / width;
  for (size_t r0 = 0, c = 0; r0 < rows; r0 += per, ++c) {
    const size_t n = std::min(per, rows - r0);
    const int sl = (int)(c % kSlots);
    PHPC_HIP(hipEventSynchronize(g_stage.done[sl]));  // the slot's previous transfer is done
    copy_rows(g_stage.slot[sl], width, (const char *)src + r0 * spitch, spitch, width, n);
    PHPC_HIP(hipMemcpy2DAsync((char *)dst + r0 * dpitch, dpitch, g_stage.slot[sl], width, width, n,
                              hipMemcpyHostToDevice, stream));
    PHPC_HIP(hipEventRecord(g_stage.done[sl], stream));
  }
  return PHPC_OK;
}

// dst (host, dpitch bytes) <- src (device, spitch bytes), rows x width bytes, after the
// work already queued on stream. Returns once dst holds the data.
int phpc_stage_d2h(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t rows,
                   hipStream_t stream) {
  if (rows == 0 || width == 0) return PHPC_OK;
  std::lock_guard<std::mutex> lk(g_stage.mu);
  if (int rc = g_stage.init()) return rc;
  if (width > kSlotBytes) {
    PHPC_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyDeviceToHost, stream));
    PHPC_HIP(hipStreamSynchronize(stream));
    return PHPC_OK;
  }
  const size_t per = kSlotBytes / width;
  const size_t nchunks = (rows + per - 1) / per;
  // DMA chunk c while the host copies chunk c - 1 out of its slot
  auto host_out = [&](size_t c) -> int {
    const size_t r0 = c * per, n = std::min(per, rows - r0);
    const int sl = (int)(c % kSlots);
    PHPC_HIP(hipEventSynchronize(g_stage.done[sl]));
    copy_rows((char *)dst + r0 * dpitch, dpitch, g_stage.slot[sl], width, width, n);
    return PHPC_OK;
  };
  for (size_t c = 0; c < nchunks; ++c) {
    const size_t r0 = c * per, n = std::min(per, rows - r0);
    const int sl = (int)(c % kSlots);
    if (c >= (size_t)kSlots) {  // the slot still holds chunk c - kSlots: drain it first
      if (int rc = host_out(c - kSlots)) return rc;
    }
    PHPC_HIP(hipMemcpy2DAsync(g_stage.slot[sl], width, (const char *)src + r0 * spitch, spitch, width, n,
                              hipMemcpyDeviceToHost, stream));
    PHPC_HIP(hipEventRecord(g_stage.done[sl], stream));
  }
  for (size_t c = nchunks > (size_t)kSlots ? nchunks - kSlots : 0; c < nchunks; ++c)
    if (int rc = host_out(c)) return rc;
  return PHPC_OK;
}

// GEMM serving loop.
//
// Requests wait in one queue. A worker (one host thread and one HIP stream per slot)
// takes the next request (the oldest, or under shortest-first the one with the least
// estimated device time less an aging credit) as soon as its operands fit in the
// operand arena, a single device allocation sized to the memory budget that is
// sub-allocated at admission, so no driver allocation sits on the request path. Requests too small to fill the GPU on
// their own (fewer output tiles than compute units) are batched: requests of the
// head's kind and load width that are waiting join its launch, one persistent grouped
// kernel over all their tiles. Large requests run alone, since they fill the GPU by
// themselves. Workers run concurrently, so a batch being prepared or finished overlaps
// the one computing. Latency is measured from submission to completion.
#include "phpc_internal.h"
#include "phpc_serve.h"
#include "arena.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;
using phpc::OffsetArena;
constexpr size_t kNone = OffsetArena::kNone;

size_t round_up(size_t x) { return OffsetArena::round_up(x); }

// One device allocation holding every admitted request's operands.
struct Arena {
  char *base = nullptr;
  size_t cap = 0;
  OffsetArena offsets;

  int init(size_t bytes) {
    cap = offsets.reset(bytes);
    if (cap == 0) return PHPC_ERR_BUDGET;
    if (hipMalloc(&base, cap) != hipSuccess) {
      base = nullptr;
      cap = 0;
      return PHPC_ERR_MEM;
    }
    return PHPC_OK;
  }
  size_t alloc(size_t bytes) { return offsets.alloc(bytes); }
  void release(size_t off, size_t bytes) { offsets.release(off, bytes); }
  void destroy() {
    if (base) (void)hipFree(base);
    base = nullptr;
    cap = 0;
    offsets.reset(0);
  }
};

// How a request is computed. kF32 and kF64Plain requests that do not fill the GPU are
// batched with waiting requests of the same kind and load width.
enum Kind { kF32 = 0, kF64Plain = 1, kF64Split = 2 };

struct Req {
  uint64_t id;
  phpc_request r;
  Kind kind;
  bool small;  // fewer output tiles than compute units: batchable
  bool heavy;  // counts against the limit on concurrent large requests
  bool wide;   // dimensions allow the 16-byte load path
  double est;  // estimated device time (s), for shortest-first scheduling
  size_t bytes[3];
  size_t off[3] = {kNone, kNone, kNone};  // A, B, C in the arena
  clk::time_point arrive, start, end;
  phpc_result res{};
  bool done = false;
};

struct Worker {
  hipStream_t stream = nullptr;
  Workspace ws;
  void *desc_dev = nullptr;   // batch descriptors (device)
  void *desc_host = nullptr;  // pinned staging copy
  size_t desc_cap = 0;
  hipEvent_t ev[3] = {};
  std::thread th;
};

}  // namespace

struct phpc_server {
  Arena arena;
  size_t reserved = 0;  // per-slot workspaces
  size_t in_use = 0, peak = 0;
  int batch_max = 16;
  int max_dim = 0;       // the per-slot workspaces are sized for m, n, k up to this
  int split_min = 4096;  // FP64 class: split CRT GEMM when min(m, n, k) >= split_min
  int policy = PHPC_SCHED_FIFO;
  double aging = 0.0;  // shortest-first: priority = est - wait / aging (0: no aging)
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req *> queue;
  std::unordered_map<uint64_t, Req *> all;
  uint64_t next_id = 1;
  int pending = 0;
  int big_running = 0;  // requests that fill the GPU alone, in flight
  int big_limit = 1;    // at most this many of them at once
  double heavy_s = 0.0;  // requests estimated shorter than this are not limited
  bool stop = false;
  std::vector<Worker *> workers;

  std::deque<Req *>::iterator pick_locked();
  bool admit_locked(Req *q);
  void release_locked(Req *q);
  void run(Worker *w);
  void process(Worker *w, std::vector<Req *> &batch);
};

// The next request to run, or queue.end() if none may start now: the oldest (FIFO), or
// the one with the least estimated device time less its waiting time divided by the
// aging factor (shortest first). Requests that fill the GPU alone start only while fewer
// than big_limit of them run (each one beyond the first only shares the GPU); small
// requests may start beside them. LANES: a free place among the GPU-filling requests
// goes to the shortest of them first; everything else runs in arrival order.
std::deque<Req *>::iterator phpc_server::pick_locked() {
  const clk::time_point now = clk::now();
  auto key_of = [&](const Req *q) {
    double key = q->est;
    if (aging > 0.0) key -= std::chrono::duration<double>(now - q->arrive).count() / aging;
    return key;
  };
  auto best = queue.end();
  double best_key = 0.0;
  if (policy == PHPC_SCHED_LANES) {
    auto first_light = queue.end();
    for (auto it = queue.begin(); it != queue.end(); ++it) {
      if (!(*it)->heavy) {
        if (first_light == queue.end()) first_light = it;
        continue;
      }
      if (big_running >= big_limit) continue;
      const double key = key_of(*it);
      if (best == queue.end() || key < best_key) {
        best = it;
        best_key = key;
      }
    }
    return best != queue.end() ? best : first_light;
  }
  for (auto it = queue.begin(); it != queue.end(); ++it) {
    if (big_running >= big_limit && (*it)->heavy) continue;
    if (policy == PHPC_SCHED_FIFO) return it;
    const double key = key_of(*it);
    if (best == queue.end() || key < best_key) {
      best = it;
      best_key = key;
    }
  }
  return best;
}

bool phpc_server::admit_locked(Req *q) {
  for (int i = 0; i < 3; ++i) {
    q->off[i] = arena.alloc(q->bytes[i]);
    if (q->off[i] == kNone) {
      release_locked(q);
      return false;
    }
  }
  for (int i = 0; i < 3; ++i) in_use += round_up(q->bytes[i]);
  peak = std::max(peak, in_use);
  return true;
}

void phpc_server::release_locked(Req *q) {
  for (int i = 0; i < 3; ++i)
    if (q->off[i] != kNone) {
      arena.release(q->off[i], q->bytes[i]);
      q->off[i] = kNone;
    }
}

void phpc_server::process(Worker *w, std::vector<Req *> &batch) {
  hipStream_t s = w->stream;
  int status = PHPC_OK;
  auto ptr = [&](Req *q, int i) { return (void *)(arena.base + q->off[i]); };
  (void)hipEventRecord(w->ev[0], s);
  for (Req *q : batch) {
    const phpc_request &r = q->r;
    if (r.accuracy == PHPC_ACC_FP64) {
      status = phpc_fill_device((double *)ptr(q, 0), r.k, r.m, r.k, r.dist, r.seed, s);
      if (!status) status = phpc_fill_device((double *)ptr(q, 1), r.n, r.k, r.n, r.dist, r.seed + 1, s);
    } else {
      status = phpc_fill_device_f32((float *)ptr(q, 0), r.k, r.m, r.k, r.dist, r.seed, s);
      if (!status) status = phpc_fill_device_f32((float *)ptr(q, 1), r.n, r.k, r.n, r.dist, r.seed + 1, s);
    }
    if (!status && hipMemsetAsync(ptr(q, 2), 0, q->bytes[2], s) != hipSuccess) status = PHPC_ERR_HIP;
    if (status) break;
  }
  (void)hipEventRecord(w->ev[1], s);
  const Kind kind = batch[0]->kind;
  const int nb = (int)batch.size();
  if (status == PHPC_OK && kind == kF64Split) {
    const phpc_request &r = batch[0]->r;
    phpc_split_info info{};
    status = phpc_dgemm_split_ws((const double *)ptr(batch[0], 0), r.k, (const double *)ptr(batch[0], 1), r.n,
                                 (double *)ptr(batch[0], 2), r.n, r.m, r.k, r.n, 0, 1e-14, PHPC_SPLIT_CRT, s, &w->ws,
                                 &info);
    batch[0]->res.error_estimate = info.error_estimate;
  } else if (status == PHPC_OK && kind == kF64Plain) {
    GemmDesc *gd = (GemmDesc *)w->desc_host;
    for (int i = 0; i < nb; ++i) {
      const phpc_request &r = batch[i]->r;
      gd[i] = GemmDesc{};
      gd[i].A = (const double *)ptr(batch[i], 0);
      gd[i].B = (const double *)ptr(batch[i], 1);
      gd[i].C = (double *)ptr(batch[i], 2);
      gd[i].lda = r.k;
      gd[i].ldb = r.n;
      gd[i].ldc = r.n;
      gd[i].m = r.m;
      gd[i].n = r.n;
      gd[i].k = r.k;
      gd[i].panel = 0;
    }
    const int cfg = phpc_f64_pick_config(gd, nb);
    phpc_f64_plan_tiles(gd, nb, cfg);
    if (nb > 1 && hipMemcpyAsync(w->desc_dev, gd, sizeof(GemmDesc) * nb, hipMemcpyHostToDevice, s) != hipSuccess)
      status = PHPC_ERR_HIP;
    if (status == PHPC_OK) status = phpc_f64_grouped_dev((const GemmDesc *)w->desc_dev, gd, nb, 0, cfg, s);
  } else if (status == PHPC_OK && !batch[0]->small && (long)batch[0]->r.m * batch[0]->r.n >= 2048L * 2048 &&
             batch[0]->r.k >= 256) {
    // a large FP32-class request: FP32-accurate 3 x BF16 slice planes, alone
    const phpc_request &r = batch[0]->r;
    status = phpc_sgemm_bf16x3_planes((const float *)ptr(batch[0], 0), r.k, (const float *)ptr(batch[0], 1), r.n,
                                      (float *)ptr(batch[0], 2), r.n, r.m, r.k, r.n, s, &w->ws);
  } else if (status == PHPC_OK) {
    F32Desc *fd = (F32Desc *)w->desc_host;
    long work = 0;
    for (int i = 0; i < nb; ++i) {
      const phpc_request &r = batch[i]->r;
      fd[i] = F32Desc{};
      fd[i].A = (const float *)ptr(batch[i], 0);
      fd[i].B = (const float *)ptr(batch[i], 1);
      fd[i].C = (float *)ptr(batch[i], 2);
      fd[i].lda = r.k;
      fd[i].ldb = r.n;
      fd[i].ldc = r.n;
      fd[i].m = r.m;
      fd[i].n = r.n;
      fd[i].k = r.k;
      work += (long)((r.m + 127) / 128) * ((r.n + 127) / 128);
    }
    const int cfg = work >= 2L * phpc_device().cus ? 2 : 0;
    const int tiles = phpc_f32_plan_tiles(fd, nb, cfg);
    if (nb > 1 && hipMemcpyAsync(w->desc_dev, fd, sizeof(F32Desc) * nb, hipMemcpyHostToDevice, s) != hipSuccess)
      status = PHPC_ERR_HIP;
    if (status == PHPC_OK)
      status = phpc_f32_launch((const F32Desc *)w->desc_dev, fd, nb, tiles, cfg, phpc_f32_vec4_ok(fd, nb), 0, s);
  }
  (void)hipEventRecord(w->ev[2], s);
  for (int i = 0; i < nb && status == PHPC_OK; ++i) {
    const phpc_request &r = batch[i]->r;
    if (r.result_host &&
        hipMemcpyAsync(r.result_host, ptr(batch[i], 2), batch[i]->bytes[2], hipMemcpyDeviceToHost, s) != hipSuccess)
      status = PHPC_ERR_HIP;
  }
  if (hipStreamSynchronize(s) != hipSuccess && status == PHPC_OK) status = PHPC_ERR_HIP;
  const clk::time_point end = clk::now();
  float fill_ms = 0.f, gemm_ms = 0.f;
  (void)hipEventElapsedTime(&fill_ms, w->ev[0], w->ev[1]);
  (void)hipEventElapsedTime(&gemm_ms, w->ev[1], w->ev[2]);
  for (Req *q : batch) {
    q->end = end;
    q->res.status = status;
    q->res.batch_size = nb;
    q->res.latency_s = std::chrono::duration<double>(q->end - q->arrive).count();
    q->res.queue_s = std::chrono::duration<double>(q->start - q->arrive).count();
    q->res.compute_s = std::chrono::duration<double>(q->end - q->start).count();
    q->res.fill_s = fill_ms * 1e-3;
    q->res.gemm_s = gemm_ms * 1e-3;
  }
}

void phpc_server::run(Worker *w) {
  (void)hipSetDevice(0);
  for (;;) {
    std::vector<Req *> batch;
    {
      std::unique_lock<std::mutex> lk(mu);
      Req *h = nullptr;
      for (;;) {
        auto it = pick_locked();
        if (it != queue.end() && admit_locked(*it)) {
          h = *it;
          queue.erase(it);
          if (h->heavy) ++big_running;
          break;
        }
        if (stop) return;  // nothing admissible left for this worker
        cv_work.wait(lk);
      }
      // small requests of the head's kind and load width that are waiting join its
      // launch, in arrival order, while their operands fit
      batch.push_back(h);
      if (h->small && h->kind != kF64Split) {
        for (auto it = queue.begin(); it != queue.end() && (int)batch.size() < batch_max;) {
          Req *q = *it;
          if (q->small && q->kind == h->kind && q->wide == h->wide && admit_locked(q)) {
            batch.push_back(q);
            it = queue.erase(it);
          } else {
            ++it;
          }
        }
      }
      const clk::time_point now = clk::now();
      for (Req *q : batch) q->start = now;
    }
    process(w, batch);
    {
      std::lock_guard<std::mutex> lk(mu);
      if (batch[0]->heavy) --big_running;
      for (Req *q : batch) {
        for (int i = 0; i < 3; ++i) in_use -= round_up(q->bytes[i]);
        release_locked(q);
        q->done = true;
      }
      pending -= (int)batch.size();
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

extern "C" int phpc_server_create(size_t budget_bytes, int slots, int max_dim, int batch_max, phpc_server_t *out) {
  if (!out) return phpc_set_error(PHPC_ERR_ARG, "null server pointer");
  *out = nullptr;
  if (phpc_init()) return phpc_last_error();
  if (slots < 1 || slots > 64 || max_dim < 1 || batch_max < 1 || batch_max > 4096)
    return phpc_set_error(PHPC_ERR_ARG, "bad server arguments");
  phpc_server *srv = new phpc_server();
  srv->batch_max = batch_max;
  srv->max_dim = max_dim;
  // per-slot split workspace for max_dim^3 at up to 20 moduli: two int8 operand copies,
  // 20 residue planes, metadata
  const size_t kp = (size_t)phpc_split_kp(max_dim);
  const size_t desc_bytes = (size_t)batch_max * std::max(sizeof(GemmDesc), sizeof(F32Desc));
  int rc = PHPC_OK;
  for (int i = 0; i < slots && rc == PHPC_OK; ++i) {
    Worker *w = new Worker();
    srv->workers.push_back(w);
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) {
      rc = phpc_set_error(PHPC_ERR_HIP, "stream creation failed");
      break;
    }
    for (auto &e : w->ev)
      if (hipEventCreate(&e) != hipSuccess) rc = phpc_set_error(PHPC_ERR_HIP, "event creation failed");
    if (rc) break;
    const size_t m = (size_t)max_dim, n = (size_t)max_dim;
    if (!w->ws.get(m * kp * 20, 0, w->stream) || !w->ws.get(n * kp * 20, 1, w->stream) ||
        !w->ws.get(20 * m * n, 4, w->stream) || !w->ws.get((m + n) * 12 + n * 16 + 64, 3, w->stream) ||
        hipMalloc(&w->desc_dev, desc_bytes) != hipSuccess ||
        hipHostMalloc(&w->desc_host, desc_bytes, hipHostMallocDefault) != hipSuccess) {
      rc = phpc_set_error(PHPC_ERR_MEM, "workspace allocation failed");
      break;
    }
    w->desc_cap = desc_bytes;
    srv->reserved += w->ws.total() + desc_bytes;
  }
  if (rc == PHPC_OK && srv->reserved >= budget_bytes)
    rc = phpc_set_error(PHPC_ERR_BUDGET, "budget %zu bytes is below the workspaces (%zu bytes)", budget_bytes,
                        srv->reserved);
  if (rc == PHPC_OK) {
    const int a = srv->arena.init(budget_bytes - srv->reserved);
    if (a) rc = phpc_set_error(a, "operand arena of %zu bytes could not be allocated", budget_bytes - srv->reserved);
  }
  if (rc) {
    phpc_server_destroy(srv);
    return rc;
  }
  for (Worker *w : srv->workers) w->th = std::thread([srv, w] { srv->run(w); });
  *out = srv;
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_server_submit(phpc_server_t srv, const phpc_request *req, uint64_t *ticket) {
  if (!srv || !req || !ticket) return phpc_set_error(PHPC_ERR_ARG, "bad submit arguments");
  if ((req->accuracy != PHPC_ACC_FP64 && req->accuracy != PHPC_ACC_FP32) || req->m < 1 || req->n < 1 || req->k < 1 ||
      req->dist < 0 || req->dist > 2)
    return phpc_set_error(PHPC_ERR_ARG, "bad request");
  const phpc_request &r = *req;
  // the per-slot workspaces were allocated once, for dimensions up to max_dim; a wider
  // request would grow them on the request path, outside the budget
  if (std::max(r.m, std::max(r.n, r.k)) > srv->max_dim)
    return phpc_set_error(PHPC_ERR_BUDGET, "request %d x %d x %d exceeds the server's max_dim %d", r.m, r.k, r.n,
                          srv->max_dim);
  Req *q = new Req();
  q->r = r;
  const size_t e = r.accuracy == PHPC_ACC_FP64 ? 8 : 4;
  q->bytes[0] = e * (size_t)r.m * r.k;
  q->bytes[1] = e * (size_t)r.k * r.n;
  q->bytes[2] = e * (size_t)r.m * r.n;
  const size_t total = round_up(q->bytes[0]) + round_up(q->bytes[1]) + round_up(q->bytes[2]);
  const long cus = phpc_device().cus;
  q->arrive = clk::now();
  std::lock_guard<std::mutex> lk(srv->mu);
  if (r.accuracy == PHPC_ACC_FP32) {
    q->kind = kF32;
    q->small = (long)((r.m + 127) / 128) * ((r.n + 127) / 128) < cus;
    q->wide = r.k % 4 == 0 && r.n % 4 == 0;
  } else {
    q->kind = std::min(r.m, std::min(r.n, r.k)) >= srv->split_min ? kF64Split : kF64Plain;
    q->small = (long)((r.m + 255) / 256) * ((r.n + 127) / 128) < cus;
    q->wide = r.k % 2 == 0 && r.n % 2 == 0;
  }
  // device-time estimate: the operand generation (HBM writes) plus the flop count at the
  // rate of the path the request will take, with that path's fixed cost (rates measured on
  // MI355X: split CRT ~140 TFLOP/s after ~0.5 ms of statistics, residues and one stream
  // synchronisation; plain FP64 ~55; 3 x BF16 slice planes ~180; grouped FP32 ~80)
  const double flops = 2.0 * r.m * (double)r.n * r.k;
  const double gen = (double)(q->bytes[0] + q->bytes[1] + q->bytes[2]) / 4e12;
  double fixed = 3e-5, rate = 80e12;
  if (q->kind == kF64Split) {
    fixed = 5e-4;
    rate = 140e12;
  } else if (q->kind == kF64Plain) {
    rate = 55e12;
  } else if (!q->small && (long)r.m * r.n >= 2048L * 2048 && r.k >= 256) {
    rate = 180e12;
  }
  q->est = fixed + gen + flops / rate;
  q->heavy = !q->small && q->est >= srv->heavy_s;
  if (total > srv->arena.cap) {
    delete q;
    return phpc_set_error(PHPC_ERR_BUDGET, "request needs %zu bytes, operand budget %zu", total, srv->arena.cap);
  }
  q->id = srv->next_id++;
  srv->all[q->id] = q;
  srv->queue.push_back(q);
  ++srv->pending;
  *ticket = q->id;
  srv->cv_work.notify_one();
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_server_wait(phpc_server_t srv, uint64_t ticket, phpc_result *result) {
  if (!srv) return phpc_set_error(PHPC_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->all.find(ticket);
  if (it == srv->all.end()) return phpc_set_error(PHPC_ERR_ARG, "unknown ticket");
  Req *q = it->second;
  srv->cv_done.wait(lk, [&] { return q->done; });
  if (result) *result = q->res;
  const int st = q->res.status;
  srv->all.erase(it);
  delete q;
  return phpc_set_error(st, st ? "request failed" : nullptr);
}

extern "C" int phpc_server_drain(phpc_server_t srv) {
  if (!srv) return phpc_set_error(PHPC_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(srv->mu);
  srv->cv_done.wait(lk, [&] { return srv->pending == 0; });
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_server_set_split_min(phpc_server_t srv, int dim) {
  if (!srv || dim < 1) return phpc_set_error(PHPC_ERR_ARG, "bad split threshold");
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->split_min = dim;
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_server_set_policy(phpc_server_t srv, int policy, double aging) {
  if (!srv || (policy != PHPC_SCHED_FIFO && policy != PHPC_SCHED_SJF && policy != PHPC_SCHED_LANES) ||
      !(aging >= 0.0))
    return phpc_set_error(PHPC_ERR_ARG, "bad scheduling policy");
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->policy = policy;
  srv->aging = aging;
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_server_set_big_limit(phpc_server_t srv, int limit, double min_seconds) {
  if (!srv || limit < 1 || !(min_seconds >= 0.0)) return phpc_set_error(PHPC_ERR_ARG, "bad limit");
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->big_limit = limit;
    srv->heavy_s = min_seconds;
  }
  srv->cv_work.notify_all();
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_server_memory(phpc_server_t srv, size_t *in_use, size_t *peak) {
  if (!srv) return phpc_set_error(PHPC_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(srv->mu);
  if (in_use) *in_use = srv->in_use;
  if (peak) *peak = srv->peak;
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" void phpc_server_destroy(phpc_server_t srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  for (Worker *w : srv->workers)
    if (w->th.joinable()) w->th.join();
  for (Worker *w : srv->workers) {
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    if (w->desc_dev) (void)hipFree(w->desc_dev);
    if (w->desc_host) (void)hipHostFree(w->desc_host);
    for (auto &e : w->ev)
      if (e) (void)hipEventDestroy(e);
    w->ws.release();
    if (w->stream) (void)hipStreamDestroy(w->stream);
    delete w;
  }
  srv->arena.destroy();
  for (auto &kv : srv->all) delete kv.second;
  delete srv;
}

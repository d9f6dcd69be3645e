// Grouped GEMM plans: device-resident item lists, autotuned tile choice, hipGraph replay.
#include "phpc_grouped.h"
#include "phpc_internal.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

struct phpc_plan {
  int dtype = PHPC_F64;
  int count = 0;
  std::vector<GemmDesc> d64;
  std::vector<F32Desc> d32;
  void *descs_dev = nullptr;
  int config = 0;
  bool vec = false;
  int total_tiles = 0;
  int grid = 0;
  double tuned_ms = 0.0;
  // transposed-A configurations: the operand copies they read (plan-owned workspace) and
  // the descriptors that point at them
  F64Staging staging;
  std::vector<GemmDesc> d64_at;
  double *at_ws = nullptr;
  // the FP32 transposed-A configuration: staged copies, their list on the device, and the
  // descriptors that point at them
  F32Staging staging32;
  std::vector<F32Desc> d32_at;
  float *at32_ws = nullptr;
  void *copies_dev = nullptr;  // the staging list (F32Staging::list_bytes)
  // 3 x BF16 plans on slice planes (configurations kF32X3PlanesConfig + g): the planes,
  // the cut list and the level launch's items
  Bf3Planes planes;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
  // last launch on each stream that used the descriptors; destroy waits for these only
  std::map<hipStream_t, hipEvent_t> last_use;
  void mark_use(hipStream_t s) {
    hipEvent_t &e = last_use[s];
    if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) e = nullptr;
    if (e) (void)hipEventRecord(e, s);
  }
};

namespace {

std::mutex g_tune_mu;
std::map<std::string, int> g_tune_cache;  // shape signature -> config

std::string signature(int dtype, const phpc_gemm_item *it, int n) {
  std::string s = std::to_string(dtype) + ":";
  for (int i = 0; i < n; ++i) {
    s += std::to_string(it[i].m) + "x" + std::to_string(it[i].n) + "x" + std::to_string(it[i].k) + "p" +
         std::to_string(it[i].panel) + "," + std::to_string(it[i].lda % 4) + std::to_string(it[i].ldb % 4) + ";";
  }
  return s;
}

// Transposed-A copies for a plan: allocated once (the plan keeps them), filled by every
// launch before the GEMM kernel.
int plan_prepare_at(phpc_plan *p) {
  if (p->at_ws) return PHPC_OK;
  p->staging.plan(p->d64.data(), p->count);
  if (hipMalloc(&p->at_ws, p->staging.total * sizeof(double)) != hipSuccess) {
    p->at_ws = nullptr;
    return phpc_set_error(PHPC_ERR_MEM, "plan transposed-A workspace allocation failed (%zu bytes)",
                          p->staging.total * 8);
  }
  p->d64_at = p->d64;
  p->staging.apply(p->at_ws, p->d64_at.data(), p->count);
  return PHPC_OK;
}

int plan_prepare_at32(phpc_plan *p) {
  if (p->at32_ws) return PHPC_OK;
  p->staging32.plan(p->d32.data(), p->count);
  const size_t ws_bytes = p->staging32.total * sizeof(float);
  const size_t list_bytes = p->staging32.list_bytes();
  if (hipMalloc(&p->at32_ws, ws_bytes) != hipSuccess) {
    p->at32_ws = nullptr;
    return phpc_set_error(PHPC_ERR_MEM, "plan FP32 staging workspace allocation failed (%zu bytes)", ws_bytes);
  }
  if (list_bytes && hipMalloc(&p->copies_dev, list_bytes) != hipSuccess) {
    p->copies_dev = nullptr;
    (void)hipFree(p->at32_ws);
    p->at32_ws = nullptr;
    return phpc_set_error(PHPC_ERR_MEM, "plan FP32 staging list allocation failed");
  }
  p->d32_at = p->d32;
  p->staging32.apply(p->at32_ws, p->d32_at.data(), p->count);
  if (list_bytes)
    if (int rc = p->staging32.upload(p->copies_dev)) return rc;
  return PHPC_OK;
}

void plan_release_at32(phpc_plan *p) {
  if (p->copies_dev) (void)hipFree(p->copies_dev);
  if (p->at32_ws) (void)hipFree(p->at32_ws);
  p->copies_dev = nullptr;
  p->at32_ws = nullptr;
  p->d32_at.clear();
}

// (re)plan tiles for config id and upload the descriptors
int plan_upload(phpc_plan *p, int id, hipStream_t s) {
  size_t bytes;
  const void *src;
  if (p->dtype == PHPC_F64 && phpc_f64_config_trans_a(id)) {
    if (int rc = plan_prepare_at(p)) return rc;
    p->total_tiles = phpc_f64_plan_tiles(p->d64_at.data(), p->count, id);
    phpc_f64_plan_tiles(p->d64.data(), p->count, id);
    bytes = sizeof(GemmDesc) * p->count;
    src = p->d64_at.data();
  } else if (p->dtype == PHPC_F64) {
    p->total_tiles = phpc_f64_plan_tiles(p->d64.data(), p->count, id);
    bytes = sizeof(GemmDesc) * p->count;
    src = p->d64.data();
  } else if (p->dtype == PHPC_F32 && id == kF32AtConfig) {
    if (int rc = plan_prepare_at32(p)) return rc;
    p->total_tiles = phpc_f32_plan_tiles(p->d32_at.data(), p->count, id);
    phpc_f32_plan_tiles(p->d32.data(), p->count, id);
    p->vec = true;
    bytes = sizeof(F32Desc) * p->count;
    src = p->d32_at.data();
  } else if (p->dtype == PHPC_F32X3 && id >= kF32X3PlanesConfig) {
    if (int rc = p->planes.build(p->d32.data(), p->count, id - kF32X3PlanesConfig)) return rc;
    p->total_tiles = p->planes.total_tiles;
    p->config = id;
    return PHPC_OK;
  } else {
    // the in-register 3 x BF16 kernel has one tile shape, 128 x 128 (config 2's)
    p->total_tiles = phpc_f32_plan_tiles(p->d32.data(), p->count, p->dtype == PHPC_F32X3 ? 2 : id);
    p->vec = phpc_f32_vec4_ok(p->d32.data(), p->count);
    bytes = sizeof(F32Desc) * p->count;
    src = p->d32.data();
  }
  p->config = id;
  PHPC_HIP(hipMemcpyAsync(p->descs_dev, src, bytes, hipMemcpyHostToDevice, s));
  PHPC_HIP(hipStreamSynchronize(s));
  return PHPC_OK;
}

int launch(phpc_plan *p, hipStream_t s) {
  if (p->count == 0 || p->total_tiles == 0) return PHPC_OK;
  if (p->dtype == PHPC_F64 && phpc_f64_config_trans_a(p->config)) {
    if (int rc = p->staging.run(p->at_ws, s)) return rc;
    return phpc_f64_grouped_dev_at((const GemmDesc *)p->descs_dev, p->d64_at.data(), p->count, p->grid, p->config,
                                   s);
  }
  if (p->dtype == PHPC_F64)
    return phpc_f64_grouped_dev((const GemmDesc *)p->descs_dev, p->d64.data(), p->count, p->grid, p->config, s);
  if (p->dtype == PHPC_F32 && p->config == kF32AtConfig) {
    if (int rc = p->staging32.run(p->copies_dev, s)) return rc;
    return phpc_f32_launch_at((const F32Desc *)p->descs_dev, p->d32_at.data(), p->count, p->total_tiles, p->grid, s);
  }
  if (p->dtype == PHPC_F32X3 && p->config >= kF32X3PlanesConfig) return p->planes.run(s, p->grid);
  if (p->dtype == PHPC_F32X3)
    return phpc_f32x3_launch((const F32Desc *)p->descs_dev, p->d32.data(), p->count, p->total_tiles, p->vec, p->grid,
                             s);
  return phpc_f32_launch((const F32Desc *)p->descs_dev, p->d32.data(), p->count, p->total_tiles, p->config, p->vec,
                         p->grid, s);
}

}  // namespace

extern "C" int phpc_plan_create(int dtype, const phpc_gemm_item *items, int count, int flags, void *stream_,
                                phpc_plan_t *out) {
  if (!out) return phpc_set_error(PHPC_ERR_ARG, "null plan pointer");
  *out = nullptr;
  if (phpc_init()) return phpc_last_error();
  if ((dtype != PHPC_F64 && dtype != PHPC_F32 && dtype != PHPC_F32X3) || count < 0 || (count > 0 && !items))
    return phpc_set_error(PHPC_ERR_ARG, "bad plan arguments");
  hipStream_t s = (hipStream_t)stream_;
  phpc_plan *p = new phpc_plan();
  p->dtype = dtype;
  std::vector<phpc_gemm_item> live;
  for (int i = 0; i < count; ++i) {
    const phpc_gemm_item &it = items[i];
    if (it.m < 0 || it.n < 0 || it.k < 0 || (it.m && it.k && (!it.A || it.lda < it.k)) ||
        (it.k && it.n && (!it.B || it.ldb < it.n)) || (it.m && it.n && (!it.C || it.ldc < it.n)) ||
        (dtype != PHPC_F64 && it.panel > 0 && it.panel < it.k)) {
      delete p;
      return phpc_set_error(PHPC_ERR_ARG, "bad GEMM item %d", i);
    }
    if (it.m == 0 || it.n == 0 || it.k == 0) continue;
    live.push_back(it);
  }
  // Items in decreasing k: a tile's time grows with k, and the persistent workgroups take
  // tiles round-robin from the flat list, so tiles of similar cost then share each round
  // and the workgroups finish together (on log-uniform k, the slowest of 512 workgroups
  // would otherwise carry about 18% more than the mean). Outputs do not depend on the order.
  std::stable_sort(live.begin(), live.end(),
                   [](const phpc_gemm_item &a, const phpc_gemm_item &b) { return a.k > b.k; });
  p->count = (int)live.size();
  bool mp = false;
  for (const auto &it : live) {
    if (dtype == PHPC_F64) {
      GemmDesc d{};
      d.A = (const double *)it.A;
      d.B = (const double *)it.B;
      d.C = (double *)it.C;
      d.lda = it.lda;
      d.ldb = it.ldb;
      d.ldc = it.ldc;
      d.m = it.m;
      d.n = it.n;
      d.k = it.k;
      d.panel = it.panel > 0 && it.panel < it.k ? it.panel : 0;
      if (d.panel) mp = true;
      p->d64.push_back(d);
    } else {
      F32Desc d{};
      d.A = (const float *)it.A;
      d.B = (const float *)it.B;
      d.C = (float *)it.C;
      d.lda = it.lda;
      d.ldb = it.ldb;
      d.ldc = it.ldc;
      d.m = it.m;
      d.n = it.n;
      d.k = it.k;
      p->d32.push_back(d);
    }
  }
  if (p->count == 0) {
    *out = p;
    return phpc_set_error(PHPC_OK, nullptr);
  }
  const size_t bytes = dtype == PHPC_F64 ? sizeof(GemmDesc) * p->count : sizeof(F32Desc) * p->count;  // F32, F32X3
  if (hipMalloc(&p->descs_dev, bytes) != hipSuccess) {
    delete p;
    return phpc_set_error(PHPC_ERR_MEM, "descriptor allocation failed");
  }
  // default choice: the planner's rule; large single-panel FP64 work takes the
  // transposed-A kernel when B allows 16-byte loads (the plan then keeps a copy of A)
  int cfg = dtype == PHPC_F64 ? phpc_f64_pick_config_at(p->d64.data(), p->count) : 0;
  const bool at_ok = dtype == PHPC_F64 && phpc_f64_trans_a_ok(p->d64.data(), p->count);
  const bool at32_ok = dtype == PHPC_F32 && phpc_f32_at_ok(p->d32.data(), p->count);
  if (dtype == PHPC_F32) cfg = phpc_f32_pick_plan_config(p->d32.data(), p->count);
  if (dtype == PHPC_F32X3) cfg = phpc_f32x3_pick_plan_config(p->d32.data(), p->count);
  if (flags & PHPC_PLAN_AUTOTUNE) {
    const std::string sig = signature(dtype, live.data(), (int)live.size());
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto hit = g_tune_cache.find(sig);
    if (hit != g_tune_cache.end()) {
      cfg = hit->second;
    } else {
      // save every C, time each candidate once (after one warm-up), restore C
      const size_t esz = dtype == PHPC_F64 ? 8 : 4;
      size_t cbytes = 0;
      for (const auto &it : live) cbytes += (size_t)it.m * it.n * esz;
      char *save = (char *)phpc_scratch(cbytes, 13);
      if (!save) {
        delete p;
        return phpc_set_error(PHPC_ERR_MEM, "autotune scratch allocation failed");
      }
      size_t off = 0;
      for (const auto &it : live) {
        if (hipMemcpy2DAsync(save + off, it.n * esz, it.C, it.ldc * esz, it.n * esz, it.m, hipMemcpyDeviceToDevice,
                             s) != hipSuccess) {
          delete p;
          return phpc_set_error(PHPC_ERR_HIP, "autotune save failed");
        }
        off += (size_t)it.m * it.n * esz;
      }
      const int ncfg = dtype == PHPC_F64    ? phpc_f64_num_configs()
                       : dtype == PHPC_F32X3 ? kF32X3PlanesConfig + phpc_bf3_num_configs()
                                             : phpc_f32_num_configs();
      double best = 1e30;
      int best_cfg = cfg;
      hipEvent_t e0, e1;
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
      for (int c = 0; c < ncfg; ++c) {
        if (dtype == PHPC_F64 && mp && !phpc_f64_config_panels(c)) continue;
        if (dtype == PHPC_F64 && phpc_f64_config_trans_a(c) && !at_ok) continue;
        if (dtype == PHPC_F32X3 && c != 2 && c < kF32X3PlanesConfig) continue;
        if (dtype == PHPC_F32 && c == kF32AtConfig && !at32_ok) continue;
        if (plan_upload(p, c, s) != PHPC_OK) continue;
        p->grid = 0;
        if (launch(p, s) != PHPC_OK) continue;
        (void)hipEventRecord(e0, s);
        if (launch(p, s) != PHPC_OK) continue;
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
          best = ms;
          best_cfg = c;
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      off = 0;
      for (const auto &it : live) {
        (void)hipMemcpy2DAsync(it.C, it.ldc * esz, save + off, it.n * esz, it.n * esz, it.m, hipMemcpyDeviceToDevice,
                               s);
        off += (size_t)it.m * it.n * esz;
      }
      (void)hipStreamSynchronize(s);
      cfg = best_cfg;
      p->tuned_ms = best;
      g_tune_cache[sig] = cfg;
    }
  }
  int rc = plan_upload(p, cfg, s);
  if (rc && dtype == PHPC_F32X3 && cfg >= kF32X3PlanesConfig) {
    // the slice planes (6 bytes per element of every operand) did not fit: the in-register
    // kernel needs no workspace
    cfg = 2;
    rc = plan_upload(p, cfg, s);
  }
  if (rc) {
    phpc_plan_destroy(p);
    return rc;
  }
  if (p->at_ws && !(dtype == PHPC_F64 && phpc_f64_config_trans_a(cfg))) {
    // the autotuner tried a transposed-A configuration but kept another one
    (void)hipStreamSynchronize(s);
    (void)hipFree(p->at_ws);
    p->at_ws = nullptr;
    p->d64_at.clear();
  }
  if (p->at32_ws && cfg != kF32AtConfig) {
    (void)hipStreamSynchronize(s);
    plan_release_at32(p);
  }
  if (p->planes.ws && cfg < kF32X3PlanesConfig) {
    (void)hipStreamSynchronize(s);
    p->planes.release();
  }
  p->grid = 0;
  *out = p;
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_plan_launch(phpc_plan_t p, void *stream) {
  if (!p) return phpc_set_error(PHPC_ERR_ARG, "null plan");
  int rc = launch(p, (hipStream_t)stream);
  if (rc) return rc;
  p->mark_use((hipStream_t)stream);
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_plan_graph_launch(phpc_plan_t p, void *stream) {
  if (!p || !stream) return phpc_set_error(PHPC_ERR_ARG, "graph launch needs a plan and a created stream");
  hipStream_t s = (hipStream_t)stream;
  if (!p->exec) {
    PHPC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = launch(p, s);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ec != hipSuccess) return phpc_set_error(PHPC_ERR_HIP, "stream capture failed: %s", hipGetErrorString(ec));
    p->graph = g;
    PHPC_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
    p->graph_stream = s;
  }
  PHPC_HIP(hipGraphLaunch(p->exec, s));
  p->mark_use(s);
  return phpc_set_error(PHPC_OK, nullptr);
}

extern "C" int phpc_plan_info(phpc_plan_t p, int *config, int *tiles, int *grid, double *tuned_ms) {
  if (!p) return phpc_set_error(PHPC_ERR_ARG, "null plan");
  if (config) *config = p->config;
  if (tiles) *tiles = p->total_tiles;
  if (grid) *grid = p->grid;
  if (tuned_ms) *tuned_ms = p->tuned_ms;
  return PHPC_OK;
}

extern "C" void phpc_plan_destroy(phpc_plan_t p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  // wait for the plan's own launches (not the whole device) before freeing what they read
  for (auto &kv : p->last_use) {
    (void)hipEventSynchronize(kv.second);
    (void)hipEventDestroy(kv.second);
  }
  if (p->descs_dev) (void)hipFree(p->descs_dev);
  if (p->at_ws) (void)hipFree(p->at_ws);
  plan_release_at32(p);
  p->planes.release();
  delete p;
}

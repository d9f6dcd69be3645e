// SUMMA k-panel loop (src/phpc_summa.c:24-122) on an emulated p x q grid, one GPU.
//
// Rank (i, j) of the grid owns the C block rows R_i x columns C_j. At step t the owner
// of A's panel t in grid row i (column t % q) and the owner of B's panel t in grid
// column j (row t % p) broadcast it (:63-89); every rank then runs the local update
// C_ij += A(R_i, t) * B(t, C_j) (:93). On one GPU every rank sees the whole of A and B
// in HBM, so:
//   FUSED     one persistent grouped launch covers all p*q rank blocks; each tile walks
//             the panels in order, summing each panel from zero and adding it to the C
//             value carried in registers, i.e. exactly the sequence of per-step updates;
//   STEPWISE  per panel, the non-owners' receive buffers are produced by a strided ->
//             packed copy (the MPI_Type_vector send of :53-56), owners read in place
//             with ld = N (:72-73, :82-83), then one grouped launch does every rank's
//             update; the C blocks stay in place (the gather of :97-110 is a no-op);
//   VENDOR    the same steps with rocBLAS dgemm as the local GEMM (phpc_gemm_cublas);
//   CHAINED   like FUSED, but each output is one FMA chain over all of k (the panels
//             are walked in the same order without rounding C between them), which
//             frees the registers that carry C and admits the 256x128 tile. The
//             result equals the reference kernel's arithmetic for a single panel.
//   SPLIT     every rank's update over all of k through the split-precision CRT GEMM
//             (FP64-accurate, err_cw <= 1e-14 against the exact product).
// FUSED and STEPWISE run the same per-element arithmetic and give bit-identical C.
#include "phpc_internal.h"
#include "phpc_summa.h"

#include <algorithm>
#include <vector>

namespace {

int gcd(int a, int b) {
  while (b) {
    int r = a % b;
    a = b;
    b = r;
  }
  return a;
}

inline int blk_size(int n, int parts, int i) { return n / parts + (i < n % parts); }
inline int blk_start(int n, int parts, int i) {
  const int q = n / parts, r = n % parts;
  return i * q + (i < r ? i : r);
}

}  // namespace

extern "C" int phpc_summa_device(int p, int q, int panel, int mode, const double *dA, const double *dB, double *dC,
                                 int N, int wg, void *stream_) {
  if (phpc_init()) return phpc_last_error();
  if (p < 1 || q < 1 || N < 0) return phpc_set_error(PHPC_ERR_ARG, "bad grid %dx%d or N=%d", p, q, N);
  if (p > N || q > N) return phpc_set_error(PHPC_ERR_ARG, "grid %dx%d larger than N=%d", p, q, N);
  if (N == 0) return phpc_set_error(PHPC_OK, nullptr);
  if (!dA || !dB || !dC) return phpc_set_error(PHPC_ERR_ARG, "null matrix");
  const int lcm = p / gcd(p, q) * q;
  const int w = panel > 0 ? panel : N / lcm;
  if (w < 1) return phpc_set_error(PHPC_ERR_ARG, "panel width N/lcm(p,q) = 0 for N=%d, grid %dx%d", N, p, q);
  hipStream_t s = (hipStream_t)stream_;
  const LaunchPolicy &pol = phpc_policy();
  const int G = wg > 0 ? wg : pol.grid_override;

  if (mode == PHPC_SUMMA_FUSED || mode == PHPC_SUMMA_CHAINED) {
    std::vector<GemmDesc> d;
    d.reserve((size_t)p * q);
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < q; ++j) {
        GemmDesc g{};
        const int r0 = blk_start(N, p, i), c0 = blk_start(N, q, j);
        g.A = dA + (long)r0 * N;
        g.B = dB + c0;
        g.C = dC + (long)r0 * N + c0;
        g.lda = N;
        g.ldb = N;
        g.ldc = N;
        g.m = blk_size(N, p, i);
        g.n = blk_size(N, q, j);
        g.k = N;
        g.panel = mode == PHPC_SUMMA_CHAINED ? 0 : w;
        d.push_back(g);
      }
    int rc = phpc_f64_grouped(d.data(), (int)d.size(), G, pol.config_id, s);
    if (rc) return rc;
    return phpc_set_error(PHPC_OK, nullptr);
  }
  if (mode == PHPC_SUMMA_SPLIT) {
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < q; ++j) {
        const int r0 = blk_start(N, p, i), c0 = blk_start(N, q, j);
        int rc = phpc_dgemm_split_device(dA + (long)r0 * N, N, dB + c0, N, dC + (long)r0 * N + c0, N,
                                         blk_size(N, p, i), N, blk_size(N, q, j), 0, 1e-14, s, nullptr);
        if (rc) return rc;
      }
    return phpc_set_error(PHPC_OK, nullptr);
  }
  if (mode != PHPC_SUMMA_STEPWISE && mode != PHPC_SUMMA_VENDOR)
    return phpc_set_error(PHPC_ERR_ARG, "unknown SUMMA mode %d", mode);

  // receive buffers: one A panel per grid row (m_i x w) and one B panel per grid
  // column (w x n_j), packed (ld = w and ld = n_j), as src/phpc_summa.c:47-48
  double *abuf = (double *)phpc_scratch((size_t)N * w * 8, 3);
  double *bbuf = (double *)phpc_scratch((size_t)N * w * 8, 4);
  if (!abuf || !bbuf) return phpc_set_error(PHPC_ERR_MEM, "panel buffers");
  const int steps = (N + w - 1) / w;
  std::vector<GemmDesc> d((size_t)p * q);
  for (int t = 0; t < steps; ++t) {
    const int k0 = t * w;
    const int kw = k0 + w <= N ? w : N - k0;
    const int a_owner_col = t % q, b_owner_row = t % p;
    // "broadcast": non-owner ranks get a packed copy of the panel
    for (int i = 0; i < p; ++i) {
      const int r0 = blk_start(N, p, i), mr = blk_size(N, p, i);
      int rc = phpc_pack_launch(dA + (long)r0 * N + k0, N, abuf + (long)r0 * w, kw, mr, kw, s);
      if (rc) return rc;
    }
    for (int j = 0; j < q; ++j) {
      const int c0 = blk_start(N, q, j), nc = blk_size(N, q, j);
      int rc = phpc_pack_launch(dB + (long)k0 * N + c0, N, bbuf + (long)c0 * w, nc, kw, nc, s);
      if (rc) return rc;
    }
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < q; ++j) {
        GemmDesc &g = d[(size_t)i * q + j];
        g = GemmDesc{};
        const int r0 = blk_start(N, p, i), c0 = blk_start(N, q, j);
        const int mr = blk_size(N, p, i), nc = blk_size(N, q, j);
        if (j == a_owner_col) {  // owner: strided block in place, lda = N
          g.A = dA + (long)r0 * N + k0;
          g.lda = N;
        } else {
          g.A = abuf + (long)r0 * w;
          g.lda = kw;
        }
        if (i == b_owner_row) {
          g.B = dB + (long)k0 * N + c0;
          g.ldb = N;
        } else {
          g.B = bbuf + (long)c0 * w;
          g.ldb = nc;
        }
        g.C = dC + (long)r0 * N + c0;
        g.ldc = N;
        g.m = mr;
        g.n = nc;
        g.k = kw;
        g.panel = 0;
      }
    if (mode == PHPC_SUMMA_STEPWISE) {
      int rc = phpc_f64_grouped(d.data(), (int)d.size(), G, pol.config_id, s);
      if (rc) return rc;
    } else {
      for (const GemmDesc &g : d) {
        int rc = phpc_dgemm_vendor_device(g.A, g.lda, g.B, g.ldb, g.C, g.ldc, g.m, g.k, g.n, s);
        if (rc) return rc;
      }
    }
  }
  return phpc_set_error(PHPC_OK, nullptr);
}

// Host pointers. FUSED and CHAINED: every output's arithmetic depends only on its own
// row and column and the panel width, so C can be computed in row bands. B goes over
// first; then the A and C rows of band i+1 travel (pinned staging, copy stream) while band
// i computes (compute stream) and band i-1 comes back. Other modes copy everything in,
// run phpc_summa_device and copy C out.
extern "C" int phpc_summa_host(int p, int q, int panel, int mode, const double *A, const double *B, double *C, int N,
                               int wg, float *kernel_seconds) {
  if (kernel_seconds) *kernel_seconds = 0.0f;
  if (phpc_init()) return phpc_last_error();
  if (N < 0 || (N > 0 && (!A || !B || !C))) return phpc_set_error(PHPC_ERR_ARG, "bad SUMMA host arguments");
  if (p < 1 || q < 1 || p > N || q > N) return phpc_set_error(PHPC_ERR_ARG, "bad grid %dx%d for N=%d", p, q, N);
  if (N == 0) return phpc_set_error(PHPC_OK, nullptr);
  const size_t bytes = (size_t)N * N * 8, row = (size_t)N * 8;
  double *dA = (double *)phpc_scratch(bytes, 5);
  double *dB = (double *)phpc_scratch(bytes, 6);
  double *dC = (double *)phpc_scratch(bytes, 7);
  if (!dA || !dB || !dC) return phpc_set_error(PHPC_ERR_MEM, "device allocation failed (3 x %zu bytes)", bytes);
  hipStream_t sx, sc;
  if (int rc = phpc_host_streams(&sx, &sc)) return rc;
  const bool banded = mode == PHPC_SUMMA_FUSED || mode == PHPC_SUMMA_CHAINED;
  if (!banded) {
    if (int rc = phpc_stage_h2d(dA, row, A, row, row, N, sx)) return rc;
    if (int rc = phpc_stage_h2d(dB, row, B, row, row, N, sx)) return rc;
    if (int rc = phpc_stage_h2d(dC, row, C, row, row, N, sx)) return rc;
    ScopedEvent e0, e1;
    if (!e0.e || !e1.e) return phpc_set_error(PHPC_ERR_HIP, "event creation failed");
    PHPC_HIP(hipEventRecord(e0.e, sx));
    int rc = phpc_summa_device(p, q, panel, mode, dA, dB, dC, N, wg, sx);
    PHPC_HIP(hipEventRecord(e1.e, sx));
    if (rc == PHPC_OK) rc = phpc_stage_d2h(C, row, dC, row, row, N, sx);
    PHPC_HIP(hipStreamSynchronize(sx));
    if (rc) return rc;
    float ms = 0;
    PHPC_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
    if (kernel_seconds) *kernel_seconds = ms / 1000.0f;
    return phpc_set_error(PHPC_OK, nullptr);
  }
  const int lcm = p / gcd(p, q) * q;
  const int w = panel > 0 ? panel : N / lcm;
  if (w < 1) return phpc_set_error(PHPC_ERR_ARG, "panel width N/lcm(p,q) = 0 for N=%d, grid %dx%d", N, p, q);
  const LaunchPolicy &pol = phpc_policy();
  const int G = wg > 0 ? wg : pol.grid_override;
  // bands of whole 256-row tiles, about 16 of them
  const int hb = std::max(256, (N / 16 + 255) / 256 * 256);
  const int nb = (N + hb - 1) / hb;
  std::vector<ScopedEvent> ev_in, ev_s, ev_e;
  ev_in.reserve(nb);
  ev_s.reserve(nb);
  ev_e.reserve(nb);
  for (int i = 0; i < nb; ++i) {
    ev_in.emplace_back(hipEventDisableTiming);
    ev_s.emplace_back();
    ev_e.emplace_back();
    if (!ev_in.back().e || !ev_s.back().e || !ev_e.back().e) return phpc_set_error(PHPC_ERR_HIP, "event creation failed");
  }
  {
    // the largest band's staging copies, allocated before the copies start
    GemmDesc d{};
    d.A = dA;
    d.B = dB;
    d.C = dC;
    d.lda = d.ldb = d.ldc = N;
    d.m = std::min(hb, N);
    d.n = d.k = N;
    d.panel = mode == PHPC_SUMMA_CHAINED ? 0 : w;
    if (pol.config_id < 0)
      if (int rc0 = phpc_f64_reserve_staging(&d, 1, sc)) return rc0;
  }
  int rc = phpc_stage_h2d(dB, row, B, row, row, N, sx);
  for (int i = 0; i <= nb && rc == PHPC_OK; ++i) {
    if (i < nb) {
      const long r0 = (long)i * hb;
      const int rows = (int)std::min<long>(hb, N - r0);
      rc = phpc_stage_h2d(dA + r0 * N, row, A + r0 * N, row, row, rows, sx);
      if (!rc) rc = phpc_stage_h2d(dC + r0 * N, row, C + r0 * N, row, row, rows, sx);
      if (rc) break;
      PHPC_HIP(hipEventRecord(ev_in[i].e, sx));
      PHPC_HIP(hipStreamWaitEvent(sc, ev_in[i].e, 0));
      GemmDesc d{};
      d.A = dA + r0 * N;
      d.B = dB;
      d.C = dC + r0 * N;
      d.lda = N;
      d.ldb = N;
      d.ldc = N;
      d.m = rows;
      d.n = N;
      d.k = N;
      d.panel = mode == PHPC_SUMMA_CHAINED ? 0 : w;
      PHPC_HIP(hipEventRecord(ev_s[i].e, sc));
      rc = phpc_f64_grouped(&d, 1, G, pol.config_id, sc);
      PHPC_HIP(hipEventRecord(ev_e[i].e, sc));
      if (rc) break;
    }
    if (i >= 1) {  // band i-1 back to the host once computed
      const long r0 = (long)(i - 1) * hb;
      const int rows = (int)std::min<long>(hb, N - r0);
      PHPC_HIP(hipStreamWaitEvent(sx, ev_e[i - 1].e, 0));
      rc = phpc_stage_d2h(C + r0 * N, row, dC + r0 * N, row, row, rows, sx);
    }
  }
  (void)hipStreamSynchronize(sc);
  (void)hipStreamSynchronize(sx);
  float total = 0;
  for (int i = 0; i < nb && rc == PHPC_OK; ++i) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev_s[i].e, ev_e[i].e) == hipSuccess) total += ms;
  }
  if (rc) return rc;
  if (kernel_seconds) *kernel_seconds = total / 1000.0f;
  return phpc_set_error(PHPC_OK, nullptr);
}

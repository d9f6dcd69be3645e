/*
 * SUMMA entry points over an emulated process grid (src/phpc_summa.c of the reference).
 *
 * The reference runs one MPI rank per grid position; rank (i, j) holds the C block
 * (i, j), receives A's panel t from grid column t % q and B's panel t from grid row
 * t % p, and calls the local GEMM once per panel (src/phpc_summa.c:63-95). Here the
 * grid is a plain descriptor and every rank runs on the one GPU:
 *   - phpc_gemm_summa_cuda / _cublas hand the whole grid to the device path
 *     (phpc_summa_host), which copies A, B, C into HBM once;
 *   - phpc_gemm_summa keeps the reference's host-side loop for any gemm_t, with the
 *     broadcasts replaced by pointer aliases into the full host matrices.
 */
#include "phpc_summa.h"

#include <stdlib.h>

/* status recorder of the library (phpc_last_error) */
int phpc_set_error(int code, const char *fmt, ...);

struct phpc_grid {
  int p, q;
  int panel; /* 0: N / lcm(p, q) */
  int mode;  /* PHPC_SUMMA_FUSED, _STEPWISE, _CHAINED or _SPLIT */
};

/* src/phpc_summa.c:9-22 */
int find_lcm(int a, int b) {
  int x = a, y = b;
  while (y != 0) {
    int r = x % y;
    x = y;
    y = r;
  }
  return x == 0 ? 0 : a / x * b;
}

/* MPI_Dims_create(P, 2): the most balanced factor pair, larger dimension first. */
static void dims_create(int nranks, int dims[2]) {
  int best = 1;
  for (int d = 1; (long)d * d <= nranks; ++d)
    if (nranks % d == 0) best = d;
  dims[0] = nranks / best;
  dims[1] = best;
}

int phpc_grid_create_dims(int p, int q, phpc_grid_t *grid) {
  if (!grid || p < 1 || q < 1) return PHPC_ERR_ARG;
  struct phpc_grid *g = (struct phpc_grid *)calloc(1, sizeof(*g));
  if (!g) return PHPC_ERR_MEM;
  g->p = p;
  g->q = q;
  g->panel = 0;
  g->mode = PHPC_SUMMA_FUSED;
  *grid = g;
  return PHPC_OK;
}

/* src/main.c:38-45: one rank -> 1x1, otherwise MPI_Dims_create(size, 2). */
int phpc_grid_create(int nranks, phpc_grid_t *grid) {
  if (nranks < 1) return PHPC_ERR_ARG;
  int dims[2] = {1, 1};
  if (nranks > 1) dims_create(nranks, dims);
  return phpc_grid_create_dims(dims[0], dims[1], grid);
}

void phpc_grid_destroy(phpc_grid_t grid) { free(grid); }

int phpc_grid_dims(phpc_grid_t grid, int dims[2]) {
  if (!grid || !dims) return PHPC_ERR_ARG;
  dims[0] = grid->p;
  dims[1] = grid->q;
  return PHPC_OK;
}

int phpc_grid_set_panel(phpc_grid_t grid, int panel) {
  if (!grid || panel < 0) return PHPC_ERR_ARG;
  grid->panel = panel;
  return PHPC_OK;
}

int phpc_grid_get_panel(phpc_grid_t grid) { return grid ? grid->panel : -1; }

int phpc_grid_set_mode(phpc_grid_t grid, int mode) {
  if (!grid || (mode != PHPC_SUMMA_FUSED && mode != PHPC_SUMMA_STEPWISE && mode != PHPC_SUMMA_CHAINED &&
                mode != PHPC_SUMMA_SPLIT))
    return PHPC_ERR_ARG;
  grid->mode = mode;
  return PHPC_OK;
}

/* src/phpc_summa.c:124-126 */
void phpc_gemm_summa_cuda(phpc_grid_t grid, const double *A, const double *B, double *C, int n, int gpu_count,
                          int grid_width, int grid_height, int block_width, float *compute_time) {
  if (compute_time) *compute_time = 0.0f;
  if (!grid || gpu_count < 1 || block_width < 1 || block_width > 32 || grid_width < 1 || grid_height < 1) {
    phpc_set_error(PHPC_ERR_ARG, "bad SUMMA arguments (gpu_count %d, grid %dx%d, block_width %d)", gpu_count,
                   grid_width, grid_height, block_width);
    return;
  }
  (void)phpc_summa_host(grid->p, grid->q, grid->panel, grid->mode, A, B, C, n, 0, compute_time);
}

/* src/phpc_summa.c:128-130 */
void phpc_gemm_summa_cublas(phpc_grid_t grid, const double *A, const double *B, double *C, int n, int gpu_count,
                            float *compute_time) {
  if (compute_time) *compute_time = 0.0f;
  if (!grid || gpu_count < 1) {
    phpc_set_error(PHPC_ERR_ARG, "bad SUMMA arguments (gpu_count %d)", gpu_count);
    return;
  }
  float t = 0.0f;
  (void)phpc_summa_host(grid->p, grid->q, grid->panel, PHPC_SUMMA_VENDOR, A, B, C, n, 0, &t);
}

static int blk_size(int n, int parts, int i) { return n / parts + (i < n % parts); }
static int blk_start(int n, int parts, int i) {
  int qq = n / parts, r = n % parts;
  return i * qq + (i < r ? i : r);
}

/* src/phpc_summa.c:24-122 with the MPI traffic replaced by aliases into A, B and C. */
int phpc_gemm_summa(gemm_t f, phpc_grid_t grid, const double *A, const double *B, double *C, int N, int gpu_count,
                    int grid_width, int grid_height, int block_width, float *gpu_time) {
  if (gpu_time) *gpu_time = 0.0f;
  if (!f || !grid || N < 0 || (N > 0 && (!A || !B || !C)))
    return phpc_set_error(PHPC_ERR_ARG, "bad SUMMA loop arguments");
  if (grid->p > N || grid->q > N)
    return phpc_set_error(PHPC_ERR_ARG, "grid %dx%d larger than N = %d", grid->p, grid->q, N);
  /* A status left by an earlier call on this thread must not stop the loop: only an
   * error raised by f during this call does. (A user-supplied gemm_t that never
   * records a status then always reads back PHPC_OK.) */
  phpc_set_error(PHPC_OK, NULL);
  if (N == 0) return PHPC_OK;
  const int lcm = find_lcm(grid->p, grid->q);
  const int w = grid->panel > 0 ? grid->panel : N / lcm;
  if (w < 1) return phpc_set_error(PHPC_ERR_ARG, "panel width %d", w);
  float total = 0.0f;
  for (int k0 = 0; k0 < N; k0 += w) {
    const int kw = k0 + w <= N ? w : N - k0;
    for (int i = 0; i < grid->p; ++i)
      for (int j = 0; j < grid->q; ++j) {
        const int r0 = blk_start(N, grid->p, i), c0 = blk_start(N, grid->q, j);
        float t = 0.0f;
        f(A + (size_t)r0 * N + k0, N, B + (size_t)k0 * N + c0, N, C + (size_t)r0 * N + c0, N,
          blk_size(N, grid->p, i), kw, blk_size(N, grid->q, j), gpu_count, grid_width, grid_height, block_width, &t);
        if (phpc_last_error() != PHPC_OK) return phpc_last_error();
        total += t;
      }
  }
  if (gpu_time) *gpu_time = total;
  return PHPC_OK;
}

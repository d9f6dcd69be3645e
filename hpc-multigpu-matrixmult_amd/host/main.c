/*
 * main.out — benchmark CLI with the reference's contract (src/main.c:17-122):
 *
 *   main.out <matrix_size> <tile_width> <grid_width> <grid_height> <test_name> [options]
 *
 * writes one row to csv/<test_name>_N%d_T%d_G%d_TW%d_GW%d_GH%d.csv in the format of
 * log_to_csv. The reference takes its rank count from `mpirun -np P`; this build
 * emulates the P ranks on one GPU and takes P from --ranks (default 1).
 *
 * Options (extensions; the five positional arguments are unchanged):
 *   --ranks P        emulated rank count; grid = 1x1 or MPI_Dims_create(P, 2)
 *   --grid PxQ       explicit grid (overrides --ranks)
 *   --panel W        k-panel width (default N / lcm(p, q), as the reference)
 *   --dist D         ramp (default, A[i] = B[i] = i as src/main.c:85-86), uniform, wide
 *   --seed S         generator seed for uniform / wide (A uses S, B uses S + 1)
 *   --mode M         fused (default), stepwise, chained or split (FP64-accurate
 *                    split-precision GEMM on int8 matrix cores)
 *   --check S        after the HIP pass, check S sampled entries of C (plus the
 *                    corners) against a double-double dot product; exit status 3
 *                    if any exceeds the FP64 dot-product bound (N + 2) * 2^-53
 *   --uneven         allow N not divisible by the grid (the reference aborts, :47-52)
 *   --no-vendor      skip the rocBLAS pass (cublas_time column = 0)
 *
 * Differences from the reference, kept deliberately: C is zeroed before each pass
 * (the reference leaves it uninitialised, :66, and its cuBLAS pass accumulates onto
 * the CUDA result); cuda_gpu_time is the device time of the whole emulated grid.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fill.h"
#include "phpc_summa.h"
#include "utils.h"

#define CLI_ASSERT(check)                                                \
  if (!(check)) {                                                        \
    fprintf(stderr, "Check at " __FILE__ " line %d failed\n", __LINE__); \
    exit(EXIT_FAILURE);                                                  \
  }

int main(int argc, char *argv[]) {
  int pos = 0;
  const char *posv[5];
  int nranks = 1, gp = 0, gq = 0, panel = 0, uneven = 0, vendor = 1, mode = PHPC_SUMMA_FUSED, check = 0;
  const char *dist = "ramp";
  uint64_t seed = 1;
  for (int i = 1; i < argc; ++i) {
    if (strncmp(argv[i], "--", 2) != 0) {
      if (pos < 5) posv[pos] = argv[i];
      ++pos;
      continue;
    }
    const char *o = argv[i];
    const char *v = i + 1 < argc ? argv[i + 1] : NULL;
    if (!strcmp(o, "--ranks") && v) {
      nranks = atoi(v);
      ++i;
    } else if (!strcmp(o, "--grid") && v) {
      if (sscanf(v, "%dx%d", &gp, &gq) != 2) gp = gq = 0;
      ++i;
    } else if (!strcmp(o, "--panel") && v) {
      panel = atoi(v);
      ++i;
    } else if (!strcmp(o, "--dist") && v) {
      dist = v;
      ++i;
    } else if (!strcmp(o, "--seed") && v) {
      seed = strtoull(v, NULL, 10);
      ++i;
    } else if (!strcmp(o, "--mode") && v) {
      if (!strcmp(v, "fused")) {
        mode = PHPC_SUMMA_FUSED;
      } else if (!strcmp(v, "stepwise")) {
        mode = PHPC_SUMMA_STEPWISE;
      } else if (!strcmp(v, "chained")) {
        mode = PHPC_SUMMA_CHAINED;
      } else if (!strcmp(v, "split")) {
        mode = PHPC_SUMMA_SPLIT;
      } else {
        fprintf(stderr, "Error: --mode must be fused, stepwise, chained or split\n");
        return EXIT_FAILURE;
      }
      ++i;
    } else if (!strcmp(o, "--check") && v) {
      check = atoi(v);
      ++i;
    } else if (!strcmp(o, "--uneven")) {
      uneven = 1;
    } else if (!strcmp(o, "--no-vendor")) {
      vendor = 0;
    } else {
      fprintf(stderr, "Unknown option %s\n", o);
      return EXIT_FAILURE;
    }
  }
  if (pos != 5) {
    fprintf(stderr, "Usage: %s <matrix_size> <tile_width> <grid_width> <grid_height> <test_name>\n", argv[0]);
    return EXIT_FAILURE;
  }
  int N = atoi(posv[0]);
  int tile_width = atoi(posv[1]);
  int grid_width = atoi(posv[2]);
  int grid_height = atoi(posv[3]);
  const char *test_name = posv[4];
  if (strcmp(dist, "ramp") && strcmp(dist, "uniform") && strcmp(dist, "wide")) {
    fprintf(stderr, "Error: --dist must be ramp, uniform or wide\n");
    return EXIT_FAILURE;
  }
  if (nranks < 1 || N < 1) {
    fprintf(stderr, "Error: matrix size and rank count must be positive\n");
    return EXIT_FAILURE;
  }

  phpc_grid_t grid = NULL;
  int rc = (gp > 0 && gq > 0) ? phpc_grid_create_dims(gp, gq, &grid) : phpc_grid_create(nranks, &grid);
  CLI_ASSERT(rc == PHPC_OK);
  int dims[2];
  phpc_grid_dims(grid, dims);
  int size = dims[0] * dims[1];

  if (!uneven && (N % dims[0] != 0 || N % dims[1] != 0)) {
    fprintf(stderr, "Error: Matrix size N (%d) must be divisible by process grid dimensions (%d x %d).\n", N, dims[0],
            dims[1]);
    return EXIT_FAILURE;
  }
  phpc_grid_set_panel(grid, panel);
  phpc_grid_set_mode(grid, mode);

  rc = phpc_init();
  if (rc != PHPC_OK) {
    fprintf(stderr, "Error: %s\n", phpc_last_error_message());
    return EXIT_FAILURE;
  }
  int gpu_count = 1; /* one GPU drives every emulated rank */

  size_t nn = (size_t)N * (size_t)N;
  double *A = (double *)malloc(nn * sizeof(double));
  double *B = (double *)malloc(nn * sizeof(double));
  double *C = (double *)malloc(nn * sizeof(double));
  CLI_ASSERT(A != NULL);
  CLI_ASSERT(B != NULL);
  CLI_ASSERT(C != NULL);

  char filename[512];
  snprintf(filename, sizeof(filename), "csv/%s_N%d_T%d_G%d_TW%d_GW%d_GH%d.csv", test_name, N, size, gpu_count,
           tile_width, grid_width, grid_height);
  FILE *csv_file = fopen(filename, "w");
  if (csv_file == NULL) {
    fprintf(stderr, "Error: Could not create CSV file %s\n", filename);
    return EXIT_FAILURE;
  }

  phpc_host_fill(A, nn, dist, seed); /* dist was validated above */
  phpc_host_fill(B, nn, dist, seed + 1);

  /* CUDA (here: HIP MFMA) pass */
  memset(C, 0, nn * sizeof(double));
  float cuda_gpu_time = 0.0f;
  double start_time = get_cur_time();
  phpc_gemm_summa_cuda(grid, A, B, C, N, gpu_count, grid_width, grid_height, tile_width, &cuda_gpu_time);
  double cuda_time = get_cur_time() - start_time;
  if (phpc_last_error() != PHPC_OK) {
    fprintf(stderr, "Error: %s\n", phpc_last_error_message());
    return EXIT_FAILURE;
  }

  int check_failed = 0;
  if (check > 0) {
    double err_cw = 0.0, err_nw = 0.0;
    rc = phpc_check_sampled_host(A, N, B, N, C, N, N, N, N, check, seed + 7, &err_cw, &err_nw);
    if (rc != PHPC_OK) {
      fprintf(stderr, "Error: %s\n", phpc_last_error_message());
      return EXIT_FAILURE;
    }
    const double bound = (N + 2.0) * 0x1.0p-53;
    check_failed = !(err_cw <= bound);
    fprintf(stderr, "check: %d samples, max componentwise error %.3e, normwise %.3e (bound %.3e) %s\n", check + 4,
            err_cw, err_nw, bound, check_failed ? "FAILED" : "ok");
  }

  /* vendor (cuBLAS in the reference, rocBLAS here) pass */
  double cublas_time = 0.0;
  if (vendor) {
    memset(C, 0, nn * sizeof(double));
    float cublas_gpu_time = 0.0f;
    start_time = get_cur_time();
    phpc_gemm_summa_cublas(grid, A, B, C, N, gpu_count, &cublas_gpu_time);
    cublas_time = get_cur_time() - start_time;
    if (phpc_last_error() != PHPC_OK) {
      fprintf(stderr, "Error: %s\n", phpc_last_error_message());
      return EXIT_FAILURE;
    }
  }

  log_to_csv(csv_file, N, size, gpu_count, grid_width * grid_height, tile_width * tile_width, cuda_time, cuda_gpu_time,
             cublas_time);
  fclose(csv_file);

  free(A);
  free(B);
  free(C);
  phpc_grid_destroy(grid);
  return check_failed ? 3 : 0;
}

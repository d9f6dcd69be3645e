/*
 * iterative.out — the CPU path of the benchmark harness (src/iterative.c:8-46, built
 * by the reference's Makefile:9 as `gcc ... -Wall`, i.e. without optimisation):
 *
 *   iterative.out <matrix_size> [--dist ramp|uniform|wide] [--seed S] [--out FILE]
 *
 * prints "n,seconds", the row the harness collects into csv/iterative.csv
 * (scripts/tests.sh:36-44) and divides by the GPU times in process() (:11-33).
 *   --dist, --seed  inputs as main.out generates them (default ramp, A[i] = B[i] = i,
 *                   the reference's own input); A uses seed S, B uses S + 1
 *   --out FILE      write C (n * n doubles, row-major, native byte order) to FILE
 *
 * The loop runs on one host thread with no GPU involved.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fill.h"
#include "utils.h"

/*
 * C[i][k] += A[i][j] * B[j][k] with j the middle loop (src/iterative.c:8-13): every
 * C[i][k] accumulates its products in ascending j, each product rounded and then added
 * (no fused multiply-add at the reference's default x86-64 flags).
 */
void phpc_gemm_iterative(const double *A, const double *B, double *C, int n) {
  for (int i = 0; i < n; ++i) {
    double *c = C + (size_t)i * n;
    for (int j = 0; j < n; ++j) {
      const double a = A[(size_t)i * n + j];
      const double *b = B + (size_t)j * n;
      for (int k = 0; k < n; ++k) {
        const double prod = a * b[k];
        c[k] = c[k] + prod;
      }
    }
  }
}

int main(int argc, char *argv[]) {
  if (argc < 2) {
    fprintf(stderr, "Usage: %s <matrix_size> [--dist ramp|uniform|wide] [--seed S] [--out FILE]\n", argv[0]);
    return 1;
  }
  const int n = atoi(argv[1]);
  const char *dist = "ramp", *out = NULL;
  unsigned long long seed = 1;
  for (int a = 2; a < argc; ++a) {
    const char *v = a + 1 < argc ? argv[a + 1] : NULL;
    if (!strcmp(argv[a], "--dist") && v) {
      dist = v;
      ++a;
    } else if (!strcmp(argv[a], "--seed") && v) {
      seed = strtoull(v, NULL, 10);
      ++a;
    } else if (!strcmp(argv[a], "--out") && v) {
      out = v;
      ++a;
    } else {
      fprintf(stderr, "Error: unknown option %s\n", argv[a]);
      return 1;
    }
  }
  if (n < 1) {
    fprintf(stderr, "Error: matrix size must be positive\n");
    return 1;
  }
  const size_t nn = (size_t)n * (size_t)n;
  double *A = (double *)malloc(nn * sizeof(double));
  double *B = (double *)malloc(nn * sizeof(double));
  double *C = (double *)malloc(nn * sizeof(double));
  if (!A || !B || !C) {
    fprintf(stderr, "Error: allocation of 3 x %zu bytes failed\n", nn * sizeof(double));
    return 1;
  }
  if (phpc_host_fill(A, nn, dist, seed) || phpc_host_fill(B, nn, dist, seed + 1)) {
    fprintf(stderr, "Error: --dist must be ramp, uniform or wide\n");
    return 1;
  }
  memset(C, 0, nn * sizeof(double));

  const double start = get_cur_time();
  phpc_gemm_iterative(A, B, C, n);
  const double elapsed = get_cur_time() - start;
  printf("%d,%lf\n", n, elapsed);

  int rc = 0;
  if (out) {
    FILE *f = fopen(out, "wb");
    if (!f || fwrite(C, sizeof(double), nn, f) != nn) {
      fprintf(stderr, "Error: could not write %s\n", out);
      rc = 1;
    }
    if (f) fclose(f);
  }
  free(C);
  free(B);
  free(A);
  return rc;
}

/*
 * CPU baseline CLI: `iterative.out <n>` prints "n,seconds" (src/iterative.c:15-46).
 * TEST / BASELINE INFRASTRUCTURE ONLY (part of the oracle; never linked by the product).
 * Same contract as the reference: ramp inputs A[i] = B[i] = i (src/iterative.c:30-31),
 * C zeroed (:33), wall time from gettimeofday (src/utils.c:9-17) around the loop (:35-37).
 * An optional second argument `u` selects uniform[-1,1) inputs from seed 1 instead.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "oracle.h"

static double wall_seconds(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + tv.tv_usec / 1000000.0;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "Usage: %s <matrix_size> [u]\n", argv[0]);
    return 1;
  }
  int n = atoi(argv[1]);
  size_t nn = (size_t)n * (size_t)n;
  double *A = malloc(nn * sizeof(double));
  double *B = malloc(nn * sizeof(double));
  double *C = malloc(nn * sizeof(double));
  if (!A || !B || !C) {
    fprintf(stderr, "allocation failed\n");
    return 1;
  }
  if (argc > 2 && argv[2][0] == 'u') {
    orc_fill_uniform(A, nn, 1);
    orc_fill_uniform(B, nn, 2);
  } else {
    orc_fill_ramp(A, nn);
    orc_fill_ramp(B, nn);
  }
  memset(C, 0, nn * sizeof(double));
  double t0 = wall_seconds();
  orc_iterative(A, B, C, n);
  double t = wall_seconds() - t0;
  printf("%d,%lf\n", n, t);
  free(C);
  free(B);
  free(A);
  return 0;
}

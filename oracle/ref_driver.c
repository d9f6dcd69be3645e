/*
 * Driver for the reference's CPU path on seeded inputs. TEST / BASELINE INFRASTRUCTURE
 * ONLY: linked with the reference's own src/iterative.c and src/utils.c (their main
 * renamed) by `make ref`, into oracle/_ref/. Never part of the product.
 *
 *   iterative_ref_O{0,2}.out <n> <seed> [out_file]
 *
 * A = uniform[-1,1) stream `seed`, B = stream seed + 1 (orc_fill_uniform, the values the
 * device fill writes), C zeroed as src/iterative.c:33 does, then the reference's
 * phpc_gemm_iterative (src/iterative.c:8-13) timed with its own get_cur_time
 * (src/utils.c:9-17), as its main times it (:35-37). Prints "n,seconds"; writes C to
 * out_file (n * n doubles, row-major) when given.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* defined by the reference's sources */
void phpc_gemm_iterative(const double *A, const double *B, double *C, int n);
double get_cur_time(void);

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "Usage: %s <n> <seed> [out_file]\n", argv[0]);
    return 1;
  }
  const int n = atoi(argv[1]);
  const unsigned long long seed = strtoull(argv[2], NULL, 10);
  if (n < 1) return 1;
  const size_t nn = (size_t)n * (size_t)n;
  double *A = malloc(nn * sizeof(double)), *B = malloc(nn * sizeof(double)), *C = malloc(nn * sizeof(double));
  if (!A || !B || !C) return 1;
  orc_fill_uniform(A, nn, seed);
  orc_fill_uniform(B, nn, seed + 1);
  memset(C, 0, nn * sizeof(double));
  const double t0 = get_cur_time();
  phpc_gemm_iterative(A, B, C, n);
  const double t = get_cur_time() - t0;
  printf("%d,%lf\n", n, t);
  int rc = 0;
  if (argc > 3) {
    FILE *f = fopen(argv[3], "wb");
    if (!f || fwrite(C, sizeof(double), nn, f) != nn) rc = 1;
    if (f) fclose(f);
  }
  free(A);
  free(B);
  free(C);
  return rc;
}

/*
 * Host generator of the CLIs' synthetic inputs (main.out, iterative.out): the same
 * values phpc_fill_device writes on the GPU, so a host run and a device run of one
 * seed see identical matrices.
 */
#ifndef PHPC_HOST_FILL_H
#define PHPC_HOST_FILL_H

#include <stddef.h>
#include <stdint.h>

/* Element i of stream `seed`, z = SplitMix64 mix of seed + (i + 1) * 0x9E3779B97F4A7C15:
 *   "ramp"    x[i] = i (the reference's input, src/main.c:85-86, src/iterative.c:30-31)
 *   "uniform" 2 ((z >> 11) 2^-53) - 1, in [-1, 1)
 *   "wide"    sign(z & 1) 2^e, e = ((z >> 32) 41 >> 32) - 20, in [-20, 20]
 * Returns 0, or -1 for an unknown distribution name (x untouched). */
int phpc_host_fill(double *x, size_t n, const char *dist, uint64_t seed);

#endif

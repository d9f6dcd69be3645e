/*
 * Timer and CSV row of the benchmark CLI. Replaces src/utils.h of the reference
 * (get_cur_time :6, log_to_csv :7) with the same signatures and output format.
 */
#ifndef PHPC_UTILS_H
#define PHPC_UTILS_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Wall-clock seconds (gettimeofday), src/utils.c:9-17. */
double get_cur_time(void);

/* One CSV row, no header (src/utils.c:19-30):
 * N,size,gpu_count,num_blocks,threads_per_block,total_threads,cuda_time,cuda_gpu_time,cublas_time
 * with total_threads = gpu_count * num_blocks * threads_per_block. */
void log_to_csv(FILE *csv_file, int N, int size, int gpu_count, int num_blocks, int threads_per_block,
                double cuda_time, float cuda_gpu_time, double cublas_time);

#ifdef __cplusplus
}
#endif
#endif

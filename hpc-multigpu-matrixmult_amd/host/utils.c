/* Timer and CSV row (src/utils.c:9-30). The row format is kept byte-for-byte so the
 * reference's scripts/tests.sh and scripts/plots.py read this build's CSV files. */
#include "utils.h"

#include <sys/time.h>

double get_cur_time(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return (double)tv.tv_sec + (double)tv.tv_usec / 1000000.0;
}

void log_to_csv(FILE *csv_file, int N, int size, int gpu_count, int num_blocks, int threads_per_block,
                double cuda_time, float cuda_gpu_time, double cublas_time) {
  const int total_threads = gpu_count * num_blocks * threads_per_block;
  fprintf(csv_file, "%d,%d,%d,%d,%d,%d,%f,%f,%f\n", N, size, gpu_count, num_blocks, threads_per_block, total_threads,
          cuda_time, cuda_gpu_time, cublas_time);
}

// generate_bit_transition_cache — writes ./transition_counts/cache_{n'}_{k'} for
// 0 <= k' < n' <= n, byte-identical to the reference generator
// (backend/generate_bit_transition_cache.cpp:46-77, cache_io.h:7-26), with the
// counts computed on the GPU.
//   generate_bit_transition_cache n
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>

#include "../../include/bdc.h"

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "Usage %s n\n", argv[0]);
    return 1;
  }
  const int n = atoi(argv[1]);
  mkdir("transition_counts", 0755);
  if (bdc_write_caches(n, "transition_counts") != BDC_OK) {
    fprintf(stderr, "Error: %s\n", bdc_last_error());
    return 2;
  }
  return 0;
}

// main_BAA — command-line driver of the MI355X Blahut–Arimoto engine with the
// reference's surface (backend/main_BAA.cu:134-170):
//
//   main_BAA n k [read_from_file=0] [a=0.005] [save_iterations=100] [precision=fp64|split16]
//
// Runs BA iterations on BDC_{n,k} until mx = max_x log2(p_{t+1}/p_t) < a
// (main_BAA.cu:73), saves distributions/dist_n_k.txt every save_iterations
// iterations and at the end (%.17g per line, ascending representatives,
// main_BAA.cu:83-96,113-120) and prints the rate I(p) in bits with the last mx
// (main_BAA.cu:129-130), followed by the FP64-certified upper bound
// max_x D(x||q).  Caches are read from ./transition_counts when the needed
// files are there (their sizes are checked) and generated on the GPU otherwise.
// Differences from the reference (checkpoint hardening):
//  - the dist file is written through a temporary file and renamed, so a crash
//    mid-save leaves the previous checkpoint intact;
//  - beside it, dist_n_k.txt.meta records n, k, precision, the iteration count
//    and the last mx; a resumed run continues the iteration count from it;
//  - a resumed file must hold one finite, non-negative value per orbit that sum
//    (with orbit sizes) to 1 within 1e-6 (the reference reads whatever is there,
//    main_BAA.cu:3-18);
//  - values are formatted and parsed on all host threads (std::to_chars /
//    std::from_chars, the same digits as printf/strtod with %.17g): the file is
//    about 1.6 GB of text at n = 28;
//  - k must be in [1, n] (the reference does not check it); save_iterations 0
//    disables periodic saves (the reference divides by it); distributions/ is
//    created when missing.
#include <sys/stat.h>

#include <algorithm>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bdc.h"
#include "dist_io.hpp"

using dist_io::load_dist;
using dist_io::load_meta;
using dist_io::Meta;
using dist_io::save_dist;
using dist_io::save_meta;

static int die(const char *what) {
  fprintf(stderr, "Error: %s: %s\n", what, bdc_last_error());
  return 1;
}

static bool file_exists(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "Usage: %s <n> <k> [read_from_file] [a] [save_iterations] [fp64|split16]\n", argv[0]);
    fprintf(stderr, "Example: %s 21 20\nExample: %s 21 20 1 0.01 200\n", argv[0], argv[0]);
    return 1;
  }
  const int n = atoi(argv[1]), k = atoi(argv[2]);
  const bool read_from_file = argc > 3 && atoi(argv[3]) != 0;
  const double a = argc > 4 ? atof(argv[4]) : 0.005;
  const long save_iterations = argc > 5 ? atol(argv[5]) : 100;
  int prec = BDC_PREC_FP64;
  if (argc > 6) {
    if (!strcmp(argv[6], "split16"))
      prec = BDC_PREC_SPLIT16;
    else if (strcmp(argv[6], "fp64")) {
      fprintf(stderr, "precision must be fp64 or split16\n");
      return 1;
    }
  }
  if (!bdc_supported(n, k)) {
    fprintf(stderr, "Error: BDC_{%d,%d} not supported: %s\n", n, k, bdc_last_error());
    return 1;
  }
  // load caches when the half-length files are present (cache_io.h:7-14 names);
  // outputs longer than 20 symbols and k = n run the deletion-chain path, which
  // reads no count caches
  const int n1 = (n + 1) / 2;
  std::string cache_dir = "transition_counts";
  bool have = k <= 20 && k < n;
  for (int j = 0; j < n1 && have; ++j)
    have = file_exists(cache_dir + "/cache_" + std::to_string(n1) + "_" + std::to_string(j)) &&
           file_exists(cache_dir + "/cache_" + std::to_string(n - n1) + "_" + std::to_string(j < n - n1 ? j : 0));
  if (!have && k <= 20 && k < n)
    printf("Cache files not found in ./%s: generating the counts on the GPU\n", cache_dir.c_str());

  bdc_ctx *ctx = nullptr;
  if (bdc_create(n, k, have ? cache_dir.c_str() : nullptr, &ctx) != BDC_OK) return die("create");
  uint64_t norb = 0;
  bdc_num_orbits(ctx, &norb);
  mkdir("distributions", 0755);
  const std::string file = "distributions/dist_" + std::to_string(n) + "_" + std::to_string(k) + ".txt";
  const std::string metafile = file + ".meta";
  Meta meta;
  meta.n = n;
  meta.k = k;
  meta.precision = prec == BDC_PREC_SPLIT16 ? "split16" : "fp64";
  long first = 1;
  if (read_from_file) {
    printf("Loading distribution from file: %s\n", file.c_str());
    std::vector<double> Q;
    if (!load_dist(file, Q)) {
      fprintf(stderr, "Error: Could not read %s.\n", file.c_str());
      return 1;
    }
    if (bdc_set_p_orbits(ctx, Q.data(), Q.size()) != BDC_OK) return die("read distribution");
    Meta old;
    if (load_meta(metafile, old) && old.n == n && old.k == k && old.iteration > 0) {
      first = old.iteration + 1;
      printf("Resuming after iteration %ld (last mx = %.6f)\n", old.iteration, old.mx);
    }
  } else {
    printf("Initializing uniform distribution\n");
    if (bdc_set_p_uniform(ctx) != BDC_OK) return die("init");
  }
  printf("Starting BAA iterations...\n");
  std::vector<double> Q(norb);
  double mx = a;
  long it = first;
  auto checkpoint = [&](long iteration) {
    if (!save_dist(file, Q)) {
      fprintf(stderr, "Warning: could not write %s\n", file.c_str());
      return false;
    }
    meta.iteration = iteration;
    meta.mx = mx;
    if (!save_meta(metafile, meta)) fprintf(stderr, "Warning: could not write %s\n", metafile.c_str());
    return true;
  };
  // Steps run in calls of up to save_iterations steps (to the next save point)
  // with the stopping rule inside the library, which stops after the first step
  // with mx < a (main_BAA.cu:73) without a host round trip per step.
  const long chunk_max = save_iterations > 0 ? save_iterations : 1000;
  std::vector<bdc_iter_stats> st((size_t)chunk_max);
  while (mx >= a) {
    const long to_save = save_iterations > 0 ? save_iterations - ((it - 1) % save_iterations) : chunk_max;
    const int want = (int)std::min(to_save, chunk_max);
    const auto t0 = std::chrono::steady_clock::now();
    int done = 0;
    if (bdc_iterate(ctx, prec, want, a, st.data(), &done) != BDC_OK) return die("iterate");
    const double per = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / std::max(done, 1);
    for (int s = 0; s < done; ++s, ++it) {
      mx = st[(size_t)s].mx_bits;
      printf("Iteration %ld: mx = %.6f, time = %.3f s\n", it, mx, per);
    }
    if (done == 0) break;
    if (save_iterations > 0 && (it - 1) % save_iterations == 0) {
      const auto s0 = std::chrono::steady_clock::now();
      if (bdc_get_p_orbits(ctx, Q.data(), norb) != BDC_OK) return die("get p");
      if (checkpoint(it - 1))
        printf("Distribution saved in %.3f seconds\n",
               std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count());
    }
  }
  if (bdc_get_p_orbits(ctx, Q.data(), norb) != BDC_OK) return die("get p");
  if (checkpoint(it - 1)) printf("Final distribution saved to %s\n", file.c_str());
  bdc_certificate cert;
  if (bdc_certify(ctx, &cert) != BDC_OK) return die("certify");
  const double ln2 = std::log(2.0);
  printf("Computed C_{%d,%d} = %.6f bits (max deviation %.6f)\n", n, k, cert.lower / ln2, mx);
  printf("Certified upper bound max_x D(x||q) = %.10f bits (I(p) + mx = %.10f)\n", cert.upper_cert / ln2,
         cert.lower / ln2 + mx);
  bdc_destroy(ctx);
  return 0;
}

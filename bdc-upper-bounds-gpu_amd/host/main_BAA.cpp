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
// Differences from the reference: the dist file is written through a temporary
// file and renamed; a resumed file must have one value per orbit; k must be
// below n; save_iterations 0 disables periodic saves (the reference divides by it).
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/bdc.h"

static int die(const char *what) {
  fprintf(stderr, "Error: %s: %s\n", what, bdc_last_error());
  return 1;
}

static bool file_exists(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

static bool save_dist(const std::string &file, const std::vector<double> &Q) {
  const std::string tmp = file + ".tmp";
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) return false;
  for (double v : Q) fprintf(f, "%.17g\n", v);
  if (fclose(f) != 0) return false;
  return rename(tmp.c_str(), file.c_str()) == 0;
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
  // load caches when the half-length files are present (cache_io.h:7-14 names)
  const int n1 = (n + 1) / 2;
  std::string cache_dir = "transition_counts";
  bool have = true;
  for (int j = 0; j < n1 && have; ++j)
    have = file_exists(cache_dir + "/cache_" + std::to_string(n1) + "_" + std::to_string(j)) &&
           file_exists(cache_dir + "/cache_" + std::to_string(n - n1) + "_" + std::to_string(j < n - n1 ? j : 0));
  if (!have) printf("Cache files not found in ./%s: generating the counts on the GPU\n", cache_dir.c_str());

  bdc_ctx *ctx = nullptr;
  if (bdc_create(n, k, have ? cache_dir.c_str() : nullptr, &ctx) != BDC_OK) return die("create");
  uint64_t norb = 0;
  bdc_num_orbits(ctx, &norb);
  const std::string file = "distributions/dist_" + std::to_string(n) + "_" + std::to_string(k) + ".txt";
  if (read_from_file) {
    printf("Loading distribution from file: %s\n", file.c_str());
    std::ifstream in(file);
    if (!in) {
      fprintf(stderr, "Error: Could not open %s for reading.\n", file.c_str());
      return 1;
    }
    std::vector<double> Q;
    double v;
    while (in >> v) Q.push_back(v);
    if (bdc_set_p_orbits(ctx, Q.data(), Q.size()) != BDC_OK) return die("read distribution");
  } else {
    printf("Initializing uniform distribution\n");
    if (bdc_set_p_uniform(ctx) != BDC_OK) return die("init");
  }
  printf("Starting BAA iterations...\n");
  std::vector<double> Q(norb);
  double mx = a;
  long it = 1;
  while (mx >= a) {
    const auto t0 = std::chrono::steady_clock::now();
    bdc_iter_stats st;
    int done = 0;
    if (bdc_iterate(ctx, prec, 1, 0.0, &st, &done) != BDC_OK) return die("iterate");
    mx = st.mx_bits;
    if (save_iterations > 0 && it % save_iterations == 0) {
      const auto s0 = std::chrono::steady_clock::now();
      if (bdc_get_p_orbits(ctx, Q.data(), norb) != BDC_OK) return die("get p");
      if (!save_dist(file, Q)) fprintf(stderr, "Warning: could not write %s\n", file.c_str());
      printf("Distribution saved in %.3f seconds\n",
             std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count());
    }
    printf("Iteration %ld: mx = %.6f, time = %.3f s\n", it, mx,
           std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    ++it;
  }
  if (bdc_get_p_orbits(ctx, Q.data(), norb) != BDC_OK) return die("get p");
  if (save_dist(file, Q))
    printf("Final distribution saved to %s\n", file.c_str());
  else
    fprintf(stderr, "Warning: could not write %s\n", file.c_str());
  bdc_certificate cert;
  if (bdc_certify(ctx, &cert) != BDC_OK) return die("certify");
  const double ln2 = std::log(2.0);
  printf("Computed C_{%d,%d} = %.6f bits (max deviation %.6f)\n", n, k, cert.lower / ln2, mx);
  printf("Certified upper bound max_x D(x||q) = %.10f bits (I(p) + mx = %.10f)\n", cert.upper_cert / ln2,
         cert.lower / ln2 + mx);
  bdc_destroy(ctx);
  return 0;
}

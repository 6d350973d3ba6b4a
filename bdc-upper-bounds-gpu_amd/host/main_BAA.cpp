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
//  - k must be below n; save_iterations 0 disables periodic saves (the
//    reference divides by it); distributions/ is created when missing.
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

static int die(const char *what) {
  fprintf(stderr, "Error: %s: %s\n", what, bdc_last_error());
  return 1;
}

static bool file_exists(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

static unsigned host_threads() { return std::max(1u, std::min(64u, std::thread::hardware_concurrency())); }

// one "%.17g\n" line per value, formatted in parallel chunks
static bool save_dist(const std::string &file, const std::vector<double> &Q) {
  const size_t nv = Q.size();
  const unsigned nt = (unsigned)std::min<size_t>(host_threads(), std::max<size_t>(1, nv / 65536));
  std::vector<std::string> parts(nt);
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      const size_t a = nv * t / nt, b = nv * (t + 1) / nt;
      std::string &out = parts[t];
      out.resize((b - a) * 26);
      char *p = out.data();
      for (size_t i = a; i < b; ++i) {
        p = std::to_chars(p, out.data() + out.size(), Q[i], std::chars_format::general, 17).ptr;
        *p++ = '\n';
      }
      out.resize(p - out.data());
    });
  for (auto &th : pool) th.join();
  const std::string tmp = file + ".tmp";
  FILE *f = fopen(tmp.c_str(), "wb");
  if (!f) return false;
  bool ok = true;
  for (const auto &s : parts) ok = ok && fwrite(s.data(), 1, s.size(), f) == s.size();
  ok = (fclose(f) == 0) && ok;
  return ok && rename(tmp.c_str(), file.c_str()) == 0;
}

// whitespace-separated values, parsed in parallel chunks split at whitespace
static bool load_dist(const std::string &file, std::vector<double> &Q) {
  FILE *f = fopen(file.c_str(), "rb");
  if (!f) return false;
  std::string buf;
  fseek(f, 0, SEEK_END);
  const long len = ftell(f);
  fseek(f, 0, SEEK_SET);
  buf.resize(len > 0 ? (size_t)len : 0);
  const bool rd = buf.empty() || fread(buf.data(), 1, buf.size(), f) == buf.size();
  fclose(f);
  if (!rd) return false;
  const unsigned nt = (unsigned)std::min<size_t>(host_threads(), std::max<size_t>(1, buf.size() / (1 << 20)));
  std::vector<size_t> cut(nt + 1, buf.size());
  cut[0] = 0;
  auto space = [](char ch) { return ch == ' ' || ch == '\n' || ch == '\r' || ch == '\t'; };
  for (unsigned t = 1; t < nt; ++t) {
    size_t c = buf.size() * t / nt;
    while (c < buf.size() && !space(buf[c])) ++c;
    cut[t] = c;
  }
  std::vector<std::vector<double>> parts(nt);
  std::vector<int> bad(nt, 0);
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      const char *p = buf.data() + cut[t], *e = buf.data() + cut[t + 1];
      while (true) {
        while (p < e && space(*p)) ++p;
        if (p >= e) break;
        double v;
        auto r = std::from_chars(p, e, v);
        if (r.ec != std::errc()) {
          bad[t] = 1;
          break;
        }
        parts[t].push_back(v);
        p = r.ptr;
      }
    });
  for (auto &th : pool) th.join();
  Q.clear();
  for (unsigned t = 0; t < nt; ++t) {
    if (bad[t]) return false;
    Q.insert(Q.end(), parts[t].begin(), parts[t].end());
  }
  return true;
}

struct Meta {
  int n = -1, k = -1;
  long iteration = 0;
  double mx = NAN;
  std::string precision;
};

static bool save_meta(const std::string &file, const Meta &m) {
  const std::string tmp = file + ".tmp";
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) return false;
  fprintf(f, "n %d\nk %d\nprecision %s\niteration %ld\nmx %.17g\n", m.n, m.k, m.precision.c_str(), m.iteration, m.mx);
  if (fclose(f) != 0) return false;
  return rename(tmp.c_str(), file.c_str()) == 0;
}

static bool load_meta(const std::string &file, Meta &m) {
  FILE *f = fopen(file.c_str(), "r");
  if (!f) return false;
  char key[32], val[64];
  while (fscanf(f, "%31s %63s", key, val) == 2) {
    if (!strcmp(key, "n")) m.n = atoi(val);
    else if (!strcmp(key, "k")) m.k = atoi(val);
    else if (!strcmp(key, "iteration")) m.iteration = atol(val);
    else if (!strcmp(key, "mx")) m.mx = strtod(val, nullptr);
    else if (!strcmp(key, "precision")) m.precision = val;
  }
  fclose(f);
  return true;
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
  while (mx >= a) {
    const auto t0 = std::chrono::steady_clock::now();
    bdc_iter_stats st;
    int done = 0;
    if (bdc_iterate(ctx, prec, 1, 0.0, &st, &done) != BDC_OK) return die("iterate");
    mx = st.mx_bits;
    if (save_iterations > 0 && it % save_iterations == 0) {
      const auto s0 = std::chrono::steady_clock::now();
      if (bdc_get_p_orbits(ctx, Q.data(), norb) != BDC_OK) return die("get p");
      if (checkpoint(it))
        printf("Distribution saved in %.3f seconds\n",
               std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count());
    }
    printf("Iteration %ld: mx = %.6f, time = %.3f s\n", it, mx,
           std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    ++it;
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

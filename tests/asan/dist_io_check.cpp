// Host-only check of main_BAA's distribution-file and sidecar I/O
// (bdc-upper-bounds-gpu_amd/host/dist_io.hpp), built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_host_sanitizers.py.  Round trip of
// %.17g values, then malformed files that must be rejected without a fault.
#include <cstdio>
#include <random>

#include "../../bdc-upper-bounds-gpu_amd/host/dist_io.hpp"

static int failures = 0;
#define CHECK(c)                                                     \
  do {                                                               \
    if (!(c)) {                                                      \
      fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      ++failures;                                                    \
    }                                                                \
  } while (0)

static void put(const std::string &path, const std::string &text) {
  FILE *f = fopen(path.c_str(), "wb");
  fwrite(text.data(), 1, text.size(), f);
  fclose(f);
}

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  const std::string file = dir + "/dist.txt";
  // round trip: bit-exact through %.17g, over enough values to use every thread
  std::mt19937_64 rng(7);
  std::vector<double> q(300000);
  for (auto &v : q) v = std::ldexp((double)(rng() >> 11), -53 - (int)(rng() % 900));
  q[0] = 0.0;
  q[1] = 4.9406564584124654e-324;
  q[2] = 1.0;
  CHECK(dist_io::save_dist(file, q));
  std::vector<double> back;
  CHECK(dist_io::load_dist(file, back));
  CHECK(back.size() == q.size());
  bool same = back.size() == q.size();
  for (size_t i = 0; same && i < q.size(); ++i) same = back[i] == q[i];
  CHECK(same);
  // accepted layouts: CRLF, tabs, no final newline, leading blanks
  put(file, "  0.5\r\n0.25\t0.25");
  CHECK(dist_io::load_dist(file, back) && back.size() == 3 && back[2] == 0.25);
  put(file, "");
  CHECK(dist_io::load_dist(file, back) && back.empty());
  // rejected: words, a dangling exponent, a sign alone, out-of-range values
  for (const char *bad : {"0.5\nx\n", "1e\n", "-\n", "0.1 0.2 zz", "1e999999\n", "nanx"}) {
    put(file, bad);
    CHECK(!dist_io::load_dist(file, back));
  }
  // a long run of digits straddling the per-thread chunk boundaries
  std::string big(3u << 20, '1');
  big += "\n";
  put(file, big);
  CHECK(!dist_io::load_dist(file, back) || back.size() == 1);
  CHECK(!dist_io::load_dist(dir + "/missing.txt", back));
  // sidecar
  dist_io::Meta m;
  m.n = 16;
  m.k = 8;
  m.iteration = 42;
  m.mx = 0.125;
  m.precision = "split16";
  CHECK(dist_io::save_meta(file + ".meta", m));
  dist_io::Meta r;
  CHECK(dist_io::load_meta(file + ".meta", r));
  CHECK(r.n == 16 && r.k == 8 && r.iteration == 42 && r.mx == 0.125 && r.precision == "split16");
  put(file + ".meta", std::string(5000, 'z') + " " + std::string(5000, 'y') + "\nn\n");
  dist_io::Meta g;
  CHECK(dist_io::load_meta(file + ".meta", g));
  printf("dist_io checks: %d failures\n", failures);
  return failures ? 1 : 0;
}

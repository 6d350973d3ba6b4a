// dist_io.hpp — distribution-file and checkpoint-sidecar I/O of main_BAA.
//
// dist_n_k.txt: one value per orbit, ascending representatives, written as
// "%.17g\n" (backend/main_BAA.cu:83-96) and read as whitespace-separated
// values (backend/main_BAA.cu:3-18).  Formatting and parsing run on all host
// threads with std::to_chars / std::from_chars (the same digits as printf and
// strtod at %.17g): the file is about 1.6 GB of text at n = 28.  Writes go
// through a temporary file and a rename.  Host-only code (no HIP), so the
// sanitizer test builds it on its own.
#pragma once
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace dist_io {

inline unsigned host_threads() { return std::max(1u, std::min(64u, std::thread::hardware_concurrency())); }

// one "%.17g\n" line per value, formatted in parallel chunks
inline bool save_dist(const std::string &file, const std::vector<double> &Q) {
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
inline bool load_dist(const std::string &file, std::vector<double> &Q) {
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

inline bool save_meta(const std::string &file, const Meta &m) {
  const std::string tmp = file + ".tmp";
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) return false;
  fprintf(f, "n %d\nk %d\nprecision %s\niteration %ld\nmx %.17g\n", m.n, m.k, m.precision.c_str(), m.iteration, m.mx);
  if (fclose(f) != 0) return false;
  return rename(tmp.c_str(), file.c_str()) == 0;
}

inline bool load_meta(const std::string &file, Meta &m) {
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

}  // namespace dist_io

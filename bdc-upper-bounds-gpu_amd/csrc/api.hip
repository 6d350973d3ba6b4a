// C ABI of the engine (include/bdc.h): context lifetime, caches, distribution
// I/O, the BA step pipeline and the bounds.  Every entry point returns a
// status; nothing exits.
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "bdc_internal.hpp"
#include "device_common.hpp"
#include "kernels.hpp"

namespace bdc {

static thread_local std::string g_err;

// bdc_iterate with a stopping rule: steps per batch, and the largest updated
// part of P (bytes) for which batching (and its snapshot copy) is used.  A
// round trip per step costs tens of microseconds; copying 256 MiB costs
// about 0.1 ms, once per batch.
constexpr int kStopBatch = 32;
constexpr size_t kStopBatchMaxBytes = (size_t)256 << 20;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}
int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// Set for the duration of a bdc_iterate_noalloc call on this thread: any lazy
// device allocation on its path fails instead of calling hipMalloc.
static thread_local bool g_forbid_alloc = false;

int DevBuf::alloc(size_t b) {
  if (g_forbid_alloc)
    return fail(BDC_ERR_STATE, "this call allocates no device memory, but %zu bytes were needed (prepare the context "
                               "with one bdc_iterate at this precision first)", b);
  release();
  if (b == 0) return BDC_OK;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    return fail(BDC_ERR_MEMORY, "hipMalloc(%zu bytes) failed: %s", b, hipGetErrorString(e));
  }
  bytes = b;
  return BDC_OK;
}
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

static int check_launch(int rc, const char *what) {
  if (rc != BDC_OK) {
    hipError_t e = hipGetLastError();
    return fail(rc, "%s: launch failed (%s)", what, hipGetErrorString(e));
  }
  return BDC_OK;
}
#define BDC_LAUNCH(expr) BDC_TRY(check_launch((expr), #expr))
// inside bdc_create: release the half-built context on failure
#define BDC_TRY_C(expr)                 \
  do {                                  \
    int r_ = (expr);                    \
    if (r_ != BDC_OK) return cleanup(r_); \
  } while (0)

// ------------------------------------------------------- no-allocation steps
// bdc_iterate_noalloc keeps its per-step statistics (4 doubles per step) in
// this file-scope device array instead of a buffer the context grows with
// hipMalloc.  Every GPU the library runs on holds one copy until the process
// ends; its address on each device is looked up once, with that device
// current, when the first context on it is created.  Calls on one GPU share
// the array: calls that overlap on one GPU race (each call writes the rows of
// its own steps before reading them, so nothing needs clearing).
constexpr int kNoAllocMaxSteps = 4096;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocMaxSteps * 4];
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices] = {};

static int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(BDC_ERR_HIP, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return BDC_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(BDC_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
  g_scratch_dev[device] = static_cast<double *>(ptr);
  return BDC_OK;
}
static double *scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch_dev[device] : nullptr;
}

static uint32_t nck_u32(uint32_t n, uint32_t k) {  // backend/BAA.cu:15-24
  if (k > n) return 0;
  if (k > n - k) k = n - k;
  uint64_t r = 1;
  for (uint32_t i = 1; i <= k; ++i) r = r * (n - i + 1) / i;
  return (uint32_t)r;
}

static uint64_t burnside_orbits(int n) {
  const uint64_t full = 1ull << n, h = 1ull << ((n + 1) / 2);
  const uint64_t pal = (n % 2 == 0) ? (1ull << (n / 2)) : 0ull;  // anti-palindromes exist only for even n
  return (full + h + pal) / 4;
}

static int timer_begin(bdc_ctx *c, Timer &t) {
  if (!c->timing || g_forbid_alloc) return BDC_OK;
  if ((int)t.start.size() <= t.used) {
    hipEvent_t a, b;
    BDC_HIP(hipEventCreate(&a));
    BDC_HIP(hipEventCreate(&b));
    t.start.push_back(a);
    t.stop.push_back(b);
  }
  BDC_HIP(hipEventRecord(t.start[t.used], c->stream));
  return BDC_OK;
}
static int timer_end(bdc_ctx *c, Timer &t) {
  if (!c->timing || g_forbid_alloc) return BDC_OK;
  BDC_HIP(hipEventRecord(t.stop[t.used], c->stream));
  ++t.used;
  return BDC_OK;
}

// ------------------------------------------------------------- caches
// cache_io.h:7-14: <dir>/cache_<n'>_<k'>
static std::string cache_path(const char *dir, int len, int m) {
  return std::string(dir) + "/cache_" + std::to_string(len) + "_" + std::to_string(m);
}

// Entries per chunk when a cache file is scanned or converted (64 MiB of u32).
constexpr size_t kCacheChunk = (size_t)1 << 24;

// The reference loads whatever bytes the file holds (cache_io.h:29-44).  Here a
// file must exist, hold exactly 2^(n'+k') raw uint32 counts (cache_io.h:17-19)
// and every count must fit the u16 device layout; checked before any device
// work, in chunks (cache_16_15 alone is 8 GiB).
static int validate_cache_file(const char *dir, int len, int m) {
  const std::string path = cache_path(dir, len, m);
  struct stat st;
  if (stat(path.c_str(), &st) != 0) return fail(BDC_ERR_IO, "cannot open %s", path.c_str());
  const size_t cnt = (size_t)1 << (len + m);
  if ((uint64_t)st.st_size != cnt * sizeof(uint32_t))
    return fail(BDC_ERR_IO, "%s: %lld bytes, expected %zu uint32 counts (%zu bytes)", path.c_str(),
                (long long)st.st_size, cnt, cnt * sizeof(uint32_t));
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return fail(BDC_ERR_IO, "cannot open %s", path.c_str());
  std::vector<uint32_t> h(std::min(cnt, kCacheChunk));
  for (size_t done = 0; done < cnt;) {
    const size_t want = std::min(h.size(), cnt - done);
    if (fread(h.data(), sizeof(uint32_t), want, f) != want) {
      fclose(f);
      return fail(BDC_ERR_IO, "%s: short read at count %zu", path.c_str(), done);
    }
    for (size_t i = 0; i < want; ++i)
      if (h[i] > 65535u) {
        fclose(f);
        return fail(BDC_ERR_IO, "%s: count %u at index %zu does not fit the u16 device layout", path.c_str(), h[i],
                    done + i);
      }
    done += want;
  }
  fclose(f);
  return BDC_OK;
}

// Half-length caches a context for (n, k) reads: cache_{n1,j} for the splits j
// (j < n1) and cache_{n2,k-j} (k - j < n2); the identity factors j = n1 and
// k - j = n2 are implicit (backend/BAA.cu:371-384).
static std::vector<std::pair<int, int>> needed_caches(int n, int k) {
  const int n1 = (n + 1) / 2, n2 = n - n1;
  std::vector<std::pair<int, int>> out;
  auto add = [&](int len, int m) {
    for (auto &p : out)
      if (p.first == len && p.second == m) return;
    out.push_back({len, m});
  };
  for (int j = std::max(0, k - n2); j <= std::min(k, n1); ++j) {
    if (j < n1) add(n1, j);
    if (k - j < n2) add(n2, k - j);
  }
  return out;
}

static int make_cache_u16(bdc_ctx *c, int len, int m, DevBuf &dst, const char *dir) {
  BDC_TRY(dst.alloc(sizeof(uint16_t) << (len + m)));
  if (dir && dir[0]) {
    // validated by bdc_create before any device work; converted chunk by chunk
    const std::string path = cache_path(dir, len, m);
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return fail(BDC_ERR_IO, "cannot open %s", path.c_str());
    const size_t cnt = (size_t)1 << (len + m);
    std::vector<uint32_t> h(std::min(cnt, kCacheChunk));
    std::vector<uint16_t> h16(h.size());
    for (size_t done = 0; done < cnt;) {
      const size_t want = std::min(h.size(), cnt - done);
      if (fread(h.data(), sizeof(uint32_t), want, f) != want) {
        fclose(f);
        return fail(BDC_ERR_IO, "%s: short read at count %zu", path.c_str(), done);
      }
      for (size_t i = 0; i < want; ++i) {
        if (h[i] > 65535u) {
          fclose(f);
          return fail(BDC_ERR_IO, "%s: count %u does not fit the u16 device layout", path.c_str(), h[i]);
        }
        h16[i] = (uint16_t)h[i];
      }
      const hipError_t e = hipMemcpyAsync(dst.as<uint16_t>() + done, h16.data(), want * sizeof(uint16_t),
                                          hipMemcpyHostToDevice, c->stream);
      const hipError_t e2 = e == hipSuccess ? hipStreamSynchronize(c->stream) : e;  // h16 is reused
      if (e2 != hipSuccess) {
        fclose(f);
        return fail(BDC_ERR_HIP, "cache upload: %s", hipGetErrorString(e2));
      }
      done += want;
    }
    fclose(f);
    return BDC_OK;
  }
  BDC_LAUNCH(launch_count_cache_u16(len, m, dst.as<uint16_t>(), c->stream));
  return BDC_OK;
}

static int g64_g1_splits(const bdc_ctx *c, int rows);

// BDC_QUARTER: unset or 1 = bit-reversed column storage and quarter-space BA
// steps; "cols" = the column order only; 0 = natural order, half-space steps
static bool column_order_enabled() {
  const char *e = getenv("BDC_QUARTER");
  return !(e && e[0] == '0');
}
static bool quarter_enabled() {
  const char *e = getenv("BDC_QUARTER");
  return !(e && (e[0] == '0' || !strcmp(e, "cols")));
}

// The deletion-chain path (no count caches, FP64) serves k close to n: above
// k = 20 the dense split contractions would hold 2^k outputs per split, and
// at k = n both halves are consumed whole.  BDC_CHAIN_PATH=1 forces it (tests).
static bool chain_path(int n, int k) {
  const char *e = getenv("BDC_CHAIN_PATH");
  if (e && e[0] == '1') return true;
  return k > kDenseMaxK || k == n;
}

// The entropy term by enumerating distinct subsequences (kernels_chain.hip)
// where they are few against the 2^k outputs of the dense sweep; forced with
// BDC_ENTROPY_DFS=1 (0 forbids it where the dense sweep applies).
static bool entropy_by_enumeration(int n, int k) {
  if (n - k > 16) return false;
  if (k == n) return true;
  const char *e = getenv("BDC_ENTROPY_DFS");
  if (e && e[0]) return e[0] == '1' || k > kDenseMaxK;
  return k > kDenseMaxK;
}

// Bit groups of the chain path, top to bottom: the bottom group (at most 13
// bits) is a contiguous row per block; the bits above it form groups of at
// most 9, each processed in tiles of 16 low columns.  Level buffer slot t
// (2^(n-t) doubles) holds the strings after t deletions.
static int chain_setup(bdc_ctx *c) {
  const int n = c->n, d = c->n - c->k;
  const int bottom = std::min(n, 13), rest = n - bottom;
  const int ntop = (rest + 8) / 9;
  c->ngrp = 0;
  int above = 0;
  for (int g = 0; g < ntop; ++g) {
    const int L = rest / ntop + (g < rest % ntop ? 1 : 0);
    above += L;
    c->grpL[c->ngrp] = L;
    c->grpB[c->ngrp] = n - above;
    ++c->ngrp;
  }
  c->grpL[c->ngrp] = bottom;
  c->grpB[c->ngrp] = 0;
  ++c->ngrp;
  c->lvl_off.assign(d + 2, 0);
  for (int t = 0; t <= d; ++t) c->lvl_off[t + 1] = c->lvl_off[t] + ((uint64_t)1 << (n - t));
  return c->lvl.alloc(c->lvl_off[d + 1] * 8);
}

}  // namespace bdc

using namespace bdc;

extern "C" {

const char *bdc_last_error(void) { return g_err.c_str(); }
int bdc_version(void) { return 1; }

int bdc_generate_cache(int n, int k, uint32_t *out) {
  if (n < 0 || n > 20 || k < 0 || k > n || n + k > 32 || !out)
    return fail(BDC_ERR_ARG, "bdc_generate_cache: need 0 <= k <= n <= 20 and n + k <= 32");
  DevBuf d;
  const size_t bytes = sizeof(uint32_t) << (n + k);
  BDC_TRY(d.alloc(bytes));
  int rc = launch_count_cache_u32(n, k, d.as<uint32_t>(), 0);
  if (rc != BDC_OK) {
    d.release();
    return fail(rc, "count cache kernel failed");
  }
  hipError_t e = hipMemcpy(out, d.p, bytes, hipMemcpyDeviceToHost);
  d.release();
  if (e != hipSuccess) return fail(BDC_ERR_HIP, "hipMemcpy: %s", hipGetErrorString(e));
  return BDC_OK;
}

int bdc_write_caches(int n_max, const char *dir) {
  // generate_bit_transition_cache.cpp:56-73: 0 <= k' < n' <= n_max
  if (n_max < 0 || n_max > 20 || !dir) return fail(BDC_ERR_ARG, "bdc_write_caches: need 0 <= n_max <= 20");
  std::vector<uint32_t> h;
  for (int n = 0; n <= n_max; ++n)
    for (int k = 0; k < n; ++k) {
      h.resize((size_t)1 << (n + k));
      BDC_TRY(bdc_generate_cache(n, k, h.data()));
      std::string path = std::string(dir) + "/cache_" + std::to_string(n) + "_" + std::to_string(k);
      std::string tmp = path + ".tmp";
      FILE *f = fopen(tmp.c_str(), "wb");
      if (!f) return fail(BDC_ERR_IO, "cannot open %s for writing", tmp.c_str());
      const size_t w = fwrite(h.data(), sizeof(uint32_t), h.size(), f);
      const int cl = fclose(f);
      if (w != h.size() || cl != 0) return fail(BDC_ERR_IO, "short write to %s", tmp.c_str());
      if (rename(tmp.c_str(), path.c_str()) != 0) return fail(BDC_ERR_IO, "rename %s failed", tmp.c_str());
    }
  return BDC_OK;
}

int bdc_supported(int n, int k) {
  // backend/config.hpp:24 caps n at 32; the reference itself overflows `1 << n`
  // at n >= 31 (main_BAA.cu:47,64), this engine indexes inputs in 64 bits
  if (n < 2 || n > 32) return fail(0, "n must be in [2, 32]"), 0;
  if (k < 1 || k > n) return fail(0, "k must be in [1, n]"), 0;
  if (k > 30) return fail(0, "k <= 30 in this build (2^k outputs, 8 GiB per FP64 output vector at k = 30)"), 0;
  return 1;
}

int bdc_create(int n, int k, const char *cache_dir, bdc_ctx **out) {
  if (!out) return fail(BDC_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (!bdc_supported(n, k)) return BDC_ERR_ARG;
  const bool chain = chain_path(n, k);
  if (cache_dir && cache_dir[0] && !chain)  // the chain path reads no count caches
    for (auto &lm : needed_caches(n, k)) BDC_TRY(validate_cache_file(cache_dir, lm.first, lm.second));
  bdc_ctx *c = new bdc_ctx();
  auto cleanup = [&](int rc) {
    bdc_destroy(c);
    return rc;
  };
  c->n = n;
  c->k = k;
  c->n1 = (n + 1) / 2;  // backend/BAA.cu:341-342
  c->n2 = n - c->n1;
  c->N1 = 1 << c->n1;
  c->N2 = 1 << c->n2;
  c->nx = 1ull << n;
  c->nck = (double)nck_u32(n, k);
  c->inv_c = 1.0 / c->nck;  // backend/BAA.cu:96-111
  c->log_c = std::log(c->nck);
  if (hipGetDevice(&c->device) != hipSuccess) return cleanup(fail(BDC_ERR_HIP, "no HIP device"));
  BDC_TRY_C(scratch_setup(c->device));
  if (hipStreamCreate(&c->stream) != hipSuccess) return cleanup(fail(BDC_ERR_HIP, "hipStreamCreate failed"));
  {
    const char *e = getenv("BDC_STAGE_FACTORED");
    c->stage_factored = !(e && e[0] == '0');
  }

  c->chain = chain;
  if (!chain) {
    // split sets
    const int jlo = std::max(0, k - c->n2), jhi = std::min(k, c->n1);
    for (int j = jlo; j <= jhi; ++j) {
      if (2 * j <= k) {
        c->offL.push_back(c->KL);
        c->JL.push_back(j);
        c->KL += 1 << j;
      } else {
        c->offR.push_back(c->KR);
        c->JR.push_back(j);
        c->KR += 1 << (k - j);
      }
    }
    for (int j : c->JL)
      if (j == c->n1) return cleanup(fail(BDC_ERR_ARG, "identity factor in the stacked operand (k too large)"));
    for (int j : c->JR)
      if (k - j == c->n2) return cleanup(fail(BDC_ERR_ARG, "identity factor in the stacked operand (k too large)"));
    c->KLp = (c->KL + 63) / 64 * 64;
    c->KRp = c->KR ? (c->KR + 63) / 64 * 64 : 0;
    c->KT = c->KL + c->KR;

    // half-length caches (only the ones a split uses; identities are implicit)
    for (int j = jlo; j <= jhi; ++j) {
      if (j < c->n1 && !c->cnt1[j].p) BDC_TRY_C(make_cache_u16(c, c->n1, j, c->cnt1[j], cache_dir));
      const int m = k - j;
      if (m < c->n2) {
        if (c->n2 == c->n1) {
          if (!c->cnt1[m].p) BDC_TRY_C(make_cache_u16(c, c->n1, m, c->cnt1[m], cache_dir));
        } else if (!c->cnt2[m].p) {
          BDC_TRY_C(make_cache_u16(c, c->n2, m, c->cnt2[m], cache_dir));
        }
      }
    }
    // stacked operands
    BDC_TRY_C(c->abar16.alloc(sizeof(uint16_t) * (size_t)c->N1 * c->KLp));
    if (hipMemsetAsync(c->abar16.p, 0, c->abar16.bytes, c->stream) != hipSuccess) return cleanup(fail(BDC_ERR_HIP, "memset"));
    for (size_t i = 0; i < c->JL.size(); ++i)
      BDC_TRY_C(check_launch(launch_stack_cols(c->cnt1[c->JL[i]].as<uint16_t>(), c->JL[i], c->N1, c->abar16.as<uint16_t>(),
                                               c->KLp, c->offL[i], c->stream), "stack A"));
    if (c->KR) {
      BDC_TRY_C(c->bbar16.alloc(sizeof(uint16_t) * (size_t)c->N2 * c->KRp));
      if (hipMemsetAsync(c->bbar16.p, 0, c->bbar16.bytes, c->stream) != hipSuccess) return cleanup(fail(BDC_ERR_HIP, "memset"));
      for (size_t i = 0; i < c->JR.size(); ++i) {
        const int m = k - c->JR[i];
        BDC_TRY_C(check_launch(launch_stack_cols(c->c2(m), m, c->N2, c->bbar16.as<uint16_t>(), c->KRp, c->offR[i],
                                                 c->stream), "stack B"));
      }
    }
    // stored column order: bit-reversed columns for even n with half length 8..14
    if (c->n1 == c->n2 && c->n1 >= 8 && c->n1 <= 14 && column_order_enabled()) c->cr = c->n2;
    if (c->cr && c->KR) {  // Bbar rows follow the stored columns of P
      DevBuf tmp;
      BDC_TRY_C(tmp.alloc(c->bbar16.bytes));
      BDC_TRY_C(check_launch(launch_permute_rows_u16(c->bbar16.as<uint16_t>(), tmp.as<uint16_t>(), c->cr, c->KRp,
                                                     c->stream), "permute Bbar"));
      if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(fail(BDC_ERR_HIP, "permute Bbar"));
      std::swap(c->bbar16.p, tmp.p);  // tmp releases the natural-order copy
      std::swap(c->bbar16.bytes, tmp.bytes);
    }
  }

  // orbits: representative mask, word ranks, ascending list (main_BAA.cu:46-50)
  {
    const uint64_t nwords = std::max<uint64_t>(1, c->nx / 64);
    DevBuf pop, bsum, total;
    BDC_TRY_C(c->repmask.alloc(nwords * 8));
    BDC_TRY_C(c->wordbase.alloc(nwords * 4));
    BDC_TRY_C(pop.alloc(nwords * 4));
    BDC_TRY_C(bsum.alloc(((nwords + 1023) / 1024 + 1) * 4));
    BDC_TRY_C(total.alloc(8));
    BDC_TRY_C(check_launch(launch_repmask(n, nwords, c->repmask.as<uint64_t>(), pop.as<uint32_t>(), c->stream), "repmask"));
    BDC_TRY_C(check_launch(launch_exclusive_scan(pop.as<uint32_t>(), nwords, c->wordbase.as<uint32_t>(),
                                                 bsum.as<uint32_t>(), total.as<uint32_t>(), c->stream), "scan"));
    uint32_t tot = 0;
    if (hipMemcpyAsync(&tot, total.p, 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      return cleanup(fail(BDC_ERR_HIP, "orbit count copy"));
    c->norb = tot;
    if (c->norb != burnside_orbits(n))
      return cleanup(fail(BDC_ERR_NUMERIC, "orbit count %llu != Burnside count %llu", (unsigned long long)c->norb,
                          (unsigned long long)burnside_orbits(n)));
    BDC_TRY_C(c->orbits.alloc(c->norb * 4));
    BDC_TRY_C(check_launch(launch_orbit_list(nwords, c->repmask.as<uint64_t>(), c->wordbase.as<uint32_t>(),
                                             c->orbits.as<uint32_t>(), c->stream), "orbit list"));
    pop.release();
    bsum.release();
    total.release();
  }

  // work buffers
  BDC_TRY_C(c->P.alloc(c->nx * 8));
  BDC_TRY_C(c->Hc.alloc(c->nx * 8));
  const size_t ny = (size_t)1 << k;
  BDC_TRY_C(c->qv.alloc(4 * ny * 8));  // q, L, q2 (scratch), grad
  BDC_TRY_C(c->qsum.alloc((ny / 256 + 1) * 8));
  if (!chain) {
    BDC_TRY_C(c->T.alloc((size_t)c->KLp * c->N2 * 8));
    if (c->KR) BDC_TRY_C(c->U.alloc((size_t)c->KRp * c->N1 * 8));
    BDC_TRY_C(c->qpart.alloc((c->JL.size() + c->JR.size()) * ny * 8));
    if (c->KR) {
      BDC_TRY_C(c->g2a.alloc((size_t)c->N1 * c->KRp * 8));
      if (hipMemsetAsync(c->g2a.p, 0, c->g2a.bytes, c->stream) != hipSuccess) return cleanup(fail(BDC_ERR_HIP, "memset"));
    }
    BDC_TRY_C(c->g2b.alloc((size_t)c->KLp * c->N2 * 8));
    if (hipMemsetAsync(c->g2b.p, 0, c->g2b.bytes, c->stream) != hipSuccess) return cleanup(fail(BDC_ERR_HIP, "memset"));
    if (c->cr) {
      BDC_TRY_C(c->T2.alloc((size_t)c->KLp * c->N2 * 8));
      BDC_TRY_C(c->V2.alloc((size_t)c->KLp * c->N2 * 8));
      if (hipMemsetAsync(c->V2.p, 0, c->V2.bytes, c->stream) != hipSuccess) return cleanup(fail(BDC_ERR_HIP, "memset"));
    }
    c->splitk_elems = (size_t)64 * ny;  // split-K partials for the stage GEMMs: up to 64 splits of 2^k outputs
    {  // split-K slabs of the aligned FP64 output law (full and half space)
      size_t need = 0;
      for (int rows : {c->N1, c->N1 / 2}) {
        const int ks = g64_g1_splits(c, rows);
        if (ks > 1 && g64_g1_fits(c->KLp, c->N2, rows, ks)) need = std::max(need, (size_t)ks * c->KLp * c->N2 * 8);
      }
      if (need) BDC_TRY_C(c->g64part.alloc(need));
    }
  } else {
    BDC_TRY_C(chain_setup(c));
    c->splitk_elems = 4096;  // the certificate's three sums
  }
  BDC_TRY_C(c->splitk.alloc(c->splitk_elems * 8));
  c->part_blocks = std::max<size_t>(4096, ((size_t)c->N1 / 64) * ((size_t)c->N2 / 64));
  BDC_TRY_C(c->part.alloc((c->part_blocks + 256) * kPartStride * 8));  // + first-level finalize sums
  BDC_TRY_C(c->scal.alloc(kScalCount * 8));
  BDC_TRY_C(c->exptab.alloc(32 * 8));
  {
    double t[32];
    for (int j = 0; j < 32; ++j) t[j] = (double)exp2l((long double)j / 32.0L);
    if (hipMemcpy(c->exptab.p, t, sizeof t, hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(BDC_ERR_HIP, "exp table"));
  }
  {
    double h[kScalCount] = {0};
    h[kScalPScale] = 1.0;
    h[kScalOne] = 1.0;
    if (hipMemcpyAsync(c->scal.p, h, sizeof h, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      return cleanup(fail(BDC_ERR_HIP, "scalars"));
  }

  // hoisted entropy term H(x)/C - ln C for every x (computed at representatives, expanded)
  {
    DevBuf hrep, ltab;
    BDC_TRY_C(hrep.alloc(c->norb * 8));
    BDC_TRY_C(ltab.alloc(130 * 8));
    {
      double t[130];
      for (int j = 0; j <= 64; ++j) {
        const long double F = 1.0L + j / 64.0L;
        t[j] = (double)logl(F);
        t[65 + j] = (double)(1.0L / F);
      }
      if (hipMemcpy(ltab.p, t, sizeof t, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(BDC_ERR_HIP, "log table"));
    }
    if (entropy_by_enumeration(n, k))
      BDC_TRY_C(check_launch(launch_entropy_dfs(n, k, c->norb, c->orbits.as<uint32_t>(), ltab.as<double>(),
                                                hrep.as<double>(), c->stream), "entropy (distinct subsequences)"));
    else
      BDC_TRY_C(check_launch(launch_entropy_reps(n, k, c->n1, c->n2, c->repmask.as<uint64_t>(),
                                                 c->wordbase.as<uint32_t>(), ltab.as<double>(), hrep.as<double>(),
                                                 c->stream), "entropy"));
    BDC_TRY_C(check_launch(launch_expand_orbits(n, hrep.as<double>(), c->repmask.as<uint64_t>(), c->wordbase.as<uint32_t>(),
                                                c->inv_c, -c->log_c, c->cr, c->Hc.as<double>(), c->stream), "expand H"));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(fail(BDC_ERR_HIP, "entropy sync"));
    hrep.release();
    ltab.release();
    c->have_h = true;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(fail(BDC_ERR_HIP, "setup sync"));
  *out = c;
  return BDC_OK;
}

int bdc_destroy(bdc_ctx *c) {
  if (!c) return BDC_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto *t : {&c->t_g1, &c->t_g2}) {
    for (auto e : t->start) (void)hipEventDestroy(e);
    for (auto e : t->stop) (void)hipEventDestroy(e);
  }
  for (auto &b : c->cnt1) b.release();
  for (auto &b : c->cnt2) b.release();
  for (bdc::DevBuf *b : {&c->abar16, &c->bbar16, &c->abarh, &c->abarTh, &c->bbarh, &c->residA, &c->rowptrA,
                         &c->entA, &c->rowptrB, &c->entB, &c->blkA, &c->blkB, &c->restiles[0], &c->restiles[1], &c->g1part, &c->repmask, &c->wordbase, &c->orbits, &c->P,
                         &c->Hc, &c->T, &c->U, &c->qpart, &c->qv, &c->g2a, &c->g2b, &c->splitk, &c->part, &c->scal, &c->exptab,
                         &c->hist, &c->ptexp, &c->qsum, &c->vlimb, &c->wlimb, &c->theta, &c->adam_m,
                         &c->adam_v, &c->grad, &c->dvec, &c->ck_p, &c->ck_s, &c->ck_t, &c->lvl, &c->g64part,
                         &c->T2, &c->V2, &c->qtiles[0], &c->qtiles[1], &c->qtiles[2]})
    b->release();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return BDC_OK;
}

}  // extern "C"

// =====================================================================
// BA step pipeline
// =====================================================================
namespace bdc {

static int pick_splits(int M, int N, int K) {
  const int tiles = ((M + 15) / 16) * ((N + 63) / 64) / (M > 16 && N > 16 ? 4 : 1);
  int s = std::max(1, 512 / std::max(1, tiles));
  s = std::min(s, std::max(1, K / 256));
  return std::min(s, 64);
}

// C (M x N, row-major, contiguous) = A B with split-K through c->splitk
static int stage_gemm(bdc_ctx *c, int M, int N, int K, const Opnd &A, const Opnd &B, double *C) {
  const Opnd none{nullptr, 0, 0, 0};
  int splits = pick_splits(M, N, K);
  if ((size_t)splits * M * N > c->splitk_elems) splits = std::max<int>(1, (int)(c->splitk_elems / ((size_t)M * N)));
  if (splits == 1) {
    BDC_LAUNCH(gemm64(M, N, A, B, K, none, none, 0, C, N, 1, 1, 0, EPI_STORE, nullptr, c->stream, nullptr));
    return BDC_OK;
  }
  double *part = c->splitk.as<double>();
  BDC_LAUNCH(gemm64(M, N, A, B, K, none, none, 0, part, N, 1, splits, (int64_t)M * N, EPI_STORE, nullptr, c->stream,
                    nullptr));
  BDC_LAUNCH(launch_reduce_splits(part, splits, (int64_t)M * N, (uint64_t)M * N, C, c->stream));
  return BDC_OK;
}

// C (M x N, row stride ldc) = A (M x K, FP64, row stride lda) x B(k, n) over a
// count matrix (transb 0: B[k * ldb + n], 1: B[n * ldb + k]) on the
// count-streaming kernel, split over K into c->splitk when the grid is small;
// false when the shape does not fit that kernel.
static int count_stage(bdc_ctx *c, int M, int N, int K, const double *A, int64_t lda, const uint16_t *B, int64_t ldb,
                       int transb, double *C, int64_t ldc, bool *done) {
  *done = false;
  if (M > 128 || N % 64 || K % 32) return BDC_OK;
  // split K until ~2048 blocks stream the count matrix (each holds one 4 KiB
  // stage in flight, so memory parallelism comes from the block count)
  int ks = 1;
  while ((N / 64) * ks < 2048 && ks < 64 && K / (ks * 2) >= 128 && (K / 32) % (ks * 2) == 0 &&
         (size_t)(ks * 2) * M * N <= c->splitk_elems)
    ks *= 2;
  if (ks == 1) {
    BDC_LAUNCH(launch_count_gemm(M, N, K, A, lda, B, ldb, transb, C, ldc, 1, 0, c->stream));
  } else {
    double *part = c->splitk.as<double>();
    BDC_LAUNCH(launch_count_gemm(M, N, K, A, lda, B, ldb, transb, part, N, ks, (int64_t)M * N, c->stream));
    if (ldc != N) return fail(BDC_ERR_STATE, "split count stage needs a contiguous output");
    BDC_LAUNCH(launch_reduce_splits(part, ks, (int64_t)M * N, (uint64_t)M * N, C, c->stream));
  }
  *done = true;
  return BDC_OK;
}

// The aligned FP64 product kernels (launch_g64_*) unless BDC_GEMM64=legacy.
static bool use_g64() {
  const char *e = getenv("BDC_GEMM64");
  return !(e && !strcmp(e, "legacy"));
}
// split-K of the aligned FP64 output law: enough blocks (>= 512) for 256 CUs
static int g64_g1_splits(const bdc_ctx *c, int rows) {
  const int base = (c->N2 / 128) * (c->KLp / 128);
  int ks = 1;
  while (ks < 8 && base * ks < 512 && rows % (16 * ks * 2) == 0 && rows / (ks * 2) >= 256) ks *= 2;
  return ks;
}

// ---- deletion-chain path (kernels_chain.hip) ----
// Tile width of a group: one contiguous row for the bottom group, 16 columns
// of the bits below otherwise.
static int grp_width(const bdc_ctx *c, int g) { return (int)std::min<uint64_t>(16, (uint64_t)1 << c->grpB[g]); }

// q = (1/C) D^d X / d! over the bit groups (top first), then L = ln q.
static int chain_output_law(bdc_ctx *c, const double *X, int scale_slot, double *qout, double *Lout) {
  const int n = c->n, d = c->n - c->k;
  double *lvl = c->lvl.as<double>();
  const uint64_t ny = (uint64_t)1 << c->k;
  if (d == 0) {  // k = n: the channel is the identity
    BDC_LAUNCH(launch_chain_q(X, ny, c->scal.as<double>(), scale_slot, c->inv_c, qout, Lout, c->stream));
    return BDC_OK;
  }
  std::vector<bool> zero(d + 1, true);
  zero[0] = false;
  for (int g = 0; g < c->ngrp; ++g) {
    const int L = c->grpL[g], B = c->grpB[g];
    for (int t = d - 1; t >= 0; --t) {
      const int H = n - t - L - B;  // high bits: the groups above after t deletions
      if (H < 0 || zero[t]) continue;
      GrpChain gc{};
      gc.src = t == 0 ? X : lvl + c->lvl_off[t];
      gc.m = std::min(d - t, L);
      double fact = 1.0;
      for (int j = 1; j <= gc.m; ++j) {
        fact *= j;
        gc.lvl[j] = lvl + c->lvl_off[t + j];
        gc.scale[j] = 1.0 / fact;
        if (!zero[t + j]) gc.acc |= 1u << j;
      }
      gc.L = L;
      gc.B = B;
      gc.W = grp_width(c, g);
      gc.nh = (uint64_t)1 << H;
      gc.ntiles = ((uint64_t)1 << B) / gc.W;
      BDC_LAUNCH(launch_grp_del(gc, c->stream));
      for (int j = 1; j <= gc.m; ++j) zero[t + j] = false;
    }
  }
  if (zero[d]) return fail(BDC_ERR_STATE, "chain output law: no path reached %d deletions", d);
  BDC_LAUNCH(launch_chain_q(lvl + c->lvl_off[d], ny, c->scal.as<double>(), scale_slot, c->inv_c, qout, Lout,
                            c->stream));
  return BDC_OK;
}

// S(x) = sum_y c(x,y) L(y) into level slot 0: the transpose of the output law
// (groups bottom first, targets ascending, Horner over the sources).
static int chain_cross(bdc_ctx *c, const double *L) {
  const int n = c->n, d = c->n - c->k;
  double *lvl = c->lvl.as<double>();
  double *top = lvl + c->lvl_off[d];
  if (L != top) BDC_HIP(hipMemcpyAsync(top, L, ((size_t)8) << c->k, hipMemcpyDeviceToDevice, c->stream));
  std::vector<bool> zero(d + 1, true);
  zero[d] = false;
  for (int g = c->ngrp - 1; g >= 0; --g) {
    const int Lg = c->grpL[g], B = c->grpB[g];
    for (int t = 0; t < d; ++t) {
      const int H = n - t - Lg - B;
      const int m = std::min(d - t, Lg);
      if (H < 0 || m == 0) continue;
      GrpChain gc{};
      bool any = false;
      for (int j = 1; j <= m; ++j) {
        gc.lvl[j] = zero[t + j] ? nullptr : lvl + c->lvl_off[t + j];
        gc.scale[j] = 1.0 / j;
        any = any || !zero[t + j];
      }
      if (!any) continue;
      gc.dst = lvl + c->lvl_off[t];
      gc.acc = zero[t] ? 0u : 1u;
      gc.m = m;
      gc.L = Lg;
      gc.B = B;
      gc.W = grp_width(c, g);
      gc.nh = (uint64_t)1 << H;
      gc.ntiles = ((uint64_t)1 << B) / gc.W;
      BDC_LAUNCH(launch_grp_ins(gc, c->stream));
      zero[t] = false;
    }
  }
  if (zero[0]) return fail(BDC_ERR_STATE, "chain cross term: no path reached the input length");
  return BDC_OK;
}

// blocks of the chain epilogue (each owns a contiguous range of inputs)
static int chain_epi_blocks(const bdc_ctx *c) {
  return (int)std::min<uint64_t>(c->part_blocks, std::max<uint64_t>(1, c->nx / 1024));
}

// q(y) and L(y) = ln q - mu from weights X (p = scal[scale_slot] * X).
// mirror: use q_j(y) = q_{k-j}(rev y) for j in JR (X reversal-symmetric, n even).
// rows = N1/2: X holds only the half space x1 < N1/2 of a complement-symmetric
// p; the contractions run over those rows and q_j(y) = R_j(y) + R_j(~y).
// defer_center: leave L = ln q uncentred for the stage contraction that
// consumes it next (cross_operands centres it, in the deletion chain when that
// runs)
static int output_law(bdc_ctx *c, int prec, const double *X, int scale_slot, bool mirror, int rows, double *qout,
                      double *Lout, bool defer_center = false) {
  const int k = c->k, N1 = c->N1, N2 = c->N2;
  const size_t ny = (size_t)1 << k;
  const Opnd none{nullptr, 0, 0, 0};
  if (c->chain) {
    BDC_TRY(timer_begin(c, c->t_g1));
    BDC_TRY(chain_output_law(c, X, scale_slot, qout, Lout));
    return timer_end(c, c->t_g1);
  }
  if (c->n1 != c->n2) mirror = false;
  if (rows != N1 && !(mirror && rows == N1 / 2)) return fail(BDC_ERR_STATE, "half-space output law needs the mirror rule");
  // G1: Tbar = Abar[rows]^T X   (KL x N2, K = rows)
  BDC_TRY(timer_begin(c, c->t_g1));
  const int g64ks = g64_g1_splits(c, rows);
  if (prec == BDC_PREC_FP64 && use_g64() && g64_g1_fits(c->KLp, N2, rows, g64ks) &&
      (g64ks == 1 || c->g64part.bytes >= (size_t)g64ks * c->KLp * N2 * 8)) {
    const size_t slab = (size_t)c->KLp * N2;
    double *dst = g64ks > 1 ? c->g64part.as<double>() : c->T.as<double>();
    BDC_LAUNCH(launch_g64_g1(c->KLp, N2, rows, c->abar16.as<uint16_t>(), c->KLp, X, N2, dst, N2, g64ks, (int64_t)slab,
                             c->stream));
    if (g64ks > 1) BDC_LAUNCH(launch_reduce_splits(dst, g64ks, (int64_t)slab, slab, c->T.as<double>(), c->stream));
  } else if (prec == BDC_PREC_FP64) {
    const Opnd A{c->abar16.p, 1, c->KLp, 1}, B{X, N2, 1, 0};
    BDC_LAUNCH(gemm64(c->KL, N2, A, B, rows, none, none, 0, c->T.as<double>(), N2, 1, 1, 0, EPI_STORE, nullptr,
                      c->stream, nullptr));
  } else {
    BDC_TRY(split_output_law_g1(c, X, rows));
  }
  BDC_TRY(timer_end(c, c->t_g1));
  // Tbar's columns follow the stored order of X: the factored stage runs on
  // them as they are (reversed coordinates, outputs written in natural order);
  // the other stage forms read a copy in natural order
  double *Tn = c->T.as<double>();
  if (!mirror && c->KR) {
    // Ubar = Bbar^T X^T  (KR x N1, K = N2)
    const Opnd A{c->bbar16.p, 1, c->KRp, 1}, B{X, 1, N2, 0};
    BDC_LAUNCH(gemm64(c->KR, N1, A, B, N2, none, none, 0, c->U.as<double>(), N1, 1, 1, 0, EPI_STORE, nullptr, c->stream,
                      nullptr));
  }
  double *qL = c->qpart.as<double>();
  double *qR = qL + c->JL.size() * ny;
  // stage 2, j in JL: q_j = T_j B_{k-j}, all j in one launch of factored
  // single-deletion steps (row y1 of T_j, 2^n2 entries, to 2^(k-j) entries)
  bool staged = false;
  if (c->stage_factored) {
    ChainJobs jb{};
    jb.count = (int)c->JL.size();
    int rows = 0;
    for (size_t i = 0; i < c->JL.size() && jb.count <= kMaxChainJobs; ++i) {
      const int j = c->JL[i], m = k - j;
      jb.row0[i] = rows;
      rows += 1 << j;
      jb.d[i] = c->n2 - m;
      jb.src_off[i] = (int64_t)c->offL[i] * N2;
      jb.src_ld[i] = N2;
      jb.dst_off[i] = (int64_t)i * ny;
      jb.dst_ld[i] = (int64_t)1 << m;
    }
    if (jb.count <= kMaxChainJobs) {
      jb.row0[jb.count] = rows;
      jb.rev = c->cr ? 1 : 0;
      const int rc = launch_chain_ins(Tn, qL, c->n2, jb, c->stream);
      if (rc == BDC_ERR_HIP) return fail(rc, "chain_ins launch failed");
      staged = rc == BDC_OK;
    }
  }
  if (!staged && c->cr) {
    BDC_LAUNCH(launch_permute_cols(c->T.as<double>(), c->T2.as<double>(), (uint64_t)c->KLp, c->cr, c->stream));
    Tn = c->T2.as<double>();
  }
  for (size_t i = 0; i < c->JL.size() && !staged; ++i) {
    const int j = c->JL[i], m = k - j;
    const double *Tj = Tn + (size_t)c->offL[i] * N2;
    double *dst = qL + i * ny;
    const int d = c->n2 - m;  // symbols B_m deletes
    if (d == 0) {
      BDC_LAUNCH(launch_copy2d(Tj, N2, 1, dst, 1 << m, 1, 1 << j, 1 << m, c->stream));
    } else if (d <= kMaxGatherInsertions) {
      BDC_LAUNCH(launch_ins_gather(Tj, N2, 1, dst, 1 << m, 1, 1 << j, c->n2, d, c->stream));
    } else {
      bool done = false;
      BDC_TRY(count_stage(c, 1 << j, 1 << m, N2, Tj, N2, c->c2(m), 1 << m, 0, dst, 1 << m, &done));
      if (!done) {
        const Opnd A{Tj, N2, 1, 0}, B{c->c2(m), 1 << m, 1, 1};
        BDC_TRY(stage_gemm(c, 1 << j, 1 << m, N2, A, B, dst));
      }
    }
  }
  QCombine qc{};
  qc.k = k;
  qc.nL = (int)c->JL.size();
  qc.qL = qL;
  qc.scale_slot = scale_slot;
  qc.comp = rows < N1 ? 1 : 0;
  if (mirror) {
    qc.mirror = 1;
    for (int j : c->JR) {
      const int m = k - j;
      for (size_t i = 0; i < c->JL.size(); ++i)
        if (c->JL[i] == m) qc.mirrorSel |= 1ull << i;
    }
  } else {
    // stage 2, j in JR: q_j = A_j^T U_{k-j}
    for (size_t i = 0; i < c->JR.size(); ++i) {
      const int j = c->JR[i], m = k - j;
      const double *Uj = c->U.as<double>() + (size_t)c->offR[i] * N1;
      double *dst = qR + i * ny;
      const int d = c->n1 - j;  // symbols A_j deletes
      if (d == 0) {
        // q_j[y1][z] = U[off + z][y1]
        BDC_LAUNCH(launch_copy2d(Uj, 1, N1, dst, 1 << m, 1, 1 << j, 1 << m, c->stream));
      } else if (d <= kMaxGatherInsertions) {
        BDC_LAUNCH(launch_ins_gather(Uj, N1, 1, dst, 1, 1 << m, 1 << m, c->n1, d, c->stream));
      } else {
        const Opnd A{c->c1(j), 1, 1 << j, 1}, B{Uj, 1, N1, 0};
        BDC_TRY(stage_gemm(c, 1 << j, 1 << m, N1, A, B, dst));
      }
    }
    qc.nR = (int)c->JR.size();
    qc.qR = qR;
  }
  BDC_LAUNCH(launch_q_combine(qc, c->inv_c, c->scal.as<double>(), qout, Lout, c->qsum.as<double>(),
                              defer_center ? 0 : 1, c->stream));
  c->l_uncentred = defer_center;
  return BDC_OK;
}

// stage 1: Vbar rows (j in JL) and Wbar columns (j in JR) from L (FP64).
// mirror: L is reversal-symmetric and n1 = n2, so Wbar is a permutation of Vbar;
// only the first `rows` rows of Wbar are formed then.  amax (may be null):
// receives one maximum of |Vbar| per row block of the deletion-chain launch,
// and *namax their number, when that launch formed Vbar and Wbar is its
// permutation (else *namax = 0).  defer_w (may be null): with those maxima,
// Wbar is not formed; its mirror description goes to *defer_w (count > 0) for
// a launch that also splits it into limbs.
static int cross_operands_nat(bdc_ctx *c, const double *L, bool mirror, int rows, double *amax, size_t amax_cap,
                              int *namax, MirrorW *defer_w);
// Vbar in the stored column order of P when the columns are bit-reversed: the
// factored stage writes it so (g2b_stored), the other stage forms write natural
// order into g2b and it is permuted into V2
static double *vbar(bdc_ctx *c) { return c->cr && !c->g2b_stored ? c->V2.as<double>() : c->g2b.as<double>(); }
static int cross_operands(bdc_ctx *c, const double *L, bool mirror = false, int rows = 0, double *amax = nullptr,
                          size_t amax_cap = 0, int *namax = nullptr, MirrorW *defer_w = nullptr) {
  BDC_TRY(cross_operands_nat(c, L, mirror, rows, amax, amax_cap, namax, defer_w));
  if (c->cr && !c->chain && !c->g2b_stored)
    BDC_LAUNCH(launch_permute_cols(c->g2b.as<double>(), c->V2.as<double>(), (uint64_t)c->KLp, c->cr, c->stream));
  return BDC_OK;
}
static int cross_operands_nat(bdc_ctx *c, const double *L, bool mirror, int rows, double *amax, size_t amax_cap,
                              int *namax, MirrorW *defer_w) {
  if (namax) *namax = 0;
  if (defer_w) defer_w->count = 0;
  c->g2b_stored = false;
  if (c->chain) return chain_cross(c, L);
  const int k = c->k, N1 = c->N1, N2 = c->N2;
  // ln q left uncentred by the output law is centred by the deletion chain on
  // load, unless W is formed from L itself below (not the mirror rule): then
  // it is centred in place first
  const bool w_from_l = !(mirror && c->n1 == c->n2) && !c->JR.empty();
  if (c->l_uncentred && w_from_l) {
    BDC_LAUNCH(launch_q_center(k, c->qsum.as<double>(), c->scal.as<double>(), const_cast<double *>(L), c->stream));
    c->l_uncentred = false;
  }
  // V_j = L_j B_{k-j}^T for all j in JL in one launch of factored
  // single-deletion steps (row y1 of L_j, 2^(k-j) entries, to 2^n2 entries)
  bool staged = false;
  if (c->stage_factored) {
    ChainJobs jb{};
    jb.count = (int)c->JL.size();
    int nrows = 0;
    for (size_t i = 0; i < c->JL.size() && jb.count <= kMaxChainJobs; ++i) {
      const int j = c->JL[i], m = k - j;
      jb.row0[i] = nrows;
      nrows += 1 << j;
      jb.d[i] = c->n2 - m;
      jb.src_off[i] = 0;
      jb.src_ld[i] = (int64_t)1 << m;
      jb.dst_off[i] = (int64_t)c->offL[i] * N2;
      jb.dst_ld[i] = N2;
    }
    if (jb.count <= kMaxChainJobs) {
      jb.row0[jb.count] = nrows;
      jb.rev = c->cr ? 1 : 0;  // Vbar straight into the stored column order
      const bool fuse_max = amax && mirror && c->n1 == c->n2 && (size_t)nrows <= amax_cap;
      const int rc = launch_chain_del(L, c->g2b.as<double>(), c->n2, jb, fuse_max ? amax : nullptr,
                                      c->l_uncentred ? c->qsum.as<double>() : nullptr, k, c->scal.as<double>(),
                                      c->stream);
      if (rc == BDC_OK) c->l_uncentred = false;
      if (rc == BDC_ERR_HIP) return fail(rc, "chain_del launch failed");
      staged = rc == BDC_OK;
      if (staged && fuse_max && namax) *namax = nrows;
      c->g2b_stored = staged && c->cr;
    }
  }
  if (c->l_uncentred) {  // another stage form reads L: centre it first
    BDC_LAUNCH(launch_q_center(k, c->qsum.as<double>(), c->scal.as<double>(), const_cast<double *>(L), c->stream));
    c->l_uncentred = false;
  }
  for (size_t i = 0; i < c->JL.size() && !staged; ++i) {
    const int j = c->JL[i], m = k - j;
    double *dst = c->g2b.as<double>() + (size_t)c->offL[i] * N2;
    const int d = c->n2 - m;
    if (d == 0) {
      BDC_LAUNCH(launch_copy2d(L, 1 << m, 1, dst, N2, 1, 1 << j, N2, c->stream));
    } else if (d <= kMaxGatherDeletions) {
      BDC_LAUNCH(launch_del_gather(L, 1 << m, 1, dst, N2, 1, 1 << j, c->n2, d, c->stream));
    } else {
      bool done = false;
      BDC_TRY(count_stage(c, 1 << j, N2, 1 << m, L, 1 << m, c->c2(m), 1 << m, 1, dst, N2, &done));
      if (!done) {
        const Opnd A{L, 1 << m, 1, 0}, B{c->c2(m), 1, 1 << m, 1};
        const Opnd none{nullptr, 0, 0, 0};
        BDC_LAUNCH(gemm64(1 << j, N2, A, B, 1 << m, none, none, 0, dst, N2, 1, 1, 0, EPI_STORE, nullptr, c->stream,
                          nullptr));
      }
    }
  }
  if (mirror && c->n1 == c->n2 && !c->JR.empty()) {
    MirrorW mw{};
    mw.count = (int)c->JR.size();
    for (size_t i = 0; i < c->JR.size(); ++i) {
      const int m = k - c->JR[i];
      mw.offR[i] = c->offR[i];
      mw.m[i] = m;
      mw.offL[i] = -1;
      for (size_t q = 0; q < c->JL.size(); ++q)
        if (c->JL[q] == m) mw.offL[i] = c->offL[q];
      if (mw.offL[i] < 0) return fail(BDC_ERR_STATE, "mirror slice for j=%d missing", c->JR[i]);
    }
    mw.stored = c->g2b_stored ? 1 : 0;
    if (defer_w && namax && *namax > 0) {
      *defer_w = mw;
      return BDC_OK;
    }
    BDC_LAUNCH(launch_w_from_v(mw, c->g2b.as<double>(), N2, c->n2, c->g2a.as<double>(), rows > 0 ? rows : N1, c->KR,
                               c->KRp, c->stream));
    return BDC_OK;
  }
  if (rows > 0 && rows != N1) return fail(BDC_ERR_STATE, "half-space cross operands need the mirror rule");
  for (size_t i = 0; i < c->JR.size(); ++i) {
    const int j = c->JR[i], m = k - j;
    double *dst = c->g2a.as<double>() + c->offR[i];
    const int d = c->n1 - j;
    if (d == 0) {
      BDC_LAUNCH(launch_copy2d(L, 1 << m, 1, dst, c->KRp, 1, N1, 1 << m, c->stream));
    } else if (d <= kMaxGatherDeletions) {
      BDC_LAUNCH(launch_del_gather(L, 1, 1 << m, dst, 1, c->KRp, 1 << m, c->n1, d, c->stream));
    } else {
      const Opnd A{c->c1(j), 1 << j, 1, 1}, B{L, 1 << m, 1, 0};
      const Opnd none{nullptr, 0, 0, 0};
      BDC_LAUNCH(gemm64(N1, 1 << m, A, B, 1 << j, none, none, 0, dst, c->KRp, 1, 1, 0, EPI_STORE, nullptr, c->stream,
                        nullptr));
    }
  }
  return BDC_OK;
}

// G2 with the given epilogue over the rows x1 < rows (N1 unless the half-space
// step); returns the number of partial blocks
static int cross_term(bdc_ctx *c, int prec, int epi, double *Dout, int *nblocks, int rows = 0) {
  if (rows <= 0) rows = c->N1;
  EpiArgs ea{};
  ea.P = c->P.as<double>();
  ea.Hc = c->Hc.as<double>();
  ea.D = Dout;
  ea.part = c->part.as<double>();
  ea.scal = c->scal.as<double>();
  ea.inv_c = c->inv_c;
  ea.ldx = c->N2;
  ea.exptab = c->exptab.as<double>();
  BDC_TRY(timer_begin(c, c->t_g2));
  if (c->chain) {
    const int nb = chain_epi_blocks(c);
    BDC_LAUNCH(launch_chain_epi(epi, c->lvl.as<double>(), ea.P, ea.Hc, Dout, c->nx, ea.scal, c->inv_c, ea.exptab,
                                ea.part, nb, c->stream));
    if (nblocks) *nblocks = nb;
  } else if (prec == BDC_PREC_FP64 && use_g64() && g64_g2_fits(rows, c->N2, c->KLp, c->KRp)) {
    BDC_LAUNCH(launch_g64_g2(rows, c->N2, c->abar16.as<uint16_t>(), c->KLp, vbar(c), c->N2, c->KLp,
                             c->g2a.as<double>(), c->KRp, c->bbar16.as<uint16_t>(), c->KRp, c->KRp, epi, ea, c->stream,
                             nblocks));
  } else if (prec == BDC_PREC_FP64) {
    const Opnd A1{c->abar16.p, c->KLp, 1, 1}, B1{vbar(c), c->N2, 1, 0};
    const Opnd A2{c->g2a.p, c->KRp, 1, 0}, B2{c->bbar16.p, 1, c->KRp, 1};
    BDC_LAUNCH(gemm64(rows, c->N2, A1, B1, c->KL, A2, B2, c->KR, nullptr, 0, 0, 1, 0, epi, &ea, c->stream, nblocks));
  } else {
    BDC_TRY(split_cross_term(c, epi, ea, nblocks, rows));
  }
  BDC_TRY(timer_end(c, c->t_g2));
  return BDC_OK;
}

static int ensure_hist(bdc_ctx *c, int n) {
  if (c->hist_cap >= n) return BDC_OK;
  BDC_TRY(c->hist.alloc((size_t)n * 4 * 8));
  c->hist_cap = n;
  return BDC_OK;
}

static int ensure_dvec(bdc_ctx *c) {
  c->have_d = false;  // every user but bdc_certify overwrites D(x||q)
  if (c->dvec.p) return BDC_OK;
  return c->dvec.alloc(c->nx * 8);
}

// Rows of P a BA step updates.  p is complement-symmetric (every setter
// expands orbits, and the update commutes with x -> ~x), so when the reversal
// mirror applies (n1 = n2) the step runs on the half space x1 < N1/2 (the top
// bit of x is 0) and the rows beyond are their complement images.  The split
// path needs 128-row tiles of that half.  BDC_HALF_SPACE=0 forces the full space.
static int ba_rows(bdc_ctx *c, int prec) {
  const char *e = getenv("BDC_HALF_SPACE");
  if (e && e[0] == '0') return c->N1;
  if (c->chain || c->n1 != c->n2 || c->N1 < 2) return c->N1;
  if (prec == BDC_PREC_SPLIT16 && (c->N1 / 2) % 128) return c->N1;
  return c->N1 / 2;
}

static bool quarter_step(const bdc_ctx *c, int rows);
static bool tri_reads();

// after quarter steps that updated only the canonical elements: the rest of
// the half space from their partners (before the complement image is written)
static int complete_mirrors(bdc_ctx *c) {
  if (!c->p_mirror_stale) return BDC_OK;
  BDC_LAUNCH(launch_fill_mirror(c->P.as<double>(), c->N2, c->stream));
  c->p_mirror_stale = false;
  return BDC_OK;
}

// one BA step on the device over the rows x1 < rows; stats land in hist[idx]
static int ba_step(bdc_ctx *c, int prec, int idx, int rows, double *hist) {
  double *q = c->qv.as<double>(), *L = q + ((size_t)1 << c->k);
  BDC_TRY(output_law(c, prec, c->P.as<double>(), kScalPScale, true, rows, q, L, !c->chain));
  if (prec == BDC_PREC_FP64) {
    BDC_TRY(cross_operands(c, L, true, rows));
  } else {
    BDC_TRY(split_cross_operands(c, L, true, rows, quarter_step(c, rows)));
  }
  int nb = 0;
  BDC_TRY(cross_term(c, prec, EPI_BA, nullptr, &nb, rows));
  if (prec == BDC_PREC_FP64) c->plimb_valid = false;  // P changed without new limbs
  BDC_LAUNCH(launch_ba_finalize(c->part.as<double>(), nb, (double)c->N1 / rows, c->scal.as<double>(),
                                hist, idx, c->part.as<double>() + c->part_blocks * kPartStride, c->stream));
  return BDC_OK;
}

static int reset_scalars(bdc_ctx *c, double pscale) {
  double h[kScalCount] = {0};
  h[kScalPScale] = pscale;
  h[kScalOne] = 1.0;
  BDC_HIP(hipMemcpyAsync(c->scal.p, h, sizeof h, hipMemcpyHostToDevice, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  return BDC_OK;
}

// an x-indexed device array (stored column order) to the host in natural order
static int copy_x_to_host(bdc_ctx *c, const double *src, double *dst) {
  if (c->cr) {
    DevBuf tmp;
    BDC_TRY(tmp.alloc(c->nx * 8));
    BDC_LAUNCH(launch_permute_cols(src, tmp.as<double>(), (uint64_t)c->N1, c->cr, c->stream));
    BDC_HIP(hipMemcpyAsync(dst, tmp.p, c->nx * 8, hipMemcpyDeviceToHost, c->stream));
    BDC_HIP(hipStreamSynchronize(c->stream));
    return BDC_OK;
  }
  BDC_HIP(hipMemcpyAsync(dst, src, c->nx * 8, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  return BDC_OK;
}

}  // namespace bdc

extern "C" {

int bdc_num_orbits(bdc_ctx *c, uint64_t *out) {
  if (!c || !out) return fail(BDC_ERR_ARG, "null argument");
  *out = c->norb;
  return BDC_OK;
}

int bdc_get_orbits(bdc_ctx *c, uint32_t *out, uint64_t len) {
  if (!c || !out || len != c->norb) return fail(BDC_ERR_ARG, "orbit buffer length must be %llu", (unsigned long long)(c ? c->norb : 0));
  BDC_HIP(hipMemcpyAsync(out, c->orbits.p, len * 4, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  return BDC_OK;
}

int bdc_set_p_uniform(bdc_ctx *c) {
  if (!c) return fail(BDC_ERR_ARG, "null context");
  BDC_LAUNCH(launch_fill(c->P.as<double>(), c->nx, 1.0 / (double)c->nx, c->stream));  // main_BAA.cu:64
  BDC_TRY(reset_scalars(c, 1.0));
  c->have_p = true;
  c->last_mx_bits = NAN;
  c->plimb_valid = false;
  return BDC_OK;
}

int bdc_set_p_orbits(bdc_ctx *c, const double *q, uint64_t len) {
  if (!c || !q) return fail(BDC_ERR_ARG, "null argument");
  // the reference reads whatever the file holds (main_BAA.cu:3-18); here the
  // length, sign and finiteness are checked
  if (len != c->norb)
    return fail(BDC_ERR_ARG, "distribution has %llu values, BDC_{%d,%d} has %llu orbits", (unsigned long long)len, c->n,
                c->k, (unsigned long long)c->norb);
  for (uint64_t i = 0; i < len; ++i)
    if (!(q[i] >= 0.0) || !std::isfinite(q[i]))
      return fail(BDC_ERR_ARG, "distribution value %llu is negative or not finite", (unsigned long long)i);
  DevBuf d, part;
  BDC_TRY(d.alloc(len * 8));
  BDC_TRY(part.alloc(1025 * 8));
  BDC_HIP(hipMemcpyAsync(d.p, q, len * 8, hipMemcpyHostToDevice, c->stream));
  BDC_LAUNCH(launch_expand_orbits(c->n, d.as<double>(), c->repmask.as<uint64_t>(), c->wordbase.as<uint32_t>(), 1.0, 0.0,
                                  c->cr, c->P.as<double>(), c->stream));
  // sum over all strings (orbit value x orbit size) must be 1
  BDC_LAUNCH(launch_sum(c->P.as<double>(), c->nx, part.as<double>(), 1024, part.as<double>() + 1024, c->stream));
  double total = 0.0;
  BDC_HIP(hipMemcpyAsync(&total, part.as<double>() + 1024, 8, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  d.release();
  part.release();
  if (!(std::fabs(total - 1.0) <= 1e-6)) {
    c->have_p = false;
    return fail(BDC_ERR_ARG, "distribution sums to %.17g over the 2^%d strings, not 1", total, c->n);
  }
  BDC_TRY(reset_scalars(c, 1.0));
  c->have_p = true;
  c->last_mx_bits = NAN;
  c->plimb_valid = false;
  return BDC_OK;
}

int bdc_get_p_orbits(bdc_ctx *c, double *q, uint64_t len) {
  if (!c || !q || len != c->norb) return fail(BDC_ERR_ARG, "buffer length must equal the orbit count");
  if (!c->have_p) return fail(BDC_ERR_STATE, "p is not set");
  DevBuf d;
  BDC_TRY(d.alloc(len * 8));
  BDC_LAUNCH(launch_gather_orbits(len, c->orbits.as<uint32_t>(), c->P.as<double>(), c->scal.as<double>(), c->cr,
                                  d.as<double>(), c->stream));
  BDC_HIP(hipMemcpyAsync(q, d.p, len * 8, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  d.release();
  return BDC_OK;
}

int bdc_set_p_dirichlet(bdc_ctx *c, double alpha, uint64_t seed) {
  if (!c || !(alpha > 0.0)) return fail(BDC_ERR_ARG, "alpha must be > 0");
  DevBuf Q, mass, part, tot;
  BDC_TRY(Q.alloc(c->norb * 8));
  BDC_TRY(mass.alloc(c->norb * 8));
  BDC_TRY(part.alloc(1024 * 8));
  BDC_TRY(tot.alloc(8));
  BDC_LAUNCH(launch_dirichlet_orbits(c->n, c->norb, c->orbits.as<uint32_t>(), alpha, seed, Q.as<double>(), mass.as<double>(),
                                     c->stream));
  BDC_LAUNCH(launch_sum(mass.as<double>(), c->norb, part.as<double>(), 1024, tot.as<double>(), c->stream));
  double total = 0.0;
  BDC_HIP(hipMemcpyAsync(&total, tot.p, 8, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  if (!(total > 0.0)) return fail(BDC_ERR_NUMERIC, "Dirichlet masses sum to %g", total);
  BDC_LAUNCH(launch_expand_orbits(c->n, Q.as<double>(), c->repmask.as<uint64_t>(), c->wordbase.as<uint32_t>(), 1.0 / total,
                                  0.0, c->cr, c->P.as<double>(), c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  BDC_TRY(reset_scalars(c, 1.0));
  c->have_p = true;
  c->last_mx_bits = NAN;
  c->plimb_valid = false;
  return BDC_OK;
}

int bdc_iterate(bdc_ctx *c, int prec, int iters, double tol, bdc_iter_stats *stats, int *done) {
  if (!c) return fail(BDC_ERR_ARG, "null context");
  if (prec != BDC_PREC_FP64 && prec != BDC_PREC_SPLIT16) return fail(BDC_ERR_ARG, "unknown precision %d", prec);
  if (iters < 0) return fail(BDC_ERR_ARG, "iters < 0");
  if (!c->have_p) return fail(BDC_ERR_STATE, "p is not set");
  if (prec == BDC_PREC_SPLIT16) BDC_TRY(split_prepare(c));
  BDC_TRY(ensure_hist(c, std::max(iters, 1)));
  const int rows = ba_rows(c, prec);
  int ran = 0;
  if (tol <= 0.0) {
    for (int it = 0; it < iters; ++it) BDC_TRY(ba_step(c, prec, it, rows, c->hist.as<double>()));
    ran = iters;
  } else {
    // The stopping rule (main_BAA.cu:73: the loop runs while mx >= a) ends at
    // the first step with mx < tol.  Reading mx back after every step costs a
    // host round trip per step, which dominates small problems; there, steps
    // run in batches from a snapshot of the step state, and a batch that
    // crosses the stopping point is rerun from the snapshot up to that step.
    // No reduction uses atomics, so the rerun is bitwise the same.
    const size_t pbytes = (size_t)rows * c->N2 * 8;
    const bool split = prec == BDC_PREC_SPLIT16;
    const int batch = pbytes <= kStopBatchMaxBytes ? kStopBatch : 1;
    std::vector<double> h;
    while (ran < iters) {
      const int nb = std::min(batch, iters - ran);
      const bool snap = nb > 1;
      bool limb_valid = c->plimb_valid, mirror_stale = c->p_mirror_stale;
      int limb_rows = c->plimb_rows;
      if (snap) {
        if (c->ck_p.bytes < pbytes) BDC_TRY(c->ck_p.alloc(pbytes));
        if (c->ck_s.bytes < c->scal.bytes) BDC_TRY(c->ck_s.alloc(c->scal.bytes));
        BDC_HIP(hipMemcpyAsync(c->ck_p.p, c->P.p, pbytes, hipMemcpyDeviceToDevice, c->stream));
        BDC_HIP(hipMemcpyAsync(c->ck_s.p, c->scal.p, c->scal.bytes, hipMemcpyDeviceToDevice, c->stream));
        if (split) {
          if (c->ck_t.bytes < c->ptexp.bytes) BDC_TRY(c->ck_t.alloc(c->ptexp.bytes));
          BDC_HIP(hipMemcpyAsync(c->ck_t.p, c->ptexp.p, c->ptexp.bytes, hipMemcpyDeviceToDevice, c->stream));
        }
      }
      for (int it = 0; it < nb; ++it) BDC_TRY(ba_step(c, prec, ran + it, rows, c->hist.as<double>()));
      h.resize((size_t)4 * nb);
      BDC_HIP(hipMemcpyAsync(h.data(), c->hist.as<double>() + 4 * ran, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
      BDC_HIP(hipStreamSynchronize(c->stream));
      int stop = -1;
      for (int it = 0; it < nb && stop < 0; ++it)
        if (h[4 * it] < tol) stop = it;
      if (stop < 0) {
        ran += nb;
        continue;
      }
      if (stop < nb - 1) {  // the batch ran past the stopping point
        BDC_HIP(hipMemcpyAsync(c->P.p, c->ck_p.p, pbytes, hipMemcpyDeviceToDevice, c->stream));
        BDC_HIP(hipMemcpyAsync(c->scal.p, c->ck_s.p, c->scal.bytes, hipMemcpyDeviceToDevice, c->stream));
        if (split) BDC_HIP(hipMemcpyAsync(c->ptexp.p, c->ck_t.p, c->ptexp.bytes, hipMemcpyDeviceToDevice, c->stream));
        c->plimb_valid = limb_valid;
        c->plimb_rows = limb_rows;
        c->p_mirror_stale = mirror_stale;
        for (int it = 0; it <= stop; ++it) BDC_TRY(ba_step(c, prec, ran + it, rows, c->hist.as<double>()));
      }
      ran += stop + 1;
      break;
    }
  }
  BDC_TRY(complete_mirrors(c));
  if (ran && rows < c->N1) BDC_LAUNCH(launch_reflect_half(c->P.as<double>(), c->nx, c->stream));
  if (stats && ran) BDC_HIP(hipMemcpyAsync(stats, c->hist.p, (size_t)ran * sizeof(bdc_iter_stats), hipMemcpyDeviceToHost, c->stream));
  double last_mx = c->last_mx_bits;
  if (ran) BDC_HIP(hipMemcpyAsync(&last_mx, c->hist.as<double>() + 4 * (ran - 1), 8, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  c->last_mx_bits = last_mx;
  if (done) *done = ran;
  return BDC_OK;
}

int bdc_iterate_noalloc(bdc_ctx *c, int prec, int iters, bdc_iter_stats *stats) {
  if (!c) return fail(BDC_ERR_ARG, "null context");
  if (prec != BDC_PREC_FP64 && prec != BDC_PREC_SPLIT16) return fail(BDC_ERR_ARG, "unknown precision %d", prec);
  if (iters < 0 || iters > kNoAllocMaxSteps)
    return fail(BDC_ERR_ARG, "bdc_iterate_noalloc runs 0..%d steps per call (got %d)", kNoAllocMaxSteps, iters);
  if (!c->have_p) return fail(BDC_ERR_STATE, "p is not set");
  if (prec == BDC_PREC_SPLIT16 && !c->split.ok)
    return fail(BDC_ERR_STATE, "split16 operands are not prepared (one bdc_iterate or bdc_split_config call does it)");
  double *hist = scratch_ptr(c->device);
  if (!hist) return fail(BDC_ERR_STATE, "no device scratch for device %d", c->device);
  const int rows = ba_rows(c, prec);
  g_forbid_alloc = true;
  int rc = BDC_OK;
  for (int it = 0; it < iters && rc == BDC_OK; ++it) rc = ba_step(c, prec, it, rows, hist);
  g_forbid_alloc = false;
  if (rc != BDC_OK) return rc;
  BDC_TRY(complete_mirrors(c));
  if (iters && rows < c->N1) BDC_LAUNCH(launch_reflect_half(c->P.as<double>(), c->nx, c->stream));
  if (stats && iters) BDC_HIP(hipMemcpyAsync(stats, hist, (size_t)iters * sizeof(bdc_iter_stats), hipMemcpyDeviceToHost, c->stream));
  double last_mx = c->last_mx_bits;
  if (iters) BDC_HIP(hipMemcpyAsync(&last_mx, hist + 4 * (iters - 1), 8, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  c->last_mx_bits = last_mx;
  return BDC_OK;
}

int bdc_certify(bdc_ctx *c, bdc_certificate *out) {
  if (!c || !out) return fail(BDC_ERR_ARG, "null argument");
  if (!c->have_p) return fail(BDC_ERR_STATE, "p is not set");
  BDC_TRY(ensure_dvec(c));
  double *q = c->qv.as<double>(), *L = q + ((size_t)1 << c->k);
  // FP64 throughout, no symmetry assumption
  BDC_TRY(output_law(c, BDC_PREC_FP64, c->P.as<double>(), kScalPScale, false, c->N1, q, L));
  BDC_TRY(cross_operands(c, L));
  int nb = 0;
  BDC_TRY(cross_term(c, BDC_PREC_FP64, EPI_D, c->dvec.as<double>(), &nb));
  c->have_d = true;
  double r[3], ps = 0.0;
  BDC_LAUNCH(launch_dot(c->P.as<double>(), c->dvec.as<double>(), c->nx, c->part.as<double>(), c->splitk.as<double>(), c->stream));
  BDC_HIP(hipMemcpyAsync(r, c->splitk.p, 24, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipMemcpyAsync(&ps, c->scal.as<double>() + kScalPScale, 8, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  out->lower = ps * r[0];
  out->upper_cert = r[1];
  out->sum_p = ps * r[2];
  // the reference's reported bound (main_BAA.cu:123-130): I(p_T) + mx_T
  out->upper_ref = out->lower + c->last_mx_bits * M_LN2;
  return BDC_OK;
}

int bdc_output_law(bdc_ctx *c, int prec, const double *w, double *qo) {
  if (!c || !w || !qo) return fail(BDC_ERR_ARG, "null argument");
  if (prec == BDC_PREC_SPLIT16) BDC_TRY(split_prepare(c));
  BDC_TRY(ensure_dvec(c));
  if (c->cr) {  // natural order in, stored order on the device
    DevBuf tmp;
    BDC_TRY(tmp.alloc(c->nx * 8));
    BDC_HIP(hipMemcpyAsync(tmp.p, w, c->nx * 8, hipMemcpyHostToDevice, c->stream));
    BDC_LAUNCH(launch_permute_cols(tmp.as<double>(), c->dvec.as<double>(), (uint64_t)c->N1, c->cr, c->stream));
    BDC_HIP(hipStreamSynchronize(c->stream));
  } else {
    BDC_HIP(hipMemcpyAsync(c->dvec.p, w, c->nx * 8, hipMemcpyHostToDevice, c->stream));
  }
  double *q = c->qv.as<double>() + 2 * ((size_t)1 << c->k), *L = c->qv.as<double>() + 3 * ((size_t)1 << c->k);
  BDC_TRY(output_law(c, prec, c->dvec.as<double>(), kScalOne, false, c->N1, q, L));
  BDC_HIP(hipMemcpyAsync(qo, q, ((size_t)1 << c->k) * 8, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  return BDC_OK;
}

int bdc_kl(bdc_ctx *c, int prec, const double *log_q, double *d) {
  if (!c || !log_q || !d) return fail(BDC_ERR_ARG, "null argument");
  if (prec == BDC_PREC_SPLIT16) BDC_TRY(split_prepare(c));
  BDC_TRY(ensure_dvec(c));
  const size_t ny = (size_t)1 << c->k;
  double *L = c->qv.as<double>() + ny;
  // centre ln q by its mean (exact: sum_y W(y|x) = 1)
  double mu = 0.0;
  for (size_t y = 0; y < ny; ++y) mu += log_q[y];
  mu /= (double)ny;
  std::vector<double> Lh(ny);
  for (size_t y = 0; y < ny; ++y) Lh[y] = log_q[y] - mu;
  BDC_HIP(hipMemcpyAsync(L, Lh.data(), ny * 8, hipMemcpyHostToDevice, c->stream));
  BDC_HIP(hipMemcpyAsync(c->scal.as<double>() + kScalMu, &mu, 8, hipMemcpyHostToDevice, c->stream));
  if (prec == BDC_PREC_FP64)
    BDC_TRY(cross_operands(c, L));
  else
    BDC_TRY(split_cross_operands(c, L, false, c->N1));
  int nb = 0;
  BDC_TRY(cross_term(c, prec, EPI_D, c->dvec.as<double>(), &nb));
  return copy_x_to_host(c, c->dvec.as<double>(), d);
}

int bdc_entropy_term(bdc_ctx *c, double *h) {
  if (!c || !h) return fail(BDC_ERR_ARG, "null argument");
  return copy_x_to_host(c, c->Hc.as<double>(), h);
}

int bdc_sample_inputs(bdc_ctx *c, int what, const uint32_t *x, uint64_t len, double *out) {
  if (!c || !x || !out) return fail(BDC_ERR_ARG, "null argument");
  if (len == 0) return BDC_OK;
  if (len > (1ull << 24)) return fail(BDC_ERR_ARG, "at most 2^24 inputs per call");
  for (uint64_t i = 0; i < len; ++i)
    if ((uint64_t)x[i] >= c->nx) return fail(BDC_ERR_ARG, "input %u out of range for n = %d", x[i], c->n);
  const double *src = nullptr;
  int slot = -1;
  if (what == BDC_SAMPLE_ENTROPY) {
    src = c->Hc.as<double>();
  } else if (what == BDC_SAMPLE_P) {
    if (!c->have_p) return fail(BDC_ERR_STATE, "p is not set");
    src = c->P.as<double>();
    slot = kScalPScale;
  } else if (what == BDC_SAMPLE_KL) {
    if (!c->have_d) return fail(BDC_ERR_STATE, "no D(x||q) pass yet (call bdc_certify)");
    src = c->dvec.as<double>();
  } else {
    return fail(BDC_ERR_ARG, "unknown sample kind %d", what);
  }
  DevBuf dx, dout;
  BDC_TRY(dx.alloc(len * 4));
  BDC_TRY(dout.alloc(len * 8));
  BDC_HIP(hipMemcpyAsync(dx.p, x, len * 4, hipMemcpyHostToDevice, c->stream));
  BDC_LAUNCH(launch_gather_at(src, dx.as<uint32_t>(), len, c->scal.as<double>(), slot, c->cr, dout.as<double>(),
                              c->stream));
  BDC_HIP(hipMemcpyAsync(out, dout.p, len * 8, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  return BDC_OK;
}

}  // extern "C"

namespace bdc {
// Is theta (2^k, host) invariant under y -> ~y and y -> rev y within a relative
// tolerance (0: exactly)?  Only meaningful when n1 = n2.
static bool theta_symmetric(const bdc_ctx *c, const double *theta, double rtol) {
  if (c->n1 != c->n2) return false;
  const int k = c->k;
  const uint32_t ny = 1u << k;
  for (uint32_t y = 0; y < ny; ++y) {
    const uint32_t o[3] = {comp_bits(y, k), rev_bits(y, k), rev_bits(comp_bits(y, k), k)};
    for (uint32_t z : o) {
      const double a = theta[y], b = theta[z];
      if (rtol == 0.0 ? a != b : !(std::fabs(a - b) <= rtol * (1.0 + std::fabs(a)))) return false;
    }
  }
  return true;
}

// U_beta(theta) and its gradient, theta resident in c->theta.  Writes U to
// devU (device) and the gradient to c->grad.  sym: theta is complement- and
// reversal-symmetric, so q and D(x||q) are too, and the D pass and W^T pi run
// on the half space x1 < N1/2 with the mirror rule, as the BA step does; the
// gradient (symmetric in exact arithmetic) is averaged over each output orbit
// so that Adam keeps theta exactly symmetric.
static int dual_device(bdc_ctx *c, int prec, double beta, double *devU, bool want_grad, bool sym) {
  const size_t ny = (size_t)1 << c->k;
  double *q = c->qv.as<double>(), *L = q + ny, *wq = q + 2 * ny, *wl = q + 3 * ny;
  const int rows = sym ? ba_rows(c, prec) : c->N1;
  const bool half = rows < c->N1;
  BDC_LAUNCH(launch_softmax_log(c->theta.as<double>(), c->k, q, L, c->scal.as<double>(), c->stream));
  if (prec == BDC_PREC_FP64)
    BDC_TRY(cross_operands(c, L, half, half ? rows : 0));
  else
    BDC_TRY(split_cross_operands(c, L, half, rows));
  int nb = 0;
  BDC_TRY(cross_term(c, prec, EPI_D, c->dvec.as<double>(), &nb, rows));
  BDC_LAUNCH(launch_dual_weights(c->dvec.as<double>(), (uint64_t)rows * c->N2, beta, (double)c->N1 / rows,
                                 c->part.as<double>(), nb, c->scal.as<double>(), devU, c->stream));
  if (!want_grad) return BDC_OK;
  // W^T pi through the output-law contraction (adjoint of the cross term)
  BDC_TRY(output_law(c, prec, c->dvec.as<double>(), kScalDualScale, half, rows, wq, wl));
  if (half) {
    BDC_LAUNCH(launch_vec_sub(q, wq, 1.0, wl, ny, c->stream));
    BDC_LAUNCH(launch_orbit_mean(wl, c->k, c->grad.as<double>(), c->stream));
  } else {
    BDC_LAUNCH(launch_vec_sub(q, wq, 1.0, c->grad.as<double>(), ny, c->stream));
  }
  return BDC_OK;
}
static int ensure_dual(bdc_ctx *c) {
  const size_t ny = (size_t)1 << c->k;
  BDC_TRY(ensure_dvec(c));
  if (!c->theta.p) BDC_TRY(c->theta.alloc(ny * 8));
  if (!c->grad.p) BDC_TRY(c->grad.alloc(ny * 8));
  if (!c->adam_m.p) BDC_TRY(c->adam_m.alloc(ny * 8));
  if (!c->adam_v.p) BDC_TRY(c->adam_v.alloc(ny * 8));
  return BDC_OK;
}
}  // namespace bdc

extern "C" {

int bdc_dual(bdc_ctx *c, int prec, const double *theta, double beta, double *u, double *grad) {
  if (!c || !theta || !u || !(beta > 0.0)) return fail(BDC_ERR_ARG, "bad argument (beta must be > 0)");
  if (prec == BDC_PREC_SPLIT16) BDC_TRY(split_prepare(c));
  BDC_TRY(ensure_dual(c));
  const size_t ny = (size_t)1 << c->k;
  BDC_HIP(hipMemcpyAsync(c->theta.p, theta, ny * 8, hipMemcpyHostToDevice, c->stream));
  // (a scalar slot of its own: the split-K scratch is reused by the gradient's output law)
  double *devU = c->scal.as<double>() + kScalDualU;
  BDC_TRY(dual_device(c, prec, beta, devU, grad != nullptr, theta_symmetric(c, theta, 0.0)));
  BDC_HIP(hipMemcpyAsync(u, devU, 8, hipMemcpyDeviceToHost, c->stream));
  if (grad) BDC_HIP(hipMemcpyAsync(grad, c->grad.p, ny * 8, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  return BDC_OK;
}

int bdc_refine_dual(bdc_ctx *c, int prec, double *theta, double beta, int steps, double lr, double *u_hist) {
  if (!c || !theta || !(beta > 0.0) || steps < 0 || !(lr > 0.0)) return fail(BDC_ERR_ARG, "bad argument");
  if (prec == BDC_PREC_SPLIT16) BDC_TRY(split_prepare(c));
  BDC_TRY(ensure_dual(c));
  const size_t ny = (size_t)1 << c->k;
  // a theta symmetric to 1e-9 relative (ln q of a BA iterate: the split path's
  // limb truncation differs between an input and its reversal) is made exactly
  // symmetric (orbit mean) and refined on the half space; the refined bound is
  // certified in FP64 whatever the starting point
  const bool sym = theta_symmetric(c, theta, 1e-9);
  BDC_HIP(hipMemcpyAsync(c->theta.p, theta, ny * 8, hipMemcpyHostToDevice, c->stream));
  if (sym) {
    BDC_HIP(hipMemcpyAsync(c->grad.p, c->theta.p, ny * 8, hipMemcpyDeviceToDevice, c->stream));
    BDC_LAUNCH(launch_orbit_mean(c->grad.as<double>(), c->k, c->theta.as<double>(), c->stream));
  }
  BDC_HIP(hipMemsetAsync(c->adam_m.p, 0, ny * 8, c->stream));
  BDC_HIP(hipMemsetAsync(c->adam_v.p, 0, ny * 8, c->stream));
  DevBuf uh;
  BDC_TRY(uh.alloc((size_t)std::max(steps, 1) * 8));
  for (int s = 0; s < steps; ++s) {
    BDC_TRY(dual_device(c, prec, beta, uh.as<double>() + s, true, sym));
    BDC_LAUNCH(launch_adam(c->theta.as<double>(), c->grad.as<double>(), c->adam_m.as<double>(), c->adam_v.as<double>(), ny,
                           lr, 0.9, 0.999, 1e-8, s + 1, c->stream));
  }
  BDC_HIP(hipMemcpyAsync(theta, c->theta.p, ny * 8, hipMemcpyDeviceToHost, c->stream));
  if (u_hist && steps) BDC_HIP(hipMemcpyAsync(u_hist, uh.p, (size_t)steps * 8, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  return BDC_OK;
}

int bdc_certify_q(bdc_ctx *c, const double *theta, double *upper) {
  if (!c || !theta || !upper) return fail(BDC_ERR_ARG, "null argument");
  BDC_TRY(ensure_dual(c));
  const size_t ny = (size_t)1 << c->k;
  BDC_HIP(hipMemcpyAsync(c->theta.p, theta, ny * 8, hipMemcpyHostToDevice, c->stream));
  double *q = c->qv.as<double>(), *L = q + ny;
  BDC_LAUNCH(launch_softmax_log(c->theta.as<double>(), c->k, q, L, c->scal.as<double>(), c->stream));
  BDC_TRY(cross_operands(c, L));
  int nb = 0;
  BDC_TRY(cross_term(c, BDC_PREC_FP64, EPI_D, c->dvec.as<double>(), &nb));
  c->have_d = true;  // D(x||softmax theta), readable through bdc_sample_inputs
  double r[3];
  BDC_LAUNCH(launch_dot(c->dvec.as<double>(), c->dvec.as<double>(), c->nx, c->part.as<double>(), c->splitk.as<double>(), c->stream));
  BDC_HIP(hipMemcpyAsync(r, c->splitk.p, 24, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  *upper = r[1];
  return BDC_OK;
}

int bdc_log_output(bdc_ctx *c, double *theta) {
  if (!c || !theta) return fail(BDC_ERR_ARG, "null argument");
  if (!c->have_p) return fail(BDC_ERR_STATE, "p is not set");
  const size_t ny = (size_t)1 << c->k;
  double *q = c->qv.as<double>(), *L = q + ny;
  BDC_TRY(output_law(c, BDC_PREC_FP64, c->P.as<double>(), kScalPScale, false, c->N1, q, L));
  std::vector<double> h(ny);
  BDC_HIP(hipMemcpyAsync(h.data(), q, ny * 8, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  for (size_t y = 0; y < ny; ++y) theta[y] = std::log(h[y]);
  return BDC_OK;
}

int bdc_sync(bdc_ctx *c) {
  if (!c) return fail(BDC_ERR_ARG, "null context");
  BDC_HIP(hipStreamSynchronize(c->stream));
  return BDC_OK;
}

int bdc_reset_kernel_timers(bdc_ctx *c, int enable) {
  if (!c) return fail(BDC_ERR_ARG, "null context");
  BDC_HIP(hipStreamSynchronize(c->stream));
  c->timing = enable != 0;
  c->t_g1.used = 0;
  c->t_g2.used = 0;
  return BDC_OK;
}

int bdc_kernel_time(bdc_ctx *c, const char *name, double *avg_ms, int *count) {
  if (!c || !name || !avg_ms) return fail(BDC_ERR_ARG, "null argument");
  Timer *t = nullptr;
  if (!strcmp(name, "g1")) t = &c->t_g1;
  if (!strcmp(name, "g2")) t = &c->t_g2;
  if (!t) return fail(BDC_ERR_ARG, "unknown timer %s", name);
  BDC_HIP(hipStreamSynchronize(c->stream));
  double tot = 0.0;
  for (int i = 0; i < t->used; ++i) {
    float ms = 0.f;
    BDC_HIP(hipEventElapsedTime(&ms, t->start[i], t->stop[i]));
    tot += ms;
  }
  *avg_ms = t->used ? tot / t->used : 0.0;
  if (count) *count = t->used;
  return BDC_OK;
}

int bdc_work_model(bdc_ctx *c, int prec, double *g1, double *g2, double *bytes) {
  if (!c) return fail(BDC_ERR_ARG, "null context");
  if (prec == BDC_PREC_SPLIT16) BDC_TRY(split_prepare(c));
  if (c->chain) {
    // chain path: additions of the deletion steps (2l per output of a step from
    // length l, l per output of a transposed step) are reported as flops; bytes:
    // the epilogue reads S, p and H and writes p (32 B per input)
    const int n = c->n, d = c->n - c->k;
    double fwd = 0.0, bwd = 0.0;
    for (int g = 0; g < c->ngrp; ++g)
      for (int t = 0; t < d; ++t) {
        const int L = c->grpL[g], H = n - t - L - c->grpB[g];
        if (H < 0) continue;
        for (int j = 1; j <= std::min(d - t, L); ++j) {
          const double outs = std::ldexp(1.0, n - t - j);
          fwd += outs * 2.0 * (L - j + 1);
          bwd += outs * 2.0 * (L - j + 1);  // the transposed step into length L - j + 1
        }
      }
    if (g1) *g1 = fwd;
    if (g2) *g2 = bwd;
    if (bytes) *bytes = 32.0 * (double)c->nx;
    return BDC_OK;
  }
  // a BA step covers the rows x1 < ba_rows of P
  const int rows = ba_rows(c, prec);
  const double nx = (double)c->nx * ((double)rows / c->N1);
  // G1 with the mirror rule contracts only Abar (KL columns); G2 has K = KL + KR.
  if (g1) *g1 = 2.0 * c->KL * nx;
  if (g2) *g2 = 2.0 * (c->KL + c->KR) * nx;
  // G2 epilogue: p read and written, H read (24 B/x) at both precisions (the
  // split16 output law forms p's limbs itself from the FP64 p)
  if (bytes) *bytes = nx * 24.0;
  if (prec == BDC_PREC_SPLIT16 && quarter_step(c, rows)) {
    // quarter mode: the listed 128 x 32 tiles are contracted and read (p and
    // H); written are the canonical elements (as many as orbits: M (M + 1) for
    // the M = N1/2 rows) or, with BDC_QUARTER_TRI=0, the whole half space (each
    // element or its mirror)
    const double tiled = 4096.0 * (c->nqtiles[0] + c->nqtiles[1]);
    const double M = (double)rows;
    if (g2) *g2 = 2.0 * (c->KL + c->KR) * tiled;
    if (bytes) *bytes = 16.0 * tiled + 8.0 * (tri_reads() ? M * (M + 1.0) : nx);
  }
  return BDC_OK;
}

int bdc_split_config(bdc_ctx *c, int *g1_limbs, int *g1_bits, int *g2_limbs, int *g2_bits, int *residuals) {
  if (!c) return fail(BDC_ERR_ARG, "null context");
  if (c->chain) {
    c->split.ok = true;  // either precision runs the FP64 chain path
    return fail(BDC_ERR_ARG, "BDC_{%d,%d} runs the FP64 deletion-chain path (k > %d or k = n): no split operands", c->n,
                c->k, kDenseMaxK);
  }
  BDC_TRY(split_prepare(c));
  const Split &sp = c->split;
  if (g1_limbs) *g1_limbs = sp.g1_limbs;
  if (g1_bits) *g1_bits = sp.g1_bits;
  if (g2_limbs) *g2_limbs = sp.g2_limbs;
  if (g2_bits) *g2_bits = sp.g2_bits;
  if (residuals) *residuals = c->n_resid_a + c->n_resid_b;
  return BDC_OK;
}

}  // extern "C"

// =====================================================================
// split-FP16 path orchestration
// =====================================================================
namespace bdc {

static int env_int(const char *name, int dflt) {
  const char *v = getenv(name);
  return v && v[0] ? atoi(v) : dflt;
}

// Largest limb width b <= 11 with mass * 2^(b-1) <= 2^24 (exact FP32 sums).
static int exact_bits(double mass) {
  for (int b = 11; b >= 2; --b)
    if (mass * std::ldexp(1.0, b - 1) <= 16777216.0) return b;
  return 0;
}

// the BA step of the split path runs on the quarter's tiles (columns of P stored
// bit-reversed, the half space, quarter tile lists built)
static bool quarter_step(const bdc_ctx *c, int rows) { return c->cr && rows == c->N1 / 2 && c->qtiles[0].p; }
// quarter steps update only the canonical elements and the output law reads
// the triangle; BDC_QUARTER_TRI=0 writes both members of each reversal pair
static bool tri_reads() {
  const char *e = getenv("BDC_QUARTER_TRI");
  return !(e && e[0] == '0');
}

// G1 split-K: enough blocks (>= 512) for the 256 CUs, each K range >= 1024 rows
static int g1_splits(int rows, int blocks) {
  int ks = 1;
  while (ks < 8 && blocks * ks < 512 && (rows / 128) % (ks * 2) == 0 && rows / (ks * 2) >= 1024) ks *= 2;
  return env_int("BDC_G1_KSPLIT", ks);
}

int split_prepare(bdc_ctx *c) {
  if (c->split.ok) return BDC_OK;
  if (c->chain) {  // the chain path runs in FP64 at either precision setting
    c->split.ok = true;
    return BDC_OK;
  }
  if (c->N1 % 128 || c->N2 % 128)
    return fail(BDC_ERR_ARG, "split-FP16 path needs n >= 14 (128 x 128 tiles of P); use FP64 for BDC_{%d,%d}", c->n,
                c->k);
  const size_t N1 = c->N1, N2 = c->N2, KLp = c->KLp, KRp = c->KRp;
  std::vector<uint16_t> a(N1 * KLp), b(N2 * KRp);
  BDC_HIP(hipMemcpyAsync(a.data(), c->abar16.p, a.size() * 2, hipMemcpyDeviceToHost, c->stream));
  if (KRp) BDC_HIP(hipMemcpyAsync(b.data(), c->bbar16.p, b.size() * 2, hipMemcpyDeviceToHost, c->stream));
  BDC_HIP(hipStreamSynchronize(c->stream));
  // counts fp16 cannot hold are rounded toward zero (the exact residual r >= 0 comes
  // from the sparse lists), so the count mass behind each exactness bound below is
  // at most the true mass (2^16 per row of [Abar | Bbar] at half length 16)
  auto round16 = [](uint16_t v) -> _Float16 {
    _Float16 h = (_Float16)(float)v;
    if ((float)h > (float)v) {  // next fp16 toward zero: one step down in the bit pattern (v > 0)
      const uint16_t b = __builtin_bit_cast(uint16_t, h);
      h = __builtin_bit_cast(_Float16, (uint16_t)(b - 1u));
    }
    return h;
  };
  // aTh: Abar^T for the output law in MFMA operand order -- 32 kappa x 16 x1
  // chunks of 64 lanes x 8 values (lane l: kappa 32 t + l % 32, x1 16 c +
  // 8 (l / 32) + e), chunk (t, c) at (t N1 / 16 + c) * 512 values, kappa
  // padded with zero rows to a multiple of 256 (the output law's block height)
  std::vector<_Float16> ah(N1 * KLp), aTh((KLp + 255) / 256 * 256 * N1, (_Float16)0.0f), bh(N2 * KRp);
  // ah: Abar for the cross term in MFMA operand order -- 32 x1 x 16 kappa
  // chunks of 64 lanes x 8 values (lane l: x1 32 t + l % 32, kappa 16 c +
  // 8 (l / 32) + e), chunk (t, c) at (t KLp / 16 + c) * 512 values
  auto aop = [&](size_t x1, size_t kk) {
    return (((x1 / 32) * (KLp / 16) + kk / 16) * 64 + (x1 % 32) + 32 * ((kk % 16) / 8)) * 8 + kk % 8;
  };
  auto opidx = [&](size_t kk, size_t x1) {
    return (((kk / 32) * (N1 / 16) + x1 / 16) * 64 + (kk % 32) + 32 * ((x1 % 16) / 8)) * 8 + x1 % 8;
  };
  std::vector<int32_t> resA, rowptrA(N1 + 1, 0), entA, rowptrB(N2 + 1, 0), entB;
  double rowA = 0, rowB = 0, chunk = 0;
  std::vector<double> csum(KLp, 0.0);
  for (size_t x1 = 0; x1 < N1; ++x1) {
    double rs = 0;
    for (size_t kk = 0; kk < KLp; ++kk) {
      const uint16_t v = a[x1 * KLp + kk];
      const _Float16 h = round16(v);
      ah[aop(x1, kk)] = h;
      aTh[opidx(kk, x1)] = h;
      const float hf = (float)h;
      rs += std::fabs(hf);
      csum[kk] += std::fabs(hf);
      const int r = (int)v - (int)hf;
      if (r) {
        resA.insert(resA.end(), {(int32_t)x1, (int32_t)kk, r});
        entA.insert(entA.end(), {(int32_t)kk, r});
      }
    }
    rowptrA[x1 + 1] = (int32_t)(entA.size() / 2);
    rowA = std::max(rowA, rs);
    if ((x1 + 1) % 128 == 0) {  // G1 accumulates 128 rows of x1 in FP32
      for (size_t kk = 0; kk < KLp; ++kk) {
        chunk = std::max(chunk, csum[kk]);
        csum[kk] = 0;
      }
    }
  }
  for (size_t x2 = 0; x2 < N2; ++x2) {
    double rs = 0;
    for (size_t kk = 0; kk < KRp; ++kk) {
      const uint16_t v = b[x2 * KRp + kk];
      const _Float16 h = round16(v);
      bh[x2 * KRp + kk] = h;
      rs += std::fabs((float)h);
      const int r = (int)v - (int)(float)h;
      if (r) entB.insert(entB.end(), {(int32_t)kk, r});
    }
    rowptrB[x2 + 1] = (int32_t)(entB.size() / 2);
    rowB = std::max(rowB, rs);
  }
  c->n_resid_a = (int)(resA.size() / 3);
  c->n_resid_b = (int)(entB.size() / 2);
  c->g1_chunk_sum = chunk;
  c->g2_row_sum = rowA + rowB;
  Split &sp = c->split;
  sp.g1_bits = exact_bits(chunk);
  sp.g2_bits = exact_bits(rowA + rowB);
  if (sp.g1_bits < 6 || sp.g2_bits < 6)
    return fail(BDC_ERR_NUMERIC, "exact FP32 accumulation needs limbs of %d / %d bits, too narrow", sp.g1_bits, sp.g2_bits);
  sp.g1_bits = std::min(sp.g1_bits, env_int("BDC_SPLIT_G1_BITS", sp.g1_bits));
  sp.g2_bits = std::min(sp.g2_bits, env_int("BDC_SPLIT_G2_BITS", sp.g2_bits));
  // G1: enough limbs for 44 bits of p relative to its tile maximum, at most 5
  // (5 x 9 = 45 bits at (28,14): the per-step error against FP64 stays ~1e-12 bits
  // over 2000 steps, against ~2e-9 with 4 x 9 = 36 bits, for ~4% of the step time);
  // G2: 4 limbs (4 x 11 = 44 bits at (28,14), 4 x 9 = 36 at half length 16)
  sp.g1_limbs = env_int("BDC_SPLIT_G1_LIMBS", std::min(5, (44 + sp.g1_bits - 1) / sp.g1_bits));
  sp.g2_limbs = env_int("BDC_SPLIT_G2_LIMBS", std::min(4, (42 + sp.g2_bits - 1) / sp.g2_bits));
  if (sp.g1_limbs < 2 || sp.g1_limbs > 5 || sp.g2_limbs < 2 || sp.g2_limbs > 4)
    return fail(BDC_ERR_ARG, "limb counts must be in [2, 5] for G1 and [2, 4] for G2 (got %d, %d)", sp.g1_limbs,
                sp.g2_limbs);
  std::vector<uint8_t> fA(N1 / 128, 0), fB(N2 / 64, 0);
  for (size_t x1 = 0; x1 < N1; ++x1)
    if (rowptrA[x1 + 1] > rowptrA[x1]) fA[x1 / 128] = 1;
  for (size_t x2 = 0; x2 < N2; ++x2)
    if (rowptrB[x2 + 1] > rowptrB[x2]) fB[x2 / 64] = 1;
  for (int h = 0; h < 2; ++h) {  // residual tile lists for the full and the half space
    const size_t rows = h ? N1 / 2 : N1;
    std::vector<int32_t> t;
    if (rows % 128 == 0)
      for (size_t ty = 0; ty < rows / 128; ++ty)
        for (size_t tx = 0; tx < N2 / 32; ++tx)
          if (fA[ty] || fB[tx / 2]) t.push_back((int32_t)(ty * (N2 / 32) + tx));
    c->nrestiles[h] = (int)t.size();
    if (t.empty()) t.push_back(0);
    BDC_TRY(c->restiles[h].alloc(t.size() * 4));
    BDC_HIP(hipMemcpyAsync(c->restiles[h].p, t.data(), t.size() * 4, hipMemcpyHostToDevice, c->stream));
  }
  if (c->cr && quarter_enabled()) {
    // quarter mode of the BA step: the 128 x 32 tiles of the half space holding a
    // representative (canonical r <= c left of the middle column, r + c <= N - 1
    // right of it), without and with residuals, in 8 x 8 patches dealt to the
    // eight XCDs in turn (block b runs on XCD b mod 8)
    const int N = (int)N2, TX = N / 32, TY = N / 2 / 128;
    std::vector<std::vector<int32_t>> lists(3);
    for (int h = 0; h < 2; ++h) {
      std::vector<std::vector<int32_t>> queues(8);
      int patch = 0;
      for (int py = 0; py < (TY + 7) / 8; ++py)
        for (int px = 0; px < (TX + 7) / 8; ++px) {
          std::vector<int32_t> &q = queues[patch % 8];
          bool any = false;
          for (int ty = py * 8; ty < std::min(TY, py * 8 + 8); ++ty)
            for (int tx = px * 8; tx < std::min(TX, px * 8 + 8); ++tx) {
              const int m0 = ty * 128, n0 = tx * 32;
              const bool need = n0 < N / 2 ? m0 <= n0 + 31 : m0 + n0 <= N - 1;
              const bool res = fA[ty] || fB[tx / 2];
              if (need && res == (h == 1)) {
                q.push_back(ty * TX + tx);
                any = true;
              }
            }
          if (any) ++patch;
        }
      for (size_t i = 0;; ++i) {
        bool more = false;
        for (auto &q : queues)
          if (i < q.size()) {
            lists[h].push_back(q[i]);
            more = true;
          }
        if (!more) break;
      }
      c->nqtiles[h] = (int)lists[h].size();
    }
    // the launch list: residual tiles first (blocks below nqtiles[1] add the
    // residual terms), then the others.  The offset moves a patch to another
    // XCD but keeps its tiles together (consecutive tiles of a patch are eight
    // list positions apart while every XCD queue still has tiles)
    lists[2] = lists[1];
    lists[2].insert(lists[2].end(), lists[0].begin(), lists[0].end());
    for (int h = 0; h < 3; ++h) {
      if (lists[h].empty()) lists[h].push_back(0);
      BDC_TRY(c->qtiles[h].alloc(lists[h].size() * 4));
      BDC_HIP(hipMemcpyAsync(c->qtiles[h].p, lists[h].data(), lists[h].size() * 4, hipMemcpyHostToDevice, c->stream));
    }
    BDC_HIP(hipStreamSynchronize(c->stream));
  }
  BDC_TRY(c->blkA.alloc(fA.size()));
  BDC_TRY(c->blkB.alloc(fB.size()));
  BDC_HIP(hipMemcpyAsync(c->blkA.p, fA.data(), fA.size(), hipMemcpyHostToDevice, c->stream));
  BDC_HIP(hipMemcpyAsync(c->blkB.p, fB.data(), fB.size(), hipMemcpyHostToDevice, c->stream));
  if (resA.empty()) resA.push_back(0);
  if (entA.empty()) entA.push_back(0);
  if (entB.empty()) entB.push_back(0);
  BDC_TRY(c->abarh.alloc(ah.size() * 2));
  BDC_TRY(c->abarTh.alloc(aTh.size() * 2));
  BDC_HIP(hipMemcpyAsync(c->abarh.p, ah.data(), ah.size() * 2, hipMemcpyHostToDevice, c->stream));
  BDC_HIP(hipMemcpyAsync(c->abarTh.p, aTh.data(), aTh.size() * 2, hipMemcpyHostToDevice, c->stream));
  if (KRp) {
    BDC_TRY(c->bbarh.alloc(bh.size() * 2));
    BDC_HIP(hipMemcpyAsync(c->bbarh.p, bh.data(), bh.size() * 2, hipMemcpyHostToDevice, c->stream));
  }
  BDC_TRY(c->residA.alloc(resA.size() * 4));
  BDC_TRY(c->rowptrA.alloc(rowptrA.size() * 4));
  BDC_TRY(c->entA.alloc(entA.size() * 4));
  BDC_TRY(c->rowptrB.alloc(rowptrB.size() * 4));
  BDC_TRY(c->entB.alloc(entB.size() * 4));
  BDC_HIP(hipMemcpyAsync(c->residA.p, resA.data(), resA.size() * 4, hipMemcpyHostToDevice, c->stream));
  BDC_HIP(hipMemcpyAsync(c->rowptrA.p, rowptrA.data(), rowptrA.size() * 4, hipMemcpyHostToDevice, c->stream));
  BDC_HIP(hipMemcpyAsync(c->entA.p, entA.data(), entA.size() * 4, hipMemcpyHostToDevice, c->stream));
  BDC_HIP(hipMemcpyAsync(c->rowptrB.p, rowptrB.data(), rowptrB.size() * 4, hipMemcpyHostToDevice, c->stream));
  BDC_HIP(hipMemcpyAsync(c->entB.p, entB.data(), entB.size() * 4, hipMemcpyHostToDevice, c->stream));
  {  // split-K slabs of G1 for every row count a call may use (full and half space)
    size_t need = 0;
    for (size_t rows : {N1, N1 / 2})
      if (rows >= 128 && rows % 128 == 0) {
        const int ks = g1_splits((int)rows, (int)((N2 / 32) * ((KLp + 255) / 256)));
        if (ks > 1) need = std::max(need, (size_t)ks * KLp * N2 * 8);
      }
    if (need > c->g1part.bytes) BDC_TRY(c->g1part.alloc(need));
  }
  BDC_TRY(c->ptexp.alloc((N1 / 128) * (N2 / 32) * 4));
  BDC_TRY(c->vlimb.alloc((size_t)sp.g2_limbs * N2 * KLp * 2));
  if (KRp) BDC_TRY(c->wlimb.alloc((size_t)sp.g2_limbs * N1 * KRp * 2));
  BDC_HIP(hipStreamSynchronize(c->stream));
  sp.ok = true;
  return BDC_OK;
}

int split_output_law_g1(bdc_ctx *c, const double *X, int rows) {
  const Split &sp = c->split;
  if (rows % 128) return fail(BDC_ERR_ARG, "split output law needs a multiple of 128 rows (%d)", rows);
  // tile exponents of X: written by the previous BA update when X is the current P
  const bool isP = X == c->P.as<double>();
  // P of a quarter step: read from the triangle whenever the mirrors are stale.
  // The exponent of a chunk of a strip (128 rows x 32 columns) is that of the
  // values read, whichever element holds them: the tile exponent of the
  // complete P here, the quarter cross term's accumulation after a step.
  const int tri = isP && quarter_step(c, rows) && (tri_reads() || c->p_mirror_stale) ? 1 : 0;
  if (!(isP && c->plimb_valid && c->plimb_rows == rows)) {
    if (c->p_mirror_stale) return fail(BDC_ERR_STATE, "tile exponents of a stale half space");
    BDC_LAUNCH(launch_tile_exp(X, rows, c->N2, sp.g1_bits, c->ptexp.as<int32_t>(), c->stream));
  }
  c->plimb_valid = isP;
  c->plimb_rows = rows;
  const int ks = g1_splits(rows, (c->N2 / 32) * ((c->KLp + 255) / 256));
  if (ks < 1 || rows % (ks * 128)) return fail(BDC_ERR_ARG, "G1 split-K %d does not divide %d rows in 128s", ks, rows);
  const size_t slab = (size_t)c->KLp * c->N2;
  if (ks > 1 && c->g1part.bytes < ks * slab * 8) BDC_TRY(c->g1part.alloc(ks * slab * 8));
  double *dstT = ks > 1 ? c->g1part.as<double>() : c->T.as<double>();
  BDC_LAUNCH(launch_split_g1(c->KLp, rows, c->N1, c->N2, c->abarTh.as<_Float16>(), X, c->ptexp.as<int32_t>(),
                             sp.g1_limbs, sp.g1_bits, ks, dstT, tri, env_int("BDC_G1_MERGE", 1),
                             env_int("BDC_G1_FIELDS", 1), env_int("BDC_G1_UNIFORM_LOOPS", 1), c->stream));
  if (ks > 1) BDC_LAUNCH(launch_reduce_splits(dstT, ks, (int64_t)slab, (uint64_t)slab, c->T.as<double>(), c->stream));
  BDC_LAUNCH(launch_resid_g1(c->residA.as<int32_t>(), c->n_resid_a, X, rows, c->N2, tri, c->T.as<double>(), c->stream));
  return BDC_OK;
}

int split_cross_operands(bdc_ctx *c, const double *L, bool mirror, int rows, bool reset_texp) {
  if (c->chain) return chain_cross(c, L);
  // V, W in FP64 (stage contractions); the deletion chain also leaves the block
  // maxima of |V| for the limb exponent when W is a permutation of V
  int namax = 0;
  MirrorW mw{};
  BDC_TRY(cross_operands(c, L, mirror, rows, c->part.as<double>(), c->part.bytes / 8, &namax,
                         env_int("BDC_VW_FUSED", 1) ? &mw : nullptr));
  const Split &sp = c->split;
  // the quarter step's tile exponents are reset on the way (its cross term
  // accumulates them with atomic maxima)
  const int nt = reset_texp ? (rows / 128) * (c->N2 / 32) : 0;
  if (mw.count > 0) {  // Wbar formed from Vbar with its limbs, beside the V limbs
    BDC_LAUNCH(launch_vw_limbs(vbar(c), c->KL, c->KLp, c->N2, mw, c->g2b.as<double>(), c->n2, c->g2a.as<double>(),
                               rows > 0 ? rows : c->N1, c->KR, c->KRp, sp.g2_limbs, sp.g2_bits,
                               c->vlimb.as<_Float16>(), c->wlimb.as<_Float16>(), c->part.as<double>(), namax,
                               c->scal.as<double>(), c->ptexp.as<int32_t>(), nt, c->stream));
  } else {
    BDC_LAUNCH(launch_split_vw(vbar(c), c->KL, c->KLp, c->N2, c->g2a.as<double>(), rows, c->KR, c->KRp,
                               sp.g2_limbs, sp.g2_bits, c->vlimb.as<_Float16>(), c->wlimb.as<_Float16>(),
                               c->part.as<double>(), namax, c->scal.as<double>(), c->ptexp.as<int32_t>(), nt,
                               c->stream));
  }
  c->texp_reset = nt > 0;
  return BDC_OK;
}

int split_cross_term(bdc_ctx *c, int epi, const EpiArgs &ea, int *nb, int rows) {
  const Split &sp = c->split;
  // the BA update writes the next step's P limbs in its epilogue
  const bool fuse = (epi == EPI_BA) && ea.P == c->P.as<double>();
  SplitG2Args a{};
  a.N1 = c->N1;
  a.N2 = c->N2;
  a.M = rows;
  a.KL = c->KL;
  a.KLp = c->KLp;
  a.KR = c->KR;
  a.KRp = c->KRp;
  a.A16 = c->abarh.as<_Float16>();
  a.vT = c->vlimb.as<_Float16>();
  a.wl = c->wlimb.as<_Float16>();
  a.B16 = c->bbarh.as<_Float16>();
  a.limbs = sp.g2_limbs;
  a.bits = sp.g2_bits;
  a.epi = epi;
  a.blkA = c->blkA.as<uint8_t>();
  a.blkB = c->blkB.as<uint8_t>();
  a.rowptrA = c->rowptrA.as<int32_t>();
  a.entA = c->entA.as<int32_t>();
  a.rowptrB = c->rowptrB.as<int32_t>();
  a.entB = c->entB.as<int32_t>();
  a.V64 = vbar(c);
  a.W64 = c->g2a.as<double>();
  a.texp_next = fuse ? c->ptexp.as<int32_t>() : nullptr;
  a.bits_p = sp.g1_bits;
  a.exptab = c->exptab.as<double>();
  const int h = rows == c->N1 ? 0 : 1;
  a.res_tiles = c->restiles[h].as<int32_t>();
  a.nres_tiles = c->nrestiles[h];
  // the half-space BA step with bit-reversed columns runs on the quarter's tiles,
  // updating the canonical elements only (the output law reads the triangle)
  a.quarter = fuse && quarter_step(c, rows) ? 1 : 0;
  a.mirror = tri_reads() ? 0 : 1;
  a.texp_reset = a.quarter && c->texp_reset ? 1 : 0;
  c->texp_reset = false;
  if (a.quarter && !a.mirror) c->p_mirror_stale = true;
  a.q_tiles = c->qtiles[0].as<int32_t>();
  a.qres_tiles = c->qtiles[1].as<int32_t>();
  a.qall_tiles = c->qtiles[2].as<int32_t>();
  a.nq_tiles = c->nqtiles[0];
  a.nqres_tiles = c->nqtiles[1];
  BDC_LAUNCH(launch_split_g2(a, ea, c->stream, nb));
  if (epi == EPI_BA) {
    c->plimb_valid = fuse;
    c->plimb_rows = rows;
  }
  return BDC_OK;
}

}  // namespace bdc

// host_io.cpp - errors, edge-list input/output, library information.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cctype>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "kcl_internal.h"

namespace kcl {

static thread_local char g_err[512] = "";

kcl_status fail(kcl_status s, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return s;
}

}  // namespace kcl

extern "C" const char *kcl_last_error(void) { return kcl::g_err; }

extern "C" const char *kcl_status_name(kcl_status s) {
  switch (s) {
    case KCL_OK: return "KCL_OK";
    case KCL_E_IO: return "KCL_E_IO";
    case KCL_E_PARSE: return "KCL_E_PARSE";
    case KCL_E_INVALID_ARG: return "KCL_E_INVALID_ARG";
    case KCL_E_OOM_HOST: return "KCL_E_OOM_HOST";
    case KCL_E_OOM_DEVICE: return "KCL_E_OOM_DEVICE";
    case KCL_E_HIP: return "KCL_E_HIP";
    case KCL_E_OVERFLOW: return "KCL_E_OVERFLOW";
    case KCL_E_UNSUPPORTED: return "KCL_E_UNSUPPORTED";
  }
  return "KCL_E_UNKNOWN";
}

extern "C" const char *kcl_version(void) { return "kclpp 0.1.0 (gfx950)"; }

extern "C" void kcl_free(void *p) { free(p); }

namespace {

// One "%u" conversion as glibc's scanf performs it: skip white space,
// optional sign, decimal digits into an unsigned long (saturating), the
// result truncated to 32 bits.  Returns false when no digit follows.
bool scan_u32(const char *&p, const char *end, uint32_t *out) {
  while (p < end && isspace((unsigned char)*p)) ++p;
  bool neg = false;
  if (p < end && (*p == '+' || *p == '-')) {
    neg = (*p == '-');
    ++p;
  }
  if (p >= end || !isdigit((unsigned char)*p)) return false;
  uint64_t v = 0;
  bool sat = false;
  while (p < end && isdigit((unsigned char)*p)) {
    uint64_t d = (uint64_t)(*p - '0');
    if (!sat) {
      if (v > (UINT64_MAX - d) / 10) {
        sat = true;
        v = UINT64_MAX;
      } else {
        v = v * 10 + d;
      }
    }
    ++p;
  }
  if (neg && !sat) v = (uint64_t)(-(int64_t)v);
  *out = (uint32_t)v;
  return true;
}

// The text of an edge list parsed by the host threads: the file is cut
// into chunks at white space (so no number straddles two chunks); each
// thread counts its chunk's numbers up to its first text that does not
// parse, the counts are summed in order to give every chunk its first
// number's index and the index where parsing stops (the first such text in
// the file, like fscanf), then each thread converts its chunk into place.
// Pairs: floor(numbers / 2), the reference's "%u %u" loop.
kcl_status parse_text(const char *base, size_t size, const char *path, uint32_t **out,
                      uint64_t *pairs_out, uint32_t *max_out) {
  const uint32_t hw = std::thread::hardware_concurrency();
  const uint32_t W = size < (1u << 22) ? 1u : std::max(1u, std::min(16u, hw ? hw : 1u));
  std::vector<const char *> cut(W + 1);
  cut[0] = base;
  cut[W] = base + size;
  for (uint32_t t = 1; t < W; ++t) {
    const char *c = std::max(cut[t - 1], base + size * t / W);
    while (c < base + size && !isspace((unsigned char)*c)) ++c;
    cut[t] = c;
  }
  struct Part {
    uint64_t numbers = 0;
    bool stopped = false;  // text that does not parse inside this chunk
    uint32_t mx = 0;
  };
  std::vector<Part> part(W);
  // Numbers of chunk t (into dst[0, limit) when dst is set).  After a
  // number, white space or a sign may follow (fscanf's next "%u" starts
  // there); any other text ends the input after that number.
  auto scan_chunk = [&](uint32_t t, uint32_t *dst, uint64_t limit) {
    const char *p = cut[t], *end = cut[t + 1];
    Part &pt = part[t];
    uint64_t n = 0;
    for (;;) {
      while (p < end && isspace((unsigned char)*p)) ++p;
      if (p >= end) break;
      uint32_t v;
      if (!scan_u32(p, end, &v)) {
        pt.stopped = true;
        break;
      }
      if (dst) {
        if (n >= limit) break;
        dst[n] = v;
        pt.mx = std::max(pt.mx, v);
      }
      ++n;
      if (p < end && !isspace((unsigned char)*p) && *p != '+' && *p != '-') {
        pt.stopped = true;
        break;
      }
    }
    if (!dst) pt.numbers = n;
  };
  auto run = [&](auto fn) {
    std::vector<std::thread> th;
    for (uint32_t t = 1; t < W; ++t) th.emplace_back(fn, t);
    fn(0);
    for (auto &x : th) x.join();
  };
  run([&](uint32_t t) { scan_chunk(t, nullptr, 0); });
  std::vector<uint64_t> first(W + 1, 0);
  uint32_t last = W;
  for (uint32_t t = 0; t < W; ++t) {
    first[t + 1] = first[t] + part[t].numbers;
    if (part[t].stopped) {
      last = t + 1;
      break;
    }
  }
  const uint64_t numbers = first[last];
  const uint64_t pairs = numbers / 2;
  uint32_t *buf = (uint32_t *)malloc(std::max<uint64_t>(2 * pairs, 2) * sizeof(uint32_t));
  if (!buf) return kcl::fail(KCL_E_OOM_HOST, "out of host memory reading %s", path);
  for (auto &pt : part) pt.mx = 0;
  run([&](uint32_t t) {
    if (t >= last || first[t] >= 2 * pairs) return;
    scan_chunk(t, buf + first[t], 2 * pairs - first[t]);
  });
  uint32_t mx = 0;
  for (auto &pt : part) mx = std::max(mx, pt.mx);
  *out = buf;
  *pairs_out = pairs;
  *max_out = mx;
  return KCL_OK;
}

bool ends_with(const char *s, const char *suf) {
  size_t a = strlen(s), b = strlen(suf);
  return a >= b && strcmp(s + a - b, suf) == 0;
}

}  // namespace

extern "C" kcl_status kcl_read_edge_list(const char *path, uint32_t **pairs, uint64_t *m,
                                         uint32_t *n_in) {
  using kcl::fail;
  if (!path || !pairs || !m || !n_in) return fail(KCL_E_INVALID_ARG, "NULL argument");
  *pairs = nullptr;
  *m = 0;
  *n_in = 0;
  int fd = open(path, O_RDONLY);
  if (fd < 0) return fail(KCL_E_IO, "cannot open %s: %s", path, strerror(errno));
  struct stat st;
  if (fstat(fd, &st) != 0) {
    close(fd);
    return fail(KCL_E_IO, "cannot stat %s", path);
  }
  size_t size = (size_t)st.st_size;
  const char *base = nullptr;
  if (size > 0) {
    void *mp = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (mp == MAP_FAILED) {
      close(fd);
      return fail(KCL_E_IO, "cannot map %s", path);
    }
    base = (const char *)mp;
    madvise(mp, size, MADV_SEQUENTIAL);
  }
  close(fd);

  uint32_t *buf = nullptr;
  uint64_t cnt = 0;
  uint32_t mx = 0;
  kcl_status rc = KCL_OK;
  if (ends_with(path, ".u32")) {
    if (size % 8 != 0) {
      rc = fail(KCL_E_PARSE, "%s: %zu bytes is not a whole number of uint32 pairs", path, size);
    } else {
      cnt = size / 8;
      buf = (uint32_t *)malloc(cnt ? size : 8);
      if (!buf) rc = fail(KCL_E_OOM_HOST, "out of host memory for %llu edges",
                          (unsigned long long)cnt);
      else {
        if (size) memcpy(buf, base, size);
        for (uint64_t i = 0; i < 2 * cnt; ++i) mx = buf[i] > mx ? buf[i] : mx;
      }
    }
  } else {
    rc = parse_text(base, size, path, &buf, &cnt, &mx);
  }
  if (base) munmap((void *)base, size);
  if (rc != KCL_OK) {
    free(buf);
    return rc;
  }
  if (cnt > 0xFFFFFFFFull - 1)
    return free(buf), fail(KCL_E_UNSUPPORTED, "%llu edges: the limit is 2^32-1",
                           (unsigned long long)cnt);
  if (mx == 0xFFFFFFFFu)
    return free(buf), fail(KCL_E_UNSUPPORTED, "node id 4294967295 leaves no room for n = max+1");
  *pairs = buf;
  *m = cnt;
  *n_in = mx + 1;
  return KCL_OK;
}

extern "C" kcl_status kcl_write_edge_list(const char *path, const uint32_t *pairs, uint64_t m) {
  using kcl::fail;
  if (!path || (!pairs && m)) return fail(KCL_E_INVALID_ARG, "NULL argument");
  if (ends_with(path, ".u32")) {
    FILE *f = fopen(path, "wb");
    if (!f) return fail(KCL_E_IO, "cannot create %s", path);
    size_t w = m ? fwrite(pairs, 8, (size_t)m, f) : 0;
    int c = fclose(f);
    if (w != m || c != 0) return fail(KCL_E_IO, "short write to %s", path);
    return KCL_OK;
  }
  FILE *f = fopen(path, "w");
  if (!f) return fail(KCL_E_IO, "cannot create %s", path);
  std::vector<char> out;
  out.reserve(1 << 22);
  char tmp[32];
  for (uint64_t i = 0; i < m; ++i) {
    int len = snprintf(tmp, sizeof(tmp), "%u %u\n", pairs[2 * i], pairs[2 * i + 1]);
    out.insert(out.end(), tmp, tmp + len);
    if (out.size() > (1u << 22)) {
      if (fwrite(out.data(), 1, out.size(), f) != out.size()) {
        fclose(f);
        return fail(KCL_E_IO, "short write to %s", path);
      }
      out.clear();
    }
  }
  if (!out.empty() && fwrite(out.data(), 1, out.size(), f) != out.size()) {
    fclose(f);
    return fail(KCL_E_IO, "short write to %s", path);
  }
  if (fclose(f) != 0) return fail(KCL_E_IO, "short write to %s", path);
  return KCL_OK;
}

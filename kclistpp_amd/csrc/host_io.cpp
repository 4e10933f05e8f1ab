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
    uint64_t cap = size / 4 + 16;  // at least 4 bytes per number: "a b\n"
    buf = (uint32_t *)malloc(cap * sizeof(uint32_t));
    if (!buf) rc = fail(KCL_E_OOM_HOST, "out of host memory reading %s", path);
    const char *p = base, *end = base + size;
    while (rc == KCL_OK) {
      uint32_t a, b;
      if (!scan_u32(p, end, &a)) break;
      if (!scan_u32(p, end, &b)) break;
      if (2 * cnt + 2 > cap) {
        cap = cap * 2 + 16;
        uint32_t *nb = (uint32_t *)realloc(buf, cap * sizeof(uint32_t));
        if (!nb) {
          rc = fail(KCL_E_OOM_HOST, "out of host memory reading %s", path);
          break;
        }
        buf = nb;
      }
      buf[2 * cnt] = a;
      buf[2 * cnt + 1] = b;
      mx = a > mx ? a : mx;
      mx = b > mx ? b : mx;
      ++cnt;
    }
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

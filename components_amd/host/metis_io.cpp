// METIS graph files, read with the semantics of GraphIO::ReadMETISFile
// (lib/io/graph_io.h:281-382) and ParseVertexFilestream (:1207-1248):
//   - lines starting with '%' are skipped (header and vertex lines);
//   - the header holds n and m; m is not trusted (the graph is built from the
//     lines) and any third field (METIS fmt/weights) is ignored;
//   - vertex line i lists 1-based neighbours; parsing of a line stops at a 0
//     or at the end of the line; an empty line is an isolated vertex; lines
//     after the n-th are ignored, missing lines are isolated vertices.
// The reference adds one directed edge per listed neighbour and relies on the
// file listing both directions; this reader symmetrises, drops self-loops and
// merges duplicate edges, which leaves the components unchanged.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "csr_build.h"
#include "kacomp.h"

namespace kc {
void set_error(const char *fmt, ...);
}

namespace {

bool read_file(const char *path, std::string &buf) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz < 0) {
    std::fclose(f);
    return false;
  }
  buf.resize(static_cast<size_t>(sz));
  size_t got = sz ? std::fread(&buf[0], 1, static_cast<size_t>(sz), f) : 0;
  std::fclose(f);
  return got == static_cast<size_t>(sz);
}

}  // namespace

extern "C" int kc_read_metis(const char *path, kc_csr *out) {
  if (!path || !out) {
    kc::set_error("kc_read_metis: null argument");
    return KC_ERR_INVALID;
  }
  std::memset(out, 0, sizeof(*out));
  std::string buf;
  if (!read_file(path, buf)) {
    kc::set_error("Error opening %s", path);
    return KC_ERR_IO;
  }
  const char *p = buf.data();
  const char *end = p + buf.size();
  auto next_line = [&](const char *&ls, const char *&le) -> bool {
    if (p >= end) return false;
    ls = p;
    const char *nl = static_cast<const char *>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
    le = nl ? nl : end;
    p = nl ? nl + 1 : end;
    return true;
  };
  const char *ls = nullptr, *le = nullptr;
  bool have = false;
  while ((have = next_line(ls, le)) && ls < le && *ls == '%') {
  }
  if (!have) {
    kc::set_error("%s: missing header", path);
    return KC_ERR_IO;
  }
  std::string header(ls, le);
  char *q = nullptr;
  errno = 0;
  unsigned long long n = std::strtoull(header.c_str(), &q, 10);
  if (q == header.c_str() || errno) {
    kc::set_error("%s: malformed header '%s'", path, header.c_str());
    return KC_ERR_IO;
  }
  if (n >= 0xFFFFFFFFull) {
    kc::set_error("%s: n=%llu exceeds 32-bit vertex ids", path, n);
    return KC_ERR_INVALID;
  }
  std::vector<uint32_t> src, dst;
  uint64_t v = 0;
  std::string line;
  while (v < n && next_line(ls, le)) {
    if (ls < le && *ls == '%') continue;
    line.assign(ls, le);
    const char *s = line.c_str();
    for (;;) {
      char *e = nullptr;
      long long t = std::strtoll(s, &e, 10);
      if (e == s || t == 0) break;
      if (t < 0 || static_cast<unsigned long long>(t) > n) {
        kc::set_error("%s: vertex %llu lists neighbour %lld outside [1,%llu]", path,
                      static_cast<unsigned long long>(v + 1), t, n);
        return KC_ERR_INVALID;
      }
      src.push_back(static_cast<uint32_t>(v));
      dst.push_back(static_cast<uint32_t>(t - 1));
      s = e;
    }
    ++v;
  }
  kc::build_symmetric_csr(n, src.data(), dst.data(), src.size(), out);
  if (!out->offsets) {
    kc::set_error("%s: out of host memory", path);
    return KC_ERR_OOM;
  }
  return KC_OK;
}

extern "C" int kc_write_metis(const char *path, const kc_csr *g) {
  if (!path || !g || (g->n && !g->offsets)) {
    kc::set_error("kc_write_metis: null argument");
    return KC_ERR_INVALID;
  }
  FILE *f = std::fopen(path, "wb");
  if (!f) {
    kc::set_error("Error opening %s for writing", path);
    return KC_ERR_IO;
  }
  const uint64_t m2 = g->n ? g->offsets[g->n] : 0;
  std::fprintf(f, "%llu %llu\n", static_cast<unsigned long long>(g->n), static_cast<unsigned long long>(m2 / 2));
  std::string line;
  char tmp[24];
  for (uint64_t v = 0; v < g->n; ++v) {
    line.clear();
    for (uint64_t e = g->offsets[v]; e < g->offsets[v + 1]; ++e) {
      int k = std::snprintf(tmp, sizeof(tmp), e + 1 < g->offsets[v + 1] ? "%u " : "%u",
                            static_cast<unsigned>(g->targets[e] + 1));
      line.append(tmp, static_cast<size_t>(k));
    }
    line.push_back('\n');
    std::fwrite(line.data(), 1, line.size(), f);
  }
  const bool ok = std::fclose(f) == 0;
  if (!ok) {
    kc::set_error("write to %s failed", path);
    return KC_ERR_IO;
  }
  return KC_OK;
}

extern "C" void kc_csr_free(kc_csr *g) {
  if (!g) return;
  std::free(g->offsets);
  std::free(g->targets);
  g->offsets = nullptr;
  g->targets = nullptr;
  g->n = g->m = 0;
}

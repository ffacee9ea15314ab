// Edge-list graph files, read with the semantics of the reference's readers:
//   - sorted edge file, GraphIO::ReadSortedEdgeFile (lib/io/graph_io.h:385-669)
//     with the line parser ParseEdgeFilestream (:1250-1300): lines starting
//     with '%' are skipped; the first other line is the header "<type> n m";
//     then every line "<type> u v" (1-based ids) is one edge, and the reader
//     stops after m of them (the header's m bounds the lines read);
//   - partitioned sorted edge files, ReadPartitionedSortedEdgeFile
//     (:672-907): rank r reads "<path>_<r>", which carries the same global
//     header and that rank's edge lines, all of them; one GPU holds every
//     part, so parts files "<path>_0" .. "<path>_<parts-1>" are read;
//   - sorted binary file, ReadSortedBinaryFile (:910-1175): n and m as u64,
//     then m (u, v) pairs of u64, 1-based.
// Each edge adds both directions (the reference's edge_list gets (u, v) and
// (v, u)); like the METIS reader this one symmetrises, drops self-loops and
// merges duplicate edges, which leaves the components unchanged.  Ids outside
// [1, n] are rejected (the reference does not check them).
#include <algorithm>
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

// Splits a buffer into lines ('\n', a trailing '\r' dropped).
struct Lines {
  const char *p, *end;
  bool next(const char *&ls, const char *&le) {
    if (p >= end) return false;
    ls = p;
    const char *nl = static_cast<const char *>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
    le = nl ? nl : end;
    p = nl ? nl + 1 : end;
    if (le > ls && le[-1] == '\r') --le;
    return true;
  }
};

const char *skip_space(const char *s, const char *e) {
  while (s < e && (*s == ' ' || *s == '\t')) ++s;
  return s;
}

// Next whitespace-separated token of [s, e) as an unsigned number; false when
// the token is missing or not a number.
bool next_number(const char *&s, const char *e, unsigned long long &v) {
  s = skip_space(s, e);
  if (s >= e || *s < '0' || *s > '9') return false;
  unsigned long long x = 0;
  while (s < e && *s >= '0' && *s <= '9') {
    const unsigned d = static_cast<unsigned>(*s - '0');
    if (x > (~0ull - d) / 10) return false;
    x = x * 10 + d;
    ++s;
  }
  v = x;
  return true;
}

// Skips the leading type token of a line ("e", "a", "p", ...).
const char *skip_token(const char *s, const char *e) {
  s = skip_space(s, e);
  while (s < e && *s != ' ' && *s != '\t') ++s;
  return s;
}

struct EdgeFile {
  unsigned long long n = 0, m = 0;
  std::vector<uint32_t> src, dst;
};

// Parses one sorted edge file into f (header checked against f.n when
// have_header is set).  Returns KC_OK or an error with the message set.
int parse_edge_file(const char *path, bool bounded, bool have_header, EdgeFile &f) {
  std::string buf;
  if (!read_file(path, buf)) {
    kc::set_error("Error opening %s", path);
    return KC_ERR_IO;
  }
  Lines lines{buf.data(), buf.data() + buf.size()};
  const char *ls = nullptr, *le = nullptr;
  bool have = false;
  while ((have = lines.next(ls, le)) && ls < le && *ls == '%') {
  }
  if (!have) {
    kc::set_error("%s: missing header", path);
    return KC_ERR_IO;
  }
  unsigned long long n = 0, m = 0;
  const char *s = skip_token(ls, le);
  if (!next_number(s, le, n) || !next_number(s, le, m)) {
    kc::set_error("%s: malformed header '%.*s' (expected '<type> n m')", path, static_cast<int>(le - ls), ls);
    return KC_ERR_IO;
  }
  if (n >= 0xFFFFFFFFull) {
    kc::set_error("%s: n=%llu exceeds 32-bit vertex ids", path, n);
    return KC_ERR_INVALID;
  }
  if (have_header && (n != f.n || m != f.m)) {
    kc::set_error("%s: header n=%llu m=%llu differs from the first part's n=%llu m=%llu", path, n, m, f.n, f.m);
    return KC_ERR_INVALID;
  }
  f.n = n;
  f.m = m;
  unsigned long long read = 0;
  while ((!bounded || read < m) && lines.next(ls, le)) {
    if (ls >= le || *ls == '%') continue;
    unsigned long long u = 0, v = 0;
    s = skip_token(ls, le);
    if (!next_number(s, le, u) || !next_number(s, le, v)) {
      kc::set_error("%s: malformed edge line '%.*s' (expected '<type> u v')", path, static_cast<int>(le - ls), ls);
      return KC_ERR_IO;
    }
    if (u == 0 || v == 0 || u > n || v > n) {
      kc::set_error("%s: edge (%llu, %llu) outside [1,%llu]", path, u, v, n);
      return KC_ERR_INVALID;
    }
    f.src.push_back(static_cast<uint32_t>(u - 1));
    f.dst.push_back(static_cast<uint32_t>(v - 1));
    ++read;
  }
  return KC_OK;
}

int finish(const char *what, EdgeFile &f, kc_csr *out) {
  kc::build_symmetric_csr(f.n, f.src.data(), f.dst.data(), f.src.size(), out);
  if (!out->offsets) {
    kc::set_error("%s: out of host memory", what);
    return KC_ERR_OOM;
  }
  return KC_OK;
}

}  // namespace

extern "C" int kc_read_edgefile(const char *path, kc_csr *out) {
  if (!path || !out) {
    kc::set_error("kc_read_edgefile: null argument");
    return KC_ERR_INVALID;
  }
  std::memset(out, 0, sizeof(*out));
  EdgeFile f;
  const int rc = parse_edge_file(path, true, false, f);
  return rc != KC_OK ? rc : finish(path, f, out);
}

extern "C" int kc_read_partitioned_edgefile(const char *path, int parts, kc_csr *out) {
  if (!path || !out || parts < 1) {
    kc::set_error("kc_read_partitioned_edgefile: null argument or parts < 1");
    return KC_ERR_INVALID;
  }
  std::memset(out, 0, sizeof(*out));
  EdgeFile f;
  for (int r = 0; r < parts; ++r) {
    const std::string part = std::string(path) + "_" + std::to_string(r);
    const int rc = parse_edge_file(part.c_str(), false, r > 0, f);
    if (rc != KC_OK) return rc;
  }
  return finish(path, f, out);
}

extern "C" int kc_read_binary(const char *path, kc_csr *out) {
  if (!path || !out) {
    kc::set_error("kc_read_binary: null argument");
    return KC_ERR_INVALID;
  }
  std::memset(out, 0, sizeof(*out));
  FILE *fp = std::fopen(path, "rb");
  if (!fp) {
    kc::set_error("Error opening %s", path);
    return KC_ERR_IO;
  }
  uint64_t head[2] = {0, 0};
  if (std::fread(head, sizeof(uint64_t), 2, fp) != 2) {
    std::fclose(fp);
    kc::set_error("%s: missing header (n and m as u64)", path);
    return KC_ERR_IO;
  }
  const uint64_t n = head[0], m = head[1];
  if (n >= 0xFFFFFFFFull) {
    std::fclose(fp);
    kc::set_error("%s: n=%llu exceeds 32-bit vertex ids", path, static_cast<unsigned long long>(n));
    return KC_ERR_INVALID;
  }
  std::fseek(fp, 0, SEEK_END);
  const long size = std::ftell(fp);
  std::fseek(fp, 2 * sizeof(uint64_t), SEEK_SET);
  const uint64_t present = size > 16 ? static_cast<uint64_t>(size - 16) / 16 : 0;
  if (present < m) {
    std::fclose(fp);
    kc::set_error("%s: %llu of %llu edges present", path, static_cast<unsigned long long>(present),
                  static_cast<unsigned long long>(m));
    return KC_ERR_IO;
  }
  EdgeFile f;
  f.n = n;
  f.m = m;
  f.src.reserve(m);
  f.dst.reserve(m);
  std::vector<uint64_t> chunk(2 * (1u << 20));
  uint64_t left = m;
  while (left) {
    const uint64_t take = std::min<uint64_t>(left, chunk.size() / 2);
    if (std::fread(chunk.data(), 2 * sizeof(uint64_t), take, fp) != take) {
      std::fclose(fp);
      kc::set_error("%s: %llu of %llu edges present", path, static_cast<unsigned long long>(m - left),
                    static_cast<unsigned long long>(m));
      return KC_ERR_IO;
    }
    for (uint64_t i = 0; i < take; ++i) {
      const uint64_t u = chunk[2 * i], v = chunk[2 * i + 1];
      if (u == 0 || v == 0 || u > n || v > n) {
        std::fclose(fp);
        kc::set_error("%s: edge (%llu, %llu) outside [1,%llu]", path, static_cast<unsigned long long>(u),
                      static_cast<unsigned long long>(v), static_cast<unsigned long long>(n));
        return KC_ERR_INVALID;
      }
      f.src.push_back(static_cast<uint32_t>(u - 1));
      f.dst.push_back(static_cast<uint32_t>(v - 1));
    }
    left -= take;
  }
  std::fclose(fp);
  return finish(path, f, out);
}

extern "C" int kc_write_binary(const char *path, const kc_csr *g) {
  if (!path || !g || (g->n && !g->offsets)) {
    kc::set_error("kc_write_binary: null argument");
    return KC_ERR_INVALID;
  }
  FILE *fp = std::fopen(path, "wb");
  if (!fp) {
    kc::set_error("Error opening %s for writing", path);
    return KC_ERR_IO;
  }
  // every undirected edge once, as (smaller, larger), sorted by its first id
  uint64_t m = 0;
  for (uint64_t v = 0; v < g->n; ++v)
    for (uint64_t e = g->offsets[v]; e < g->offsets[v + 1]; ++e) m += g->targets[e] > v;
  const uint64_t head[2] = {g->n, m};
  bool ok = std::fwrite(head, sizeof(uint64_t), 2, fp) == 2;
  std::vector<uint64_t> pairs;
  for (uint64_t v = 0; v < g->n && ok; ++v) {
    pairs.clear();
    for (uint64_t e = g->offsets[v]; e < g->offsets[v + 1]; ++e)
      if (g->targets[e] > v) {
        pairs.push_back(v + 1);
        pairs.push_back(static_cast<uint64_t>(g->targets[e]) + 1);
      }
    if (!pairs.empty()) ok = std::fwrite(pairs.data(), sizeof(uint64_t), pairs.size(), fp) == pairs.size();
  }
  ok = std::fclose(fp) == 0 && ok;
  if (!ok) {
    kc::set_error("write to %s failed", path);
    return KC_ERR_IO;
  }
  return KC_OK;
}

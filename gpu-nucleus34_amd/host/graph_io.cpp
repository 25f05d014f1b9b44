#include "graph_io.hpp"

#include <algorithm>
#include <cstring>
#include <fcntl.h>
#include <fstream>
#include <iterator>
#include <limits>
#include <omp.h>
#include <sstream>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

namespace nuc34 {

std::string file_ext(const std::string& path) {
  const auto p = path.find_last_of('.');
  if (p == std::string::npos) throw std::runtime_error("Unknown dataset format: (no extension)");
  return path.substr(p);
}

std::string strip_path(const std::string& path) {
  return path.substr(path.find_last_of("/\\") + 1);
}

namespace {

inline bool is_space(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

// One unsigned extraction the way `istream >> unsigned` behaves in libstdc++:
// optional sign, decimal digits; a '-' negates modulo 2^32; overflow saturates.
// Returns false when no digit was found.
inline bool parse_u32(const char*& p, const char* end, uint32_t& out, bool& overflow) {
  bool neg = false;
  overflow = false;
  if (p < end && (*p == '+' || *p == '-')) {
    neg = (*p == '-');
    ++p;
  }
  if (p >= end || *p < '0' || *p > '9') return false;
  uint64_t v = 0;
  while (p < end && *p >= '0' && *p <= '9') {
    if (!overflow) {
      v = v * 10 + static_cast<uint64_t>(*p - '0');
      if (v > std::numeric_limits<uint32_t>::max()) overflow = true;
    }
    ++p;
  }
  if (overflow) {
    out = std::numeric_limits<uint32_t>::max();
  } else {
    out = static_cast<uint32_t>(neg ? (0u - static_cast<uint32_t>(v)) : static_cast<uint32_t>(v));
  }
  return true;
}

unsigned long stoul_like(const std::string& tok) {
  return std::stoul(tok, nullptr, 0);  // same call as graph.hpp:328-335
}

}  // namespace

namespace {

// Edges of one slice of the input, one (src << 32 | dst) word per kept edge.
struct EdgeSlice {
  std::vector<uint64_t> pairs;
  uint32_t max_node = 0;
  uint64_t lines = 0;
};

void scan_edge_lines(const char* b, const char* end, EdgeSlice& out) {
  out.pairs.reserve(static_cast<size_t>(end - b) / 12);
  while (b < end) {
    const void* nl = std::memchr(b, '\n', static_cast<size_t>(end - b));
    const char* le = nl ? static_cast<const char*>(nl) : end;
    const char* q = b;
    b = nl ? le + 1 : end;
    ++out.lines;
    while (q < le && is_space(*q)) ++q;
    uint32_t s = 0, d = 0;
    bool ovf = false;
    // A line without a leading number, or without a second token, leaves the
    // reference's destination uninitialised; such lines are skipped here.
    if (!parse_u32(q, le, s, ovf) || ovf) continue;
    while (q < le && is_space(*q)) ++q;
    if (q >= le) continue;
    if (!parse_u32(q, le, d, ovf)) d = 0;  // failed conversion stores 0
    if (s == d) continue;
    out.max_node = std::max(out.max_node, std::max(s, d));
    out.pairs.push_back((static_cast<uint64_t>(s) << 32) | d);
  }
}

// Undirected CSR over max_node + 1 vertices from edge slices. Both
// directions of every edge become (src << 32 | dst) keys, partitioned by
// src >> shift into at most 4096 buckets of consecutive sources (per-slice
// histograms, so the scatter needs no atomics); each bucket is small enough
// to sort and de-duplicate in cache, and the buckets' vertex ranges are
// disjoint, so offsets and targets are written bucket by bucket. The result
// equals sorting and de-duplicating all keys at once.
Csr csr_from_slices(std::vector<EdgeSlice>& slices, uint32_t max_node) {
  if (max_node == std::numeric_limits<uint32_t>::max())
    throw std::runtime_error("Vertices id are larger than defined type.");
  const int64_t ns = static_cast<int64_t>(slices.size());
  const uint64_t n = static_cast<uint64_t>(max_node) + 1;
  int shift = 0;
  while ((n - 1) >> shift >= 4096) ++shift;
  const int64_t nb = static_cast<int64_t>(((n - 1) >> shift) + 1);
  // at[j * (nb + 1) + b]: first key slot of slice j in bucket b.
  std::vector<uint64_t> at(static_cast<size_t>(ns) * (nb + 1), 0);
#pragma omp parallel for schedule(static, 1)
  for (int64_t j = 0; j < ns; ++j) {
    uint64_t* h = at.data() + j * (nb + 1);
    for (uint64_t e : slices[j].pairs) {
      ++h[(e >> 32) >> shift];
      ++h[(e & 0xffffffffu) >> shift];
    }
  }
  std::vector<uint64_t> bucket(static_cast<size_t>(nb) + 1, 0);
  for (int64_t b = 0; b < nb; ++b) {
    uint64_t pos = bucket[b];
    for (int64_t j = 0; j < ns; ++j) {
      const uint64_t c = at[j * (nb + 1) + b];
      at[j * (nb + 1) + b] = pos;
      pos += c;
    }
    bucket[b + 1] = pos;
  }
  std::vector<uint64_t> keys(bucket[nb]);
#pragma omp parallel for schedule(static, 1)
  for (int64_t j = 0; j < ns; ++j) {
    uint64_t* cur = at.data() + j * (nb + 1);
    for (uint64_t e : slices[j].pairs) {
      const uint64_t s = e >> 32, d = e & 0xffffffffu;
      keys[cur[s >> shift]++] = e;
      keys[cur[d >> shift]++] = (d << 32) | s;
    }
    std::vector<uint64_t>().swap(slices[j].pairs);
  }
  std::vector<uint64_t> kept(static_cast<size_t>(nb) + 1, 0);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t b = 0; b < nb; ++b) {
    uint64_t* lo = keys.data() + bucket[b];
    uint64_t* hi = keys.data() + bucket[b + 1];
    std::sort(lo, hi);
    kept[b + 1] = static_cast<uint64_t>(std::unique(lo, hi) - lo);
  }
  for (int64_t b = 0; b < nb; ++b) kept[b + 1] += kept[b];
  if (kept[nb] > std::numeric_limits<uint32_t>::max())
    throw std::runtime_error("Edges size are larger than defined type.");
  Csr g;
  g.n = static_cast<uint32_t>(n);
  g.O.resize(n + 1);
  g.E.resize(kept[nb]);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t b = 0; b < nb; ++b) {
    const uint64_t* k = keys.data() + bucket[b];
    const uint64_t cnt = kept[b + 1] - kept[b];
    const uint64_t v0 = static_cast<uint64_t>(b) << shift;
    const uint64_t v1 = std::min(n, v0 + (uint64_t(1) << shift));
    uint64_t i = 0;
    for (uint64_t v = v0; v < v1; ++v) {
      g.O[v] = static_cast<uint32_t>(kept[b] + i);
      while (i < cnt && (k[i] >> 32) == v) g.E[kept[b] + i] = static_cast<uint32_t>(k[i]), ++i;
    }
  }
  g.O[n] = static_cast<uint32_t>(kept[nb]);
  return g;
}

}  // namespace

LoadedGraph parse_edge_text(const char* data, size_t len, bool mtx_header) {
  const char* p = data;
  const char* end = data + len;
  auto next_line = [&](const char*& b, const char*& e) -> bool {
    if (p >= end) return false;
    b = p;
    const void* nl = std::memchr(p, '\n', static_cast<size_t>(end - p));
    e = nl ? static_cast<const char*>(nl) : end;
    p = nl ? e + 1 : end;
    return true;
  };
  // Leading comment lines (first char '#' or '%'); the first other line is the
  // mtx header or, for bare lists, dropped (graph.hpp:307-314).
  const char* lb = nullptr;
  const char* le = nullptr;
  std::string header;
  while (next_line(lb, le)) {
    header.assign(lb, le);
    if (lb == le || (*lb != '#' && *lb != '%')) break;
  }
  if (mtx_header) {
    std::istringstream iss(header);
    std::vector<std::string> tok{std::istream_iterator<std::string>{iss},
                                 std::istream_iterator<std::string>{}};
    uint64_t nV = 0, nE = 0;
    if (tok.size() == 2) {
      nV = stoul_like(tok[0]);
      nE = stoul_like(tok[1]);
    } else if (tok.size() == 3) {
      nV = std::max(stoul_like(tok[0]), stoul_like(tok[1]));
      nE = stoul_like(tok[2]);
    } else {
      throw std::runtime_error("Wrong header for mtx format!");
    }
    if (nE > std::numeric_limits<uint32_t>::max())
      throw std::runtime_error("Edges size are larger than defined type.");
    if (nV > std::numeric_limits<uint32_t>::max())
      throw std::runtime_error("Vertices id are larger than defined type.");
  }
  // The body is cut into slices that each start right after a newline, one
  // per thread (slices of at least 1 MiB), and scanned independently; a line
  // is wholly inside one slice, so the counts add up to a serial scan's.
  const size_t body = static_cast<size_t>(end - p);
  const size_t min_slice = size_t(1) << 20;
  const int threads = std::max(1, std::min<int>(omp_get_max_threads(),
                                                static_cast<int>(body / min_slice)));
  std::vector<const char*> cut(static_cast<size_t>(threads) + 1, end);
  cut[0] = p;
  for (int j = 1; j < threads; ++j) {
    const char* at = std::max(cut[j - 1], p + body / threads * j);
    const void* nl = at < end ? std::memchr(at, '\n', static_cast<size_t>(end - at)) : nullptr;
    cut[j] = nl ? static_cast<const char*>(nl) + 1 : end;
  }
  std::vector<EdgeSlice> slices(static_cast<size_t>(threads));
#pragma omp parallel for schedule(static, 1) num_threads(threads)
  for (int j = 0; j < threads; ++j) scan_edge_lines(cut[j], cut[j + 1], slices[j]);

  LoadedGraph out;
  uint32_t max_node = 0;
  for (const EdgeSlice& s : slices) {
    out.lines += s.lines;
    max_node = std::max(max_node, s.max_node);
  }
  out.g = csr_from_slices(slices, max_node);
  return out;
}

namespace {

struct MappedFile {
  int fd = -1;
  void* ptr = nullptr;
  size_t len = 0;
  explicit MappedFile(const std::string& path) {
    fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) throw std::runtime_error("Dataset " + path + " is not accessible!");
    struct stat st {};
    if (fstat(fd, &st) != 0) {
      ::close(fd);
      throw std::runtime_error("Dataset " + path + " is not accessible!");
    }
    len = static_cast<size_t>(st.st_size);
    if (len) {
      ptr = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
      if (ptr == MAP_FAILED) {
        ::close(fd);
        throw std::runtime_error("Dataset " + path + " is not accessible!");
      }
    }
  }
  ~MappedFile() {
    if (ptr && ptr != MAP_FAILED) munmap(ptr, len);
    if (fd >= 0) ::close(fd);
  }
};

LoadedGraph read_grh(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) throw std::runtime_error("Can't open file for reading.");
  uint64_t sv = 0, nV = 0, nE = 0;
  f.read(reinterpret_cast<char*>(&sv), 8);
  if (sv != sizeof(uint32_t)) throw std::runtime_error("Incompatible types in dataset.");
  f.read(reinterpret_cast<char*>(&nV), 8);
  f.read(reinterpret_cast<char*>(&nE), 8);
  if (!f) throw std::runtime_error("Truncated .grh header.");
  if (nV > std::numeric_limits<uint32_t>::max() || nE > std::numeric_limits<uint32_t>::max())
    throw std::runtime_error("Edges size are larger than defined type.");
  // the body must be in the file before anything is sized from the header
  f.seekg(0, std::ios::end);
  const uint64_t file_bytes = static_cast<uint64_t>(f.tellg());
  if (file_bytes < 24 + 4 * (nV + nE)) throw std::runtime_error("Truncated .grh body.");
  f.seekg(24, std::ios::beg);
  LoadedGraph out;
  out.oriented = true;
  Csr& g = out.g;
  g.n = static_cast<uint32_t>(nV);
  std::vector<uint32_t> D(nV);
  g.E.resize(nE);
  f.read(reinterpret_cast<char*>(D.data()), static_cast<std::streamsize>(nV * 4));
  f.read(reinterpret_cast<char*>(g.E.data()), static_cast<std::streamsize>(nE * 4));
  if (!f) throw std::runtime_error("Truncated .grh body.");
  g.O.assign(nV + 1, 0);
  uint64_t run = 0;  // in 64 bits: degrees that wrap a 32-bit sum are rejected, not folded
  for (size_t v = 0; v < nV; ++v) {
    run += D[v];
    if (run > nE) throw std::runtime_error("Degree sum of .grh does not match its edge count.");
    g.O[v + 1] = static_cast<uint32_t>(run);
  }
  if (run != nE) throw std::runtime_error("Degree sum of .grh does not match its edge count.");
  out.lines = nE;
  return out;
}

}  // namespace

LoadedGraph load_dataset(const std::string& path) {
  const std::string ext = file_ext(path);
  if (ext == ".grh") return read_grh(path);
  if (ext != ".mtx" && ext != ".edges" && ext != ".txt")
    throw std::runtime_error("Unknown dataset format: " + ext);
  MappedFile mf(path);
  return parse_edge_text(static_cast<const char*>(mf.ptr), mf.len, ext == ".mtx");
}

Csr csr_from_pairs(const uint32_t* s, const uint32_t* d, size_t count) {
  const size_t per = size_t(1) << 20;
  const size_t parts = std::max<size_t>(1, std::min<size_t>(omp_get_max_threads(), count / per));
  std::vector<EdgeSlice> slices(parts);
#pragma omp parallel for schedule(static, 1)
  for (int64_t j = 0; j < static_cast<int64_t>(parts); ++j) {
    EdgeSlice& sl = slices[j];
    const size_t lo = count * j / parts, hi = count * (j + 1) / parts;
    sl.pairs.reserve(hi - lo);
    for (size_t i = lo; i < hi; ++i) {
      if (s[i] == d[i]) continue;
      sl.max_node = std::max(sl.max_node, std::max(s[i], d[i]));
      sl.pairs.push_back((static_cast<uint64_t>(s[i]) << 32) | d[i]);
    }
  }
  uint32_t max_node = 0;
  for (const EdgeSlice& sl : slices) max_node = std::max(max_node, sl.max_node);
  return csr_from_slices(slices, max_node);
}

Csr orient_host(const Csr& und) {
  Csr o;
  o.n = und.n;
  o.O.assign(static_cast<size_t>(und.n) + 1, 0);
  o.E.reserve(und.E.size() / 2);
  auto deg = [&](uint32_t v) { return und.O[v + 1] - und.O[v]; };
  for (uint32_t u = 0; u < und.n; ++u) {
    const uint32_t du = deg(u);
    for (uint32_t i = und.O[u]; i < und.O[u + 1]; ++i) {
      const uint32_t v = und.E[i];
      const uint32_t dv = deg(v);
      if (du < dv || (du == dv && u < v)) o.E.push_back(v);
    }
    o.O[u + 1] = static_cast<uint32_t>(o.E.size());
  }
  return o;
}

void write_grh(const std::string& path, const Csr& g) {
  std::ofstream f(path, std::ios::binary);
  if (!f.is_open()) throw std::runtime_error("Can't open file for serializing.");
  uint64_t sv = 4, nV = g.n, nE = g.E.size();
  f.write(reinterpret_cast<const char*>(&sv), 8);
  f.write(reinterpret_cast<const char*>(&nV), 8);
  f.write(reinterpret_cast<const char*>(&nE), 8);
  std::vector<uint32_t> D(g.n);
  for (uint32_t v = 0; v < g.n; ++v) D[v] = g.O[v + 1] - g.O[v];
  f.write(reinterpret_cast<const char*>(D.data()), static_cast<std::streamsize>(D.size() * 4));
  f.write(reinterpret_cast<const char*>(g.E.data()), static_cast<std::streamsize>(g.E.size() * 4));
}

std::vector<uint64_t> split_parts(const Csr& g, size_t parts) {
  if (parts == 0) parts = 1;
  // Segment boundaries over vertices: boundary j is the first vertex whose
  // edge prefix O[v] exceeds j/parts of all edges; unused boundaries stay at n-1.
  std::vector<uint64_t> sp(parts, g.n ? g.n - 1 : 0);
  sp[0] = 0;
  const double scale = 1.0 / static_cast<double>(parts);
  const double total = static_cast<double>(g.O.empty() ? 0 : g.O[g.n > 0 ? g.n - 1 : 0]);
  for (size_t i = 0, j = 1; i < g.n && j < parts; ++i) {
    if (static_cast<double>(g.O[i]) / total > static_cast<double>(j) * scale) sp[j++] = i;
  }
  return sp;
}

std::string check_oriented(const Csr& g) {
  if (g.O.size() != static_cast<size_t>(g.n) + 1) return "offset array has wrong length";
  if (g.O[0] != 0 || g.O[g.n] != g.E.size()) return "offsets do not span the edge array";
  for (uint32_t u = 0; u < g.n; ++u) {
    if (g.O[u] > g.O[u + 1]) return "offsets not monotone at vertex " + std::to_string(u);
    for (uint32_t i = g.O[u]; i < g.O[u + 1]; ++i) {
      const uint32_t v = g.E[i];
      if (v >= g.n) return "vertex id out of range in list of " + std::to_string(u);
      if (v == u) return "self-loop at vertex " + std::to_string(u);
      if (i > g.O[u] && g.E[i - 1] >= v) return "list of vertex " + std::to_string(u) + " not strictly increasing";
      // reciprocal pair: v's list must not hold u
      const uint32_t* b = g.E.data() + g.O[v];
      const uint32_t* e = g.E.data() + g.O[v + 1];
      if (std::binary_search(b, e, u)) return "reciprocal edge between " + std::to_string(u) + " and " + std::to_string(v);
    }
  }
  return std::string();
}

}  // namespace nuc34

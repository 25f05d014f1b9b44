// Writes a raw edge list as a bare ".edges" file: a neutral first line "0 0"
// (the loader drops the first non-comment line of bare lists, and a self-loop
// is ignored either way), then one "u v" line per pair.
#pragma once
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "gen.hpp"

namespace nuc34 {

inline void write_edges_file(const std::string& path, const gen::Edges& e) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot open " + path + " for writing");
  std::vector<char> buf;
  buf.reserve(1 << 22);
  auto put_u32 = [&](uint32_t x) {
    char tmp[12];
    int n = 0;
    do {
      tmp[n++] = static_cast<char>('0' + x % 10);
      x /= 10;
    } while (x);
    while (n) buf.push_back(tmp[--n]);
  };
  buf.push_back('0');
  buf.push_back(' ');
  buf.push_back('0');
  buf.push_back('\n');
  for (const auto& p : e) {
    put_u32(p.first);
    buf.push_back(' ');
    put_u32(p.second);
    buf.push_back('\n');
    if (buf.size() > (1u << 22) - 32) {
      std::fwrite(buf.data(), 1, buf.size(), f);
      buf.clear();
    }
  }
  std::fwrite(buf.data(), 1, buf.size(), f);
  std::fclose(f);
}

}  // namespace nuc34

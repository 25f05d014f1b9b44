// Shared pieces of the two command-line front ends: environment facts for the
// JSON record (CPU model and cores from /proc/cpuinfo, affinity, host name,
// device properties), the k-file writer and the "(a,b,c)" tuple format
// (nucleus/utils.cuh:138-149, nucleus/utility/cpuinfo.hpp:43-160).
#pragma once
#include <sched.h>
#include <unistd.h>

#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/nuc34.h"

namespace nuc34cli {

inline std::string env_or_empty(const char* k) {
  const char* v = std::getenv(k);
  return v ? std::string(v) : std::string();
}

inline std::string hostname() {
  char buf[256] = {0};
  if (gethostname(buf, sizeof(buf) - 1) != 0) return "";
  return buf;
}

// One map per logical CPU from /proc/cpuinfo.
inline std::vector<std::map<std::string, std::string>> cpuinfo() {
  std::vector<std::map<std::string, std::string>> out;
  std::ifstream f("/proc/cpuinfo");
  std::string line;
  std::map<std::string, std::string> cur;
  while (std::getline(f, line)) {
    if (line.empty()) {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
      continue;
    }
    const auto p = line.find(':');
    if (p == std::string::npos) continue;
    std::string k = line.substr(0, p), v = line.substr(p + 1);
    while (!k.empty() && (k.back() == ' ' || k.back() == '\t')) k.pop_back();
    while (!v.empty() && v.front() == ' ') v.erase(v.begin());
    cur[k] = v;
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

inline std::string cpu_affinity() {
  cpu_set_t set;
  CPU_ZERO(&set);
  if (sched_getaffinity(0, sizeof(set), &set) != 0) return "";
  std::string s;
  for (int i = 0; i < CPU_SETSIZE; ++i)
    if (CPU_ISSET(i, &set)) s += (s.empty() ? "" : ",") + std::to_string(i);
  return s;
}

inline std::vector<std::pair<std::string, std::string>> device_props(int dev) {
  std::vector<std::pair<std::string, std::string>> out;
  char buf[4096];
  if (nuc34_device_info(dev, buf, sizeof(buf)) != NUC34_OK) return out;
  std::istringstream is(buf);
  std::string line;
  while (std::getline(is, line)) {
    const auto p = line.find('=');
    if (p != std::string::npos) out.emplace_back(line.substr(0, p), line.substr(p + 1));
  }
  return out;
}

// "[mem] <tag> <detail> | gpu_free=<free>/<total> MiB" on stderr: the device
// memory milestones of the reference's CLI (nucleus/cuNucleus34.cu:72-79).
inline void mem_note(const char* tag, const std::string& detail) {
  unsigned long long free_b = 0, total_b = 0;
  for (const auto& p : device_props(0)) {
    if (p.first == "free_memory_bytes") free_b = std::stoull(p.second);
    if (p.first == "global_memory_bytes") total_b = std::stoull(p.second);
  }
  std::fprintf(stderr, "[mem] %s %s | gpu_free=%llu/%llu MiB\n", tag, detail.c_str(), free_b >> 20, total_b >> 20);
  std::fflush(stderr);
}

inline std::string tuple3(const uint32_t* t) {
  std::ostringstream os;
  os << '(' << t[0] << ',' << t[1] << ',' << t[2] << ')';
  return os.str();
}

inline std::string tuple4(const uint32_t* t) {
  std::ostringstream os;
  os << '(' << t[0] << ',' << t[1] << ',' << t[2] << ',' << t[3] << ')';
  return os.str();
}

// "(a,b,c)\tk\n" per triangle, (a,b) order.
inline void write_kfile(const std::string& path, const std::vector<uint32_t>& abc, const std::vector<int32_t>& k) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("Failed to open output file: " + path);
  std::vector<char> buf(1 << 20);
  size_t used = 0;
  for (size_t i = 0; i < k.size(); ++i) {
    if (used + 64 > buf.size()) {
      std::fwrite(buf.data(), 1, used, f);
      used = 0;
    }
    used += static_cast<size_t>(std::snprintf(buf.data() + used, 64, "(%u,%u,%u)\t%d\n", abc[3 * i],
                                              abc[3 * i + 1], abc[3 * i + 2], k[i]));
  }
  std::fwrite(buf.data(), 1, used, f);
  std::fclose(f);
}

inline std::string smi_dump() {
  std::string out;
  FILE* p = popen("rocm-smi 2>&1", "r");
  if (!p) return out;
  char buf[512];
  while (std::fgets(buf, sizeof(buf), p)) out += buf;
  pclose(p);
  return out;
}

}  // namespace nuc34cli

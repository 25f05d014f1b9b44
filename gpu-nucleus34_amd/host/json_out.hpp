// JSON record writer with the reference's layout: a two-level object whose
// members open with "{"/"," + "\n\t\"key\":" and whose fields are written as
// "{"/"," + "\n\t\t\"key\": value", closed by "\n\t}" and "\n}"
// (nucleus/utility/json.h:40-92). Numbers use default ostream formatting.
#pragma once
#include <cstdio>
#include <sstream>
#include <string>

namespace nuc34 {

class JsonOut {
 public:
  static std::string escape(const std::string& s) {
    std::string out;
    out.reserve(s.size() + 2);
    out += '"';
    for (unsigned char c : s) {
      switch (c) {
        case '"': out += "\\\""; break;
        case '\\': out += "\\\\"; break;
        case '\b': out += "\\b"; break;
        case '\f': out += "\\f"; break;
        case '\n': out += "\\n"; break;
        case '\r': out += "\\r"; break;
        case '\t': out += "\\t"; break;
        default:
          if (c < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof(buf), "\\u%04x", c);
            out += buf;
          } else {
            out += static_cast<char>(c);
          }
      }
    }
    out += '"';
    return out;
  }
  void begin_object(bool first, const std::string& key) {
    ss_ << (first ? "{" : ",") << "\n\t\"" << key << "\":";
    first_field_ = true;
  }
  void end_object() { ss_ << "\n\t}"; }
  void last_object() { ss_ << "\n}"; }
  template <typename T>
  void field(const std::string& key, const T& v) {
    open(key);
    ss_ << v;
  }
  void field(const std::string& key, const std::string& v) {
    open(key);
    ss_ << escape(v);
  }
  void field(const std::string& key, const char* v) { field(key, std::string(v ? v : "")); }
  void field(const std::string& key, bool v) {
    open(key);
    ss_ << (v ? "true" : "false");
  }
  void raw(const std::string& key, const std::string& v) {
    open(key);
    ss_ << v;
  }
  template <typename It, typename Fmt>
  void array(const std::string& key, It b, It e, Fmt fmt) {
    open(key);
    ss_ << "[";
    bool first = true;
    for (auto it = b; it != e; ++it) {
      if (!first) ss_ << ", ";
      fmt(ss_, *it);
      first = false;
    }
    ss_ << "]";
  }
  bool empty() const { return ss_.str().empty(); }
  std::string str() const { return ss_.str(); }

 private:
  void open(const std::string& key) {
    ss_ << (first_field_ ? "{" : ",") << "\n\t\t\"" << key << "\": ";
    first_field_ = false;
  }
  std::ostringstream ss_;
  bool first_field_ = true;
};

}  // namespace nuc34

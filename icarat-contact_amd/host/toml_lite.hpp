// Minimal reader for the TOML subset the reference's job files use
// (opt-mma-HelmholtzFilter/2d.toml, 3d.toml, fdm.toml; opt-macro-oc/init.toml):
// [section] headers, key = value with numbers, strings, booleans and nested
// arrays, '#' comments.  Values are kept as a small variant tree.
#pragma once

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace icarat {
namespace toml_lite {

struct Value {
  enum Kind { NUM, STR, BOOL, ARR } kind = NUM;
  double num = 0.0;
  std::string str;
  bool b = false;
  std::vector<Value> arr;

  double as_double() const {
    if (kind != NUM) throw std::runtime_error("toml: number expected");
    return num;
  }
  int as_int() const {
    const double v = as_double();
    if (!(v >= -2147483648.0 && v <= 2147483647.0)) throw std::runtime_error("toml: integer out of range");
    return (int)v;
  }
  const std::string &as_string() const {
    if (kind != STR) throw std::runtime_error("toml: string expected");
    return str;
  }
  bool as_bool() const {
    if (kind != BOOL) throw std::runtime_error("toml: boolean expected");
    return b;
  }
  const std::vector<Value> &as_array() const {
    if (kind != ARR) throw std::runtime_error("toml: array expected");
    return arr;
  }
};

class Table {
 public:
  const Value &at(const std::string &section, const std::string &key) const {
    auto s = data_.find(section);
    if (s == data_.end()) throw std::runtime_error("toml: missing section [" + section + "]");
    auto k = s->second.find(key);
    if (k == s->second.end()) throw std::runtime_error("toml: missing key " + section + "." + key);
    return k->second;
  }
  bool has(const std::string &section, const std::string &key) const {
    auto s = data_.find(section);
    return s != data_.end() && s->second.count(key);
  }
  std::map<std::string, std::map<std::string, Value>> data_;
};

namespace detail {
struct Parser {
  const std::string &s;
  size_t i = 0;
  explicit Parser(const std::string &src) : s(src) {}
  void skip_ws() {
    for (;;) {
      while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
      if (i < s.size() && s[i] == '#') {
        while (i < s.size() && s[i] != '\n') ++i;
        continue;
      }
      break;
    }
  }
  // the next character after white space and comments; throws at the end of the text
  char peek() {
    skip_ws();
    if (i >= s.size()) throw std::runtime_error("toml: unexpected end");
    return s[i];
  }
  Value value(int depth = 0) {
    if (depth > 32) throw std::runtime_error("toml: arrays nested too deeply");
    const char c0 = peek();
    Value v;
    if (c0 == '[') {
      v.kind = Value::ARR;
      ++i;
      for (;;) {
        if (peek() == ']') { ++i; break; }
        v.arr.push_back(value(depth + 1));
        const char c = peek();
        if (c == ',') { ++i; continue; }
        if (c == ']') { ++i; break; }
        throw std::runtime_error("toml: ',' or ']' expected");
      }
    } else if (c0 == '"') {
      v.kind = Value::STR;
      const size_t j = s.find('"', i + 1);
      if (j == std::string::npos) throw std::runtime_error("toml: unterminated string");
      v.str = s.substr(i + 1, j - i - 1);
      i = j + 1;
    } else if (s.compare(i, 4, "true") == 0) {
      v.kind = Value::BOOL; v.b = true; i += 4;
    } else if (s.compare(i, 5, "false") == 0) {
      v.kind = Value::BOOL; v.b = false; i += 5;
    } else {
      const char *b = s.c_str() + i;
      char *e = nullptr;
      v.num = std::strtod(b, &e);
      if (e == b) throw std::runtime_error("toml: bad value near " + s.substr(i, 16));
      i += (size_t)(e - b);
    }
    return v;
  }
};
}  // namespace detail

inline Table parse_string(const std::string &src) {
  Table t;
  detail::Parser p(src);
  std::string section;
  for (;;) {
    p.skip_ws();
    if (p.i >= src.size()) break;
    if (src[p.i] == '[') {
      const size_t j = src.find(']', p.i);
      if (j == std::string::npos) throw std::runtime_error("toml: unterminated section header");
      section = src.substr(p.i + 1, j - p.i - 1);
      if (section.find('\n') != std::string::npos) throw std::runtime_error("toml: bad section header");
      p.i = j + 1;
      continue;
    }
    const size_t eq = src.find('=', p.i);
    if (eq == std::string::npos) throw std::runtime_error("toml: key = value expected");
    std::string key = src.substr(p.i, eq - p.i);
    if (key.find('\n') != std::string::npos) throw std::runtime_error("toml: key = value expected near " + key);
    while (!key.empty() && std::isspace((unsigned char)key.back())) key.pop_back();
    p.i = eq + 1;
    t.data_[section][key] = p.value();
  }
  return t;
}

inline Table parse(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("toml: cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse_string(ss.str());
}

}  // namespace toml_lite
}  // namespace icarat

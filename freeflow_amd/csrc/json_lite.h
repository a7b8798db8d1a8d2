// json_lite.h -- a small JSON reader for the simulator configuration files.
//
// Parses RFC 8259 JSON (objects, arrays, strings with escapes, numbers,
// true/false/null) into a tree of Value nodes.  Only what the configuration
// schema needs is provided: typed getters with defaults, like the
// `j.value(key, default)` lookups of the reference's Config.hpp.
#pragma once

#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace ffjson {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;

  bool is_object() const { return kind == Object; }
  bool is_array() const { return kind == Array; }
  bool is_number() const { return kind == Number; }
  bool is_string() const { return kind == String; }
  bool contains(const std::string& k) const { return kind == Object && obj.count(k) != 0; }
  const Value& at(const std::string& k) const {
    auto it = obj.find(k);
    if (kind != Object || it == obj.end()) throw std::runtime_error("missing key '" + k + "'");
    return it->second;
  }
  double number_or(const std::string& k, double d) const {
    if (!contains(k)) return d;
    const Value& v = at(k);
    if (v.kind == Number) return v.num;
    if (v.kind == Bool) return v.b ? 1.0 : 0.0;
    throw std::runtime_error("key '" + k + "' is not a number");
  }
  int int_or(const std::string& k, int d) const { return (int)number_or(k, d); }
  bool bool_or(const std::string& k, bool d) const {
    if (!contains(k)) return d;
    const Value& v = at(k);
    if (v.kind == Bool) return v.b;
    if (v.kind == Number) return v.num != 0.0;
    throw std::runtime_error("key '" + k + "' is not a boolean");
  }
  std::string string_or(const std::string& k, const std::string& d) const {
    if (!contains(k)) return d;
    const Value& v = at(k);
    if (v.kind != String) throw std::runtime_error("key '" + k + "' is not a string");
    return v.str;
  }
  std::vector<double> numbers_or(const std::string& k, const std::vector<double>& d) const {
    if (!contains(k)) return d;
    const Value& v = at(k);
    if (v.kind != Array) throw std::runtime_error("key '" + k + "' is not an array");
    std::vector<double> out;
    for (const Value& e : v.arr) {
      if (e.kind != Number) throw std::runtime_error("key '" + k + "' holds a non-number");
      out.push_back(e.num);
    }
    return out;
  }
};

class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s) {}
  Value parse() {
    Value v = value();
    ws();
    if (p_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;

  [[noreturn]] void fail(const std::string& what) {
    size_t line = 1;
    for (size_t i = 0; i < p_ && i < s_.size(); i++) line += s_[i] == '\n';
    throw std::runtime_error("JSON parse error at line " + std::to_string(line) + ": " + what);
  }
  void ws() {
    while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) p_++;
  }
  char peek() {
    ws();
    if (p_ >= s_.size()) fail("unexpected end of input");
    return s_[p_];
  }
  void expect(char c) {
    if (peek() != c) fail(std::string("expected '") + c + "'");
    p_++;
  }
  bool literal(const char* w) {
    size_t n = std::char_traits<char>::length(w);
    if (s_.compare(p_, n, w) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  Value value() {
    char c = peek();
    Value v;
    if (c == '{') {
      p_++;
      v.kind = Value::Object;
      if (peek() == '}') {
        p_++;
        return v;
      }
      for (;;) {
        if (peek() != '"') fail("expected a key string");
        std::string k = string();
        expect(':');
        v.obj[k] = value();
        char d = peek();
        p_++;
        if (d == '}') break;
        if (d != ',') fail("expected ',' or '}'");
      }
    } else if (c == '[') {
      p_++;
      v.kind = Value::Array;
      if (peek() == ']') {
        p_++;
        return v;
      }
      for (;;) {
        v.arr.push_back(value());
        char d = peek();
        p_++;
        if (d == ']') break;
        if (d != ',') fail("expected ',' or ']'");
      }
    } else if (c == '"') {
      v.kind = Value::String;
      v.str = string();
    } else if (literal("true")) {
      v.kind = Value::Bool;
      v.b = true;
    } else if (literal("false")) {
      v.kind = Value::Bool;
      v.b = false;
    } else if (literal("null")) {
      v.kind = Value::Null;
    } else if (c == '-' || (c >= '0' && c <= '9')) {
      const char* start = s_.c_str() + p_;
      char* end = nullptr;
      v.kind = Value::Number;
      v.num = std::strtod(start, &end);
      if (end == start) fail("bad number");
      p_ += (size_t)(end - start);
    } else {
      fail(std::string("unexpected character '") + c + "'");
    }
    return v;
  }
  static void utf8(std::string& out, unsigned cp) {
    if (cp < 0x80) {
      out += (char)cp;
    } else if (cp < 0x800) {
      out += (char)(0xC0 | (cp >> 6));
      out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += (char)(0xE0 | (cp >> 12));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    } else {
      out += (char)(0xF0 | (cp >> 18));
      out += (char)(0x80 | ((cp >> 12) & 0x3F));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    }
  }
  unsigned hex4() {
    if (p_ + 4 > s_.size()) fail("short \\u escape");
    unsigned v = 0;
    for (int i = 0; i < 4; i++) {
      char h = s_[p_++];
      v <<= 4;
      if (h >= '0' && h <= '9') v |= (unsigned)(h - '0');
      else if (h >= 'a' && h <= 'f') v |= (unsigned)(h - 'a' + 10);
      else if (h >= 'A' && h <= 'F') v |= (unsigned)(h - 'A' + 10);
      else fail("bad \\u escape");
    }
    return v;
  }
  std::string string() {
    expect('"');
    std::string out;
    while (p_ < s_.size() && s_[p_] != '"') {
      char c = s_[p_++];
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p_ >= s_.size()) fail("bad escape");
      char e = s_[p_++];
      switch (e) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          unsigned cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00 && p_ + 6 <= s_.size() && s_[p_] == '\\' && s_[p_ + 1] == 'u') {
            p_ += 2;
            unsigned lo = hex4();
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(out, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
    if (p_ >= s_.size()) fail("unterminated string");
    p_++;
    return out;
  }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace ffjson

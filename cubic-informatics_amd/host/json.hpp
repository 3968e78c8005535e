// json.hpp -- a small JSON reader for the HDoG parameter files.
//
// Covers what the parameter schema uses: objects (later duplicate keys win, which is
// how the "#" comment keys of param/param_example_HDoG_FW.json behave), arrays,
// strings with escapes, numbers, true / false / null.  Errors throw JsonError with
// the byte offset.
#pragma once
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace hjson {

struct JsonError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;

  bool has(const std::string& k) const { return kind == Object && obj.count(k) > 0; }
  const Value& at(const std::string& k) const {
    if (kind != Object) throw JsonError("not an object while looking up key '" + k + "'");
    auto it = obj.find(k);
    if (it == obj.end()) throw JsonError("missing key '" + k + "'");
    return it->second;
  }
  double number() const {
    if (kind != Number) throw JsonError("expected a number");
    return num;
  }
  int integer() const { return (int)number(); }
  bool boolean() const {
    if (kind == Bool) return b;
    if (kind == Number) return num != 0.0;
    throw JsonError("expected a boolean");
  }
  const std::string& string() const {
    if (kind != String) throw JsonError("expected a string");
    return str;
  }
  const std::vector<Value>& array() const {
    if (kind != Array) throw JsonError("expected an array");
    return arr;
  }
};

class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s) {}
  Value parse() {
    Value v = value();
    ws();
    if (i_ != s_.size()) err("trailing characters");
    return v;
  }

 private:
  const std::string& s_;
  size_t i_ = 0;

  [[noreturn]] void err(const char* what) {
    throw JsonError(std::string("JSON parse error at byte ") + std::to_string(i_) + ": " + what);
  }
  void ws() {
    while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r'))
      ++i_;
  }
  bool lit(const char* w) {
    size_t n = 0;
    while (w[n]) ++n;
    if (s_.compare(i_, n, w) == 0) { i_ += n; return true; }
    return false;
  }
  Value value() {
    ws();
    if (i_ >= s_.size()) err("unexpected end");
    const char c = s_[i_];
    Value v;
    if (c == '{') {
      v.kind = Value::Object;
      ++i_;
      ws();
      if (i_ < s_.size() && s_[i_] == '}') { ++i_; return v; }
      while (true) {
        ws();
        if (i_ >= s_.size() || s_[i_] != '"') err("expected a key");
        std::string k = str();
        ws();
        if (i_ >= s_.size() || s_[i_] != ':') err("expected ':'");
        ++i_;
        v.obj[k] = value();
        ws();
        if (i_ < s_.size() && s_[i_] == ',') { ++i_; continue; }
        if (i_ < s_.size() && s_[i_] == '}') { ++i_; return v; }
        err("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.kind = Value::Array;
      ++i_;
      ws();
      if (i_ < s_.size() && s_[i_] == ']') { ++i_; return v; }
      while (true) {
        v.arr.push_back(value());
        ws();
        if (i_ < s_.size() && s_[i_] == ',') { ++i_; continue; }
        if (i_ < s_.size() && s_[i_] == ']') { ++i_; return v; }
        err("expected ',' or ']'");
      }
    }
    if (c == '"') { v.kind = Value::String; v.str = str(); return v; }
    if (lit("true")) { v.kind = Value::Bool; v.b = true; return v; }
    if (lit("false")) { v.kind = Value::Bool; v.b = false; return v; }
    if (lit("null")) return v;
    const char* p = s_.c_str() + i_;
    char* end = nullptr;
    v.num = std::strtod(p, &end);
    if (end == p) err("unexpected character");
    v.kind = Value::Number;
    i_ += (size_t)(end - p);
    return v;
  }
  static void utf8(std::string& o, unsigned cp) {
    if (cp < 0x80) o += (char)cp;
    else if (cp < 0x800) { o += (char)(0xC0 | (cp >> 6)); o += (char)(0x80 | (cp & 0x3F)); }
    else if (cp < 0x10000) {
      o += (char)(0xE0 | (cp >> 12)); o += (char)(0x80 | ((cp >> 6) & 0x3F));
      o += (char)(0x80 | (cp & 0x3F));
    } else {
      o += (char)(0xF0 | (cp >> 18)); o += (char)(0x80 | ((cp >> 12) & 0x3F));
      o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F));
    }
  }
  unsigned hex4() {
    if (i_ + 4 > s_.size()) err("short \\u escape");
    unsigned v = 0;
    for (int k = 0; k < 4; ++k) {
      const char h = s_[i_++];
      v <<= 4;
      if (h >= '0' && h <= '9') v |= h - '0';
      else if (h >= 'a' && h <= 'f') v |= h - 'a' + 10;
      else if (h >= 'A' && h <= 'F') v |= h - 'A' + 10;
      else err("bad \\u escape");
    }
    return v;
  }
  std::string str() {
    std::string o;
    ++i_;  // opening quote
    while (true) {
      if (i_ >= s_.size()) err("unterminated string");
      const char c = s_[i_++];
      if (c == '"') return o;
      if (c != '\\') { o += c; continue; }
      if (i_ >= s_.size()) err("unterminated escape");
      const char e = s_[i_++];
      switch (e) {
        case '"': o += '"'; break;
        case '\\': o += '\\'; break;
        case '/': o += '/'; break;
        case 'b': o += '\b'; break;
        case 'f': o += '\f'; break;
        case 'n': o += '\n'; break;
        case 'r': o += '\r'; break;
        case 't': o += '\t'; break;
        case 'u': {
          unsigned cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00 && i_ + 6 <= s_.size() && s_[i_] == '\\' &&
              s_[i_ + 1] == 'u') {
            i_ += 2;
            const unsigned lo = hex4();
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(o, cp);
          break;
        }
        default: err("bad escape");
      }
    }
  }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace hjson

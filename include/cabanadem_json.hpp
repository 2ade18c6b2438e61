// cabanadem_json.hpp -- a small JSON reader for the solver input files.
//
// Covers what the input files of the reference use (objects, numbers,
// strings, booleans, null, arrays) with implicit conversion of a value to
// double / int / std::string / bool, so that
//   double rho = inputs["particle_density"];
// reads like the reference drivers (they read through nlohmann::json).
// Errors throw cabanadem_json::Error (parse errors, wrong type, missing key).
#pragma once

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace cabanadem_json {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Value {
 public:
  enum class Kind { Null, Bool, Number, String, Array, Object };

  Value() = default;
  Value(double v) : kind_(Kind::Number), num_(v) {}
  Value(int v) : kind_(Kind::Number), num_(v) {}
  Value(bool v) : kind_(Kind::Bool), b_(v) {}
  Value(std::string s) : kind_(Kind::String), str_(std::move(s)) {}

  static Value object() { Value v; v.kind_ = Kind::Object; return v; }
  static Value array() { Value v; v.kind_ = Kind::Array; return v; }

  Kind kind() const { return kind_; }
  bool is_null() const { return kind_ == Kind::Null; }
  bool is_number() const { return kind_ == Kind::Number; }
  bool is_object() const { return kind_ == Kind::Object; }

  // Object access; a missing key reads as null (and throws on conversion).
  Value& operator[](const std::string& key) {
    if (kind_ == Kind::Null) kind_ = Kind::Object;
    if (kind_ != Kind::Object) throw Error("not an object");
    return obj_[key];
  }
  const Value& operator[](const std::string& key) const {
    static const Value null_value;
    if (kind_ != Kind::Object) throw Error("not an object");
    auto it = obj_.find(key);
    return it == obj_.end() ? null_value : it->second;
  }
  // exact matches for string literals (otherwise the implicit numeric
  // conversions would make `v["key"]` ambiguous with built-in indexing)
  Value& operator[](const char* key) { return (*this)[std::string(key)]; }
  const Value& operator[](const char* key) const { return (*this)[std::string(key)]; }
  const Value& at(size_t i) const {
    if (kind_ != Kind::Array || i >= arr_.size()) throw Error("array index out of range");
    return arr_[i];
  }
  size_t size() const {
    return kind_ == Kind::Array ? arr_.size() : (kind_ == Kind::Object ? obj_.size() : 0);
  }
  bool contains(const std::string& key) const {
    return kind_ == Kind::Object && obj_.count(key) != 0;
  }
  void push_back(Value v) {
    if (kind_ != Kind::Array) throw Error("not an array");
    arr_.push_back(std::move(v));
  }

  double as_double() const {
    if (kind_ != Kind::Number) throw Error("value is not a number");
    return num_;
  }
  int as_int() const { return (int)as_double(); } // truncation, as an int conversion
  bool as_bool() const {
    if (kind_ != Kind::Bool) throw Error("value is not a boolean");
    return b_;
  }
  const std::string& as_string() const {
    if (kind_ != Kind::String) throw Error("value is not a string");
    return str_;
  }
  operator double() const { return as_double(); }
  operator float() const { return (float)as_double(); }
  operator int() const { return as_int(); }
  operator long() const { return (long)as_double(); }
  operator bool() const { return as_bool(); }
  operator std::string() const { return as_string(); }

  Value& operator=(double v) { *this = Value(v); return *this; }
  Value& operator=(int v) { *this = Value(v); return *this; }

 private:
  Kind kind_ = Kind::Null;
  double num_ = 0.;
  bool b_ = false;
  std::string str_;
  std::vector<Value> arr_;
  std::map<std::string, Value> obj_;
};

namespace detail {

class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s) {}
  Value parse_document() {
    Value v = parse_value();
    skip_ws();
    if (p_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;

  [[noreturn]] void fail(const char* what) {
    std::ostringstream os;
    os << "JSON parse error at offset " << p_ << ": " << what;
    throw Error(os.str());
  }
  void skip_ws() {
    while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
  }
  char peek() {
    skip_ws();
    if (p_ >= s_.size()) fail("unexpected end");
    return s_[p_];
  }
  void expect(char c) {
    if (peek() != c) fail("unexpected character");
    ++p_;
  }
  Value parse_value() {
    char c = peek();
    if (c == '{') return parse_object();
    if (c == '[') return parse_array();
    if (c == '"') return Value(parse_string());
    if (s_.compare(p_, 4, "true") == 0) { p_ += 4; return Value(true); }
    if (s_.compare(p_, 5, "false") == 0) { p_ += 5; return Value(false); }
    if (s_.compare(p_, 4, "null") == 0) { p_ += 4; return Value(); }
    return parse_number();
  }
  Value parse_object() {
    expect('{');
    Value v = Value::object();
    if (peek() == '}') { ++p_; return v; }
    for (;;) {
      std::string key = parse_string();
      expect(':');
      v[key] = parse_value();
      char c = peek();
      ++p_;
      if (c == '}') return v;
      if (c != ',') fail("expected , or }");
    }
  }
  Value parse_array() {
    expect('[');
    Value v = Value::array();
    if (peek() == ']') { ++p_; return v; }
    for (;;) {
      v.push_back(parse_value());
      char c = peek();
      ++p_;
      if (c == ']') return v;
      if (c != ',') fail("expected , or ]");
    }
  }
  std::string parse_string() {
    if (peek() != '"') fail("expected string");
    ++p_;
    std::string out;
    while (p_ < s_.size() && s_[p_] != '"') {
      char c = s_[p_++];
      if (c == '\\') {
        if (p_ >= s_.size()) fail("bad escape");
        char e = s_[p_++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (p_ + 4 > s_.size()) fail("bad unicode escape");
            unsigned code = std::stoul(s_.substr(p_, 4), nullptr, 16);
            p_ += 4;
            if (code < 0x80) out += (char)code;
            else if (code < 0x800) { out += (char)(0xC0 | (code >> 6)); out += (char)(0x80 | (code & 0x3F)); }
            else { out += (char)(0xE0 | (code >> 12)); out += (char)(0x80 | ((code >> 6) & 0x3F)); out += (char)(0x80 | (code & 0x3F)); }
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (p_ >= s_.size()) fail("unterminated string");
    ++p_;
    return out;
  }
  Value parse_number() {
    skip_ws();
    const char* begin = s_.c_str() + p_;
    char* end = nullptr;
    double v = std::strtod(begin, &end);
    if (end == begin) fail("expected a value");
    p_ += (size_t)(end - begin);
    return Value(v);
  }
};

} // namespace detail

inline Value parse(const std::string& text) { return detail::Parser(text).parse_document(); }

inline Value parse_file(const std::string& filename) {
  std::ifstream in(filename);
  if (!in) throw Error("cannot open " + filename);
  std::stringstream ss;
  ss << in.rdbuf();
  return parse(ss.str());
}

} // namespace cabanadem_json

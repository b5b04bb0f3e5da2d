// Minimal JSON value for node settings and tensor descriptors (the reference carries its
// per-node settings as nlohmann::json, tasks.hh:260-334; that library is not part of this
// build, so this small self-contained type stands in with the subset the nodes use:
// objects, arrays, numbers, strings, booleans, null; parse and dump).
#pragma once

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace holoflow::core {

class Json {
 public:
  enum class Type { Null, Bool, Number, String, Array, Object };

  Json() = default;
  Json(std::nullptr_t) {}
  Json(bool b) : type_(Type::Bool), b_(b) {}
  Json(int v) : type_(Type::Number), n_(v) {}
  Json(long v) : type_(Type::Number), n_((double)v) {}
  Json(long long v) : type_(Type::Number), n_((double)v) {}
  Json(unsigned v) : type_(Type::Number), n_(v) {}
  Json(unsigned long v) : type_(Type::Number), n_((double)v) {}
  Json(unsigned long long v) : type_(Type::Number), n_((double)v) {}
  Json(double v) : type_(Type::Number), n_(v) {}
  Json(float v) : type_(Type::Number), n_(v) {}
  Json(const char *s) : type_(Type::String), s_(s) {}
  Json(std::string s) : type_(Type::String), s_(std::move(s)) {}
  template <typename T> Json(const std::vector<T> &v) : type_(Type::Array) {
    for (const auto &x : v) a_.emplace_back(x);
  }

  static Json object() {
    Json j;
    j.type_ = Type::Object;
    return j;
  }
  static Json array() {
    Json j;
    j.type_ = Type::Array;
    return j;
  }

  Type type() const { return type_; }
  bool is_null() const { return type_ == Type::Null; }
  bool is_object() const { return type_ == Type::Object; }
  bool is_array() const { return type_ == Type::Array; }
  bool is_number() const { return type_ == Type::Number; }
  bool is_string() const { return type_ == Type::String; }
  bool is_bool() const { return type_ == Type::Bool; }

  bool contains(const std::string &k) const { return type_ == Type::Object && o_.count(k) != 0; }
  const Json &at(const std::string &k) const {
    if (!contains(k)) throw std::invalid_argument("missing JSON key '" + k + "'");
    return o_.at(k);
  }
  Json &operator[](const std::string &k) {
    if (type_ == Type::Null) type_ = Type::Object;
    if (type_ != Type::Object) throw std::invalid_argument("JSON value is not an object");
    return o_[k];
  }
  const Json &operator[](size_t i) const {
    if (type_ != Type::Array || i >= a_.size()) throw std::invalid_argument("JSON index out of range");
    return a_[i];
  }
  void push_back(Json v) {
    if (type_ == Type::Null) type_ = Type::Array;
    if (type_ != Type::Array) throw std::invalid_argument("JSON value is not an array");
    a_.push_back(std::move(v));
  }
  size_t size() const { return type_ == Type::Array ? a_.size() : type_ == Type::Object ? o_.size() : 0; }
  const std::vector<Json> &items() const { return a_; }
  const std::map<std::string, Json> &members() const { return o_; }

  double number() const {
    if (type_ == Type::Bool) return b_ ? 1.0 : 0.0;
    if (type_ != Type::Number) throw std::invalid_argument("JSON value is not a number");
    return n_;
  }
  long long integer() const {
    const double v = number();
    if (v != std::floor(v)) throw std::invalid_argument("JSON number is not an integer");
    return (long long)v;
  }
  bool boolean() const {
    if (type_ != Type::Bool) throw std::invalid_argument("JSON value is not a boolean");
    return b_;
  }
  const std::string &str() const {
    if (type_ != Type::String) throw std::invalid_argument("JSON value is not a string");
    return s_;
  }

  // typed accessors with a default when the key is absent
  double get(const std::string &k, double d) const { return contains(k) ? at(k).number() : d; }
  long long get_int(const std::string &k, long long d) const { return contains(k) ? at(k).integer() : d; }
  bool get_bool(const std::string &k, bool d) const { return contains(k) ? at(k).boolean() : d; }

  bool operator==(const Json &o) const {
    if (type_ != o.type_) return false;
    switch (type_) {
      case Type::Null: return true;
      case Type::Bool: return b_ == o.b_;
      case Type::Number: return n_ == o.n_;
      case Type::String: return s_ == o.s_;
      case Type::Array: return a_ == o.a_;
      case Type::Object: return o_ == o.o_;
    }
    return false;
  }
  bool operator!=(const Json &o) const { return !(*this == o); }

  std::string dump() const {
    std::string out;
    dump_to(out);
    return out;
  }

  static Json parse(const std::string &text) {
    size_t i = 0;
    Json v = parse_value(text, i);
    skip_ws(text, i);
    if (i != text.size()) throw std::invalid_argument("trailing characters in JSON");
    return v;
  }

 private:
  Type type_ = Type::Null;
  bool b_ = false;
  double n_ = 0.0;
  std::string s_;
  std::vector<Json> a_;
  std::map<std::string, Json> o_;

  void dump_to(std::string &out) const {
    switch (type_) {
      case Type::Null: out += "null"; break;
      case Type::Bool: out += b_ ? "true" : "false"; break;
      case Type::Number: {
        char buf[64];
        if (n_ == std::floor(n_) && std::fabs(n_) < 9.0e15)
          std::snprintf(buf, sizeof buf, "%lld", (long long)n_);
        else
          std::snprintf(buf, sizeof buf, "%.17g", n_);
        out += buf;
        break;
      }
      case Type::String: dump_string(out, s_); break;
      case Type::Array: {
        out += '[';
        for (size_t i = 0; i < a_.size(); ++i) {
          if (i) out += ',';
          a_[i].dump_to(out);
        }
        out += ']';
        break;
      }
      case Type::Object: {
        out += '{';
        bool first = true;
        for (const auto &kv : o_) {
          if (!first) out += ',';
          first = false;
          dump_string(out, kv.first);
          out += ':';
          kv.second.dump_to(out);
        }
        out += '}';
        break;
      }
    }
  }
  static void dump_string(std::string &out, const std::string &s) {
    out += '"';
    for (char c : s) {
      if (c == '"' || c == '\\') {
        out += '\\';
        out += c;
      } else if ((unsigned char)c < 0x20) {
        char buf[8];
        std::snprintf(buf, sizeof buf, "\\u%04x", c);
        out += buf;
      } else {
        out += c;
      }
    }
    out += '"';
  }
  static void skip_ws(const std::string &t, size_t &i) {
    while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\t' || t[i] == '\r')) ++i;
  }
  static Json parse_value(const std::string &t, size_t &i) {
    skip_ws(t, i);
    if (i >= t.size()) throw std::invalid_argument("unexpected end of JSON");
    const char c = t[i];
    if (c == '{') {
      Json o = object();
      ++i;
      skip_ws(t, i);
      if (i < t.size() && t[i] == '}') {
        ++i;
        return o;
      }
      while (true) {
        skip_ws(t, i);
        if (i >= t.size() || t[i] != '"') throw std::invalid_argument("expected a JSON key");
        std::string k = parse_string(t, i);
        skip_ws(t, i);
        if (i >= t.size() || t[i] != ':') throw std::invalid_argument("expected ':' in JSON");
        ++i;
        o.o_[k] = parse_value(t, i);
        skip_ws(t, i);
        if (i < t.size() && t[i] == ',') {
          ++i;
          continue;
        }
        if (i < t.size() && t[i] == '}') {
          ++i;
          return o;
        }
        throw std::invalid_argument("expected ',' or '}' in JSON");
      }
    }
    if (c == '[') {
      Json a = array();
      ++i;
      skip_ws(t, i);
      if (i < t.size() && t[i] == ']') {
        ++i;
        return a;
      }
      while (true) {
        a.a_.push_back(parse_value(t, i));
        skip_ws(t, i);
        if (i < t.size() && t[i] == ',') {
          ++i;
          continue;
        }
        if (i < t.size() && t[i] == ']') {
          ++i;
          return a;
        }
        throw std::invalid_argument("expected ',' or ']' in JSON");
      }
    }
    if (c == '"') return Json(parse_string(t, i));
    if (t.compare(i, 4, "true") == 0) {
      i += 4;
      return Json(true);
    }
    if (t.compare(i, 5, "false") == 0) {
      i += 5;
      return Json(false);
    }
    if (t.compare(i, 4, "null") == 0) {
      i += 4;
      return Json();
    }
    size_t end = i;
    while (end < t.size() && (std::isdigit((unsigned char)t[end]) || t[end] == '-' || t[end] == '+' ||
                              t[end] == '.' || t[end] == 'e' || t[end] == 'E'))
      ++end;
    if (end == i) throw std::invalid_argument("invalid JSON value");
    const std::string num = t.substr(i, end - i);
    char *stop = nullptr;
    const double v = std::strtod(num.c_str(), &stop);
    if (stop != num.c_str() + num.size()) throw std::invalid_argument("invalid JSON number");
    i = end;
    return Json(v);
  }
  static std::string parse_string(const std::string &t, size_t &i) {
    std::string out;
    ++i;  // opening quote
    while (i < t.size() && t[i] != '"') {
      if (t[i] == '\\' && i + 1 < t.size()) {
        const char e = t[i + 1];
        i += 2;
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'u': {
            if (i + 4 > t.size()) throw std::invalid_argument("bad JSON escape");
            out += (char)std::stoi(t.substr(i, 4), nullptr, 16);
            i += 4;
            break;
          }
          default: out += e;
        }
      } else {
        out += t[i++];
      }
    }
    if (i >= t.size()) throw std::invalid_argument("unterminated JSON string");
    ++i;
    return out;
  }
};

}  // namespace holoflow::core

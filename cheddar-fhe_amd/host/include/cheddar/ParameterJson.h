// Loads the reference's parameter JSON files (parameters/bootparam_*.json,
// the schema the reference's test harness reads, unittest/Testbed.h:96-170)
// into a Parameter<uint32_t>: log_degree, log_default_scale (base scale =
// 2^log_default_scale), default_encryption_level, main_primes,
// terminal_primes (optional), auxiliary_primes, level_config. A small
// recursive-descent reader of our own: objects, arrays, numbers, strings,
// true/false/null.
#pragma once

#include <cctype>
#include <cstdint>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "cheddar/Parameter.h"

namespace cheddar {

namespace json_detail {

struct Value {
  enum Kind { kNull, kBool, kNumber, kString, kArray, kObject } kind = kNull;
  double num = 0.0;
  long long inum = 0;
  bool is_int = false;
  bool b = false;
  std::string str;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;
};

class Reader {
 public:
  explicit Reader(const std::string &s) : s_(s) {}
  Value Parse() {
    Value v = ParseValue();
    Skip();
    if (i_ != s_.size()) Fail("trailing characters");
    return v;
  }

 private:
  void Fail(const std::string &what) {
    throw Error("parameter JSON: " + what + " at offset " + std::to_string(i_));
  }
  void Skip() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) i_++;
  }
  bool Eat(char c) {
    Skip();
    if (i_ < s_.size() && s_[i_] == c) {
      i_++;
      return true;
    }
    return false;
  }
  void Expect(char c) {
    if (!Eat(c)) Fail(std::string("expected '") + c + "'");
  }
  std::string ParseString() {
    Expect('"');
    std::string out;
    while (i_ < s_.size() && s_[i_] != '"') {
      if (s_[i_] == '\\' && i_ + 1 < s_.size()) i_++;
      out += s_[i_++];
    }
    Expect('"');
    return out;
  }
  Value ParseValue() {
    Skip();
    if (i_ >= s_.size()) Fail("unexpected end");
    Value v;
    const char c = s_[i_];
    if (c == '{') {
      v.kind = Value::kObject;
      i_++;
      if (Eat('}')) return v;
      do {
        Skip();
        std::string k = ParseString();
        Expect(':');
        v.obj[k] = ParseValue();
      } while (Eat(','));
      Expect('}');
    } else if (c == '[') {
      v.kind = Value::kArray;
      i_++;
      if (Eat(']')) return v;
      do v.arr.push_back(ParseValue());
      while (Eat(','));
      Expect(']');
    } else if (c == '"') {
      v.kind = Value::kString;
      v.str = ParseString();
    } else if (s_.compare(i_, 4, "true") == 0) {
      v.kind = Value::kBool;
      v.b = true;
      i_ += 4;
    } else if (s_.compare(i_, 5, "false") == 0) {
      v.kind = Value::kBool;
      i_ += 5;
    } else if (s_.compare(i_, 4, "null") == 0) {
      i_ += 4;
    } else {
      size_t j = i_;
      while (j < s_.size() && (std::isdigit((unsigned char)s_[j]) || s_[j] == '-' ||
                               s_[j] == '+' || s_[j] == '.' || s_[j] == 'e' ||
                               s_[j] == 'E'))
        j++;
      if (j == i_) Fail("unexpected character");
      const std::string t = s_.substr(i_, j - i_);
      v.kind = Value::kNumber;
      v.num = std::stod(t);
      v.is_int = t.find_first_of(".eE") == std::string::npos;
      if (v.is_int) v.inum = std::stoll(t);
      i_ = j;
    }
    return v;
  }
  const std::string &s_;
  size_t i_ = 0;
};

inline const Value &Field(const Value &o, const char *key) {
  auto it = o.obj.find(key);
  if (it == o.obj.end()) throw Error(std::string("Missing ") + key + " in JSON file");
  return it->second;
}

inline long long Int(const Value &v, const char *what) {
  if (v.kind != Value::kNumber || !v.is_int)
    throw Error(std::string(what) + " should be an integer");
  return v.inum;
}

inline std::vector<uint32_t> Primes(const Value &o, const char *key, bool required) {
  std::vector<uint32_t> out;
  auto it = o.obj.find(key);
  if (it == o.obj.end()) {
    if (required) throw Error(std::string("Missing ") + key + " in JSON file");
    return out;
  }
  if (it->second.kind != Value::kArray)
    throw Error(std::string(key) + " should be an array");
  for (const Value &p : it->second.arr) {
    const long long x = Int(p, key);
    if (x <= 0 || x >= (1ll << 32))
      throw Error(std::string(key) + " holds a value outside 32 bits");
    out.push_back((uint32_t)x);
  }
  return out;
}

}  // namespace json_detail

inline std::unique_ptr<Parameter<uint32_t>> ParameterFromJsonText(
    const std::string &text) {
  using namespace json_detail;
  const Value root = Reader(text).Parse();
  if (root.kind != Value::kObject) throw Error("parameter JSON is not an object");
  const int log_degree = (int)Int(Field(root, "log_degree"), "log_degree");
  const int log_scale = (int)Int(Field(root, "log_default_scale"), "log_default_scale");
  const int enc_level =
      (int)Int(Field(root, "default_encryption_level"), "default_encryption_level");
  const Value &lc = Field(root, "level_config");
  if (lc.kind != Value::kArray) throw Error("level_config should be an array");
  std::vector<std::pair<int, int>> levels;
  for (const Value &e : lc.arr) {
    if (e.kind != Value::kArray || e.arr.size() != 2)
      throw Error("level_config entries should be [num_main, num_ter] pairs");
    levels.emplace_back((int)Int(e.arr[0], "level_config"),
                        (int)Int(e.arr[1], "level_config"));
  }
  std::pair<int, int> additional_base{0, 0};
  auto ab = root.obj.find("additional_base");
  if (ab != root.obj.end()) {
    if (ab->second.kind != Value::kArray || ab->second.arr.size() != 2)
      throw Error("additional_base should be an array of pairs");
    additional_base = {(int)Int(ab->second.arr[0], "additional_base"),
                       (int)Int(ab->second.arr[1], "additional_base")};
  }
  auto param = std::make_unique<Parameter<uint32_t>>(
      log_degree, std::ldexp(1.0, log_scale), enc_level, levels,
      Primes(root, "main_primes", true), Primes(root, "auxiliary_primes", true),
      Primes(root, "terminal_primes", false), additional_base);
  auto dh = root.obj.find("dense_hamming_weight");
  if (dh != root.obj.end())
    param->SetDenseHammingWeight((int)Int(dh->second, "dense_hamming_weight"));
  auto sh = root.obj.find("sparse_hamming_weight");
  if (sh != root.obj.end())
    param->SetSparseHammingWeight((int)Int(sh->second, "sparse_hamming_weight"));
  return param;
}

// Bootstrapping fields of the same JSON ("boot", "num_cts_levels",
// "num_stc_levels", read as the reference's test bed does,
// unittest/Testbed.h): returns false when "boot" is absent or false.
inline bool BootLevelsFromJsonText(const std::string &text, int *num_cts_levels,
                                   int *num_stc_levels) {
  using namespace json_detail;
  const Value root = Reader(text).Parse();
  auto b = root.obj.find("boot");
  if (b == root.obj.end() || b->second.kind != Value::kBool || !b->second.b)
    return false;
  *num_cts_levels = (int)Int(Field(root, "num_cts_levels"), "num_cts_levels");
  *num_stc_levels = (int)Int(Field(root, "num_stc_levels"), "num_stc_levels");
  return true;
}

inline std::unique_ptr<Parameter<uint32_t>> ParameterFromJsonFile(
    const std::string &path) {
  std::ifstream f(path);
  if (!f.is_open()) throw Error("Failed to open JSON file: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ParameterFromJsonText(ss.str());
}

}  // namespace cheddar

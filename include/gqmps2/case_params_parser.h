// Simulation case parameters read from a JSON file (GraceQ/MPS2 0.2-alpha.0,
// include/gqmps2/case_params_parser.h:27-101, consts.h:26).
//
// The reference reads the file with a vendored JSON library; this header has a
// small reader of its own (RFC 8259 values: objects, arrays, strings with
// escapes, numbers, true/false/null).  Behaviour kept from the reference: the
// parameters are the members of the top-level "CaseParams" object, whose
// absence prints a message and exits with status 1; ParseChar returns the
// first character of a string value; an integer parameter may be written as
// any JSON number (converted as the reference's library converts it, by
// truncation).  A missing key or a value of another JSON type throws
// std::runtime_error naming the key (the reference's library throws its own
// exception type there), as does a file that is not valid JSON.
#ifndef MPS2_AMD_GQMPS2_CASE_PARAMS_PARSER_H
#define MPS2_AMD_GQMPS2_CASE_PARAMS_PARSER_H

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <locale>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace gqmps2 {

const std::string kCaseParamsJsonObjName = "CaseParams";

namespace json_detail {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0;
  std::string str;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;
};

class Reader {
 public:
  explicit Reader(const std::string &text) : s_(text) {}
  Value Document() {
    Value v = ParseValue();
    SkipWs();
    if (p_ != s_.size()) Fail("trailing characters");
    return v;
  }

 private:
  const std::string &s_;
  size_t p_ = 0;

  [[noreturn]] void Fail(const std::string &what) const {
    throw std::runtime_error("JSON parse error at offset " + std::to_string(p_) + ": " + what);
  }
  void SkipWs() {
    while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) ++p_;
  }
  char Peek() {
    SkipWs();
    if (p_ >= s_.size()) Fail("unexpected end of input");
    return s_[p_];
  }
  void Expect(char c) {
    if (Peek() != c) Fail(std::string("expected '") + c + "'");
    ++p_;
  }
  bool Literal(const char *w) {
    size_t n = std::char_traits<char>::length(w);
    if (s_.compare(p_, n, w) != 0) return false;
    p_ += n;
    return true;
  }
  static void PutUtf8(std::string &out, unsigned cp) {
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
  unsigned Hex4() {
    if (p_ + 4 > s_.size()) Fail("short \\u escape");
    unsigned v = 0;
    for (int i = 0; i < 4; ++i) {
      char c = s_[p_++];
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
      else Fail("bad \\u escape");
    }
    return v;
  }
  std::string ParseString() {
    Expect('"');
    std::string out;
    while (true) {
      if (p_ >= s_.size()) Fail("unterminated string");
      char c = s_[p_++];
      if (c == '"') return out;
      if ((unsigned char)c < 0x20) Fail("control character in string");
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p_ >= s_.size()) Fail("unterminated escape");
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
          unsigned cp = Hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {  // surrogate pair
            if (!(Literal("\\u"))) Fail("unpaired surrogate");
            unsigned lo = Hex4();
            if (lo < 0xDC00 || lo >= 0xE000) Fail("unpaired surrogate");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          PutUtf8(out, cp);
          break;
        }
        default: Fail("bad escape");
      }
    }
  }
  double ParseNumber() {
    size_t b = p_;
    if (p_ < s_.size() && s_[p_] == '-') ++p_;
    auto digits = [&] {
      size_t d = p_;
      while (p_ < s_.size() && s_[p_] >= '0' && s_[p_] <= '9') ++p_;
      return p_ - d;
    };
    if (p_ < s_.size() && s_[p_] == '0') ++p_;
    else if (digits() == 0) Fail("bad number");
    if (p_ < s_.size() && s_[p_] == '.') {
      ++p_;
      if (digits() == 0) Fail("bad fraction");
    }
    if (p_ < s_.size() && (s_[p_] == 'e' || s_[p_] == 'E')) {
      ++p_;
      if (p_ < s_.size() && (s_[p_] == '+' || s_[p_] == '-')) ++p_;
      if (digits() == 0) Fail("bad exponent");
    }
    std::istringstream is(s_.substr(b, p_ - b));
    is.imbue(std::locale::classic());
    double v = 0;
    is >> v;
    return v;
  }
  Value ParseValue() {
    Value v;
    char c = Peek();
    if (c == '{') {
      ++p_;
      v.kind = Value::Object;
      if (Peek() == '}') {
        ++p_;
        return v;
      }
      while (true) {
        std::string k = ParseString();
        Expect(':');
        v.obj[k] = ParseValue();
        if (Peek() == ',') {
          ++p_;
          continue;
        }
        Expect('}');
        return v;
      }
    }
    if (c == '[') {
      ++p_;
      v.kind = Value::Array;
      if (Peek() == ']') {
        ++p_;
        return v;
      }
      while (true) {
        v.arr.push_back(ParseValue());
        if (Peek() == ',') {
          ++p_;
          continue;
        }
        Expect(']');
        return v;
      }
    }
    if (c == '"') {
      v.kind = Value::String;
      v.str = ParseString();
      return v;
    }
    if (Literal("true")) {
      v.kind = Value::Bool;
      v.b = true;
      return v;
    }
    if (Literal("false")) {
      v.kind = Value::Bool;
      return v;
    }
    if (Literal("null")) return v;
    v.kind = Value::Number;
    v.num = ParseNumber();
    return v;
  }
};

}  // namespace json_detail

class CaseParamsParserBasic {
 public:
  explicit CaseParamsParserBasic(const char *file) {
    std::ifstream ifs(file);
    if (!ifs) throw std::runtime_error(std::string("cannot open ") + file);
    std::stringstream ss;
    ss << ifs.rdbuf();
    std::string text = ss.str();
    json_detail::Value doc = json_detail::Reader(text).Document();
    auto it = doc.obj.find(kCaseParamsJsonObjName);
    if (doc.kind != json_detail::Value::Object || it == doc.obj.end()) {
      std::cout << kCaseParamsJsonObjName << " object not found, exit!" << std::endl;
      exit(1);
    }
    case_params_ = it->second;
  }

  int ParseInt(const std::string &item) { return (int)Get(item, json_detail::Value::Number, "number").num; }
  double ParseDouble(const std::string &item) { return Get(item, json_detail::Value::Number, "number").num; }
  char ParseChar(const std::string &item) { return Get(item, json_detail::Value::String, "string").str.at(0); }
  std::string ParseStr(const std::string &item) { return Get(item, json_detail::Value::String, "string").str; }
  bool ParseBool(const std::string &item) { return Get(item, json_detail::Value::Bool, "boolean").b; }

 private:
  json_detail::Value case_params_;

  const json_detail::Value &Get(const std::string &item, json_detail::Value::Kind kind, const char *kind_name) const {
    auto it = case_params_.obj.find(item);
    if (it == case_params_.obj.end())
      throw std::runtime_error(kCaseParamsJsonObjName + ": no parameter \"" + item + "\"");
    if (it->second.kind != kind)
      throw std::runtime_error(kCaseParamsJsonObjName + ": parameter \"" + item + "\" is not a " + kind_name);
    return it->second;
  }
};

}  // namespace gqmps2

#endif

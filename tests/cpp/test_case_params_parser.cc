// tests/test_case_params_parser.cc:17-53 (GraceQ/MPS2 0.2-alpha.0) against the
// gqmps2-compatible headers: a CaseParamsParserBasic subclass reads the
// reference's fixture (testdata/test-params.json, argv[1]); then strings with
// escapes, numbers with exponents, nested values around the parameters, the
// errors for a missing key, a wrong type and malformed JSON.  With argv[2] ==
// "missing" it parses a file without a CaseParams object (the caller expects
// exit status 1).  Runs on the CPU.  Exit status 0 = all cases pass.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>

#include "gqmps2/gqmps2.h"

using namespace gqmps2;

static int g_fail = 0, g_checks = 0;
static void Check(bool ok, const char *what) {
  ++g_checks;
  if (!ok) {
    ++g_fail;
    std::printf("FAIL %s\n", what);
  }
}

struct CustomCaseParamsParser : public CaseParamsParserBasic {
  CustomCaseParamsParser(const char *f) : CaseParamsParserBasic(f) {
    case_int = ParseInt("Int");
    case_double = ParseDouble("Double");
    case_char = ParseChar("Char");
    case_str = ParseStr("String");
    case_bool = ParseBool("Boolean");
  }
  int case_int;
  double case_double;
  char case_char;
  std::string case_str;
  bool case_bool;
};

template <typename F>
static bool Throws(F f) {
  try {
    f();
  } catch (const std::runtime_error &) {
    return true;
  }
  return false;
}

static void Write(const char *file, const std::string &text) { std::ofstream(file) << text; }

int main(int argc, char *argv[]) {
  if (argc < 2) {
    std::printf("usage: %s test-params.json [missing]\n", argv[0]);
    return 2;
  }
  if (argc > 2 && std::string(argv[2]) == "missing") {
    Write("no_case.json", "{\"Other\": {\"Int\": 1}}");
    CaseParamsParserBasic p("no_case.json");  // exits with status 1
    return 0;
  }
  {  // the reference's case (:39-46)
    CustomCaseParamsParser params(argv[1]);
    Check(params.case_int == 1, "Int");
    Check(params.case_double == 2.33, "Double");
    Check(params.case_char == 'c', "Char");
    Check(params.case_str == "string", "String");
    Check(params.case_bool == false, "Boolean");
  }
  {
    Write("edge.json",
          "{ \"Before\": [1, {\"x\": [true, null]}, \"]}\"],\n"
          "  \"CaseParams\" : {\"Neg\": -12, \"Exp\": -1.5E+2, \"Small\": 6.25e-3, \"Trunc\": 7.9,"
          " \"Esc\": \"a\\\"b\\\\c\\/d\\n\\u00e9\\ud83d\\ude00\", \"T\": true, \"Zero\": 0,"
          " \"Obj\": {\"Int\": 5}, \"Dup\": 1, \"Dup\": 2, \"Empty\": \"\"},\n"
          "  \"After\": {} }");
    CaseParamsParserBasic p("edge.json");
    Check(p.ParseInt("Neg") == -12, "negative integer");
    Check(p.ParseDouble("Exp") == -150.0, "exponent");
    Check(p.ParseDouble("Small") == 6.25e-3, "negative exponent");
    Check(p.ParseInt("Trunc") == 7, "integer from a fractional number truncates");
    Check(p.ParseStr("Esc") == "a\"b\\c/d\n\xc3\xa9\xf0\x9f\x98\x80", "string escapes and UTF-16 surrogates");
    Check(p.ParseBool("T"), "true");
    Check(p.ParseInt("Zero") == 0 && p.ParseDouble("Zero") == 0.0, "zero");
    Check(p.ParseInt("Dup") == 2, "a repeated key keeps the last value");
    Check(Throws([&] { p.ParseInt("Int"); }), "a key of a nested object is not a parameter");
    Check(Throws([&] { p.ParseInt("Missing"); }), "missing key throws");
    Check(Throws([&] { p.ParseStr("Neg"); }), "number read as a string throws");
    Check(Throws([&] { p.ParseBool("Zero"); }), "number read as a boolean throws");
    Check(Throws([&] { p.ParseDouble("Esc"); }), "string read as a number throws");
    bool oor = false;
    try {
      p.ParseChar("Empty");
    } catch (const std::out_of_range &) {
      oor = true;
    }
    Check(oor, "ParseChar of an empty string throws");
  }
  for (const char *bad : {"{\"CaseParams\": {\"a\": 1,}}", "{\"CaseParams\": {\"a\": 01}}",
                          "{\"CaseParams\": {\"a\": \"x}}", "{\"CaseParams\": {\"a\": 1}} x",
                          "{\"CaseParams\": {\"a\": .5}}", "{\"CaseParams\": {\"a\": tru}}"}) {
    Write("bad.json", bad);
    Check(Throws([] { CaseParamsParserBasic p("bad.json"); }), bad);
  }
  Check(Throws([] { CaseParamsParserBasic p("does_not_exist.json"); }), "missing file throws");
  std::printf("%s: %d checks, %d failures\n", g_fail ? "FAILED" : "PASSED", g_checks, g_fail);
  return g_fail ? 1 : 0;
}

// Checks the planner's hand-written literal matchers against std::regex with the
// patterns they replace, on every string up to length 6 over a small alphabet and on
// random longer strings (numbers and dates). Built and run by
// tests/test_literal_matchers.py; prints "ok" or the first string that differs.
#include <cstdio>
#include <random>
#include <regex>
#include <string>

#include "planner.cpp"

// The planner's type inference calls these (defined with the program compiler, which
// needs HIP); the matchers under test never reach them.
namespace blz {
int unary_output_dtype(int, int) { return 0; }
int binary_output_dtype(int, int, int) { return 0; }
}  // namespace blz

using blz::planner::match_date;
using blz::planner::match_float;
using blz::planner::match_int;

static const std::regex re_int{R"(^[-+]?[0-9]+$)"};
static const std::regex re_float{R"(^[-+]?[0-9]*\.?[0-9]+([eE][-+]?[0-9]+)?$)"};
static const std::regex re_date{R"([12]\d{3}-(0[1-9]|1[0-2])-(0[1-9]|[12]\d|3[01]))"};

static bool check(const std::string &t) {
  if (match_int(t) != std::regex_match(t, re_int)) { std::printf("int differs: '%s'\n", t.c_str()); return false; }
  if (match_float(t) != std::regex_match(t, re_float)) { std::printf("float differs: '%s'\n", t.c_str()); return false; }
  if (match_date(t) != std::regex_match(t, re_date)) { std::printf("date differs: '%s'\n", t.c_str()); return false; }
  return true;
}

int main() {
  const std::string alpha = "019+-.eEx";
  long n = 0;
  std::string t;
  // every string of length 0..6 over the alphabet
  for (int len = 0; len <= 6; len++) {
    std::vector<int> d(len, 0);
    while (true) {
      t.assign(len, ' ');
      for (int i = 0; i < len; i++) t[i] = alpha[d[i]];
      if (!check(t)) return 1;
      n++;
      int k = len - 1;
      while (k >= 0 && ++d[k] == (int)alpha.size()) d[k--] = 0;
      if (k < 0) break;
    }
  }
  std::mt19937_64 rng(12345);
  const std::string num = "0123456789+-.eE";
  for (int i = 0; i < 300000; i++) {  // longer number-like strings
    const int len = 1 + (int)(rng() % 14);
    t.assign(len, ' ');
    for (int k = 0; k < len; k++) t[k] = num[rng() % (rng() % 4 == 0 ? num.size() : 10)];
    if (!check(t)) return 1;
    n++;
  }
  const std::string dch = "0123456789-";
  for (int i = 0; i < 300000; i++) {  // date-shaped strings, near-misses included
    t = "YYYY-MM-DD";
    for (char &c : t) c = (c == '-' && rng() % 8) ? '-' : dch[rng() % dch.size()];
    if (rng() % 2) { t[0] = "12"[rng() % 2]; t[4] = '-'; t[7] = '-'; }
    if (rng() % 4 == 0) t += dch[rng() % dch.size()];
    if (!check(t)) return 1;
    n++;
  }
  std::printf("ok %ld strings\n", n);
  return 0;
}

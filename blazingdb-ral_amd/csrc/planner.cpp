/*
 * planner.cpp — Calcite expression text -> register-file plan.
 *
 *   clean_calcite_expression        CalciteExpressionParsing.cpp:705-754
 *   expand_if_logical_op            CalciteExpressionParsing.cpp:629-694
 *   get_expressions_from_expression_list  CalciteExpressionParsing.cpp:814-895
 *   find_closing_char               CalciteExpressionParsing.cpp:771-811
 *   token predicates, get_index     CalciteExpressionParsing.cpp:568-608, 272-276
 *   literal typing                  CalciteExpressionParsing.cpp:323-417
 *   get_output_type_expression      CalciteExpressionParsing.cpp:420-481
 *   add_expression_to_plan          LogicalFilter.cpp:44-267
 *   evaluate_expression (plan)      LogicalFilter.cpp:271-371
 *   parse_project_plan              CalciteInterpreter.cpp:123-309
 *   get_named_expression            CalciteInterpreter.cpp:424-436
 *
 * The blazingdb-io StringUtil helpers the reference calls (not part of the
 * reference tree) are plain replace-all, split outside single quotes (empty
 * pieces dropped) and first-match outside single quotes.
 * Plans the reference would emit as misaligned vectors (binary op on two
 * literals, unary op on a literal, unknown operator) and string literals
 * (NVCategory) are rejected with an error.
 * Date literals parse as UTC "YYYY-MM-DD[ HH:MM:SS]" (the reference uses mktime,
 * i.e. the host time zone, CalciteExpressionParsing.cpp:279-305).
 */
#include "planner.h"

#include <algorithm>
#include <cstdio>
#include <map>
#include <cctype>
#include <cstring>
#include <regex>
#include <stdexcept>

namespace blz {
namespace planner {
namespace {

void fail(const std::string &m) { throw std::runtime_error(m); }

void replace_all(std::string &s, const std::string &from, const std::string &to) {
  size_t pos = 0;
  while ((pos = s.find(from, pos)) != std::string::npos) {
    s.replace(pos, from.size(), to);
    pos += to.size();
  }
}

std::vector<bool> quoted_vector(const std::string &s) {
  std::vector<bool> q(s.size());
  bool in = false;
  for (size_t i = 0; i < s.size(); i++) {
    if (s[i] == '\'') { in = !in; q[i] = true; }
    else q[i] = in;
  }
  return q;
}

int find_first_not_in_quotes(const std::string &s, const std::string &needle, int start,
                             const std::vector<bool> &q) {
  size_t i = s.find(needle, (size_t)std::max(start, 0));
  while (i != std::string::npos) {
    if (!q[i]) return (int)i;
    i = s.find(needle, i + 1);
  }
  return -1;
}

std::vector<std::string> split_not_in_quotes(const std::string &s, const std::string &delim) {
  std::vector<std::string> out;
  std::string cur;
  bool in = false;
  size_t i = 0;
  while (i < s.size()) {
    if (s[i] == '\'') in = !in;
    if (!in && s.compare(i, delim.size(), delim) == 0) {
      out.push_back(cur);
      cur.clear();
      i += delim.size();
      continue;
    }
    cur.push_back(s[i]);
    i++;
  }
  out.push_back(cur);
  std::vector<std::string> r;
  for (auto &t : out)
    if (!t.empty()) r.push_back(t);
  return r;
}

std::string ltrim(std::string s) {
  size_t p = s.find_first_not_of(" \t\n\r");
  return p == std::string::npos ? std::string() : s.substr(p);
}
std::string trim(std::string s) {
  s = ltrim(s);
  size_t p = s.find_last_not_of(" \t\n\r");
  return p == std::string::npos ? std::string() : s.substr(0, p + 1);
}

const std::regex &coalesce_re() {
  static const std::regex re{R"(CASE\(IS NOT NULL\((\W\(.+?\)|.+)\), \1, (\W\(.+?\)|.+)\))",
                             std::regex_constants::icase};
  return re;
}
const std::regex &count_re() {
  static const std::regex re{R"(COUNT\(DISTINCT (\W\(.+?\)|.+)\))", std::regex_constants::icase};
  return re;
}

// The regular expressions below run on every query and token and std::regex is
// slow (a C1 filter took 0.23 ms to plan); each is tried only when a text the
// pattern needs is present, which leaves every result as the pattern gives it.
bool contains_icase(const std::string &s, const char *lit) {
  const size_t m = std::strlen(lit);
  if (s.size() < m) return false;
  for (size_t i = 0; i + m <= s.size(); i++) {
    size_t k = 0;
    while (k < m && std::tolower((unsigned char)s[i + k]) == std::tolower((unsigned char)lit[k])) k++;
    if (k == m) return true;
  }
  return false;
}
std::string rewrite_coalesce(const std::string &s) {
  return contains_icase(s, "CASE(IS NOT NULL(") ? std::regex_replace(s, coalesce_re(), "COALESCE($1, $2)") : s;
}

std::string common_rewrites(std::string s) {
  s = rewrite_coalesce(s);
  if (contains_icase(s, "COUNT(DISTINCT ")) s = std::regex_replace(s, count_re(), "COUNT_DISTINCT($1)");
  replace_all(s, " NOT NULL", "");
  replace_all(s, "):DOUBLE", "");
  replace_all(s, "CAST(", "");
  replace_all(s, "EXTRACT(FLAG(YEAR), ", "BL_YEAR(");
  replace_all(s, "EXTRACT(FLAG(MONTH), ", "BL_MONTH(");
  replace_all(s, "EXTRACT(FLAG(DAY), ", "BL_DAY(");
  replace_all(s, "FLOOR(", "BL_FLOUR(");
  return s;
}

const std::map<std::string, int> &unary_map() {
  static const std::map<std::string, int> m = {
      {"NOT", BLZ_NOT}, {"SIN", BLZ_SIN}, {"ASIN", BLZ_ASIN}, {"COS", BLZ_COS},
      {"ACOS", BLZ_ACOS}, {"TAN", BLZ_TAN}, {"ATAN", BLZ_ATAN}, {"BL_FLOUR", BLZ_FLOOR},
      {"CEIL", BLZ_CEIL}, {"ABS", BLZ_ABS}, {"LOG10", BLZ_LOG}, {"LN", BLZ_LN},
      {"BL_YEAR", BLZ_YEAR}, {"BL_MONTH", BLZ_MONTH}, {"BL_DAY", BLZ_DAY},
      {"BL_HOUR", BLZ_HOUR}, {"BL_MINUTE", BLZ_MINUTE}, {"BL_SECOND", BLZ_SECOND},
      {"IS_NULL", BLZ_IS_NULL}, {"IS_NOT_NULL", BLZ_IS_NOT_NULL},
      // extension token without a reference counterpart: exact int64 -> float64
      {"BL_CAST_DOUBLE", BLZ_CAST_FLOAT64}};
  return m;
}
// AND / OR map to MUL / ADD as in CalciteExpressionParsing.cpp:553-554.
const std::map<std::string, int> &binary_map() {
  static const std::map<std::string, int> m = {
      {"=", GDF_EQUAL}, {"<>", GDF_NOT_EQUAL}, {">", GDF_GREATER}, {">=", GDF_GREATER_EQUAL},
      {"<", GDF_LESS}, {"<=", GDF_LESS_EQUAL}, {"+", GDF_ADD}, {"-", GDF_SUB},
      {"*", GDF_MUL}, {"/", GDF_DIV}, {"POWER", GDF_POW}, {"MOD", GDF_MOD},
      {"AND", GDF_MUL}, {"OR", GDF_ADD}, {"COALESCE", GDF_COALESCE}};
  return m;
}

bool is_binary_token(const std::string &t) { return binary_map().count(t) > 0; }
bool is_unary_token(const std::string &t) { return unary_map().count(t) > 0; }

bool is_float_t(int t) { return t == GDF_FLOAT32 || t == GDF_FLOAT64; }

}  // namespace

int find_closing_char(const std::string &e, int start) {
  const char open = e[start];
  char close;
  if (open == '(') close = ')';
  else if (open == '[') close = ']';
  else return -1;
  int depth = 1;
  bool in = false;
  for (int i = start + 1; i < (int)e.size(); i++) {
    if (in) {
      if (e[i] == '\'' && !(i + 1 < (int)e.size() && e[i + 1] == '\'')) in = false;
    } else {
      if (e[i] == '\'') in = true;
      else if (e[i] == open) depth++;
      else if (e[i] == close) {
        if (--depth == 0) return i;
      }
    }
  }
  return -1;
}

std::vector<std::string> get_expressions_from_expression_list(std::string combined, bool do_trim) {
  combined = common_rewrites(combined);
  std::vector<std::string> exprs;
  int start = 0, pd = 0, sd = 0;
  bool in = false;
  int cur = 0;
  for (; cur < (int)combined.size(); cur++) {
    const char ch = combined[cur];
    if (in) {
      if (ch == '\'' && !(cur + 1 < (int)combined.size() && combined[cur + 1] == '\'')) in = false;
    } else {
      if (ch == '\'') in = true;
      else if (ch == '(') pd++;
      else if (ch == ')') pd--;
      else if (ch == '[') sd++;
      else if (ch == ']') sd--;
      else if (ch == ',' && pd == 0 && sd == 0) {
        std::string e = combined.substr(start, cur - start);
        exprs.push_back(do_trim ? ltrim(e) : e);
        start = cur + 1;
      }
    }
  }
  if (start < (int)combined.size()) {
    std::string e = combined.substr(start, cur - start);
    exprs.push_back(do_trim ? trim(e) : e);
  }
  for (auto &e : exprs) e = common_rewrites(e);
  return exprs;
}

std::string expand_if_logical_op(std::string expression) {
  std::string output = expression;
  int start_pos = 0;
  while (start_pos < (int)expression.size()) {
    const auto q = quoted_vector(expression);
    const int fa = find_first_not_in_quotes(expression, "AND(", start_pos, q);
    const int fo = find_first_not_in_quotes(expression, "OR(", start_pos, q);
    int first;
    std::string op;
    if (fa >= 0) {
      if (fo >= 0 && fo < fa) { first = fo; op = "OR("; }
      else { first = fa; op = "AND("; }
    } else { first = fo; op = "OR("; }
    if (first < 0) return output;
    const int es = first + (int)op.size() - 1;
    const int ee = find_closing_char(expression, es);
    if (ee < 0) fail("unbalanced parentheses in expression");
    const std::string rest = expression.substr(es + 1, ee - (es + 1));
    std::vector<std::string> processed = get_expressions_from_expression_list(rest, false);
    if (processed.size() == 2) { start_pos = es; continue; }
    if (processed.empty()) fail("empty logical operator");
    start_pos = first;
    output = expression.substr(0, first);
    for (size_t i = 0; i + 1 < processed.size(); i++) { output += op; start_pos += (int)op.size(); }
    output += processed[0] + ",";
    for (size_t i = 1; i + 1 < processed.size(); i++) output += processed[i] + "),";
    output += processed.back() + ")";
    if (ee < (int)expression.size() - 1) output += expression.substr(ee + 1);
    expression = output;
  }
  return output;
}

std::string clean_calcite_expression(std::string e) {
  e = rewrite_coalesce(e);
  replace_all(e, "IS NOT NULL", "IS_NOT_NULL");
  replace_all(e, "IS NULL", "IS_NULL");
  replace_all(e, " NOT NULL", "");
  replace_all(e, "):DOUBLE", "");
  replace_all(e, "CAST(", "");
  replace_all(e, "EXTRACT(FLAG(YEAR), ", "BL_YEAR(");
  replace_all(e, "EXTRACT(FLAG(MONTH), ", "BL_MONTH(");
  replace_all(e, "EXTRACT(FLAG(DAY), ", "BL_DAY(");
  replace_all(e, "EXTRACT(FLAG(HOUR), ", "BL_HOUR(");
  replace_all(e, "EXTRACT(FLAG(MINUTE), ", "BL_MINUTE(");
  replace_all(e, "EXTRACT(FLAG(SECOND), ", "BL_SECOND(");
  replace_all(e, "FLOOR(", "BL_FLOUR(");
  e = expand_if_logical_op(e);
  std::string out;
  out.reserve(e.size());
  for (char ch : e) {
    if (ch == '(') out.push_back(' ');
    else if (ch != ',' && ch != ')') out.push_back(ch);
  }
  return out;
}

bool is_string(const std::string &t) { return !t.empty() && t[0] == '\'' && t.back() == '\''; }
bool is_literal(const std::string &t) { return t.empty() || t[0] != '$'; }
// Exact matchers of the literal grammars the reference's planner tests with regular
// expressions (CalciteExpressionParsing.cpp:274, :324-325, :585): ^[-+]?[0-9]+$ (match_int),
// ^[-+]?[0-9]*\.?[0-9]+([eE][-+]?[0-9]+)?$ (match_float) and
// [12]\d{3}-(0[1-9]|1[0-2])-(0[1-9]|[12]\d|3[01]) (match_date), written out because
// std::regex costs microseconds per token.
bool is_digit(char c) { return c >= '0' && c <= '9'; }
bool match_int(const std::string &t) {
  size_t i = 0;
  if (i < t.size() && (t[i] == '+' || t[i] == '-')) i++;
  const size_t d = i;
  while (i < t.size() && is_digit(t[i])) i++;
  return i > d && i == t.size();
}
bool match_float(const std::string &t) {
  const size_t n = t.size();
  size_t i = 0;
  if (i < n && (t[i] == '+' || t[i] == '-')) i++;
  const size_t a = i;
  while (i < n && is_digit(t[i])) i++;
  if (i < n && t[i] == '.') {  // [0-9]*\.[0-9]+
    const size_t b = ++i;
    while (i < n && is_digit(t[i])) i++;
    if (i == b) return false;
  } else if (i == a) {  // [0-9]+
    return false;
  }
  if (i < n && (t[i] == 'e' || t[i] == 'E')) {
    i++;
    if (i < n && (t[i] == '+' || t[i] == '-')) i++;
    const size_t c = i;
    while (i < n && is_digit(t[i])) i++;
    if (i == c) return false;
  }
  return i == n;
}
bool match_date(const std::string &t) {
  if (t.size() != 10 || (t[0] != '1' && t[0] != '2') || !is_digit(t[1]) || !is_digit(t[2]) || !is_digit(t[3]) ||
      t[4] != '-' || t[7] != '-')
    return false;
  const char m0 = t[5], m1 = t[6], d0 = t[8], d1 = t[9];
  const bool month = (m0 == '0' && m1 >= '1' && m1 <= '9') || (m0 == '1' && m1 >= '0' && m1 <= '2');
  const bool day = (d0 == '0' && d1 >= '1' && d1 <= '9') || ((d0 == '1' || d0 == '2') && is_digit(d1)) ||
                   (d0 == '3' && (d1 == '0' || d1 == '1'));
  return month && day;
}

bool is_number(const std::string &t) { return match_float(t); }
bool is_date(const std::string &t) { return match_date(t); }
bool is_operator_token(const std::string &t) {
  return !t.empty() && t[0] != '$' && !is_number(t) && !is_date(t) && !is_string(t);
}

std::vector<std::string> get_tokens_in_reverse_order(const std::string &e) {
  std::vector<std::string> t = split_not_in_quotes(e, " ");
  std::reverse(t.begin(), t.end());
  return t;
}

size_t get_index(const std::string &operand) {
  const std::string c = clean_calcite_expression(operand);
  if (c.empty()) return 0;
  const std::string digits = is_literal(c) ? c : c.substr(1);
  size_t i = 0;
  while (i < digits.size() && (digits[i] == ' ' || digits[i] == '\t')) i++;
  if (i < digits.size() && digits[i] == '+') i++;
  size_t j = i;
  while (j < digits.size() && digits[j] >= '0' && digits[j] <= '9') j++;
  if (j == i) fail("not a column index: " + operand);
  return std::stoull(digits.substr(i, j - i));  // stoull stops at the first non-digit
}

int get_type_from_string(const std::string &s) {
  if (match_int(s)) return GDF_INT64;
  if (match_float(s)) return GDF_FLOAT64;
  return GDF_DATE64;
}

namespace {
int64_t days_from_civil(int64_t y, unsigned m, unsigned d) {
  y -= m <= 2;
  const int64_t era = (y >= 0 ? y : y - 399) / 400;
  const unsigned yoe = (unsigned)(y - era * 400);
  const unsigned doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
  const unsigned doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return era * 146097 + (int64_t)doe - 719468;
}
}  // namespace

gdf_scalar get_scalar_from_string(const std::string &s, int dtype) {
  gdf_scalar r{};
  if (s == "null") { r.dtype = GDF_INT8; r.is_valid = 0; return r; }
  r.is_valid = 1;
  if (dtype == GDF_INT64) {
    try { r.data.si64 = std::stoll(s); } catch (...) { fail("integer literal out of range: " + s); }
    r.dtype = GDF_INT64;
  } else if (dtype == GDF_FLOAT64) {
    r.data.fp64 = std::stod(s);
    r.dtype = GDF_FLOAT64;
  } else {
    int y, mo, d, hh = 0, mm = 0, ss = 0;
    char tail;
    int n = std::sscanf(s.c_str(), "%4d-%2d-%2d %2d:%2d:%2d%c", &y, &mo, &d, &hh, &mm, &ss, &tail);
    if (!(n == 3 || n == 6)) fail("Invalid datetime format: " + s);
    const int64_t days = days_from_civil(y, (unsigned)mo, (unsigned)d);
    r.data.dt64 = ((days * 24 + hh) * 60 + mm) * 60000 + (int64_t)ss * 1000;
    r.dtype = GDF_DATE64;
  }
  return r;
}

int get_output_type_expression(const std::vector<int> &dtypes, const std::string &expression) {
  const auto toks = get_tokens_in_reverse_order(clean_calcite_expression(expression));
  std::vector<int> st;
  for (const auto &t : toks) {
    if (is_operator_token(t)) {
      if (is_binary_token(t)) {
        if (st.size() < 2) fail("operator " + t + " needs two operands");
        int l = st.back(); st.pop_back();
        int r = st.back(); st.pop_back();
        if (l == GDF_invalid) {
          if (r == GDF_invalid) fail("invalid operands for " + t);
          l = r;
        } else if (r == GDF_invalid) {
          r = l;
        }
        st.push_back(binary_output_dtype(l, r, binary_map().at(t)));
      } else if (is_unary_token(t)) {
        if (st.empty()) fail("operator " + t + " needs an operand");
        int l = st.back(); st.pop_back();
        st.push_back(unary_output_dtype(l, unary_map().at(t)));
      } else {
        fail("unsupported operator token, " + t);
      }
    } else {
      if (is_literal(t) || is_string(t)) st.push_back(GDF_invalid);
      else {
        size_t i = get_index(t);
        if (i >= dtypes.size()) fail("column index out of range: " + t);
        st.push_back(dtypes[i]);
      }
    }
  }
  if (st.empty()) fail("empty expression");
  return st.back();
}

// LogicalFilter.cpp:44-267
void add_expression_to_plan(PlanSpec &p, const std::string &expression, int expression_position,
                            int num_outputs, int num_inputs, const std::vector<int> &new_input_indices) {
  const int start = num_inputs + num_outputs;
  const auto toks = get_tokens_in_reverse_order(clean_calcite_expression(expression));
  std::vector<bool> free_pos(512, true);
  for (int i = 0; i < start && i < 512; i++) free_pos[i] = false;
  struct Operand { std::string token; int position; };
  std::vector<Operand> st;
  gdf_scalar dummy{};
  dummy.dtype = GDF_INT32;
  dummy.is_valid = 1;
  auto pop = [&]() {
    if (st.empty()) fail("malformed expression: missing operand");
    Operand o = st.back();
    st.pop_back();
    if (!is_literal(o.token) && o.position >= start) free_pos[o.position] = true;
    return o;
  };
  for (size_t ti = 0; ti < toks.size(); ti++) {
    const std::string &t = toks[ti];
    if (is_operator_token(t)) {
      if (is_binary_token(t)) {
        Operand lo = pop(), ro = pop();
        p.binops.push_back(binary_map().at(t));
        p.unops.push_back(BLZ_INVALID_UNARY);
        const bool ll = is_literal(lo.token), rl = is_literal(ro.token);
        if (ll && rl) fail("binary operator with two literal operands is not supported");
        if (ll) {
          if (is_string(lo.token)) fail("string literals are not supported");
          gdf_scalar s = get_scalar_from_string(lo.token, get_type_from_string(lo.token));
          p.lsc.push_back(s);
          p.rsc.push_back(dummy);
          p.left.push_back(s.is_valid ? SCALAR_INDEX : SCALAR_NULL_INDEX);
          p.right.push_back((int16_t)get_index(ro.token));
        } else if (rl) {
          if (is_string(ro.token)) fail("string literals are not supported");
          gdf_scalar s = get_scalar_from_string(ro.token, get_type_from_string(ro.token));
          p.rsc.push_back(s);
          p.lsc.push_back(dummy);
          p.right.push_back(s.is_valid ? SCALAR_INDEX : SCALAR_NULL_INDEX);
          p.left.push_back((int16_t)get_index(lo.token));
        } else {
          p.left.push_back((int16_t)get_index(lo.token));
          p.right.push_back((int16_t)get_index(ro.token));
          p.lsc.push_back(dummy);
          p.rsc.push_back(dummy);
        }
      } else if (is_unary_token(t)) {
        Operand lo = pop();
        p.binops.push_back(GDF_INVALID_BINARY);
        p.unops.push_back(unary_map().at(t));
        if (is_literal(lo.token)) fail("unary operator on a literal operand is not supported");
        p.left.push_back((int16_t)get_index(lo.token));
        p.right.push_back(BLZ_UNARY_INDEX);
        p.lsc.push_back(dummy);
        p.rsc.push_back(dummy);
      } else {
        fail("unsupported operator, " + t);
      }
      if (ti == toks.size() - 1) {
        p.outs.push_back((int16_t)(expression_position + num_inputs));
      } else {
        int pos = -1;
        for (int i = start; i < 512; i++)
          if (free_pos[i]) { free_pos[i] = false; pos = i; break; }
        if (pos < 0) fail("out of plan registers");
        p.outs.push_back((int16_t)pos);
        st.push_back({"$" + std::to_string(pos), pos});
      }
    } else {
      if (is_literal(t) || is_string(t)) {
        st.push_back({t, SCALAR_INDEX});
      } else {
        size_t i = get_index(t);
        if (i >= new_input_indices.size() || new_input_indices[i] < 0) fail("column index out of range: " + t);
        int idx = new_input_indices[i];
        st.push_back({"$" + std::to_string(idx), idx});
      }
    }
  }
}

namespace {

void used_columns(const std::vector<std::string> &exprs, int ncols, std::vector<int> &new_idx,
                  std::vector<int> &cols) {
  std::vector<bool> used(ncols, false);
  for (const auto &e : exprs)
    for (const auto &t : get_tokens_in_reverse_order(clean_calcite_expression(e)))
      if (!is_operator_token(t) && !is_literal(t)) {
        size_t i = get_index(t);
        if (i >= (size_t)ncols) fail("column index out of range: " + t);
        used[i] = true;
      }
  new_idx.assign(ncols, -1);
  cols.clear();
  for (int i = 0; i < ncols; i++)
    if (used[i]) { new_idx[i] = (int)cols.size(); cols.push_back(i); }
}

}  // namespace

ExprPlan plan_expression(const std::vector<int> &dtypes, const std::string &expression) {
  ExprPlan r;
  if (expression.empty()) fail("empty expression");
  if (expression[0] == '$') {  // LogicalFilter.cpp:281-287
    r.bare = true;
    r.bare_col = (int)get_index(expression);
    if (r.bare_col >= (int)dtypes.size()) fail("column index out of range: " + expression);
    return r;
  }
  std::vector<int> new_idx, cols;
  used_columns({expression}, (int)dtypes.size(), new_idx, cols);
  r.spec.in_dev = cols;
  r.spec.final_pos = {(int16_t)cols.size()};
  add_expression_to_plan(r.spec, expression, 0, 1, (int)cols.size(), new_idx);
  return r;
}

std::string get_named_expression(const std::string &qp, const std::string &name) {
  if (qp.find(name + "=[") == std::string::npos) return "";
  size_t start;
  if (qp.find(name + "=[[") != std::string::npos) start = qp.find(name + "=[[") + 3 + name.size();
  else start = qp.find(name + "=[") + 2 + name.size();
  size_t end = qp.find(']', start);
  return qp.substr(start, end == std::string::npos ? std::string::npos : end - start);
}

bool contains_evaluation(const std::string &expression) {
  std::string e = expression;
  replace_all(e, "CAST(", "");
  replace_all(e, "):BIGINT", "");
  return e.find('(') != std::string::npos;
}

ProjectPlan parse_project(const std::vector<int> &dtypes, const std::string &qp) {
  ProjectPlan r;
  const size_t a = qp.find('('), b = qp.rfind(')');
  if (a == std::string::npos || b == std::string::npos || b <= a) fail("malformed LogicalProject");
  const auto exprs = get_expressions_from_expression_list(qp.substr(a + 1, b - a - 1));
  std::vector<std::string> bodies, names, evals;
  for (const auto &e : exprs) {
    const size_t p = e.find("=[");
    if (p == std::string::npos || e.size() < p + 3) fail("malformed projection: " + e);
    bodies.push_back(e.substr(p + 2, (e.size() - p) - 3));
    names.push_back(e.substr(0, p));
    if (contains_evaluation(bodies.back())) evals.push_back(bodies.back());
  }
  std::vector<int> new_idx, cols;
  used_columns(evals, (int)dtypes.size(), new_idx, cols);
  const int n_out = (int)evals.size();
  r.spec.in_dev = cols;
  for (int k = 0; k < n_out; k++) r.spec.final_pos.push_back((int16_t)(cols.size() + k));
  int k = 0;
  for (size_t i = 0; i < bodies.size(); i++) {
    if (contains_evaluation(bodies[i])) {
      const int dt = get_output_type_expression(dtypes, bodies[i]);
      add_expression_to_plan(r.spec, bodies[i], k, n_out, (int)cols.size(), new_idx);
      r.outputs.push_back({0, k, dt, names[i]});
      k++;
    } else {
      const size_t c = get_index(bodies[i]);
      if (c >= dtypes.size()) fail("column index out of range: " + bodies[i]);
      r.outputs.push_back({1, (int)c, dtypes[c], names[i]});
    }
  }
  r.has_eval = n_out > 0;
  return r;
}

}  // namespace planner
}  // namespace blz

"""Planner restatement: Calcite expression text -> register-file plan.

TEST INFRASTRUCTURE ONLY: the parity tests use this to check the product's C++
planner (blazingdb-ral_amd/csrc/planner.cpp) and to build plans for the oracle.

Restates, in plain Python:
  clean_calcite_expression        CalciteExpressionParsing.cpp:705-754
  expand_if_logical_op            CalciteExpressionParsing.cpp:629-694
  get_expressions_from_expression_list  CalciteExpressionParsing.cpp:814-895
  find_closing_char               CalciteExpressionParsing.cpp:771-811
  token predicates / get_index    CalciteExpressionParsing.cpp:568-608, 272-276
  get_type_from_string / get_scalar_from_string  CalciteExpressionParsing.cpp:323-417
  get_output_type(_expression)    CalciteExpressionParsing.cpp:177-270, 420-481
  add_expression_to_plan          LogicalFilter.cpp:44-267
  evaluate_expression (plan part) LogicalFilter.cpp:271-371
  parse_project_plan              CalciteInterpreter.cpp:123-309
  get_named_expression            CalciteInterpreter.cpp:424-440

The string helpers of the un-vendored blazingdb-io StringUtil (findAndReplaceAll,
splitNotInQuotes, findFirstNotInQuotes) are restated from their call sites: plain
replace-all, split on the delimiter outside single quotes (empty pieces dropped),
first occurrence at or after a start index outside single quotes.

Plans that the reference would push as misaligned vectors (a binary op with two
literals, a unary op on a literal, an unknown operator token) raise PlanError;
string literals (NVCategory) raise PlanError as well.
"""
import calendar
import re
from dataclasses import dataclass, field
from typing import List, Optional

# dtype / operator ordinals of include/blz/gdf_types.h
GDF_invalid, GDF_INT8, GDF_INT16, GDF_INT32, GDF_INT64, GDF_FLOAT32, GDF_FLOAT64, \
    GDF_BOOL8, GDF_DATE32, GDF_DATE64, GDF_TIMESTAMP, GDF_CATEGORY, GDF_STRING, \
    GDF_STRING_CATEGORY = range(14)
(GDF_ADD, GDF_SUB, GDF_MUL, GDF_DIV, GDF_TRUE_DIV, GDF_FLOOR_DIV, GDF_MOD, GDF_POW,
 GDF_EQUAL, GDF_NOT_EQUAL, GDF_LESS, GDF_GREATER, GDF_LESS_EQUAL, GDF_GREATER_EQUAL,
 GDF_BITWISE_AND, GDF_BITWISE_OR, GDF_BITWISE_XOR, GDF_LOGICAL_AND, GDF_LOGICAL_OR,
 GDF_COALESCE, GDF_INVALID_BINARY) = range(21)
(BLZ_FLOOR, BLZ_CEIL, BLZ_SIN, BLZ_COS, BLZ_ASIN, BLZ_ACOS, BLZ_TAN, BLZ_COTAN, BLZ_ATAN,
 BLZ_ABS, BLZ_NOT, BLZ_LN, BLZ_LOG, BLZ_YEAR, BLZ_MONTH, BLZ_DAY, BLZ_HOUR, BLZ_MINUTE,
 BLZ_SECOND, BLZ_IS_NULL, BLZ_IS_NOT_NULL, BLZ_INVALID_UNARY, BLZ_CAST_FLOAT64) = range(23)

SCALAR_INDEX = -2
SCALAR_NULL_INDEX = -3

WIDTH = {GDF_INT8: 1, GDF_BOOL8: 1, GDF_INT16: 2, GDF_INT32: 4, GDF_FLOAT32: 4,
         GDF_DATE32: 4, GDF_CATEGORY: 4, GDF_STRING_CATEGORY: 4, GDF_INT64: 8,
         GDF_FLOAT64: 8, GDF_DATE64: 8, GDF_TIMESTAMP: 8}

# CalciteExpressionParsing.cpp:509-556 (AND -> MUL and OR -> ADD, as the reference maps them)
UNARY_OPS = {"NOT": BLZ_NOT, "SIN": BLZ_SIN, "ASIN": BLZ_ASIN, "COS": BLZ_COS,
             "ACOS": BLZ_ACOS, "TAN": BLZ_TAN, "ATAN": BLZ_ATAN, "BL_FLOUR": BLZ_FLOOR,
             "CEIL": BLZ_CEIL, "ABS": BLZ_ABS, "LOG10": BLZ_LOG, "LN": BLZ_LN,
             "BL_YEAR": BLZ_YEAR, "BL_MONTH": BLZ_MONTH, "BL_DAY": BLZ_DAY,
             "BL_HOUR": BLZ_HOUR, "BL_MINUTE": BLZ_MINUTE, "BL_SECOND": BLZ_SECOND,
             "IS_NULL": BLZ_IS_NULL, "IS_NOT_NULL": BLZ_IS_NOT_NULL,
             # extension token (no reference counterpart): exact int64 -> float64
             "BL_CAST_DOUBLE": BLZ_CAST_FLOAT64}
BINARY_OPS = {"=": GDF_EQUAL, "<>": GDF_NOT_EQUAL, ">": GDF_GREATER,
              ">=": GDF_GREATER_EQUAL, "<": GDF_LESS, "<=": GDF_LESS_EQUAL, "+": GDF_ADD,
              "-": GDF_SUB, "*": GDF_MUL, "/": GDF_DIV, "POWER": GDF_POW, "MOD": GDF_MOD,
              "AND": GDF_MUL, "OR": GDF_ADD, "COALESCE": GDF_COALESCE}


class PlanError(ValueError):
    pass


@dataclass
class Scalar:
    dtype: int = GDF_INT32
    value: float = 0      # int or float
    is_valid: bool = True


DUMMY = Scalar(GDF_INT32, 0, True)


@dataclass
class Plan:
    left: List[int] = field(default_factory=list)
    right: List[int] = field(default_factory=list)
    outputs: List[int] = field(default_factory=list)
    binops: List[int] = field(default_factory=list)
    unops: List[int] = field(default_factory=list)
    lscalars: List[Scalar] = field(default_factory=list)
    rscalars: List[Scalar] = field(default_factory=list)
    final_positions: List[int] = field(default_factory=list)
    input_columns: List[int] = field(default_factory=list)   # frame column ids, slot order
    output_dtypes: List[int] = field(default_factory=list)


# --- StringUtil-like helpers -------------------------------------------------------

def _quoted(s):
    q, inq = [], False
    for ch in s:
        if ch == "'":
            inq = not inq
            q.append(True)
        else:
            q.append(inq)
    return q


def _find_first_not_in_quotes(s, needle, start, quoted):
    i = s.find(needle, start)
    while i >= 0:
        if not quoted[i]:
            return i
        i = s.find(needle, i + 1)
    return -1


def _split_not_in_quotes(s, delim):
    out, cur, inq = [], [], False
    i = 0
    while i < len(s):
        if s[i] == "'":
            inq = not inq
        if not inq and s.startswith(delim, i):
            out.append("".join(cur))
            cur = []
            i += len(delim)
            continue
        cur.append(s[i])
        i += 1
    out.append("".join(cur))
    return [t for t in out if t != ""]


_COALESCE_RE = re.compile(r"CASE\(IS NOT NULL\((\W\(.+?\)|.+)\), \1, (\W\(.+?\)|.+)\)", re.I)
_COUNT_RE = re.compile(r"COUNT\(DISTINCT (\W\(.+?\)|.+)\)", re.I)


def find_closing_char(expression, start):
    open_c = expression[start]
    close_c = {"(": ")", "[": "]"}.get(open_c)
    if close_c is None:
        return -1
    depth, inq = 1, False
    i = start + 1
    while i < len(expression):
        ch = expression[i]
        if inq:
            if ch == "'" and not (i + 1 < len(expression) and expression[i + 1] == "'"):
                inq = False
        else:
            if ch == "'":
                inq = True
            elif ch == open_c:
                depth += 1
            elif ch == close_c:
                depth -= 1
                if depth == 0:
                    return i
        i += 1
    return -1


def _common_rewrites(s):
    s = _COALESCE_RE.sub(r"COALESCE(\1, \2)", s)
    s = _COUNT_RE.sub(r"COUNT_DISTINCT(\1)", s)
    for a, b in ((" NOT NULL", ""), ("):DOUBLE", ""), ("CAST(", ""),
                 ("EXTRACT(FLAG(YEAR), ", "BL_YEAR("), ("EXTRACT(FLAG(MONTH), ", "BL_MONTH("),
                 ("EXTRACT(FLAG(DAY), ", "BL_DAY("), ("FLOOR(", "BL_FLOUR(")):
        s = s.replace(a, b)
    return s


def get_expressions_from_expression_list(combined, trim=True):
    combined = _common_rewrites(combined)
    exprs = []
    cur, start, inq, pd, sd = 0, 0, False, 0, 0
    while cur < len(combined):
        ch = combined[cur]
        if inq:
            if ch == "'" and not (cur + 1 < len(combined) and combined[cur + 1] == "'"):
                inq = False
        else:
            if ch == "'":
                inq = True
            elif ch == "(":
                pd += 1
            elif ch == ")":
                pd -= 1
            elif ch == "[":
                sd += 1
            elif ch == "]":
                sd -= 1
            elif ch == "," and pd == 0 and sd == 0:
                e = combined[start:cur]
                exprs.append(e.lstrip() if trim else e)
                start = cur + 1
        cur += 1
    if start < len(combined):
        e = combined[start:cur]
        exprs.append(e.strip() if trim else e)
    return [_common_rewrites(e) for e in exprs]


def expand_if_logical_op(expression):
    output = expression
    start_pos = 0
    while start_pos < len(expression):
        q = _quoted(expression)
        fa = _find_first_not_in_quotes(expression, "AND(", start_pos, q)
        fo = _find_first_not_in_quotes(expression, "OR(", start_pos, q)
        if fa >= 0:
            if 0 <= fo < fa:
                first, op = fo, "OR("
            else:
                first, op = fa, "AND("
        else:
            first, op = fo, "OR("
        if first < 0:
            return output
        es = first + len(op) - 1
        ee = find_closing_char(expression, es)
        rest = expression[es + 1:ee]
        processed = get_expressions_from_expression_list(rest, False)
        if len(processed) == 2:
            start_pos = es
            continue
        start_pos = first
        output = expression[:first]
        for _ in range(len(processed) - 1):
            output += op
            start_pos += len(op)
        output += processed[0] + ","
        for i in range(1, len(processed) - 1):
            output += processed[i] + "),"
        output += processed[-1] + ")"
        if ee < len(expression) - 1:
            output += expression[ee + 1:]
        expression = output
    return output


def clean_calcite_expression(expression):
    s = _COALESCE_RE.sub(r"COALESCE(\1, \2)", expression)
    for a, b in (("IS NOT NULL", "IS_NOT_NULL"), ("IS NULL", "IS_NULL"), (" NOT NULL", ""),
                 ("):DOUBLE", ""), ("CAST(", ""),
                 ("EXTRACT(FLAG(YEAR), ", "BL_YEAR("), ("EXTRACT(FLAG(MONTH), ", "BL_MONTH("),
                 ("EXTRACT(FLAG(DAY), ", "BL_DAY("), ("EXTRACT(FLAG(HOUR), ", "BL_HOUR("),
                 ("EXTRACT(FLAG(MINUTE), ", "BL_MINUTE("), ("EXTRACT(FLAG(SECOND), ", "BL_SECOND("),
                 ("FLOOR(", "BL_FLOUR(")):
        s = s.replace(a, b)
    s = expand_if_logical_op(s)
    return "".join(" " if ch == "(" else ch for ch in s if ch not in ",)")


_NUM_RE = re.compile(r"^[-+]?[0-9]*\.?[0-9]+([eE][-+]?[0-9]+)?$")
_INT_RE = re.compile(r"^[-+]?[0-9]+$")
_DATE_RE = re.compile(r"[12]\d{3}-(0[1-9]|1[0-2])-(0[1-9]|[12]\d|3[01])")


def is_string(t):
    return len(t) > 0 and t[0] == "'" and t[-1] == "'"


def is_literal(t):
    return t[0] != "$"


def is_number(t):
    return _NUM_RE.match(t) is not None


def is_date(t):
    return _DATE_RE.fullmatch(t) is not None


def is_operator_token(t):
    return t[0] != "$" and not is_number(t) and not is_date(t) and not is_string(t)


def get_tokens_in_reverse_order(expr):
    return list(reversed(_split_not_in_quotes(expr, " ")))


def _stoull_prefix(s):
    m = re.match(r"\s*\+?([0-9]+)", s)
    if not m:
        raise PlanError(f"not an index: {s!r}")
    return int(m.group(1))


def get_index(operand):
    c = clean_calcite_expression(operand)
    if len(c) == 0:
        return 0
    return _stoull_prefix(c if is_literal(c) else c[1:])


def get_type_from_string(s):
    if _INT_RE.match(s):
        return GDF_INT64
    if _NUM_RE.match(s):
        return GDF_FLOAT64
    return GDF_DATE64


def _date64_ms(s):
    m = re.fullmatch(r"(\d{4})-(\d{2})-(\d{2})(?: (\d{2}):(\d{2}):(\d{2}))?", s)
    if not m:
        raise PlanError(f"Invalid datetime format: {s!r}")
    y, mo, d, hh, mm, ss = (int(g) if g else 0 for g in m.groups())
    return calendar.timegm((y, mo, d, hh, mm, ss, 0, 0, 0)) * 1000


def get_scalar_from_string(s, dtype):
    if s == "null":
        return Scalar(GDF_INT8, 0, False)
    if dtype == GDF_INT64:
        v = int(s)
        if not -(1 << 63) <= v < (1 << 63):
            raise PlanError(f"integer literal out of range: {s}")
        return Scalar(GDF_INT64, v, True)
    if dtype == GDF_FLOAT64:
        return Scalar(GDF_FLOAT64, float(s), True)
    return Scalar(GDF_DATE64, _date64_ms(s), True)


# --- output dtypes -------------------------------------------------------------------

def _is_signed(t):
    return t in (GDF_INT8, GDF_INT16, GDF_INT32, GDF_INT64, GDF_FLOAT32, GDF_FLOAT64,
                 GDF_DATE32, GDF_DATE64, GDF_TIMESTAMP)


def _is_float(t):
    return t in (GDF_FLOAT32, GDF_FLOAT64)


def _is_date(t):
    return t in (GDF_DATE32, GDF_DATE64, GDF_TIMESTAMP)


def unary_output_type(l, op):
    if op == BLZ_CAST_FLOAT64:
        return GDF_FLOAT64
    if _is_date(l):
        return GDF_INT16
    if op in (BLZ_SIN, BLZ_COS, BLZ_ASIN, BLZ_ACOS, BLZ_TAN, BLZ_COTAN, BLZ_ATAN, BLZ_LOG, BLZ_LN):
        return l if _is_float(l) else GDF_FLOAT64
    if op in (BLZ_IS_NULL, BLZ_IS_NOT_NULL):
        return GDF_INT8
    return l


def binary_output_type(l, r, op):
    if not _is_signed(l) or not _is_signed(r):
        return GDF_invalid
    w = lambda t: WIDTH.get(t, 0)
    if op in (GDF_ADD, GDF_SUB, GDF_MUL, GDF_DIV, GDF_MOD):
        if _is_float(l) or _is_float(r):
            if _is_float(l) and _is_float(r):
                return l if w(l) >= w(r) else r
            return l if _is_float(l) else r
        return l if w(l) >= w(r) else r
    if op in (GDF_EQUAL, GDF_NOT_EQUAL, GDF_GREATER, GDF_GREATER_EQUAL, GDF_LESS, GDF_LESS_EQUAL):
        return GDF_INT8
    if op == GDF_POW:
        return GDF_FLOAT64 if (_is_float(l) or _is_float(r)) else GDF_INT64
    if op == GDF_COALESCE:
        return l
    return GDF_invalid


def get_output_type_expression(column_dtypes, expression):
    toks = get_tokens_in_reverse_order(clean_calcite_expression(expression))
    st = []
    for t in toks:
        if is_operator_token(t):
            if t in BINARY_OPS:
                if len(st) < 2:
                    raise PlanError("operator needs two operands")
                l, r = st.pop(), st.pop()
                if l == GDF_invalid:
                    if r == GDF_invalid:
                        raise PlanError("invalid operands")
                    l = r
                elif r == GDF_invalid:
                    r = l
                st.append(binary_output_type(l, r, BINARY_OPS[t]))
            elif t in UNARY_OPS:
                st.append(unary_output_type(st.pop(), UNARY_OPS[t]))
            else:
                raise PlanError(f"unsupported operator token, {t}")
        else:
            st.append(GDF_invalid if (is_literal(t) or is_string(t)) else column_dtypes[get_index(t)])
    return st[-1]


# --- plan construction -------------------------------------------------------------

def add_expression_to_plan(plan, expression, expression_position, num_outputs, num_inputs,
                           new_input_indices):
    start = num_inputs + num_outputs
    toks = get_tokens_in_reverse_order(clean_calcite_expression(expression))
    free = [True] * 512
    for i in range(start):
        free[i] = False
    st = []
    for ti, t in enumerate(toks):
        if is_operator_token(t):
            if t in BINARY_OPS:
                (lt, lp) = st.pop()
                if not is_literal(lt) and lp >= start:
                    free[lp] = True
                (rt, rp) = st.pop()
                if not is_literal(rt) and rp >= start:
                    free[rp] = True
                plan.binops.append(BINARY_OPS[t])
                plan.unops.append(BLZ_INVALID_UNARY)
                if is_literal(lt) and is_literal(rt):
                    raise PlanError("binary operator with two literal operands")
                elif is_literal(lt):
                    if is_string(lt):
                        raise PlanError("string literals are not supported")
                    s = get_scalar_from_string(lt, get_type_from_string(lt))
                    plan.lscalars.append(s)
                    plan.rscalars.append(DUMMY)
                    plan.left.append(SCALAR_INDEX if s.is_valid else SCALAR_NULL_INDEX)
                    plan.right.append(get_index(rt))
                elif is_literal(rt):
                    if is_string(rt):
                        raise PlanError("string literals are not supported")
                    s = get_scalar_from_string(rt, get_type_from_string(rt))
                    plan.rscalars.append(s)
                    plan.lscalars.append(DUMMY)
                    plan.right.append(SCALAR_INDEX if s.is_valid else SCALAR_NULL_INDEX)
                    plan.left.append(get_index(lt))
                else:
                    plan.left.append(get_index(lt))
                    plan.right.append(get_index(rt))
                    plan.lscalars.append(DUMMY)
                    plan.rscalars.append(DUMMY)
            elif t in UNARY_OPS:
                (lt, lp) = st.pop()
                if not is_literal(lt) and lp >= start:
                    free[lp] = True
                plan.binops.append(GDF_INVALID_BINARY)
                plan.unops.append(UNARY_OPS[t])
                if is_literal(lt):
                    raise PlanError("unary operator on a literal operand")
                plan.left.append(get_index(lt))
                plan.right.append(-1)
                plan.lscalars.append(DUMMY)
                plan.rscalars.append(DUMMY)
            else:
                raise PlanError(f"unsupported operator, {t}")
            if ti == len(toks) - 1:
                plan.outputs.append(expression_position + num_inputs)
            else:
                pos = next((i for i in range(start, 512) if free[i]), -1)
                if pos < 0:
                    raise PlanError("out of registers")
                free[pos] = False
                plan.outputs.append(pos)
                st.append(("$" + str(pos), pos))
        else:
            if is_literal(t) or is_string(t):
                st.append((t, SCALAR_INDEX))
            else:
                idx = new_input_indices[get_index(t)]
                st.append(("$" + str(idx), idx))


def _used_columns(expressions, ncols):
    used = [False] * ncols
    for e in expressions:
        for t in get_tokens_in_reverse_order(clean_calcite_expression(e)):
            if not is_operator_token(t) and not is_literal(t):
                used[get_index(t)] = True
    new_idx, cols = [-1] * ncols, []
    for i in range(ncols):
        if used[i]:
            new_idx[i] = len(cols)
            cols.append(i)
    return new_idx, cols


def plan_expression(column_dtypes, expression, output_dtype=GDF_INT8) -> Optional[Plan]:
    """evaluate_expression's plan (LogicalFilter.cpp:271-371); None for a bare "$k"."""
    if expression[0] == "$":
        return None
    new_idx, cols = _used_columns([expression], len(column_dtypes))
    plan = Plan(input_columns=cols, final_positions=[len(cols)], output_dtypes=[output_dtype])
    add_expression_to_plan(plan, expression, 0, 1, len(cols), new_idx)
    return plan


def named_expression(query_part, name):
    """get_named_expression (CalciteInterpreter.cpp:424-436)."""
    if (name + "=[") not in query_part:
        return ""
    if (name + "=[[") in query_part:
        start = query_part.find(name + "=[[") + 3 + len(name)
    else:
        start = query_part.find(name + "=[") + 2 + len(name)
    return query_part[start:query_part.find("]", start)]


def contains_evaluation(expression):
    return "(" in expression.replace("CAST(", "").replace("):BIGINT", "")


def parse_project(column_dtypes, query_part):
    """parse_project_plan (CalciteInterpreter.cpp:123-309).

    Returns (plan or None, outputs) where outputs[i] is ("expr", k) for the k-th
    evaluated expression or ("alias", column) for a "$k" projection, plus names."""
    combined = query_part[query_part.find("(") + 1:query_part.rfind(")")]
    exprs = get_expressions_from_expression_list(combined)
    bodies, names = [], []
    for e in exprs:
        p = e.find("=[")
        bodies.append(e[p + 2:p + 2 + (len(e) - p) - 3])
        names.append(e[:p])
    evals = [b for b in bodies if contains_evaluation(b)]
    new_idx, cols = _used_columns(evals, len(column_dtypes))
    n_out = len(evals)
    plan = Plan(input_columns=cols, final_positions=[len(cols) + k for k in range(n_out)])
    outputs, k = [], 0
    for b in bodies:
        if contains_evaluation(b):
            plan.output_dtypes.append(get_output_type_expression(column_dtypes, b))
            add_expression_to_plan(plan, b, k, n_out, len(cols), new_idx)
            outputs.append(("expr", k))
            k += 1
        else:
            outputs.append(("alias", get_index(b)))
    return (plan if n_out > 0 else None), outputs, names


# ---- LogicalAggregate (no GROUP BY) ----
AGG_OPS = ("SUM", "MIN", "MAX", "AVG", "COUNT")


def parse_aggregate(query_part):
    """process_aggregate's parsing (GroupBy.cpp:543-569): the group list (get_group_columns
    :41-56) and, for every other `name=[OP(input)]`, the operation
    (get_aggregation_operation, CalciteExpressionParsing.cpp:483-507) and the input
    between the outer parentheses (get_string_between_outer_parentheses :756-767).
    Returns (group column indices, [(op, input)])."""
    a, b = query_part.find("("), query_part.rfind(")")
    if a < 0 or b < a:
        raise PlanError("not a LogicalAggregate line")
    combined = query_part[a + 1:b]
    group = named_expression(combined, "group").strip()
    groups = [] if len(group) <= 2 else [int(x) for x in group[1:-1].split(",")]
    aggs = []
    for e in get_expressions_from_expression_list(combined):
        e = e.strip()
        if "group=" in e:
            continue
        eq = e.find("=[")
        lp, rp = e.find("(", eq), e.rfind(")")
        op = e[eq + 2:lp].strip()
        if op not in AGG_OPS:
            raise PlanError(f"aggregation type not supported, {op}")
        aggs.append((op, e[lp + 1:rp].strip()))
    return groups, aggs


def aggregation_output_type(input_type, op):
    """get_aggregation_output_type (CalciteExpressionParsing.cpp:92-117) without GROUP BY."""
    if op == "COUNT":
        return GDF_INT64
    if op == "SUM":
        return GDF_FLOAT64 if _is_float(input_type) else GDF_INT64
    if op == "AVG":
        return GDF_FLOAT64
    return input_type

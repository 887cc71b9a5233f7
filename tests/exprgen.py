"""Seeded random Calcite expressions over a frame (for parity tests)."""
import numpy as np

from oracle import planner as P

ARITH = ["+", "-", "*", "/", "MOD"]
CMP = ["=", "<>", "<", ">", "<=", ">="]
LOGIC = ["AND", "OR"]
UNARY_EXACT = ["NOT", "ABS", "FLOOR", "CEIL"]


def _literal(rng, want_float):
    if want_float or rng.random() < 0.3:
        return repr(round(float(rng.uniform(-50, 50)), 2))
    return str(int(rng.integers(-20, 21)))


def rand_expr(rng, dtypes, depth, allow_pow=True, nullable=()):
    """Returns a Calcite expression string (never a bare literal)."""
    n = len(dtypes)
    if depth <= 0 or rng.random() < 0.25:
        return f"${int(rng.integers(0, n))}"
    r = rng.random()
    if r < 0.12:
        op = UNARY_EXACT[int(rng.integers(0, len(UNARY_EXACT)))]
        return f"{op}({rand_expr(rng, dtypes, depth - 1, allow_pow, nullable)})"
    if r < 0.18 and nullable:
        k = int(rng.choice(nullable))
        if rng.random() < 0.5:
            return f"IS NULL(${k})" if rng.random() < 0.5 else f"IS NOT NULL(${k})"
        return f"CASE(IS NOT NULL(${k}), ${k}, {rand_expr(rng, dtypes, depth - 1, allow_pow, nullable)})"
    ops = ARITH + CMP + LOGIC + (["POWER"] if allow_pow else [])
    op = ops[int(rng.integers(0, len(ops)))]
    left = rand_expr(rng, dtypes, depth - 1, allow_pow, nullable)
    if op == "POWER":
        return f"POWER({left}, {int(rng.integers(0, 4))})"
    if rng.random() < 0.35:
        lit = _literal(rng, False)
        return f"{op}({left}, {lit})" if rng.random() < 0.7 else f"{op}({lit}, {left})"
    right = rand_expr(rng, dtypes, depth - 1, allow_pow, nullable)
    return f"{op}({left}, {right})"


def rand_frame_dtypes(rng, k):
    pool = [P.GDF_INT8, P.GDF_INT16, P.GDF_INT32, P.GDF_INT64, P.GDF_FLOAT32, P.GDF_FLOAT64]
    return [pool[int(rng.integers(0, len(pool)))] for _ in range(k)]

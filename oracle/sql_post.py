"""SQL post-stages of the reference, restated in Python — TEST INFRASTRUCTURE ONLY.

Used by oracle.run() as the checker of the device's HAVING, multi-column ORDER
BY and DISTINCT stages. Each function cites the reference lines it follows.
Values are Python ints (the reference's bigints and safe-integer numbers),
floats (numbers) or None (NULL).
"""
from __future__ import annotations

import math

TWO63 = 2.0 ** 63
MAX_SAFE = 2 ** 53 - 1


def _is_int(v) -> bool:
    return isinstance(v, int) and not isinstance(v, bool)


def _safe_int(x: float) -> bool:
    return x == math.trunc(x) and abs(x) <= MAX_SAFE if math.isfinite(x) else False


def compare(a, b) -> float:
    """compare() (src/sql/evaluator.ts:169-200) for two non-null numbers."""
    if _is_int(a) and _is_int(b):
        return -1 if a < b else (1 if a > b else 0)
    if _is_int(a):  # bigint a, number b
        b = float(b)
        if math.isnan(b):
            return -1
        if math.isinf(b):
            return -1 if b > 0 else 1
        bb = math.trunc(b)
        if a < bb:
            return -1
        if a > bb:
            return 1
        fr = b - bb
        return -1 if fr > 0 else (1 if fr < 0 else 0)
    if _is_int(b):  # number a, bigint b
        a = float(a)
        if math.isnan(a):
            return 1
        if math.isinf(a):
            return 1 if a > 0 else -1
        ab = math.trunc(a)
        if ab < b:
            return -1
        if ab > b:
            return 1
        fr = a - ab
        return 1 if fr > 0 else (-1 if fr < 0 else 0)
    a, b = float(a), float(b)
    if math.isnan(a) and math.isnan(b):
        return 0
    if math.isnan(a):
        return 1
    if math.isnan(b):
        return -1
    if math.isinf(a) and math.isinf(b) and a == b:
        return math.nan  # Infinity - Infinity
    return a - b


def loose_equal(a, b) -> bool:
    """looseEqual() (src/sql/evaluator.ts:157-167) for two non-null numbers."""
    if _is_int(a) and _is_int(b):
        return a == b
    if _is_int(a):
        b = float(b)
        return a == int(b) if _safe_int(b) else float(a) == b
    if _is_int(b):
        a = float(a)
        return int(a) == b if _safe_int(a) else a == float(b)
    return float(a) == float(b)


def _js(v):
    """A descriptor literal as the reference sees it: a bigint (plan.BigInt, or
    an int beyond 2^53) stays an int; any other number is a float."""
    if v is None or isinstance(v, float):
        return v
    if type(v).__name__ == "BigInt" or abs(int(v)) > MAX_SAFE:
        return int(v)
    return float(v)


def _cmp_test(op: str, a, b):
    if a is None or b is None:
        return None
    if op == "eq":
        return loose_equal(a, b)
    if op == "ne":
        return not loose_equal(a, b)
    d = compare(a, b)
    return {"lt": d < 0, "le": d <= 0, "gt": d > 0, "ge": d >= 0}[op]


def evaluate(expr, row: dict):
    """The HAVING tree (plan.QueryDescriptor.having) on one row with the
    evaluator's three-valued logic (src/sql/evaluator.ts:7-155): True, False
    or None (NULL)."""
    k = expr[0]
    if k == "cmp":
        _, op, col, lit = expr
        return _cmp_test(op, row.get(col), _js(lit))
    if k == "cmp_col":
        _, op, a, b = expr
        return _cmp_test(op, row.get(a), row.get(b))
    if k == "between":
        _, col, lo, hi, neg = expr
        v = row.get(col)
        if v is None:
            return None
        r = compare(v, _js(lo)) >= 0 and compare(v, _js(hi)) <= 0
        return (not r) if neg else r
    if k == "in":
        _, col, vals, neg = expr
        v = row.get(col)
        if v is None:
            return None
        found = has_null = False
        for x in vals:
            if x is None:
                has_null = True
                continue
            if loose_equal(v, _js(x)):
                found = True
                break
        if found:
            return not neg
        if has_null:
            return None
        return neg
    if k == "isnull":
        _, col, neg = expr
        return (row.get(col) is not None) if neg else (row.get(col) is None)
    if k == "and":
        a, b = evaluate(expr[1], row), evaluate(expr[2], row)
        if a is False or b is False:
            return False
        if a is None or b is None:
            return None
        return True
    if k == "or":
        a, b = evaluate(expr[1], row), evaluate(expr[2], row)
        if a is True or b is True:
            return True
        if a is None or b is None:
            return None
        return False
    if k == "not":
        a = evaluate(expr[1], row)
        return None if a is None else (not a)
    raise ValueError(f"unknown HAVING node {k}")


def having(rows: list[dict], expr) -> list[dict]:
    """HAVING filter (src/sql/executor.ts:108-112): rows whose value is truthy."""
    return [r for r in rows if evaluate(expr, r) is True]


def _multi_cmp(a: dict, b: dict, order) -> int:
    """The comparator of the multi-column ORDER BY (src/sql/executor.ts:116-139)."""
    for col, direction in order:
        av, bv = a.get(col), b.get(col)
        if (av is None and bv is None) or (av is not None and bv is not None and av == bv):
            continue  # av === bv (NaN never, -0 and 0 always)
        if av is None:
            return 1
        if bv is None:
            return -1
        if isinstance(av, float) and isinstance(bv, float):
            an, bn = math.isnan(av), math.isnan(bv)
            cmp = 0 if (an and bn) else (1 if an else (-1 if bn else (-1 if av < bv else (1 if av > bv else 0))))
        else:
            cmp = compare(av, bv)
            cmp = -1 if cmp < 0 else (1 if cmp > 0 else 0)
        if cmp != 0:
            return -cmp if direction == "desc" else cmp
    return 0


def order_by(rows: list[dict], order) -> list[dict]:
    """Stable sort (Array.prototype.sort is stable) by the multi-column comparator."""
    import functools
    return sorted(rows, key=functools.cmp_to_key(lambda a, b: _multi_cmp(a, b, order)))


def _key_str(v) -> str:
    """String(value) of the DISTINCT key (src/operators.ts:1213-1217): NULLs
    one sentinel, every NaN "NaN", -0 "0"."""
    if v is None:
        return "\x00null"
    if isinstance(v, float):
        if math.isnan(v):
            return "NaN"
        if v == 0:
            return "0"
        return repr(v)
    return str(v)


def distinct(rows: list[dict], cols: list[str]) -> list[dict]:
    """DistinctOperator (src/operators.ts:1188-1227): the first row of every
    key, in input order."""
    seen = set()
    out = []
    for r in rows:
        names = cols if cols else list(r.keys())
        key = "\x00".join(_key_str(r.get(c)) for c in names)
        if key not in seen:
            seen.add(key)
            out.append(r)
    return out

"""Result comparison helpers: integers / keys bit-exact, FP64 aggregates
within 1e-12 relative (the path's stated tolerance)."""
from __future__ import annotations

import math

REL_TOL = 1e-12


def _norm_key(v):
    if v is None:
        return (2, 0)
    if isinstance(v, float) and math.isnan(v):
        return (1, 0)
    return (0, v)


def same_value(g, w, rel=REL_TOL) -> bool:
    if w is None or g is None:
        return g is None and w is None
    if isinstance(w, float) or isinstance(g, float):
        g, w = float(g), float(w)
        if math.isnan(w):
            return math.isnan(g)
        if math.isinf(w):
            return g == w
        return math.isclose(g, w, rel_tol=rel, abs_tol=0.0) or g == w
    return g == w


def assert_rows_equal(got: list[dict], want: list[dict], float_cols=(), rel=REL_TOL, ctx=""):
    assert len(got) == len(want), f"{ctx}: {len(got)} rows vs oracle {len(want)}"
    for i, (g, w) in enumerate(zip(got, want)):
        assert set(g.keys()) == set(w.keys()), f"{ctx}: row {i} columns {sorted(g)} vs {sorted(w)}"
        for k in w:
            if k in float_cols:
                assert same_value(g[k], w[k], rel), f"{ctx}: row {i} col {k}: {g[k]!r} vs {w[k]!r}"
            else:
                gv, wv = g[k], w[k]
                if isinstance(wv, float) and isinstance(gv, float) and math.isnan(wv):
                    assert math.isnan(gv), f"{ctx}: row {i} col {k}: {gv!r} vs NaN"
                else:
                    assert gv == wv, f"{ctx}: row {i} col {k}: {gv!r} vs {wv!r}"


def assert_groups_equal(got: list[dict], want: list[dict], keys, float_cols=(), rel=REL_TOL, ctx=""):
    """Group-by results as multisets: the reference's group order is Map
    insertion order, the device's is table-slot order."""
    kf = lambda r: tuple(_norm_key(r[k]) for k in keys)  # noqa: E731
    assert_rows_equal(sorted(got, key=kf), sorted(want, key=kf), float_cols, rel, ctx)

"""The knob search of querymode_amd.autotune (no device: a synthetic cost)."""
from __future__ import annotations

import json

from querymode_amd.autotune import coordinate_descent, load_cache, save_choice


def test_descent_finds_the_separable_minimum():
    space = {"a": [1, 2, 4, 8, 16], "b": [0, 1, 2]}
    cost = lambda c: (c["a"] - 4) ** 2 + 10 * (c["b"] - 2) ** 2 + 1.0  # noqa: E731
    best, ms, trials = coordinate_descent(cost, {"a": 16, "b": 0}, space)
    assert best == {"a": 4, "b": 2} and ms == 1.0
    assert trials[0] == ({"a": 16, "b": 0}, cost({"a": 16, "b": 0}))


def test_descent_skips_rejected_configs():
    space = {"a": [1, 2, 3]}
    cost = lambda c: None if c["a"] == 3 else 10.0 - c["a"]  # noqa: E731  (3 would be fastest but is rejected)
    best, ms, _ = coordinate_descent(cost, {"a": 1}, space)
    assert best == {"a": 2} and ms == 8.0


def test_descent_ignores_noise_level_gains():
    space = {"a": [1, 2]}
    best, _, _ = coordinate_descent(lambda c: 100.0 - 0.1 * c["a"], {"a": 1}, space)
    assert best == {"a": 1}  # 0.1% faster is within noise


def test_cache_roundtrip(tmp_path):
    f = tmp_path / "tune.json"
    save_choice(f, "gfx950|rows=10|n=2|abc", {"runs_chunks_per_cu": 32})
    save_choice(f, "gfx950|rows=20|n=2|def", {"runs_win": 2048})
    data = load_cache(f)
    assert data["gfx950|rows=10|n=2|abc"] == {"runs_chunks_per_cu": 32}
    assert json.loads(f.read_text())["gfx950|rows=20|n=2|def"] == {"runs_win": 2048}

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


def test_space_follows_the_strategies():
    from querymode_amd.autotune import BATCH_SPACES, QUERY_SPACES, space_for
    sp = space_for(["part_runs", "shared_interval", "part_runs"], BATCH_SPACES)
    assert list(sp) == ["pr_scatter_bpc", "pr_b2", "pr_max_s", "shared_waves"]
    assert space_for(["sorted_runs"], BATCH_SPACES) == BATCH_SPACES["sorted_runs"]
    assert space_for(["direct_lds"], BATCH_SPACES) == {}        # a fused scan has no knob here
    assert "topk_bpc" in space_for(["topk"], QUERY_SPACES)
    # every value of every space lies in the knob's range (qmgpu.h clamps; the tuner must not rely on it)
    ranges = {"pr_scatter_bpc": (1, 8), "pr_b2": (0, 10), "pr_max_s": (1, 6), "shared_waves": (1, 64),
              "runs_chunks_per_cu": (1, 1024), "runs_qblock": (0, 64), "runs_win": (1024, 2048),
              "runs_min_len": (0, 1e12), "scatter_rpb": (0, 64), "part_load": (0.1, 0.9), "hot": (0, 1),
              "topk_bpc": (1, 16)}
    for spaces in (BATCH_SPACES, QUERY_SPACES):
        for sp in spaces.values():
            for k, vals in sp.items():
                lo, hi = ranges[k]
                assert all(lo <= v <= hi for v in vals), k

"""Writes pylance_filters.json from the reference's committed PyLance SQL
fixtures (wasm/tests/conformance/fixtures/*.expected.json), read as JSON data.

The fixtures hold each query's result rows. The 10-row input table is
recovered from those results alone: the distinct (a, b, c, d, e) tuples, in
the table order implied by the fixtures (each result lists rows in table
order; the orders are merged topologically). Each vector stores the WHERE
text and the ids of the rows it returned (or that the SQL engine rejected it).

Run in a checkout that has the reference mounted:
    python tests/golden/make_pylance_vectors.py /root/reference
"""
from __future__ import annotations

import glob
import json
import os
import sys
from graphlib import TopologicalSorter

COLS = ["a", "b", "c", "d", "e"]


def main(ref_root: str) -> None:
    fdir = os.path.join(ref_root, "wasm", "tests", "conformance", "fixtures")
    fixtures = []
    for path in sorted(glob.glob(os.path.join(fdir, "*.expected.json"))):
        with open(path) as f:
            fixtures.append(json.load(f))
    results = []
    for fx in fixtures:
        r = fx["result"]
        if r.get("success"):
            rows = [tuple(r["data"][c][i] for c in COLS) for i in range(r["row_count"])]
        else:
            rows = None
        results.append((fx["name"], fx["query"], rows, r.get("error")))
    distinct = sorted({t for _, _, rows, _ in results if rows for t in rows}, key=repr)
    ts = TopologicalSorter({t: set() for t in distinct})
    for _, _, rows, _ in results:
        for a, b in zip(rows or [], (rows or [])[1:]):
            ts.add(b, a)
    order = list(ts.static_order())
    index = {t: i for i, t in enumerate(order)}
    table = {c: [t[k] for t in order] for k, c in enumerate(COLS)}
    vectors = []
    for name, query, rows, err in results:
        v = {"name": name, "where": query}
        if rows is None:
            v["engine_error"] = (err or "")[:120]
        else:
            v["rows"] = sorted(index[t] for t in rows)
        vectors.append(v)
    out = {"source": "wasm/tests/conformance/fixtures/*.expected.json (PyLance / DataFusion SQL results)",
           "table": table, "vectors": vectors}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pylance_filters.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {dst}: {len(order)} table rows, {len(vectors)} vectors")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

"""Extract the reference's committed golden vectors for the search path into compact
fixtures under tests/golden/.

Run in a checkout that has the reference mounted (read as data only; nothing is
imported or executed from it):

    python tests/golden/make_fixtures.py /root/reference

Sources (reference test files that generated them):
  diskann/test/generated/graph/test/cases/grid_search/search_*.json
      (diskann/src/graph/test/cases/grid_search.rs:88-218)
  diskann/test/generated/graph/test/cases/grid_insert/*/*.json
      (diskann/src/graph/test/cases/grid_insert.rs)
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _search_rows(payload):
    rows = []
    for s in payload:
        rows.append({
            "query": s["query"],
            "beam_width": s["beam_width"],
            "comparisons": s["comparisons"],
            "hops": s["hops"],
            "num_results": s["num_results"],
            "results": s["results"],
        })
    return rows


def main(ref):
    base = os.path.join(ref, "diskann/test/generated/graph/test/cases")
    out = {"grid_search": {}, "grid_insert": {}}
    gs = os.path.join(base, "grid_search")
    for name in sorted(os.listdir(gs)):
        with open(os.path.join(gs, name)) as f:
            d = json.load(f)
        p = d["payload"]
        out["grid_search"][name[:-5]] = {
            "grid_dims": p[0]["grid_dims"],
            "grid_size": p[0]["grid_size"],
            "k": 10,
            "L": 10,
            "searches": _search_rows(p),
        }
    gi = os.path.join(base, "grid_insert")
    for case in sorted(os.listdir(gi)):
        for name in sorted(os.listdir(os.path.join(gi, case))):
            with open(os.path.join(gi, case, name)) as f:
                d = json.load(f)
            p = d["payload"]
            out["grid_insert"][case + "/" + name[:-5]] = {
                "grid_dims": p["grid_dims"],
                "grid_size": p["grid_size"],
                "num_inserted": p["num_inserted"],
                "insert_metrics": p["insert_metrics"],
                "searches": _search_rows(p["searches"]),
            }
    with open(os.path.join(HERE, "grid_cases.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

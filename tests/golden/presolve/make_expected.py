"""Writes expected_counts.json: the reduction counts, status and reduced sizes the
reference's own presolve (src/presolve/presolve.cpp, compiled from its sources into
oracle/_ref by oracle/Makefile) reports for every model fixture. Run in the build container:

    python tests/golden/presolve/make_expected.py
"""
import glob
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

GOLDEN = os.path.dirname(HERE)


def fixtures():
    files = sorted(glob.glob(os.path.join(GOLDEN, "netlib", "*.mps"))) + sorted(glob.glob(os.path.join(GOLDEN, "*.mps")))
    files += [os.path.join(GOLDEN, "crude_blend.lp")] + sorted(glob.glob(os.path.join(HERE, "*.mps")))
    return files


def summary(dump_text):
    out = {}
    for line in dump_text.splitlines():
        k, _, rest = line.partition(" ")
        if k == "counts":
            v = rest.split()
            out["counts"] = [int(x) for x in v[:8]]
        elif k == "infeasible":
            v = rest.split()
            out["infeasible"], out["unbounded"] = int(v[0]), int(v[2])
        elif k == "message":
            out["message"] = rest
        elif k == "reduced":
            v = rest.split()
            out["reduced"] = [int(v[0]), int(v[1]), int(v[3])]
    return out


def main():
    res = {}
    with tempfile.TemporaryDirectory() as d:
        for f in fixtures():
            out = os.path.join(d, "dump.txt")
            O.ref_presolve_dump(f, out)
            res[os.path.relpath(f, GOLDEN)] = summary(open(out).read())
    json.dump({"what": "reference PresolveEngine::apply on each fixture: counts in the order empty rows, empty "
                       "columns, fixed columns, singleton rows, forcing rows, redundant rows, free column "
                       "singletons, doubleton rows; reduced = rows, columns, nonzeros",
               "models": res}, open(os.path.join(HERE, "expected_counts.json"), "w"), indent=1)
    print("wrote", len(res), "models")


if __name__ == "__main__":
    main()

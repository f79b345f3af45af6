"""Write ambiguity_cases.json: the track patterns, p-values, configurations
and expected surviving patterns of traccc's greedy ambiguity resolution
unit tests (tests/cuda/test_ambiguity_resolution.cpp:103-772, which repeat
tests/cpu/test_ambiguity_resolution.cpp:79-424 and add Test7-Test17).

The patterns are measurement identifiers 0..100 (each test fills identifiers
0..100, identifier = index).  Run from the repository root, with the
reference checked out at /root/reference:

    python tests/golden/make_ambiguity_cases.py
"""
import json
import os
import re
import sys

SRC = "/root/reference/tests/cuda/test_ambiguity_resolution.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ambiguity_cases.json")


def ints(s):
    return [int(x) for x in re.findall(r"\d+", s)]


def main() -> int:
    text = open(SRC).read()
    blocks = re.split(r"\nTEST\(CUDAAmbiguitySolverTests, (GreedyResolverTest\d+)\)", text)
    cases = []
    for name, body in zip(blocks[1::2], blocks[2::2]):
        tracks = [{"pval": float(p), "ids": ints(ids)} for p, ids in
                  re.findall(r"fill_pattern\(trk_cands, ([0-9.]+)f, \{([^}]*)\}\)", body)]
        # a test body runs the algorithm once per scope that asserts a size;
        # a min_meas_per_track assignment before it applies to that run
        runs = []
        min_meas = 3
        for m in re.finditer(r"min_meas_per_track = (\d+)|tracks\.size\(\), (\d+)u\)|"
                             r"find_pattern\(res_trk_cands, \{([^}]*)\}\)", body):
            if m.group(1):
                min_meas = int(m.group(1))
            elif m.group(2):
                runs.append({"min_meas_per_track": min_meas, "n_expected": int(m.group(2)),
                             "expected": []})
            else:
                runs[-1]["expected"].append(ints(m.group(3)))
        for k, r in enumerate(runs):
            cases.append({"name": f"{name}" + (f"_run{k}" if len(runs) > 1 else ""),
                          "tracks": tracks, **r})
    if not cases:
        print("no cases found", file=sys.stderr)
        return 1
    with open(OUT, "w") as fh:
        json.dump({"source": "tests/cuda/test_ambiguity_resolution.cpp", "cases": cases}, fh,
                  indent=1)
    print(f"{len(cases)} cases -> {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Writes ultralong50k_100k_biwfa.json: the oracle's BiWFA (ultralow memory)
results on the first pairs of the 1k x 50-100 kbp workload (seed 0,
divergence 5-15%, 20% of the edits in indel runs of geometric mean 20):
edit distance on 4 pairs and gap-affine (0,4,6,2) on 2, full CIGAR.

For each pair it records the lengths, score, status, the CIGAR's length,
per-operation counts and gap openings, and a SHA-256 of the CIGAR's operation bytes, so the
device path can be compared at full size without shipping 100 kbp CIGARs.
The oracle follows the reference's align_biwfa (src/aligner.rs:2563-2650,
src/bialign.rs); it takes about a minute on one core.

    python tests/golden/make_ultralong_fixture.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

CASES = (("edit", (0, 0, 0, 0, 0, 0), 4), ("gap_affine", (0, 4, 6, 2, 0, 0), 2))
OUT = os.path.join(HERE, "ultralong50k_100k_biwfa.json")


def digest(ops: bytes) -> dict:
    """CIGAR length, per-operation counts, maximal I and D runs (gap
    openings) and SHA-256 of the operation bytes."""
    runs = sum(1 for i, c in enumerate(ops) if c in b"ID" and (i == 0 or ops[i - 1] != c))
    return {"cigar_len": len(ops),
            "counts": {c: ops.count(c.encode()) for c in "MXID"},
            "gap_runs": runs,
            "sha256": hashlib.sha256(ops).hexdigest()}


def main():
    import __graft_entry__ as entry
    entry.load_package()
    from conftest import attr_for
    import wfa2lib_rs_amd._native as N
    from wfa2lib_rs_amd.synth import WORKLOADS, generate
    oracle = entry.load_oracle()
    w = WORKLOADS["ultralong50k_100k"]
    pairs = generate(w, num_pairs=max(n for _, _, n in CASES))
    rows = []
    for metric, pen, n in CASES:
        a = attr_for(N, metric, pen, "alignment")
        a.memory_mode = N.MEMORY_ULTRALOW
        A = oracle.OracleAligner(a)
        for i in range(n):
            s, st, ops = A.align(pairs.pattern(i), pairs.text(i))
            row = {"metric": metric, "penalties": list(pen), "pair": i,
                   "plen": int(pairs.plen[i]), "tlen": int(pairs.tlen[i]),
                   "score": int(s), "status": int(st)}
            row.update(digest(ops))
            rows.append(row)
            print(metric, i, row["score"], row["cigar_len"], flush=True)
    with open(OUT, "w") as f:
        json.dump({"workload": w.name, "seed": w.seed, "memory_mode": "ultralow",
                   "scope": "alignment", "rows": rows}, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()

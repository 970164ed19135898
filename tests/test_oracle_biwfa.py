"""The oracle's BiWFA (ultralow memory) against the reference's own BiWFA
tests, an exact DP and the unidirectional aligner.

Literal vectors (tests/cli_output.rs:280-345, src/aligner.rs:4480-4507) are
in golden/reference_vectors.json and run in test_oracle.py."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, attr_for
from test_oracle import _cigar_wf_score, _dp_score

BIWFA_PENALTIES = [
    ("edit", (0, 0, 0, 0, 0, 0)),
    ("indel", (0, 0, 0, 0, 0, 0)),
    ("gap_affine", (0, 5, 3, 2, 0, 0)),
    ("gap_affine", (-3, 5, 3, 2, 0, 0)),
    ("gap_affine_2p", (0, 4, 6, 2, 24, 1)),
    ("gap_affine_2p", (-1, 4, 6, 2, 24, 1)),
    ("gap_linear", (0, 4, 2, 0, 0, 0)),
    ("gap_linear", (-1, 4, 2, 0, 0, 0)),
]


def _biwfa(N, oracle, metric, pen, scope="alignment", heuristic=None):
    attr = attr_for(N, metric, pen, scope)
    attr.memory_mode = N.MEMORY_ULTRALOW
    if heuristic:
        attr.heuristic = heuristic[0]
        for i, v in enumerate(heuristic[1:]):
            attr.heuristic_p[i] = v
    return oracle.OracleAligner(attr)


def _biwfa_test_pair():
    """biwfa_test_pair (src/aligner.rs:4568-4584): 300 bp, three edits."""
    x = 12345
    pat = bytearray()
    for _ in range(300):
        x = (x * 1103515245 + 12345) % 2 ** 32
        pat.append(b"ACGT"[(x >> 16) % 4])
    txt = bytearray(pat)
    txt[50] = ord("C") if txt[50] == ord("A") else ord("A")
    del txt[120]
    txt.insert(200, ord("G"))
    return bytes(pat), bytes(txt)


@pytest.mark.parametrize("metric,pen", BIWFA_PENALTIES)
def test_biwfa_scores_match_end2end(metric, pen, pkg, N, oracle):
    """check_biwfa_scores (src/aligner.rs:4431-4477): score, CIGAR score and
    status equal unidirectional WFA's, also with an empty sequence."""
    uni = oracle.OracleAligner(attr_for(N, metric, pen, "alignment"))
    bi = _biwfa(N, oracle, metric, pen)
    for p, t in [(b"ACGT", b""), (b"", b"TG"), (b"ACGT", b"AGT"), (b"GATTACA", b"GATCACAT")]:
        want = uni.align_end2end(p, t)[0]
        score, status, ops = bi.align(p, t)
        assert score == want, (p, t)
        assert bi.cigar_score() == want, (p, t)
        assert status == N.STATUS_COMPLETED
        assert pkg.Cigar(ops).check_alignment(p, t) is None


def test_biwfa_score_only_matches_end2end(N, oracle):
    """test_biwfa_score_matches_end2end (src/aligner.rs:4628-4642)."""
    p, t = _biwfa_test_pair()
    pen = (0, 4, 6, 2, 0, 0)
    score, status, _ = _biwfa(N, oracle, "gap_affine", pen, "score").align(p, t)
    assert status == N.STATUS_COMPLETED
    assert score == oracle.OracleAligner(attr_for(N, "gap_affine", pen, "score")).align(p, t)[0]


def test_biwfa_band_excluding_end_is_unreachable(pkg, N, oracle):
    """test_biwfa_band_excluding_end_is_unreachable (src/aligner.rs:4644-4679)."""
    p, _ = _biwfa_test_pair()
    t = p[:270]
    pen = (0, 4, 6, 2, 0, 0)
    band = (N.HEURISTIC_BANDED_STATIC, -5, 5)
    _, status, ops = _biwfa(N, oracle, "gap_affine", pen, heuristic=band).align(p, t)
    assert status == N.STATUS_END_UNREACHABLE and ops == b""
    _, status, _ = _biwfa(N, oracle, "gap_affine", pen, "score", heuristic=band).align(p, t)
    assert status == N.STATUS_END_UNREACHABLE
    _, status, ops = _biwfa(N, oracle, "gap_affine", pen,
                            heuristic=(N.HEURISTIC_BANDED_STATIC, -40, 5)).align(p, t)
    assert status == N.STATUS_COMPLETED
    assert pkg.Cigar(ops).check_alignment(p, t) is None


def _recursive_pairs(n, seed):
    """Pairs long and divergent enough that BiWFA splits them (score above
    the 250 fallback) at least once."""
    from wfa2lib_rs_amd.synth import Workload, generate
    import wfa2lib_rs_amd._native as Nn
    w = Workload("biwfa_mix", n, 1, 300, 2500, 0.04, 0.20, 3.0, run_frac=0.1, run_mean=12.0,
                 seq_format=Nn.SEQ_ASCII, seed=seed)
    hp = generate(w)
    return [(hp.pattern(i), hp.text(i)) for i in range(n)]


@pytest.mark.parametrize("metric,pen", BIWFA_PENALTIES)
def test_biwfa_recursion_optimal_and_valid(metric, pen, pkg, N, oracle):
    """Split alignments: the BiWFA score is optimal (exact DP where it
    applies, else the unidirectional aligner), the CIGAR is valid and scores
    what was returned, and score-only BiWFA returns the same score."""
    pairs = _recursive_pairs(12, seed=3)
    bi = _biwfa(N, oracle, metric, pen)
    bis = _biwfa(N, oracle, metric, pen, "score")
    uni = oracle.OracleAligner(attr_for(N, metric, pen, "score"))
    m = pen[0]
    split = 0
    for p, t in pairs:
        score, status, ops = bi.align(p, t)
        assert status == N.STATUS_COMPLETED
        if m == 0 or metric == "gap_linear":
            want = _dp_score(oracle, metric, pen, p, t)
        else:
            want = uni.align(p, t)[0]
        assert score == want
        assert bis.align(p, t)[0] == want
        cig = pkg.Cigar(ops)
        assert cig.check_alignment(p, t) is None
        assert bi.cigar_score() == score
        assert _cigar_wf_score(pkg, metric, pen, cig, len(p), len(t)) == score
        split += score > 250
    assert split >= 4, "the pairs must exercise the breakpoint search"


def test_biwfa_rejects_ends_free_and_zdrop(N, oracle):
    """BiWFA is end-to-end, and Z-drop is refused with it (C since 7110ee3,
    and the driver's checks, src/bin/align_benchmark.rs:529-560); X-drop
    sub-alignments run (tests/xdrop_steps.rs:57-92)."""
    attr = attr_for(N, "gap_affine", (0, 4, 6, 2, 0, 0), "alignment", span="extension")
    attr.memory_mode = N.MEMORY_ULTRALOW
    with pytest.raises(ValueError):
        oracle.OracleAligner(attr)
    attr = attr_for(N, "gap_affine", (0, 4, 6, 2, 0, 0), "alignment")
    attr.memory_mode = N.MEMORY_ULTRALOW
    attr.heuristic = N.HEURISTIC_ZDROP
    with pytest.raises(ValueError):
        oracle.OracleAligner(attr)
    attr.heuristic = N.HEURISTIC_XDROP
    oracle.OracleAligner(attr)


def _ultralong_fixture():
    import json
    with open(os.path.join(GOLDEN, "ultralong50k_100k_biwfa.json")) as f:
        return json.load(f)


def test_ultralong_fixture_is_consistent(pkg, N):
    """tests/golden/ultralong50k_100k_biwfa.json (make_ultralong_fixture.py):
    the pairs are the generator's (lengths), each CIGAR spans its pair
    (M+X+D = pattern, M+X+I = text) and scores its score under the metric
    (edit: X+I+D; gap-affine (0,4,6,2): 4X + 6 openings + 2 gap bases)."""
    from wfa2lib_rs_amd.synth import WORKLOADS, generate
    d = _ultralong_fixture()
    n = max(r["pair"] for r in d["rows"]) + 1
    pairs = generate(WORKLOADS[d["workload"]], num_pairs=n)
    for r in d["rows"]:
        i, c = r["pair"], r["counts"]
        assert (r["plen"], r["tlen"]) == (pairs.plen[i], pairs.tlen[i])
        assert c["M"] + c["X"] + c["D"] == r["plen"] and c["M"] + c["X"] + c["I"] == r["tlen"]
        assert sum(c.values()) == r["cigar_len"]
        gaps = c["I"] + c["D"]
        want = c["X"] + gaps if r["metric"] == "edit" else 4 * c["X"] + 6 * r["gap_runs"] + 2 * gaps
        assert want == r["score"], r
        assert r["status"] == 5  # COMPLETED, as align_biwfa reports


def test_ultralong_fixture_first_pair_on_oracle(pkg, N, oracle):
    """The oracle reproduces the fixture's first edit-distance row (a 60 kbp
    pair, about 3 s of BiWFA on one core)."""
    import hashlib
    from wfa2lib_rs_amd.synth import WORKLOADS, generate
    d = _ultralong_fixture()
    r = d["rows"][0]
    pairs = generate(WORKLOADS[d["workload"]], num_pairs=1)
    a = attr_for(N, r["metric"], r["penalties"], "alignment")
    a.memory_mode = N.MEMORY_ULTRALOW
    s, st, ops = oracle.OracleAligner(a).align(pairs.pattern(0), pairs.text(0))
    assert (s, st, len(ops)) == (r["score"], r["status"], r["cigar_len"])
    assert hashlib.sha256(ops).hexdigest() == r["sha256"]


@pytest.mark.parametrize("metric,pen", [BIWFA_PENALTIES[2], BIWFA_PENALTIES[3], BIWFA_PENALTIES[4],
                                        BIWFA_PENALTIES[7]])
def test_biwfa_xdrop_reports_cigar_score(metric, pen, pkg, N, oracle):
    """X-drop sub-alignments in BiWFA (tests/xdrop_steps.rs:57-92 runs
    align_biwfa with X-drop): a completed alignment has a valid CIGAR and
    reports that CIGAR's score (C 5713de8, src/aligner.rs:2607-2616);
    score-only BiWFA ends with the same status."""
    pairs = _recursive_pairs(12, seed=3)
    bi = _biwfa(N, oracle, metric, pen, heuristic=(N.HEURISTIC_XDROP, 50, 1))
    done = 0
    for p, t in pairs:
        score, status, ops = bi.align(p, t)
        if status != N.STATUS_COMPLETED:
            continue
        cig = pkg.Cigar(ops)
        assert cig.check_alignment(p, t) is None
        assert _cigar_wf_score(pkg, metric, pen, cig, len(p), len(t)) == score
        done += 1
    assert done >= 6

"""Heuristic trajectories of the oracle against the reference's own
integration tests (CPU only).

* tests/heuristic_scores.rs:104-158: 120 LCG pairs (seed 0x68657572) x 3
  penalty sets x {WF-adaptive, X-drop, Z-drop, adaptive band} x {high,
  medium}: a completed heuristic alignment reports its CIGAR's WFA score; a
  Z-dropped one scores its max-trimmed kept part; BiWFA with WF-adaptive and
  the adaptive band reports its CIGAR's score too.
* tests/xdrop_steps.rs:57-92 (HPCBio/wfa2lib-rs#21, the pair in
  tests/data/issue21_xdrop.seq): X-drop 20 every 2 steps gives the same
  valid CIGAR and score in both compat modes, unidirectional and BiWFA,
  and the CIGAR scores at least the score's Smith-Waterman equivalent.
"""
import os

import pytest

from conftest import GOLDEN, attr_for

STATUS_COMPLETED, STATUS_PARTIAL = 5, 6
WFADAPTIVE, XDROP, ZDROP, BANDED_ADAPTIVE = 1, 2, 3, 5


class Lcg:
    """The generator of tests/heuristic_scores.rs:13-43."""

    def __init__(self, seed):
        self.s = seed

    def next(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        return self.s >> 33

    def below(self, n):
        return self.next() % n

    def base(self):
        return b"ACGT"[self.below(4)]

    def pair(self, max_len):
        n = 20 + self.below(max_len)
        pattern = bytes(self.base() for _ in range(n))
        text = bytearray()
        for b in pattern:
            r = self.below(20)
            if r == 0:
                text.append(self.base())
            elif r == 1:
                text += bytes([self.base(), b])
            elif r == 2:
                pass
            else:
                text.append(b)
        return pattern, bytes(text)


PENALTY_SETS = [("gap_affine", (0, 4, 6, 2, 0, 0)), ("gap_affine", (-3, 5, 3, 2, 0, 0)),
                ("gap_affine_2p", (0, 4, 6, 2, 24, 1))]
HEURISTICS = [(WFADAPTIVE, (10, 50, 1)), (XDROP, (20, 1, 0)), (ZDROP, (20, 1, 0)),
              (BANDED_ADAPTIVE, (-5, 5, 1))]


def wfa_score(pkg, metric, pen, ops, v, h):
    """The WFA score of a CIGAR spanning v pattern and h text bases
    (tests/heuristic_scores.rs:91-103)."""
    m, x, o1, e1, o2, e2 = pen
    c = pkg.Cigar(ops)
    if metric == "gap_affine":
        classic = c.score_gap_affine(pkg.AffinePenalties(m, x, o1, e1))
    else:
        classic = c.score_gap_affine2p(pkg.Affine2pPenalties(m, x, o1, e1, o2, e2))
    return -m * (v + h) - 2 * classic if m < 0 else -classic


def spans(ops):
    return sum(op != ord("I") for op in ops), sum(op != ord("D") for op in ops)


def with_heuristic(N, metric, pen, heuristic, hp, memory):
    a = attr_for(N, metric, pen, "alignment")
    a.heuristic = heuristic
    for i, v in enumerate(hp):
        a.heuristic_p[i] = v
    a.memory_mode = memory
    return a


def test_heuristic_scores_are_cigar_scores(pkg, N, oracle):
    rng = Lcg(0x68657572)
    pairs = [rng.pair(2000 if i % 4 == 0 else 300) for i in range(120)]
    aligners = {}
    for metric, pen in PENALTY_SETS:
        for h, hp in HEURISTICS:
            for memory in (0, 1, 3):
                if memory == 3 and h in (XDROP, ZDROP):
                    continue  # not run with BiWFA (tests/heuristic_scores.rs:132-134)
                aligners[(metric, pen, h, memory)] = oracle.OracleAligner(
                    with_heuristic(N, metric, pen, h, hp, memory))
    checked = partial = 0
    for p, t in pairs:
        for (metric, pen, h, memory), A in aligners.items():
            s, st, ops = A.align(p, t)
            ctx = (metric, pen, h, memory, len(p), len(t))
            if st == STATUS_COMPLETED:
                expected = wfa_score(pkg, metric, pen, ops, len(p), len(t))
                assert s == expected, ctx
                assert A.cigar_score() == expected, ctx
                checked += 1
            elif st == STATUS_PARTIAL and h == ZDROP and memory != 3 and ops:
                v, hh = spans(ops)
                assert A.cigar_score() == wfa_score(pkg, metric, pen, ops, v, hh), ctx
                partial += 1
    assert checked > 3000  # (Z-drop 20 never fires on these pairs)


def _issue21_pair():
    with open(os.path.join(GOLDEN, "issue21_xdrop.seq"), "rb") as f:
        lines = f.read().split(b"\n")
    return lines[0].lstrip(b">"), lines[1].lstrip(b"<")


@pytest.mark.parametrize("metric,pen", [("gap_affine", (-3, 5, 3, 2, 0, 0)),
                                        ("gap_linear", (-1, 3, 2, 0, 0, 0)),
                                        ("gap_affine_2p", (-1, 4, 6, 2, 24, 1))])
def test_issue21_xdrop_steps(metric, pen, pkg, N, oracle):
    p, t = _issue21_pair()
    assert (len(p), len(t)) == (78, 100)
    m, x, o1, e1, o2, e2 = pen
    if metric == "gap_affine":
        rescore = lambda c: c.score_gap_affine(pkg.AffinePenalties(m, x, o1, e1))
    elif metric == "gap_linear":
        rescore = lambda c: c.score_gap_linear(pkg.LinearPenalties(m, x, o1))
    else:
        rescore = lambda c: c.score_gap_affine2p(pkg.Affine2pPenalties(m, x, o1, e1, o2, e2))
    for memory in (0, 3):  # unidirectional (high memory) and BiWFA
        out = []
        for compat in (0, 1):
            a = with_heuristic(N, metric, pen, XDROP, (20, 2, 0), memory)
            a.compat = compat
            s, st, ops = oracle.OracleAligner(a).align(p, t)
            assert oracle.lib(N.Attr).wfo_cigar_check(ops, len(ops), p, len(p), t,
                                                      len(t)) == 0, (memory, compat)
            out.append((s, pkg.Cigar(ops).to_string_rle(True)))
            # wf_score_to_sw_score (src/penalties.rs:300-308)
            sw = (-m * (len(p) + len(t)) - s) // 2 if -m * (len(p) + len(t)) - s >= 0 else \
                -((s + m * (len(p) + len(t))) // 2)
            assert rescore(pkg.Cigar(ops)) >= sw, (memory, compat)
        assert out[0] == out[1], memory

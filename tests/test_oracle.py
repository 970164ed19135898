"""CPU tests: the oracle pinned against the reference's literal vectors, an
exact Gotoh DP, CIGAR validity/rescoring invariants, and the generator."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, attr_for, attr_for_vec, expand, load_vectors

STATUS = {"OK": 0, "END_REACHED": 1, "END_UNREACHABLE": 2, "MAX_STEPS_REACHED": 3, "OOM": 4,
          "COMPLETED": 5, "PARTIAL": 6}


def _penalty_obj(pkg, metric, pen):
    m, x, o1, e1, o2, e2 = pen
    if metric == "gap_affine":
        return pkg.WavefrontPenalties.new_affine(pkg.AffinePenalties(m, x, o1, e1))
    if metric == "gap_affine_2p":
        return pkg.WavefrontPenalties.new_affine2p(pkg.Affine2pPenalties(m, x, o1, e1, o2, e2))
    if metric == "gap_linear":
        return pkg.WavefrontPenalties.new_linear(pkg.LinearPenalties(m, x, o1))
    if metric == "edit":
        return pkg.WavefrontPenalties.new_edit()
    return pkg.WavefrontPenalties.new_indel()


def _dp_score(oracle, metric, pen, p, t):
    """The optimal score from the exact DP, in the aligner's internal scale:
    with a match reward m < 0 the penalties are Eizenga-transformed
    (src/penalties.rs:156-296) and the DP charges matches nothing."""
    m, x, o1, e1, o2, e2 = pen
    if metric == "edit":
        return oracle.dp_edit(p, t)
    if metric == "indel":
        return oracle.dp_edit(p, t, indel=True)
    if metric == "gap_linear":
        # a gap of length l costs l * indel: affine with o = 0, e = indel
        if m < 0:
            return oracle.dp_affine(p, t, 2 * x - 2 * m, 0, 2 * o1 - m)
        return oracle.dp_affine(p, t, x, 0, o1)
    if m < 0:
        return oracle.dp_affine(p, t, 2 * x - 2 * m, 2 * o1, 2 * e1 - m, 2 * o2, 2 * e2 - m,
                                metric == "gap_affine_2p")
    return oracle.dp_affine(p, t, x, o1, e1, o2, e2, metric == "gap_affine_2p")


def _cigar_wf_score(pkg, metric, pen, cig, plen, tlen):
    """A CIGAR's score on the user penalties, in the aligner's internal scale
    (cigar_wfa_score, src/aligner.rs:2700-2719)."""
    m, x, o1, e1, o2, e2 = pen
    if metric in ("edit", "indel"):
        return cig.score_edit()
    if metric == "gap_affine":
        classic = cig.score_gap_affine(pkg.AffinePenalties(m, x, o1, e1))
    elif metric == "gap_affine_2p":
        classic = cig.score_gap_affine2p(pkg.Affine2pPenalties(m, x, o1, e1, o2, e2))
    else:
        classic = cig.score_gap_linear(pkg.LinearPenalties(m, x, o1))
    return -m * (plen + tlen) - 2 * classic if m < 0 else -classic


@pytest.mark.parametrize("vec", load_vectors(), ids=lambda v: v["id"])
def test_oracle_reference_vectors(vec, pkg, N, oracle):
    attr = attr_for_vec(N, vec)
    A = oracle.OracleAligner(attr)
    p, t = expand(vec["pattern"]), expand(vec["text"])
    score, status, ops = A.align(p, t)
    cig = pkg.Cigar(ops)
    if "score" in vec:
        assert score == vec["score"]
    if "score_at_most" in vec:
        assert score <= vec["score_at_most"]
    if "status" in vec:
        assert status == STATUS[vec["status"]]
    if "cigar" in vec:
        assert cig.to_string_rle(True) == vec["cigar"]
    if vec.get("valid_cigar"):
        assert cig.check_alignment(p, t) is None
    if "cli" in vec:
        pen = _penalty_obj(pkg, vec["metric"], vec["penalties"])
        score_only = vec["scope"] == "score"
        out = pkg.compute_output_score(cig, score, pen, score_only, len(p), len(t))
        rle = "-" if score_only or cig.is_null() else cig.to_string_rle(True)
        assert f"{out}\t{rle}" == vec["cli"]


def _lcg_pairs(n, max_len, seed=7):
    """The reference's singletrack LCG pair recipe (tests/singletrack.rs:12-51)."""
    state = [seed]

    def nxt():
        state[0] = (state[0] * 6364136223846793005 + 1442695040888963407) % 2 ** 64
        return state[0] >> 33

    def below(k):
        return nxt() % k

    def base():
        return b"ACGT"[below(4)]

    pairs = []
    for _ in range(n):
        ln = 1 + below(max_len)
        pat = bytes(base() for _ in range(ln))
        txt = bytearray()
        i = 0
        while i < len(pat):
            r = below(40)
            if r == 0:
                txt.append(base())
            elif r == 1:
                mx = 12 if below(4) == 0 else 3
                for _ in range(1 + below(mx)):
                    txt.append(base())
                txt.append(pat[i])
            elif r == 2:
                mx = 12 if below(4) == 0 else 3
                i += below(mx)
            else:
                txt.append(pat[i])
            i += 1
        pairs.append((pat, bytes(txt)))
    return pairs


PENALTY_SETS = [
    ("gap_affine", (0, 4, 6, 2, 0, 0)),
    ("gap_affine", (0, 1, 0, 1, 0, 0)),
    ("gap_affine", (0, 3, 5, 1, 0, 0)),
    ("gap_affine_2p", (0, 4, 6, 2, 24, 1)),
    ("gap_affine_2p", (0, 5, 4, 3, 12, 1)),
    ("edit", (0, 0, 0, 0, 0, 0)),
    ("indel", (0, 0, 0, 0, 0, 0)),
    ("gap_linear", (0, 4, 2, 0, 0, 0)),
    ("gap_linear", (0, 3, 5, 0, 0, 0)),
    ("gap_linear", (-1, 3, 2, 0, 0, 0)),
    ("gap_affine", (-3, 5, 3, 2, 0, 0)),
    ("gap_affine_2p", (-1, 4, 6, 2, 24, 1)),
]


@pytest.mark.parametrize("metric,pen", PENALTY_SETS)
def test_oracle_scores_match_exact_dp(metric, pen, pkg, N, oracle):
    """Optimal scores from an independent Gotoh / Levenshtein DP."""
    attr = attr_for(N, metric, pen, "alignment")
    A = oracle.OracleAligner(attr)
    for p, t in _lcg_pairs(60, 120) + [(b"", b""), (b"A", b""), (b"", b"ACG"), (b"AAAA", b"TTTT")]:
        score, status, ops = A.align_end2end(p, t)
        assert status == 5
        assert score == _dp_score(oracle, metric, pen, p, t), (p, t)
        cig = pkg.Cigar(ops)
        if p or t:
            assert cig.check_alignment(p, t) is None
        # rescoring the CIGAR gives the score back (tests/singletrack.rs:116-121)
        assert _cigar_wf_score(pkg, metric, pen, cig, len(p), len(t)) == score


def test_oracle_score_only_equals_cigar_mode(N, oracle):
    """Modular (score-only) storage gives the high-memory score."""
    pairs = _lcg_pairs(40, 200, seed=11)
    for metric, pen in PENALTY_SETS:
        a_s = oracle.OracleAligner(attr_for(N, metric, pen, "score"))
        a_c = oracle.OracleAligner(attr_for(N, metric, pen, "alignment"))
        for p, t in pairs:
            assert a_s.align_end2end(p, t)[0] == a_c.align_end2end(p, t)[0]


def test_plausible_mistakes_are_caught(pkg, N, oracle):
    """The DP check fails for a wrong gap cost: pins o + l*e, not o + (l-1)*e."""
    attr = attr_for(N, "gap_affine", (0, 4, 6, 2, 0, 0), "score")
    A = oracle.OracleAligner(attr)
    p, t = b"ACGTACGTAC", b"ACGTTTTACGTAC"
    score = A.align_end2end(p, t)[0]
    assert score == oracle.dp_affine(p, t, 4, 6, 2)
    assert score != oracle.dp_affine(p, t, 4, 4, 2)
    # gap-linear: a gap costs indel per base, never an extra opening, and the
    # mismatch and indel inputs are not interchangeable
    A = oracle.OracleAligner(attr_for(N, "gap_linear", (0, 4, 2, 0, 0, 0), "score"))
    score = A.align_end2end(p, t)[0]
    assert score == oracle.dp_affine(p, t, 4, 0, 2)
    assert score != oracle.dp_affine(p, t, 4, 2, 2)
    assert A.align_end2end(b"ACGTACGT", b"ACCTACCT")[0] == 8
    assert oracle.dp_affine(b"ACGTACGT", b"ACCTACCT", 2, 0, 4) != 8


def test_generator_deterministic_and_shaped(pkg):
    from wfa2lib_rs_amd.synth import WORKLOADS, generate
    w = WORKLOADS["reads100_300"]
    a = generate(w, num_pairs=2000)
    b = generate(w, num_pairs=2000, threads=1)
    assert np.array_equal(a.seq, b.seq) and np.array_equal(a.plen, b.plen)
    assert a.plen.min() >= 100 and a.plen.max() <= 300
    # the first pairs do not depend on how many are generated
    c = generate(w, num_pairs=10)
    assert all(c.pattern(i) == a.pattern(i) and c.text(i) == a.text(i) for i in range(10))
    s = generate(WORKLOADS["short150"], num_pairs=500)
    assert (s.plen == 150).all()
    # ~5% divergence: text length within a few % of the pattern's
    assert abs(float(s.tlen.mean()) - 150) < 6


def test_packed_and_ascii_layouts_agree(pkg):
    from wfa2lib_rs_amd import _native as N
    from wfa2lib_rs_amd.synth import WORKLOADS, generate
    w = WORKLOADS["reads100_300"]
    packed = generate(w, num_pairs=50)
    ascii_ = generate(w.__class__(**{**w.__dict__, "seq_format": N.SEQ_ASCII}), num_pairs=50)
    for i in range(50):
        assert packed.pattern(i) == ascii_.pattern(i)
        assert packed.text(i) == ascii_.text(i)


def test_oracle_batch_threads_agree(N, oracle, pkg):
    from wfa2lib_rs_amd.synth import WORKLOADS, generate
    pairs = generate(WORKLOADS["short150"], num_pairs=300)
    attr = attr_for(N, "gap_affine", (0, 4, 6, 2, 0, 0), "alignment")
    s1, st1, c1 = oracle.align_batch(attr, pairs, cigar=True, threads=1)
    s4, st4, c4 = oracle.align_batch(attr, pairs, cigar=True, threads=4)
    assert np.array_equal(s1, s4) and np.array_equal(st1, st4) and c1 == c4


@pytest.mark.parametrize("metric,pen", PENALTY_SETS)
def test_oracle_med_low_memory(metric, pen, pkg, N, oracle):
    """Piggyback backtrace (med/low): same score as high memory, a valid CIGAR
    that rescores to it, and med == low (offloading never changes the op
    chain, only where it is stored)."""
    a_h = oracle.OracleAligner(attr_for(N, metric, pen, "alignment"))
    am = attr_for(N, metric, pen, "alignment")
    am.memory_mode = N.MEMORY_MED
    al = attr_for(N, metric, pen, "alignment")
    al.memory_mode = N.MEMORY_LOW
    a_m, a_l = oracle.OracleAligner(am), oracle.OracleAligner(al)
    for p, t in _lcg_pairs(80, 250, seed=5) + [(b"", b""), (b"ACGT", b""), (b"", b"GG")]:
        sh = a_h.align_end2end(p, t)[0]
        sm, stm, cm = a_m.align_end2end(p, t)
        sl, stl, cl = a_l.align_end2end(p, t)
        assert sm == sh == sl and stm == stl == 5
        assert cm == cl
        cig = pkg.Cigar(cm)
        if p or t:
            assert cig.check_alignment(p, t) is None
        assert _cigar_wf_score(pkg, metric, pen, cig, len(p), len(t)) == sm


def test_oracle_indel_literal_med_low(pkg, N, oracle):
    """src/aligner.rs:4481-4493: 3000 and 1500I1500D in every memory mode."""
    for mem in (N.MEMORY_HIGH, N.MEMORY_MED, N.MEMORY_LOW):
        a = attr_for(N, "indel", (0,) * 6, "alignment")
        a.memory_mode = mem
        s, st, ops = oracle.OracleAligner(a).align_end2end(b"A" * 1500, b"C" * 1500)
        assert (s, st, pkg.Cigar(ops).to_string_rle(True)) == (3000, 5, "1500I1500D")


def _issue22_pair():
    """tests/data/issue22_zdrop.seq (transcribed to tests/golden/)."""
    with open(os.path.join(GOLDEN, "issue22_zdrop.seq")) as f:
        lines = f.read().split("\n")
    return lines[0].lstrip(">").encode(), lines[1].lstrip("<").encode()


def valid_prefix(ops, p, t):
    """tests/corrected_zdrop.rs:56-75: every M a match and every X a
    mismatch, within both sequences (trailing bases may be left)."""
    i = j = 0
    for op in ops:
        if op in b"MX":
            if i >= len(p) or j >= len(t) or (p[i] == t[j]) != (op == ord("M")):
                return False
            i, j = i + 1, j + 1
        elif op == ord("D"):
            i += 1
        elif op == ord("I"):
            j += 1
        else:
            return False
    return i <= len(p) and j <= len(t)


def _zdrop_attr(N, metric, pen, scope, zdrop, steps, span=None):
    a = attr_for(N, metric, pen, scope, span=span)
    a.heuristic = N.HEURISTIC_ZDROP
    a.heuristic_p[0], a.heuristic_p[1] = zdrop, steps
    return a


def test_oracle_zdrop_issue22_literal(pkg, N, oracle):
    """tests/corrected_zdrop.rs:77-89 (Compat::C, high memory): Z-drop 50,10
    with (-3,5,3,2) ends early in a valid partial alignment whose gap-affine
    score is 4498; score-only ends at the same score with the same status."""
    p, t = _issue22_pair()
    pen = (-3, 5, 3, 2, 0, 0)
    A = oracle.OracleAligner(_zdrop_attr(N, "gap_affine", pen, "alignment", 50, 10))
    s, st, ops = A.align_end2end(p, t)
    assert st == STATUS["PARTIAL"] and ops
    assert valid_prefix(ops, p, t)
    assert pkg.Cigar(ops).score_gap_affine(pkg.AffinePenalties(-3, 5, 3, 2)) == 4498
    best = A.end_position()
    assert 0 <= best[0] < s and 0 <= best[2] - best[1] <= len(p) and 0 <= best[2] <= len(t)
    S = oracle.OracleAligner(_zdrop_attr(N, "gap_affine", pen, "score", 50, 10))
    assert S.align_end2end(p, t)[:2] == (s, st)
    assert S.end_position() == best


@pytest.mark.parametrize("metric,pen", [m for m in PENALTY_SETS if m[0] != "indel"])
def test_oracle_zdrop_invariants(metric, pen, pkg, N, oracle):
    """A Z-drop that never fires changes nothing; one that fires leaves a
    valid partial alignment whose end is the recorded best cell (or, max
    trimmed, a prefix of the path to it), and score-only agrees on the score,
    status and end position. (With a mismatch of 1, or a gap extension of 1
    and no opening, the Smith-Waterman score of unrelated sequence rises, so
    Z-drop cannot fire: those sets only check the first property.)"""
    from wfa2lib_rs_amd.synth import Workload, generate
    import wfa2lib_rs_amd._native as Nn
    w = Workload("zdrop_mix", 60, 1, 60, 600, 0.0, 0.35, 3.0, run_frac=0.2, run_mean=30.0,
                 seq_format=Nn.SEQ_ASCII, seed=23)
    hp = generate(w)
    plain = oracle.OracleAligner(attr_for(N, metric, pen, "alignment"))
    never = oracle.OracleAligner(_zdrop_attr(N, metric, pen, "alignment", 10 ** 8, 1))
    zd = oracle.OracleAligner(_zdrop_attr(N, metric, pen, "alignment", 5, 1))
    zs = oracle.OracleAligner(_zdrop_attr(N, metric, pen, "score", 5, 1))
    fired = 0
    for i in range(len(hp)):
        p, t = hp.pattern(i), hp.text(i)
        assert never.align_end2end(p, t) == plain.align_end2end(p, t)
        s, st, ops = zd.align_end2end(p, t)
        assert zs.align_end2end(p, t)[:2] == (s, st)
        assert zs.end_position() == zd.end_position()
        if st == STATUS["PARTIAL"]:
            fired += 1
            assert valid_prefix(ops, p, t)
            bs, bk, boff = zd.end_position()
            assert bs < s
            v = sum(op != ord("I") for op in ops)
            h = sum(op != ord("D") for op in ops)
            assert v <= boff - bk and h <= boff
            if ops and metric != "edit":
                # max-trimmed: the kept part's WFA score (tests/heuristic_scores.rs:124-127)
                assert zd.cigar_score() == _cigar_wf_score(pkg, metric, pen, pkg.Cigar(ops), v, h)
        else:
            assert st == STATUS["COMPLETED"]
    if metric != "edit" and pen[1] >= 4:
        assert fired >= 5, "the pairs must make Z-drop fire"


def _endsfree_dp(oracle, metric, pen, p, t, ends):
    m, x, o1, e1, o2, e2 = pen
    if metric == "gap_linear":
        return oracle.dp_endsfree(p, t, m, x, 0, o1, 0, o1, False, ends)
    return oracle.dp_endsfree(p, t, m, x, o1, e1, o2, e2, metric == "gap_affine_2p", ends)


def test_oracle_endsfree_match0_equals_dp(N, oracle):
    """Ends-free with match == 0 is exact (tests/corrected_endsfree.rs:
    333-354 checks C == corrected there, and corrected against this DP): the
    WFA score equals the optimal ends-free penalty on random short pairs and
    free ends, for every gap metric."""
    import random
    rng = random.Random(7)
    seq = lambda n: bytes(rng.choice(b"ACGT") for _ in range(n))
    for _ in range(400):
        metric = rng.choice(["gap_affine", "gap_affine_2p", "gap_linear"])
        pen = (0, 1 + rng.randrange(6), rng.randrange(6), 1 + rng.randrange(4),
               rng.randrange(12), 1 + rng.randrange(2))
        if metric == "gap_linear":
            pen = (0, pen[1], 1 + rng.randrange(5), 0, 0, 0)
        p, t = seq(rng.randrange(16)), seq(rng.randrange(16))
        ends = [rng.randrange(len(p) + 1), rng.randrange(len(p) + 1),
                rng.randrange(len(t) + 1), rng.randrange(len(t) + 1)]
        s, st, _ = oracle.OracleAligner(attr_for(N, metric, pen, "score", span=ends)).align(p, t)
        assert st == STATUS["COMPLETED"]
        assert s == _endsfree_dp(oracle, metric, pen, p, t, ends), (metric, pen, p, t, ends)


def test_oracle_endsfree_match_reward_literals(pkg, N, oracle):
    """Ends-free with a match reward in the reference's default (C) mode:
    the begin-free cells are reached at score d * -match during compute.
    tests/corrected_endsfree.rs:293-301 (issue 37: IIIMMMMMMMMDDDD) and
    :317-324 (issue 38: the all-gap alignment), src/aligner.rs:3945-3967
    (2-piece, score only: completes)."""
    a = attr_for(N, "gap_affine", (-5, 4, 0, 8, 0, 0), "alignment", span=[12, 12, 11, 11])
    for mem in (N.MEMORY_HIGH, N.MEMORY_MED, N.MEMORY_LOW):
        a.memory_mode = mem
        s, st, ops = oracle.OracleAligner(a).align(b"ACGTACGTACGT", b"ACGACGTACGT")
        assert (st, ops) == (STATUS["COMPLETED"], b"IIIMMMMMMMMDDDD")
    pat = b"AAGAAGCACCGGCTAACTCCGTGCCAGCAGCCGCGGT"
    txt = b"AGTCGTGACTCTTTTTGCATGTACCTTATGAATAAGGACCGGCTAATTCCG"
    a = attr_for(N, "gap_affine", (-5, 4, 0, 8, 0, 0), "alignment",
                 span=[len(pat), len(pat), len(txt), len(txt)])
    s, st, ops = oracle.OracleAligner(a).align(pat, txt)
    assert ops and set(ops) <= set(b"ID") and pkg.Cigar(ops).check_alignment(pat, txt) is None
    a = attr_for(N, "gap_affine_2p", (-3, 6, 1, 5, 9, 1), "score", span=[0, 33, 0, 35])
    s, st, _ = oracle.OracleAligner(a).align(b"TGGATTGAGATTCCCCGGTATGCTGACAGAAGC",
                                             b"ACATTCCGACATAGAGCCATATGATAAAAGTTAGT")
    assert st == STATUS["COMPLETED"]


@pytest.mark.parametrize("metric,pen", [("gap_affine", (-3, 5, 3, 2, 0, 0)),
                                        ("gap_affine", (-5, 4, 0, 8, 0, 0)),
                                        ("gap_affine_2p", (-1, 4, 6, 2, 24, 1)),
                                        ("gap_linear", (-1, 3, 2, 0, 0, 0))])
def test_oracle_endsfree_match_reward_invariants(metric, pen, pkg, N, oracle):
    """Ends-free with a match reward (C mode): a valid CIGAR whose ends-free
    penalty is no better than the optimum (the mode is not always optimal:
    issues 37/38); medium and low memory give the high-memory score and
    each other's CIGAR; score-only gives the same score."""
    import random
    rng = random.Random(11)
    seq = lambda n: bytes(rng.choice(b"ACGT") for _ in range(n))
    m = pen[0]
    for _ in range(150):
        p = seq(1 + rng.randrange(40))
        t = bytes(c if rng.random() > 0.2 else rng.choice(b"ACGT") for c in p)
        t = seq(rng.randrange(6)) + t[rng.randrange(4):] + seq(rng.randrange(6))
        ends = [rng.randrange(len(p) + 1), rng.randrange(len(p) + 1),
                rng.randrange(len(t) + 1), rng.randrange(len(t) + 1)]
        res = {}
        for mem, scope in ((N.MEMORY_HIGH, "alignment"), (N.MEMORY_MED, "alignment"),
                           (N.MEMORY_LOW, "alignment"), (N.MEMORY_HIGH, "score")):
            a = attr_for(N, metric, pen, scope, span=ends)
            a.memory_mode = mem
            res[(mem, scope)] = oracle.OracleAligner(a).align(p, t)
        s, st, ops = res[(N.MEMORY_HIGH, "alignment")]
        assert st == STATUS["COMPLETED"]
        assert pkg.Cigar(ops).check_alignment(p, t) is None
        assert res[(N.MEMORY_HIGH, "score")][:2] == (s, st)
        assert res[(N.MEMORY_MED, "alignment")][0] == s
        assert res[(N.MEMORY_MED, "alignment")] == res[(N.MEMORY_LOW, "alignment")]
        for _, _, c in res.values():
            if c:
                pb, pe, tb, te = ends
                # the CIGAR's ends-free penalty: leading/trailing gap runs
                # free up to the limits (tests/corrected_endsfree.rs:109-146)
                runs = []
                for op in c:
                    if runs and runs[-1][0] == op:
                        runs[-1][1] += 1
                    else:
                        runs.append([op, 1])
                pen_total = 0
                for i, (op, ln) in enumerate(runs):
                    ch = ln
                    if i == 0:
                        ch -= min(ch, pb if op == ord("D") else tb if op == ord("I") else 0)
                    if i == len(runs) - 1:
                        ch -= min(ch, pe if op == ord("D") else te if op == ord("I") else 0)
                    if op == ord("M"):
                        pen_total += ln * m
                    elif op == ord("X"):
                        pen_total += ln * pen[1]
                    elif ch > 0:
                        if metric == "gap_linear":
                            pen_total += ch * pen[2]
                        elif metric == "gap_affine":
                            pen_total += pen[2] + ch * pen[3]
                        else:
                            pen_total += min(pen[2] + ch * pen[3], pen[4] + ch * pen[5])
                assert pen_total >= _endsfree_dp(oracle, metric, pen, p, t, ends)


def _penalty_of(w, m, plen, tlen):
    """The real penalty of a WFA score (wf_score_to_sw_score,
    src/penalties.rs; tests/corrected_endsfree.rs:178-190)."""
    return -((-m * (plen + tlen) - w) // 2) if m < 0 else w


def _corrected(N, metric, pen, scope, span, mem=None):
    a = attr_for(N, metric, pen, scope, span=span)
    a.compat = N.COMPAT_CORRECTED
    if mem is not None:
        a.memory_mode = mem
    return a


def test_oracle_corrected_literals(pkg, N, oracle):
    """Compat::Corrected: tests/corrected_endsfree.rs:293-312 (issue 37:
    penalty -47, MMMDMMMMMMMM) and :314-330 (issue 38: -78, the DP
    optimum); tests/corrected_zdrop.rs:91-102 (Z-drop 50,10: high, med and
    low memory give valid partial alignments of the same score)."""
    pen = (-5, 4, 0, 8, 0, 0)
    s, st, ops = oracle.OracleAligner(_corrected(N, "gap_affine", pen, "alignment",
                                                 [12, 12, 11, 11])).align(b"ACGTACGTACGT",
                                                                          b"ACGACGTACGT")
    assert (_penalty_of(s, -5, 12, 11), ops) == (-47, b"MMMDMMMMMMMM")
    pat = b"AAGAAGCACCGGCTAACTCCGTGCCAGCAGCCGCGGT"
    txt = b"AGTCGTGACTCTTTTTGCATGTACCTTATGAATAAGGACCGGCTAATTCCG"
    ends = [len(pat), len(pat), len(txt), len(txt)]
    assert _endsfree_dp(oracle, "gap_affine", pen, pat, txt, ends) == -78
    for scope in ("score", "alignment"):
        s, st, _ = oracle.OracleAligner(_corrected(N, "gap_affine", pen, scope, ends)).align(pat, txt)
        assert _penalty_of(s, -5, len(pat), len(txt)) == -78
    p, t = _issue22_pair()
    scores = set()
    for mem in (N.MEMORY_HIGH, N.MEMORY_MED, N.MEMORY_LOW):
        a = _zdrop_attr(N, "gap_affine", (-3, 5, 3, 2, 0, 0), "alignment", 50, 10)
        a.compat, a.memory_mode = N.COMPAT_CORRECTED, mem
        s, st, ops = oracle.OracleAligner(a).align_end2end(p, t)
        assert st == STATUS["PARTIAL"] and ops and valid_prefix(ops, p, t)
        scores.add(pkg.Cigar(ops).score_gap_affine(pkg.AffinePenalties(-3, 5, 3, 2)))
    assert len(scores) == 1


def test_oracle_corrected_endsfree_is_optimal(pkg, N, oracle):
    """Corrected ends-free (and extension) with a match reward returns the
    optimal alignment (tests/corrected_endsfree.rs:356-420 checks the same
    against this DP): on random short pairs, free ends and penalties
    (dominance pruning included), every memory mode's penalty equals the
    DP optimum and its CIGAR is valid (a valid prefix for extension)."""
    import random
    rng = random.Random(5)
    seq = lambda n: bytes(rng.choice(b"ACGT") for _ in range(n))
    for _ in range(300):
        metric = rng.choice(["gap_affine", "gap_affine_2p", "gap_linear"])
        m = -(1 + rng.randrange(5))
        pen = (m, 1 + rng.randrange(6), rng.randrange(6), 1 + rng.randrange(4),
               rng.randrange(12), 1 + rng.randrange(2))
        if metric == "gap_linear":
            pen = (m, pen[1], 1 + rng.randrange(5), 0, 0, 0)
        p, t = seq(rng.randrange(18)), seq(rng.randrange(18))
        ends = [rng.randrange(len(p) + 1), rng.randrange(len(p) + 1),
                rng.randrange(len(t) + 1), rng.randrange(len(t) + 1)]
        if rng.random() < 0.2:
            ends = "extension"
        e = [0, len(p), 0, len(t)] if ends == "extension" else ends
        want = _endsfree_dp(oracle, metric, pen, p, t, e)
        for mem in (N.MEMORY_HIGH, N.MEMORY_MED, N.MEMORY_LOW):
            s, st, ops = oracle.OracleAligner(_corrected(N, metric, pen, "alignment", ends,
                                                         mem)).align(p, t)
            assert _penalty_of(s, m, len(p), len(t)) == want, (metric, pen, p, t, ends, mem)
            if ends == "extension":
                assert valid_prefix(ops, p, t)
            elif ops:
                assert pkg.Cigar(ops).check_alignment(p, t) is None


def seq_pair(name):
    """The first pair of a tests/data/*.seq fixture (copied to tests/golden/)."""
    with open(os.path.join(GOLDEN, name)) as f:
        lines = f.read().split("\n")
    return lines[0].lstrip(">").encode(), lines[1].lstrip("<").encode()


def test_oracle_corrected_fixtures(pkg, N, oracle):
    """tests/corrected_endsfree.rs:523-554: gap-linear (-2,4,3), banded-static
    -10..200 off the end diagonal, whole free ends: the optimum is a
    begin-free seed reached after every wavefront emptied, so null steps
    must not end the search while a candidate is pending; every memory mode
    gives a valid alignment of the DP optimum. tests/corrected_endsfree.rs:
    611-624 (issue 38 pair B, 16S reads, (-5,4,0,8), whole free ends):
    Compat::C returns an all-gap alignment, Corrected the DP optimum."""
    pat, txt = seq_pair("corrected_banded_seeds.seq")
    pen = (-2, 4, 3, 0, 0, 0)
    ends = [len(pat), len(pat), len(txt), len(txt)]
    want = _endsfree_dp(oracle, "gap_linear", pen, pat, txt, ends)
    for mem in (N.MEMORY_HIGH, N.MEMORY_MED, N.MEMORY_LOW):
        a = _corrected(N, "gap_linear", pen, "alignment", ends, mem)
        a.heuristic = N.HEURISTIC_BANDED_STATIC
        a.heuristic_p[0], a.heuristic_p[1] = -10, 200
        s, st, ops = oracle.OracleAligner(a).align(pat, txt)
        assert _penalty_of(s, -2, len(pat), len(txt)) == want, mem
        assert pkg.Cigar(ops).check_alignment(pat, txt) is None
    pat, txt = seq_pair("issue38_pair_b.seq")
    pen = (-5, 4, 0, 8, 0, 0)
    ends = [len(pat), len(pat), len(txt), len(txt)]
    _, _, ops = oracle.OracleAligner(attr_for(N, "gap_affine", pen, "alignment",
                                              span=ends)).align(pat, txt)
    assert ops and set(ops) <= set(b"ID")
    want = _endsfree_dp(oracle, "gap_affine", pen, pat, txt, ends)
    s, st, ops = oracle.OracleAligner(_corrected(N, "gap_affine", pen, "alignment",
                                                 ends)).align(pat, txt)
    assert _penalty_of(s, -5, len(pat), len(txt)) == want < 0
    assert pkg.Cigar(ops).check_alignment(pat, txt) is None


def test_oracle_corrected_equals_c_with_match0(N, oracle):
    """tests/corrected_endsfree.rs:333-354: with match == 0 both modes give
    identical results."""
    import random
    rng = random.Random(7)
    seq = lambda n: bytes(rng.choice(b"ACGT") for _ in range(n))
    for _ in range(150):
        metric = rng.choice(["gap_affine", "gap_affine_2p", "gap_linear"])
        pen = (0, 1 + rng.randrange(6), rng.randrange(6), 1 + rng.randrange(4),
               rng.randrange(12), 1 + rng.randrange(2))
        if metric == "gap_linear":
            pen = (0, pen[1], 1 + rng.randrange(5), 0, 0, 0)
        p, t = seq(rng.randrange(16)), seq(rng.randrange(16))
        ends = [rng.randrange(len(p) + 1), rng.randrange(len(p) + 1),
                rng.randrange(len(t) + 1), rng.randrange(len(t) + 1)]
        for mem in (N.MEMORY_HIGH, N.MEMORY_MED):
            c = attr_for(N, metric, pen, "alignment", span=ends)
            c.memory_mode = mem
            assert oracle.OracleAligner(c).align(p, t) == \
                oracle.OracleAligner(_corrected(N, metric, pen, "alignment", ends, mem)).align(p, t)


def test_oracle_zdrop_refusals(N, oracle):
    """Z-drop with BiWFA is refused (src/bin/align_benchmark.rs:529-560);
    every other memory mode and scope runs."""
    a = _zdrop_attr(N, "gap_affine", (0, 4, 6, 2, 0, 0), "alignment", 20, 1)
    a.memory_mode = N.MEMORY_ULTRALOW
    with pytest.raises(ValueError):
        oracle.OracleAligner(a)
    for scope in ("alignment", "score"):
        for mem in (N.MEMORY_HIGH, N.MEMORY_MED, N.MEMORY_LOW):
            for compat in (N.COMPAT_C, N.COMPAT_CORRECTED):
                a = _zdrop_attr(N, "gap_affine", (0, 4, 6, 2, 0, 0), scope, 20, 1)
                a.memory_mode, a.compat = mem, compat
                oracle.OracleAligner(a)


def test_oracle_zdrop_issue22_c_med_low_literal(pkg, N, oracle):
    """tests/corrected_zdrop.rs:79-89 (Compat::C): Z-drop 50,10 with
    (-3,5,3,2) stops at the same score in every memory mode; medium and low
    memory backtrace the best cell's diagonal at the stop score, a cell that
    is no part of the best alignment (HPCBio/wfa2lib-rs#22), and their CIGAR
    scores 4502 where high memory's scores 4498."""
    p, t = _issue22_pair()
    pen = (-3, 5, 3, 2, 0, 0)
    high = oracle.OracleAligner(_zdrop_attr(N, "gap_affine", pen, "alignment", 50, 10))
    hs, hst, hops = high.align_end2end(p, t)
    for mem in (N.MEMORY_MED, N.MEMORY_LOW):
        a = _zdrop_attr(N, "gap_affine", pen, "alignment", 50, 10)
        a.memory_mode = mem
        A = oracle.OracleAligner(a)
        s, st, ops = A.align_end2end(p, t)
        assert (s, st) == (hs, hst) == (s, STATUS["PARTIAL"])
        assert pkg.Cigar(ops).score_gap_affine(pkg.AffinePenalties(-3, 5, 3, 2)) == 4502
        assert A.end_position() == high.end_position()
        assert ops != hops


def test_oracle_sentinel_bytes_are_ordinary_symbols(pkg, N, oracle):
    """Sequence ends are hard bounds here; the reference pads the pattern
    with '!' and the text with '?' (src/sequences.rs:16-19, 178-191), so a
    text holding '!' or a pattern holding '?' can let its extension run past
    a sequence end. In the oracle those bytes are ordinary symbols: pairs
    with '!' and '?' at and near both ends score the exact gap-affine
    optimum, with valid CIGARs."""
    from wfa2lib_rs_amd.synth import from_ascii
    rng = np.random.default_rng(11)
    alpha = np.frombuffer(b"ACGT!?", dtype=np.uint8)
    pats, txts = [], []
    for i in range(200):
        n = int(rng.integers(1, 120))
        p = rng.choice(alpha, n).tobytes()
        t = bytearray(p)
        for _ in range(int(rng.integers(0, 6))):
            j = int(rng.integers(0, len(t) + 1))
            t[j:j] = rng.choice(alpha, 1).tobytes()
        t = bytes(t)
        if i % 3 == 0:
            p, t = p + b"?", t + b"!"
        elif i % 3 == 1:
            p, t = p + b"!" * 3, t + b"?" * 2
        pats.append(p)
        txts.append(t)
    attr = attr_for(N, "gap_affine", (0, 4, 6, 2, 0, 0), "alignment")
    s, st, c = oracle.align_batch(attr, from_ascii(pats, txts, N.SEQ_ASCII), cigar=True,
                                  threads=8)
    for i in range(len(pats)):
        assert s[i] == oracle.dp_affine(pats[i], txts[i], 4, 6, 2)
        assert pkg.Cigar(c[i]).check_alignment(pats[i], txts[i]) is None

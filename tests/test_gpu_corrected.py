"""GPU parity of Compat::Corrected (ends-free and extension with a match
reward: candidate ends, band and dominance pruning, corrected Z-drop): the
HIP engine against the oracle, bit for bit, through the C ABI. The oracle's
corrected mode is itself pinned against the ends-free DP optimum and the
reference's literals (tests/corrected_endsfree.rs, tests/corrected_zdrop.rs)
in test_oracle.py."""
import random
import zlib

import pytest

from conftest import attr_for, have_gpu

pytestmark = pytest.mark.gpu

STATUS_PARTIAL = 6


@pytest.fixture(scope="module")
def engine(pkg):
    if not have_gpu():
        pytest.skip("no GPU")
    from wfa2lib_rs_amd.engine import BatchEngine
    return BatchEngine(0)


def _random_pairs(N, n, seed, max_len=60):
    """Short random pairs of every length up to max_len, half of them an
    edited copy of the other sequence (so that long matching runs reach the
    free ends), including empty sequences."""
    from wfa2lib_rs_amd.synth import from_ascii
    rng = random.Random(seed)
    seq = lambda k: bytes(rng.choice(b"ACGT") for _ in range(k))
    pats, txts = [], []
    for i in range(n):
        p = seq(rng.randrange(max_len + 1))
        if i % 2 and p:
            t = bytearray(p)
            for _ in range(rng.randrange(1 + len(p) // 5)):
                j = rng.randrange(len(t) + 1)
                r = rng.random()
                if r < 0.4 and j < len(t):
                    t[j] = rng.choice(b"ACGT")
                elif r < 0.7:
                    t[j:j] = seq(1 + rng.randrange(4))
                elif j < len(t):
                    del t[j:j + 1 + rng.randrange(4)]
            t = bytes(seq(rng.randrange(6))) + bytes(t) + seq(rng.randrange(6))
        else:
            t = seq(rng.randrange(max_len + 1))
        pats.append(p)
        txts.append(t)
    return from_ascii(pats, txts, N.SEQ_ASCII)


def _check(engine, oracle, attr, pairs, cigar, label):
    """Scores, statuses, CIGARs and end positions equal the oracle's."""
    import torch
    from wfa2lib_rs_amd.engine import DeviceBatch
    b = DeviceBatch(pairs, device="cuda:0", cigar=cigar, end_position=True)
    engine.align(attr, b)
    torch.cuda.synchronize()
    s, st, c = b.results()
    ends = b.end_positions()
    A = oracle.OracleAligner(attr)
    # free begin gaps longer than their sequence are outside the reference's
    # contract (its driver refuses them, src/bin/align_benchmark.rs:606-622;
    # its aligner then returns alignments that do not span the sequences):
    # those pairs run, memory-safely, but are not compared
    bf = attr.span == 1 and not attr.extension
    pb, tb = (attr.pattern_begin_free, attr.text_begin_free) if bf else (0, 0)
    compared = 0
    for i in range(len(pairs)):
        if pairs.plen[i] < pb or pairs.tlen[i] < tb:
            continue
        compared += 1
        os_, ost, oc = A.align(pairs.pattern(i), pairs.text(i))
        got = (int(s[i]), int(st[i]), c[i] if cigar else b"", tuple(int(x) for x in ends[i]))
        want = (os_, ost, oc if cigar else b"", tuple(A.end_position()))
        assert got == want, f"{label} pair {i}: gpu {got} oracle {want}"
    assert compared >= min(len(pairs), 10), f"{label}: only {compared} pairs compared"


CORRECTED_CONFIGS = [
    ("gap_affine", (-2, 4, 6, 2, 0, 0)),
    ("gap_affine", (-5, 4, 0, 8, 0, 0)),
    ("gap_affine_2p", (-1, 4, 6, 2, 24, 1)),
    ("gap_linear", (-1, 3, 2, 0, 0, 0)),
    ("gap_linear", (-3, 5, 4, 0, 0, 0)),
]
# begin-free gaps of at most 4 turn dominance pruning on for gap-affine;
# gap-linear prunes with any begin-free gap
CORRECTED_SPANS = [("extension", "extension"), ("endsfree_3_3", [3, 3, 3, 3]),
                   ("endsfree_5_5", [5, 5, 5, 5]), ("begin_free", [15, 0, 15, 0]),
                   ("end_free", [0, 20, 0, 20]), ("wide", [40, 10, 30, 200])]


@pytest.mark.parametrize("sname,span", CORRECTED_SPANS)
@pytest.mark.parametrize("metric,pen", CORRECTED_CONFIGS)
@pytest.mark.parametrize("memory", ["high", "med", "low"])
def test_corrected_endsfree_matches_oracle(sname, span, metric, pen, memory, pkg, N, engine,
                                           oracle):
    """Corrected ends-free/extension, score only and CIGARs in every
    memory mode: the candidate's total, its CIGAR and its end cell."""
    pairs = _random_pairs(N, 160, seed=zlib.crc32(f"{sname}/{metric}/{pen}".encode()))
    mem = {"high": N.MEMORY_HIGH, "med": N.MEMORY_MED, "low": N.MEMORY_LOW}[memory]
    for scope in ("score", "alignment"):
        attr = attr_for(N, metric, pen, scope, span=span)
        attr.compat, attr.memory_mode = N.COMPAT_CORRECTED, mem
        _check(engine, oracle, attr, pairs, scope == "alignment",
               f"{sname}/{metric}{pen}/{memory}/{scope}")


CORRECTED_HEURISTICS = [("wfadaptive_4_6_2", (1, 4, 6, 2)), ("xdrop_15_1", (2, 15, 1, 0)),
                        ("zdrop_5_1", (3, 5, 1, 0)), ("banded_static_m12_9", (4, -12, 9, 0)),
                        ("banded_adaptive_m6_7_1", (5, -6, 7, 1))]


@pytest.mark.parametrize("hname,h", CORRECTED_HEURISTICS)
@pytest.mark.parametrize("sname,span", [("extension", "extension"), ("endsfree_3_3", [3, 3, 3, 3]),
                                        ("wide", [80, 10, 60, 200])])
@pytest.mark.parametrize("memory", ["high", "med"])
def test_corrected_heuristics_match_oracle(hname, h, sname, span, memory, pkg, N, engine, oracle):
    """Corrected mode under every heuristic (band pruning only with none or
    banded-static; a cut-off with a candidate ends at the candidate;
    corrected Z-drop also with a medium-memory CIGAR)."""
    from wfa2lib_rs_amd.synth import Workload, generate
    w = Workload("heuristic_mix", 120, 1, 40, 300, 0.0, 0.30, 3.0, run_frac=0.1, run_mean=15.0,
                 seq_format=N.SEQ_ASCII, seed=11)
    pairs = generate(w)
    mem = N.MEMORY_HIGH if memory == "high" else N.MEMORY_MED
    for metric, pen in (("gap_affine", (-2, 4, 6, 2, 0, 0)), ("gap_linear", (-1, 3, 2, 0, 0, 0))):
        for scope in ("score", "alignment"):
            attr = attr_for(N, metric, pen, scope, span=span)
            attr.compat, attr.memory_mode = N.COMPAT_CORRECTED, mem
            attr.heuristic = h[0]
            for i in range(3):
                attr.heuristic_p[i] = h[i + 1]
            _check(engine, oracle, attr, pairs, scope == "alignment",
                   f"{hname}/{sname}/{metric}/{memory}/{scope}")


def test_corrected_with_match0_equals_c(pkg, N, engine):
    """tests/corrected_endsfree.rs:333-354 on the device: with match == 0
    both modes give identical results."""
    import torch
    from wfa2lib_rs_amd.engine import DeviceBatch
    pairs = _random_pairs(N, 300, seed=3)
    for metric, pen in (("gap_affine", (0, 4, 6, 2, 0, 0)), ("gap_linear", (0, 4, 2, 0, 0, 0))):
        out = []
        for compat in (N.COMPAT_C, N.COMPAT_CORRECTED):
            attr = attr_for(N, metric, pen, "alignment", span=[5, 5, 5, 5])
            attr.compat = compat
            b = DeviceBatch(pairs, device="cuda:0", cigar=True)
            engine.align(attr, b)
            torch.cuda.synchronize()
            s, st, c = b.results()
            out.append((s.tolist(), st.tolist(), c))
        assert out[0] == out[1]


def test_corrected_issue_literals_through_aligner(pkg, N):
    """tests/corrected_endsfree.rs:293-330 through the aligner API: issue 37
    (penalty -47, MMMDMMMMMMMM, every memory mode) and issue 38 (-78)."""
    from test_oracle import _penalty_of
    if not have_gpu():
        pytest.skip("no GPU")
    pen = pkg.AffinePenalties(-5, 4, 0, 8)
    for mode in (pkg.MemoryMode.HIGH, pkg.MemoryMode.MED, pkg.MemoryMode.LOW):
        A = pkg.AffineAligner(pkg.WavefrontPenalties.new_affine(pen))
        A.set_alignment_scope(pkg.AlignmentScope.COMPUTE_ALIGNMENT)
        A.set_memory_mode(mode)
        A.set_compat(N.COMPAT_CORRECTED)
        A.set_alignment_free_ends(12, 12, 11, 11)
        s = A.align_endsfree(b"ACGTACGTACGT", b"ACGACGTACGT")
        assert bytes(A.cigar().operations) == b"MMMDMMMMMMMM"
        assert _penalty_of(s, -5, 12, 11) == -47
    pat = b"AAGAAGCACCGGCTAACTCCGTGCCAGCAGCCGCGGT"
    txt = b"AGTCGTGACTCTTTTTGCATGTACCTTATGAATAAGGACCGGCTAATTCCG"
    A = pkg.AffineAligner(pkg.WavefrontPenalties.new_affine(pen))
    A.set_compat(N.COMPAT_CORRECTED)
    A.set_alignment_free_ends(len(pat), len(pat), len(txt), len(txt))
    s = A.align_endsfree(pat, txt)
    assert _penalty_of(s, -5, len(pat), len(txt)) == -78


def test_corrected_zdrop_med_low_on_gpu(pkg, N, engine, oracle):
    """tests/corrected_zdrop.rs:91-102 on the device: Z-drop 50,10 with
    (-3,5,3,2) in Compat::Corrected gives valid partial alignments of one
    score in high, medium and low memory, each equal to the oracle's."""
    from test_oracle import _issue22_pair, _zdrop_attr, valid_prefix
    from wfa2lib_rs_amd.synth import from_ascii
    p, t = _issue22_pair()
    pairs = from_ascii([p], [t], N.SEQ_ASCII)
    scores = set()
    for mem in (N.MEMORY_HIGH, N.MEMORY_MED, N.MEMORY_LOW):
        a = _zdrop_attr(N, "gap_affine", (-3, 5, 3, 2, 0, 0), "alignment", 50, 10)
        a.compat, a.memory_mode = N.COMPAT_CORRECTED, mem
        _check(engine, oracle, a, pairs, True, f"issue22/{mem}")
        A = oracle.OracleAligner(a)
        s, st, ops = A.align_end2end(p, t)
        assert st == STATUS_PARTIAL and valid_prefix(ops, p, t)
        scores.add(pkg.Cigar(ops).score_gap_affine(pkg.AffinePenalties(-3, 5, 3, 2)))
    assert len(scores) == 1


def test_corrected_reference_fixtures_on_gpu(pkg, N, engine, oracle):
    """tests/corrected_endsfree.rs:523-554 (banded-static -10..200 with a
    pending begin-free seed, gap-linear (-2,4,3)) and :611-624 (issue 38
    pair B, two 16S reads, (-5,4,0,8)), whole free ends, on the device:
    equal to the oracle, whose results those tests pin to the DP optimum."""
    from test_oracle import seq_pair
    from wfa2lib_rs_amd.synth import from_ascii
    pat, txt = seq_pair("corrected_banded_seeds.seq")
    pairs = from_ascii([pat], [txt], N.SEQ_ASCII)
    ends = [len(pat), len(pat), len(txt), len(txt)]
    for mem in (N.MEMORY_HIGH, N.MEMORY_MED, N.MEMORY_LOW):
        for scope in ("score", "alignment"):
            a = attr_for(N, "gap_linear", (-2, 4, 3, 0, 0, 0), scope, span=ends)
            a.compat, a.memory_mode = N.COMPAT_CORRECTED, mem
            a.heuristic = N.HEURISTIC_BANDED_STATIC
            a.heuristic_p[0], a.heuristic_p[1] = -10, 200
            _check(engine, oracle, a, pairs, scope == "alignment", f"banded_seeds/{mem}/{scope}")
    pat, txt = seq_pair("issue38_pair_b.seq")
    pairs = from_ascii([pat], [txt], N.SEQ_ASCII)
    ends = [len(pat), len(pat), len(txt), len(txt)]
    for mem in (N.MEMORY_HIGH, N.MEMORY_MED):
        for scope in ("score", "alignment"):
            a = attr_for(N, "gap_affine", (-5, 4, 0, 8, 0, 0), scope, span=ends)
            a.compat, a.memory_mode = N.COMPAT_CORRECTED, mem
            _check(engine, oracle, a, pairs, scope == "alignment", f"issue38b/{mem}/{scope}")

"""GPU BiWFA (ultralow memory) against the oracle, bit for bit: scores,
statuses and CIGARs of split alignments, base cases, heuristics and the
wide-row retry path."""
import numpy as np
import pytest

from conftest import attr_for, have_gpu
from test_gpu_parity import _compare, _run_gpu
from test_oracle_biwfa import BIWFA_PENALTIES, _recursive_pairs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine(pkg):
    if not have_gpu():
        pytest.skip("no GPU")
    from wfa2lib_rs_amd.engine import BatchEngine
    return BatchEngine(0)


def _attr(N, metric, pen, scope, heuristic=None):
    attr = attr_for(N, metric, pen, scope)
    attr.memory_mode = N.MEMORY_ULTRALOW
    if heuristic:
        attr.heuristic = heuristic[0]
        for i, v in enumerate(heuristic[1:]):
            attr.heuristic_p[i] = v
    return attr


def _pairs(n, seed, fmt=None):
    from wfa2lib_rs_amd.synth import from_ascii
    import wfa2lib_rs_amd._native as Nn
    ps = _recursive_pairs(n, seed)
    # short pairs (top-level fallback) and empty sequences ride along
    ps += [(b"ACGTTGCA" * 10, b"ACGTTGGA" * 11), (b"", b"ACGT"), (b"GATTACA", b""), (b"", b""),
           (b"A" * 300, b"A" * 300)]
    return from_ascii([p for p, _ in ps], [t for _, t in ps], fmt or Nn.SEQ_ASCII)


@pytest.mark.parametrize("metric,pen", BIWFA_PENALTIES)
@pytest.mark.parametrize("scope", ["score", "alignment"])
def test_biwfa_matches_oracle(metric, pen, scope, pkg, N, engine, oracle):
    pairs = _pairs(24, seed=5)
    _compare(engine, oracle, _attr(N, metric, pen, scope), pairs, scope == "alignment",
             label=f"biwfa/{metric}/{scope}")


def test_biwfa_packed_input(pkg, N, engine, oracle):
    pairs = _pairs(16, seed=9, fmt=N.SEQ_PACKED2BITS)
    _compare(engine, oracle, _attr(N, "gap_affine", (0, 4, 6, 2, 0, 0), "alignment"), pairs,
             True, label="biwfa/packed")


@pytest.mark.parametrize("h", [(1, 10, 50, 1), (1, 4, 6, 2), (4, -40, 5, 0), (4, -3, 3, 0),
                               (5, -50, 50, 1), (5, -5, 5, 1)],
                         ids=["wfadaptive_10_50_1", "wfadaptive_4_6_2", "banded_m40_5",
                              "banded_m3_3", "badaptive_m50_50_1", "badaptive_m5_5_1"])
@pytest.mark.parametrize("metric,pen", [BIWFA_PENALTIES[0], BIWFA_PENALTIES[2],
                                        BIWFA_PENALTIES[4], BIWFA_PENALTIES[6]])
def test_biwfa_heuristics_match_oracle(h, metric, pen, pkg, N, engine, oracle):
    pairs = _pairs(12, seed=13)
    for scope in ("score", "alignment"):
        _compare(engine, oracle, _attr(N, metric, pen, scope, h), pairs, scope == "alignment",
                 label=f"biwfa/{metric}/{h}/{scope}")


@pytest.mark.parametrize("h", [(2, 50, 1, 0), (2, 20, 2, 0), (2, 200, 1, 0)],
                         ids=["xdrop_50_1", "xdrop_20_2", "xdrop_200_1"])
@pytest.mark.parametrize("metric,pen", [BIWFA_PENALTIES[2], BIWFA_PENALTIES[3],
                                        BIWFA_PENALTIES[4], BIWFA_PENALTIES[7]])
def test_biwfa_xdrop_matches_oracle(h, metric, pen, pkg, N, engine, oracle):
    """X-drop sub-alignments (tests/xdrop_steps.rs:57-92 runs align_biwfa
    with X-drop), with the pair of tests/data/issue21_xdrop.seq riding
    along."""
    from wfa2lib_rs_amd.synth import from_ascii
    from test_oracle_heuristics import _issue21_pair
    ps = _recursive_pairs(12, seed=13) + [_issue21_pair()]
    pairs = from_ascii([p for p, _ in ps], [t for _, t in ps], N.SEQ_ASCII)
    for scope in ("score", "alignment"):
        _compare(engine, oracle, _attr(N, metric, pen, scope, h), pairs, scope == "alignment",
                 label=f"biwfa/{metric}/{h}/{scope}")


def test_biwfa_wide_rows_take_the_retry_path(pkg, N, engine, oracle):
    """Unrelated 40 kbp sequences: rows wider than the first launch's ring."""
    from wfa2lib_rs_amd.synth import from_ascii
    rng = np.random.default_rng(3)
    acgt = np.frombuffer(b"ACGT", dtype=np.uint8)
    ps = [(rng.choice(acgt, 40000).tobytes(), rng.choice(acgt, 38000).tobytes())
          for _ in range(2)]
    pairs = from_ascii([p for p, _ in ps], [t for _, t in ps], N.SEQ_ASCII)
    _compare(engine, oracle, _attr(N, "edit", (0,) * 6, "alignment"), pairs, True,
             label="biwfa/wide")


def test_biwfa_per_pair_equals_batch(pkg, N, engine, oracle):
    pairs = _pairs(6, seed=21)
    attr = _attr(N, "gap_affine_2p", (0, 4, 6, 2, 24, 1), "alignment")
    _compare(engine, oracle, attr, pairs, True, per_pair=True, label="biwfa/per_pair")


def test_biwfa_ultralong_matches_fixture(pkg, N, engine):
    """The 1k x 50-100 kbp workload's first pairs (2-bit packed, as the
    benchmark feeds them) at full size on the device: edit distance on 4
    pairs, gap-affine (0,4,6,2) on 2, full CIGAR; score, status and CIGAR
    digest equal the oracle's (tests/golden/ultralong50k_100k_biwfa.json)."""
    import hashlib
    from test_oracle_biwfa import _ultralong_fixture
    from wfa2lib_rs_amd.synth import WORKLOADS, generate
    d = _ultralong_fixture()
    n = max(r["pair"] for r in d["rows"]) + 1
    pairs = generate(WORKLOADS[d["workload"]], num_pairs=n)
    for metric in ("edit", "gap_affine"):
        rows = [r for r in d["rows"] if r["metric"] == metric]
        sub = pairs.subset([r["pair"] for r in rows])
        s, st, c = _run_gpu(engine, _attr(N, metric, rows[0]["penalties"], "alignment"), sub,
                            True)
        for j, r in enumerate(rows):
            got = (int(s[j]), int(st[j]), len(c[j]), hashlib.sha256(bytes(c[j])).hexdigest())
            assert got == (r["score"], r["status"], r["cigar_len"], r["sha256"]), (metric, j)

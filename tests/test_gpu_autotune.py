"""Launch shapes of long-read launches and their autotuner on the GPU: every
shape (waves per pair, workgroups per CU) gives the same results, equal to
the oracle's; the tuner keeps one of the shapes it timed."""
import numpy as np
import pytest

from conftest import attr_for, have_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine(pkg):
    if not have_gpu():
        pytest.skip("no GPU")
    from wfa2lib_rs_amd.engine import BatchEngine
    return BatchEngine(0)


def _pairs(n, seed=3):
    from wfa2lib_rs_amd.synth import WORKLOADS, generate
    w = WORKLOADS["long1k_10k"]
    from dataclasses import replace
    return generate(replace(w, len_hi=4000, seed=seed), num_pairs=n)


def _run(engine, attr, pairs, cigar):
    import torch
    from wfa2lib_rs_amd.engine import DeviceBatch
    b = DeviceBatch(pairs, device="cuda:0", cigar=cigar)
    engine.align(attr, b)
    torch.cuda.synchronize()
    return b.results()


SHAPES = [(4, 1), (4, 3), (8, 1), (8, 2), (16, 1), (16, 2)]


@pytest.mark.parametrize("metric,pen,scope,memory", [
    ("gap_affine_2p", (0, 4, 6, 2, 24, 1), "alignment", "med"),
    ("gap_affine_2p", (0, 4, 6, 2, 24, 1), "score", "high"),
    ("gap_affine", (0, 4, 6, 2, 0, 0), "alignment", "low"),
    ("edit", (0, 0, 0, 0, 0, 0), "alignment", "med"),
    ("gap_linear", (0, 4, 2, 0, 0, 0), "score", "high"),
], ids=["affine2p_med", "affine2p_score", "affine_low", "edit_med", "linear_score"])
def test_every_shape_gives_the_oracle_results(metric, pen, scope, memory, pkg, N, engine,
                                              oracle):
    pairs = _pairs(96)
    attr = attr_for(N, metric, pen, scope)
    attr.memory_mode = {"high": 0, "med": 1, "low": 2}[memory]
    cigar = scope == "alignment"
    maxlen = int(max(pairs.plen.max(), pairs.tlen.max()))
    want = oracle.align_batch(attr, pairs, cigar=cigar, threads=16)
    for shape in SHAPES:
        engine.set_launch_shape(attr, maxlen, *shape)
        got = _run(engine, attr, pairs, cigar)
        assert np.array_equal(got[0], want[0]), shape
        assert np.array_equal(got[1], want[1]), shape
        if cigar:
            assert got[2] == want[2], shape
    engine.set_launch_shape(attr, maxlen, 0, 0)
    assert engine.launch_shape(attr, maxlen) == (0, 0)


def test_autotune_keeps_a_timed_shape(pkg, N, engine, oracle):
    import torch
    from wfa2lib_rs_amd.engine import DeviceBatch
    pairs = _pairs(256, seed=5)
    attr = attr_for(N, "gap_affine_2p", (0, 4, 6, 2, 24, 1), "alignment")
    attr.memory_mode = 1
    maxlen = int(max(pairs.plen.max(), pairs.tlen.max()))
    b = DeviceBatch(pairs, device="cuda:0", cigar=True)
    rep = engine.autotune(attr, b, sample_pairs=128)
    torch.cuda.synchronize()
    assert rep["sample_pairs"] == 128 and len(rep["trials"]) >= 3
    kept = (rep["waves_per_pair"], rep["groups_per_cu"])
    assert kept[0] in (4, 8, 16)
    assert any(t[0] == kept[0] for t in rep["trials"])
    assert engine.launch_shape(attr, maxlen) == kept
    # the tuned launch gives the oracle's results
    engine.align(attr, b)
    torch.cuda.synchronize()
    got = b.results()
    want = oracle.align_batch(attr, pairs, cigar=True, threads=16)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert got[2] == want[2]
    # short reads and heuristic launches have nothing to tune
    short = DeviceBatch(pairs.subset(range(4)), device="cuda:0", cigar=True)
    h = attr_for(N, "gap_affine", (0, 4, 6, 2, 0, 0), "alignment")
    h.heuristic, h.heuristic_p[0], h.heuristic_p[1] = N.HEURISTIC_XDROP, 200, 1
    assert engine.autotune(h, short)["waves_per_pair"] == 0

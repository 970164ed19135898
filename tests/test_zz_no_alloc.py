"""wfa_align_batch_noalloc against wfa_engine_align_batch, through the C ABI.

The no-allocation entry point carves every buffer from one fixed device
array; its results must equal the engine's bit for bit (score, status,
CIGAR) on the same batch, on consecutive calls whose sizes shrink (the
array keeps the previous call's contents), and it must refuse what is
beyond its limits.
"""
import ctypes

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


def _noalloc(N, attr, batch):
    import torch
    b = batch.c_batch()
    return N.lib().wfa_align_batch_noalloc(ctypes.byref(attr), ctypes.byref(b),
                                           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def _both(N, engine, attr, pairs, cigar, label):
    """Align `pairs` with the engine and with the no-allocation entry point;
    assert identical outputs."""
    import torch
    from wfa2lib_rs_amd.engine import DeviceBatch
    print(f"[no-alloc] {label}: {len(pairs)} pairs, cigar={cigar}", flush=True)
    ref = DeviceBatch(pairs, device="cuda:0", cigar=cigar)
    engine.align(attr, ref)
    got = DeviceBatch(pairs, device="cuda:0", cigar=cigar)
    # stale values that a missing write would leave visible
    got.score.fill_(-7)
    got.status.fill_(-7)
    rc = _noalloc(N, attr, got)
    torch.cuda.synchronize()
    assert rc == 0, f"{label}: wfa_align_batch_noalloc returned {rc} ({N.ERRORS.get(rc)})"
    s0, st0, c0 = ref.results()
    s1, st1, c1 = got.results()
    bad = np.nonzero((s0 != s1) | (st0 != st1))[0]
    assert len(bad) == 0, f"{label}: {len(bad)} pairs differ, first {bad[:5]}"
    if cigar:
        diff = [i for i in range(len(pairs)) if c0[i] != c1[i]]
        assert not diff, f"{label}: {len(diff)} CIGARs differ, first {diff[:5]}"
    return st1


def test_noalloc_short_reads_two_calls(pkg, N, engine):
    """gap-affine CIGAR (high memory) on 100-300 bp pairs: a larger call,
    then a smaller one."""
    from wfa2lib_rs_amd.synth import WORKLOADS, generate
    attr = attr_for(N, "gap_affine", (0, 4, 6, 2, 0, 0), "alignment")
    pairs = generate(WORKLOADS["reads100_300"], num_pairs=30000)
    st = _both(N, engine, attr, pairs, True, "reads100_300 x30000")
    assert (st == N.STATUS_COMPLETED).all()
    _both(N, engine, attr, pairs.subset(range(2500)), True, "reads100_300 x2500")


def test_noalloc_score_only_and_medium_memory(pkg, N, engine):
    from wfa2lib_rs_amd.synth import WORKLOADS, generate
    pairs = generate(WORKLOADS["reads100_300"], num_pairs=8000)
    attr = attr_for(N, "gap_affine", (0, 4, 6, 2, 0, 0), "score")
    _both(N, engine, attr, pairs, False, "score only x8000")
    attr = attr_for(N, "gap_affine_2p", (0, 4, 6, 2, 24, 1), "alignment")
    attr.memory_mode = N.MEMORY_MED
    _both(N, engine, attr, pairs.subset(range(3000)), True, "affine2p medium memory x3000")


def test_noalloc_long_reads_two_calls(pkg, N, engine):
    """1-10 kbp pairs on the wide kernels (HBM rings, longest-first order,
    retry passes), then a smaller call."""
    from wfa2lib_rs_amd.synth import WORKLOADS, generate
    attr = attr_for(N, "gap_affine_2p", (0, 4, 6, 2, 24, 1), "alignment")
    attr.memory_mode = N.MEMORY_MED
    pairs = generate(WORKLOADS["long1k_10k"], num_pairs=96)
    _both(N, engine, attr, pairs, True, "long1k_10k x96")
    _both(N, engine, attr, pairs.subset(range(17)), True, "long1k_10k x17")


def test_noalloc_limits(pkg, N, engine):
    """Beyond the fixed limits: WFA_ERR_LIMIT, and nothing runs."""
    from wfa2lib_rs_amd.engine import DeviceBatch
    from wfa2lib_rs_amd.synth import from_ascii
    attr = attr_for(N, "gap_affine", (0, 4, 6, 2, 0, 0), "score")
    batch = DeviceBatch(from_ascii([b"ACGT"], [b"ACGA"]), device="cuda:0")
    b = batch.c_batch()
    b.num_pairs = N.NOALLOC_MAX_PAIRS + 1
    rc = N.lib().wfa_align_batch_noalloc(ctypes.byref(attr), ctypes.byref(b), None)
    assert rc == N.ERR_LIMIT
    b = batch.c_batch()
    b.max_length = N.NOALLOC_MAX_LENGTH + 1
    assert N.lib().wfa_align_batch_noalloc(ctypes.byref(attr), ctypes.byref(b), None) == N.ERR_LIMIT
    attr.memory_mode = N.MEMORY_ULTRALOW
    b = batch.c_batch()
    assert N.lib().wfa_align_batch_noalloc(ctypes.byref(attr), ctypes.byref(b), None) == N.ERR_LIMIT

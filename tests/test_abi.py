"""CPU tests of the boundary: the library loads, exports every entry point the
header declares, and validates penalties as the reference does. No compute
call is made without a GPU."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def _declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            for m in re.finditer(r"\b(wfa_[a-z0-9_]+)\s*\(", text):
                names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(N):
    lib = ctypes.CDLL(N.LIB_PATH)
    declared = _declared_symbols()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes signature table covers all of them
    assert declared == set(N.SIGNATURES), declared ^ set(N.SIGNATURES)


def test_version_and_defaults(N):
    L = N.lib()
    assert b"gfx950" in L.wfa_version()
    a = N.Attr()
    L.wfa_attr_default(ctypes.byref(a), N.METRIC_GAP_AFFINE)
    assert (a.penalties.match, a.penalties.mismatch, a.penalties.gap_opening1,
            a.penalties.gap_extension1) == (0, 4, 6, 2)
    assert a.max_alignment_steps == 2 ** 31 - 1
    assert a.alignment_scope == N.SCOPE_SCORE and a.memory_mode == N.MEMORY_HIGH


@pytest.mark.parametrize("pen,ok", [
    ((3, 0, 4, 6, 2, 0, 0), True),
    ((3, 1, 4, 6, 2, 0, 0), False),   # match > 0
    ((3, 0, 0, 6, 2, 0, 0), False),   # mismatch <= 0
    ((3, 0, 4, -1, 2, 0, 0), False),  # gap opening < 0
    ((3, 0, 4, 6, 0, 0, 0), False),   # gap extension <= 0
    ((4, 0, 4, 6, 2, 24, 1), True),
    ((4, 0, 4, 6, 2, 24, 0), False),
    ((2, 0, 4, 2, 0, 0, 0), True),
    ((2, 0, 4, 0, 0, 0, 0), False),
    ((1, 0, 0, 0, 0, 0, 0), True),
])
def test_penalty_validation_matches_reference(N, pkg, pen, ok):
    """src/penalties.rs:157-167,198-209,245-262 panic on these; the ABI
    returns WFA_ERR_INVALID_PENALTIES and Python raises ValueError."""
    p = N.Penalties(*pen)
    code = N.lib().wfa_penalties_check(ctypes.byref(p))
    assert (code == 0) == ok
    if not ok:
        with pytest.raises(ValueError):
            pkg.WavefrontPenalties(p, pen)


def test_engine_refuses_without_device(N):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except Exception:
        pass
    err = ctypes.c_int32(0)
    h = N.lib().wfa_engine_new(0, ctypes.byref(err))
    assert not h and err.value in (-5, -4)


def test_cigar_helpers(pkg):
    c = pkg.Cigar(b"MMMXMMMM")
    assert c.to_string_rle(True) == "3M1X4M"
    assert c.to_string_rle(False) == "1X"
    assert c.score_gap_affine(pkg.AffinePenalties(0, 4, 6, 2)) == -4
    assert pkg.Cigar(b"MMIIM").score_gap_affine(pkg.AffinePenalties(0, 4, 6, 2)) == -10
    assert pkg.encode_packed2bits(b"ACGTTGCA") == bytes([0b11100100, 0b00011011])


def test_server_refuses_without_device(N):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except Exception:
        pass
    cfg = N.ServerConfig()
    N.lib().wfa_server_config_default(ctypes.byref(cfg))
    assert (cfg.num_engines, cfg.max_batch_pairs, cfg.memory_budget) == (4, 65536, 64 << 30)
    err = ctypes.c_int32(0)
    assert not N.lib().wfa_server_new(ctypes.byref(cfg), ctypes.byref(err))
    assert err.value in (-5, -4)
    cfg.num_engines = 0
    assert not N.lib().wfa_server_new(ctypes.byref(cfg), ctypes.byref(err))
    assert err.value == -2


def test_serving_mix_shape(pkg):
    """The serving mix: deterministic, lengths/spans/heuristics in their
    stated shares, arrivals increasing."""
    from wfa2lib_rs_amd.synth import SERVING_HEURISTICS, serving_mix
    lengths = ((0.7, 0, 150, 150), (0.3, 2, 1000, 2000))
    a = serving_mix(400, 1000.0, seed=1, lengths=lengths)
    b = serving_mix(400, 1000.0, seed=1, lengths=lengths)
    assert [(r.pattern, r.text, r.span, r.heuristic) for r in a] == \
        [(r.pattern, r.text, r.span, r.heuristic) for r in b]
    short = sum(len(r.pattern) == 150 for r in a)
    assert 240 <= short <= 320
    assert all(1000 <= len(r.pattern) <= 2000 for r in a if len(r.pattern) != 150)
    spans = [r.span for r in a]
    assert 150 <= spans.count("end2end") <= 250 and spans.count("extension") > 60
    assert set(r.heuristic for r in a) == set(SERVING_HEURISTICS)
    t = [r.arrival_s for r in a]
    assert all(x < y for x, y in zip(t, t[1:]))
    assert 0.25 < t[-1] < 0.6  # 400 arrivals at 1000/s


def test_max_memory_compact_null_aligner(N):
    """set_max_memory_compact (src/aligner.rs:469) on the C ABI: a null
    aligner is refused and reads back 0 (the aligner itself needs a GPU;
    tests/test_gpu_parity.py covers the stored value and unchanged results)."""
    L = N.lib()
    assert L.wfa_aligner_set_max_memory_compact(None, 1 << 20) == N.ERR_INVALID_ARGUMENT
    assert L.wfa_aligner_max_memory_compact(None) == 0

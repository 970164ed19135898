import ctypes
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as entry  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")

METRICS = {"indel": 0, "edit": 1, "gap_linear": 2, "gap_affine": 3, "gap_affine_2p": 4}


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def pkg():
    return entry.load_package()


@pytest.fixture(scope="session")
def oracle():
    return entry.load_oracle()


@pytest.fixture(scope="session")
def N(pkg):
    from wfa2lib_rs_amd import _native
    return _native


def load_vectors():
    with open(os.path.join(GOLDEN, "reference_vectors.json")) as f:
        return json.load(f)["vectors"]


def expand(seq):
    """'A*1500' → 1500 A's; anything else verbatim."""
    if "*" in seq:
        ch, n = seq.split("*")
        return (ch * int(n)).encode()
    return seq.encode()


HEURISTICS = {"wfadaptive": 1, "xdrop": 2, "zdrop": 3, "banded_static": 4, "banded_adaptive": 5}
MEMORY = {"high": 0, "med": 1, "low": 2, "ultralow": 3}


def attr_for_vec(N, vec):
    """The aligner configuration a golden vector describes."""
    a = attr_for(N, vec["metric"], vec["penalties"], vec["scope"], vec.get("max_steps"),
                 vec.get("span"))
    if "heuristic" in vec:
        h = vec["heuristic"]
        a.heuristic = HEURISTICS[h[0]]
        for i, v in enumerate(h[1:]):
            a.heuristic_p[i] = v
    if "memory" in vec:
        a.memory_mode = MEMORY[vec["memory"]]
    return a


def attr_for(N, metric, penalties, scope, max_steps=None, span=None):
    a = N.Attr()
    N.lib().wfa_attr_default(ctypes.byref(a), METRICS[metric] if isinstance(metric, str) else metric)
    m, x, o1, e1, o2, e2 = penalties
    if a.penalties.distance_metric >= 2:
        a.penalties.match, a.penalties.mismatch = m, x
        a.penalties.gap_opening1, a.penalties.gap_extension1 = o1, e1
        a.penalties.gap_opening2, a.penalties.gap_extension2 = o2, e2
    a.alignment_scope = 1 if scope in ("alignment", 1) else 0
    if max_steps is not None:
        a.max_alignment_steps = max_steps
    if span == "extension":
        a.span, a.extension = N.SPAN_ENDSFREE, 1
    elif span is not None:
        a.span = N.SPAN_ENDSFREE
        (a.pattern_begin_free, a.pattern_end_free, a.text_begin_free,
         a.text_end_free) = span
    return a


def have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False

"""Serving loop: continuous batching of alignment requests on one device.

``AlignmentServer`` wraps ``wfa_server_*`` (include/wfa_mi355x.h). Each
request carries its own aligner configuration (an ``Attr``: penalties,
span, heuristic, memory mode, scope) and its sequences; the library queues
it by class (configuration, format, length bucket), admits it against the
predicted wavefront memory of the work already on the device, and runs it
in a shared persistent launch on one of several engines/streams. A result
is the same as the one the per-call aligner gives for that request alone.
"""
import ctypes
from dataclasses import dataclass

from . import _native as N


@dataclass
class ServerResult:
    tag: int
    error: int
    score: int
    status: int
    cigar: bytes
    latency_us: float


class AlignmentServer:
    def __init__(self, device=0, num_engines=4, max_batch_pairs=1 << 16,
                 memory_budget=64 << 30):
        cfg = N.ServerConfig()
        N.lib().wfa_server_config_default(ctypes.byref(cfg))
        cfg.device = device
        cfg.num_engines = num_engines
        cfg.max_batch_pairs = max_batch_pairs
        cfg.memory_budget = memory_budget
        err = ctypes.c_int32(0)
        self._h = N.lib().wfa_server_new(ctypes.byref(cfg), ctypes.byref(err))
        if not self._h:
            N.check(err.value, "wfa_server_new")
        self._buf = (N.ServerResult * 4096)()

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().wfa_server_free(h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, attr, pattern: bytes, text: bytes, tag: int, seq_format=N.SEQ_ASCII,
               pattern_length=None, text_length=None):
        """Queue one request; lengths default to the byte counts (ASCII)."""
        r = N.ServerRequest()
        r.attr = attr
        r.seq_format = seq_format
        r.pattern_length = len(pattern) if pattern_length is None else pattern_length
        r.text_length = len(text) if text_length is None else text_length
        r.pattern = pattern
        r.text = text
        r.tag = tag
        N.check(N.lib().wfa_server_submit(self._h, ctypes.byref(r)), "wfa_server_submit")

    def poll(self, timeout_us=0, max_results=4096):
        """Finished requests (completion order); waits up to timeout_us
        (negative: until one is ready or nothing is pending)."""
        n = N.lib().wfa_server_poll(self._h, self._buf, min(max_results, len(self._buf)),
                                    int(timeout_us))
        if n < 0:
            N.check(n, "wfa_server_poll")
        out = []
        for i in range(n):
            r = self._buf[i]
            cig = ctypes.string_at(r.cigar_ops, r.cigar_length) if r.cigar_length > 0 else b""
            out.append(ServerResult(r.tag, r.error, r.score, r.status, cig, r.latency_us))
        return out

    def stats(self):
        v = [ctypes.c_int64(0) for _ in range(4)]
        N.check(N.lib().wfa_server_stats(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("queued", "running", "launches", "launched_pairs"),
                        (x.value for x in v)))

    def drain(self, expected, timeout_s=600.0):
        """Poll until `expected` results arrived (raises on timeout)."""
        import time
        got = []
        t_end = time.monotonic() + timeout_s
        while len(got) < expected:
            if time.monotonic() > t_end:
                raise TimeoutError(f"{len(got)}/{expected} results after {timeout_s} s")
            got += self.poll(timeout_us=100_000)
        return got


_HEURISTIC_KIND = {"none": N.HEURISTIC_NONE, "wfadaptive": N.HEURISTIC_WFADAPTIVE,
                   "xdrop": N.HEURISTIC_XDROP, "banded_static": N.HEURISTIC_BANDED_STATIC,
                   "banded_adaptive": N.HEURISTIC_BANDED_ADAPTIVE}


def request_attr(span, heuristic, plen, tlen, penalties=(0, 4, 6, 2), scope=N.SCOPE_ALIGNMENT,
                 memory_mode=N.MEMORY_HIGH):
    """The aligner configuration of one serving request: gap-affine
    penalties, `span` "end2end" / "endsfree" (the driver's nominal 0.1 of
    each length free at every end, src/bin/align_benchmark.rs:255-263,
    600-607) / "extension", and a heuristic tuple (name, params...)."""
    a = N.Attr()
    N.lib().wfa_attr_default(ctypes.byref(a), N.METRIC_GAP_AFFINE)
    (a.penalties.match, a.penalties.mismatch, a.penalties.gap_opening1,
     a.penalties.gap_extension1) = penalties
    a.alignment_scope = scope
    a.memory_mode = memory_mode
    if span == "extension":
        a.span, a.extension = N.SPAN_ENDSFREE, 1
    elif span == "endsfree":
        a.span = N.SPAN_ENDSFREE
        a.pattern_begin_free = a.pattern_end_free = int(plen * 0.1)
        a.text_begin_free = a.text_end_free = int(tlen * 0.1)
    elif span != "end2end":
        raise ValueError(f"unknown span {span!r}")
    a.heuristic = _HEURISTIC_KIND[heuristic[0]]
    for i, v in enumerate(heuristic[1:]):
        a.heuristic_p[i] = v
    return a

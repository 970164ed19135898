"""Deterministic synthetic workloads (csrc/synth.c) as numpy arrays.

``WORKLOADS`` holds the benchmark configurations: name → generator
parameters, sequence format and the number of pairs.
"""
import ctypes
import os
from dataclasses import dataclass, replace

import numpy as np

from . import _native as N


@dataclass(frozen=True)
class Workload:
    name: str
    num_pairs: int
    len_mode: int          # 0 fixed, 1 uniform, 2 log-uniform
    len_lo: int
    len_hi: int
    div_lo: float
    div_hi: float
    indel_mean: float
    run_frac: float = 0.0
    run_mean: float = 1.0
    seq_format: int = N.SEQ_PACKED2BITS
    seed: int = 0

    def with_pairs(self, n):
        return replace(self, num_pairs=n)


WORKLOADS = {
    # 10k x 150 bp, 5% divergence, equal sub/ins/del, indels geometric mean 2, ASCII
    "short150": Workload("short150", 10_000, 0, 150, 150, 0.05, 0.05, 2.0,
                         seq_format=N.SEQ_ASCII),
    # 1M pairs, lengths uniform 100-300 bp, divergence uniform 0-10%, 2-bit packed
    "reads100_300": Workload("reads100_300", 1_000_000, 1, 100, 300, 0.0, 0.10, 2.0),
    # 100k pairs, lengths log-uniform 1-10 kbp, divergence 1-10%, indels mean 5
    "long1k_10k": Workload("long1k_10k", 100_000, 2, 1000, 10000, 0.01, 0.10, 5.0),
    # 1k pairs, 50-100 kbp, divergence 5-15%, 20% of events indel runs (mean 20)
    "ultralong50k_100k": Workload("ultralong50k_100k", 1000, 1, 50_000, 100_000, 0.05, 0.15,
                                  1.0, run_frac=0.2, run_mean=20.0),
}


@dataclass
class HostPairs:
    """Pairs in one contiguous byte buffer, offsets in bytes."""
    seq: np.ndarray      # uint8
    poff: np.ndarray     # int64
    plen: np.ndarray     # int32
    toff: np.ndarray     # int64
    tlen: np.ndarray     # int32
    seq_format: int

    def __len__(self):
        return len(self.plen)

    def max_length(self):
        if len(self.plen) == 0:
            return 0
        return int(max(self.plen.max(), self.tlen.max()))

    def pattern(self, i):
        """Pair i's pattern as ASCII bytes."""
        return _ascii(self.seq, int(self.poff[i]), int(self.plen[i]), self.seq_format)

    def text(self, i):
        return _ascii(self.seq, int(self.toff[i]), int(self.tlen[i]), self.seq_format)

    def subset(self, idx):
        idx = np.asarray(idx)
        pats = [self.pattern(i) for i in idx]
        txts = [self.text(i) for i in idx]
        return from_ascii(pats, txts, self.seq_format)


def _ascii(seq, off, n, fmt):
    if fmt == N.SEQ_ASCII:
        return bytes(seq[off:off + n])
    nb = (n + 3) // 4
    b = seq[off:off + nb].astype(np.uint8)
    sym = np.stack([(b >> (2 * j)) & 3 for j in range(4)], axis=1).reshape(-1)[:n]
    return bytes(np.frombuffer(b"ACGT", dtype=np.uint8)[sym])


def layout(plen, tlen, fmt, align=1):
    """Byte offsets for pattern i then text i, back to back."""
    pb = (plen.astype(np.int64) + 3) // 4 if fmt == N.SEQ_PACKED2BITS else plen.astype(np.int64)
    tb = (tlen.astype(np.int64) + 3) // 4 if fmt == N.SEQ_PACKED2BITS else tlen.astype(np.int64)
    if align > 1:
        pb = (pb + align - 1) // align * align
        tb = (tb + align - 1) // align * align
    sizes = np.empty(2 * len(plen), dtype=np.int64)
    sizes[0::2] = pb
    sizes[1::2] = tb
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]) if len(sizes) else np.zeros(0, np.int64)
    total = int(sizes.sum()) if len(sizes) else 0
    return starts[0::2].copy(), starts[1::2].copy(), total


def generate(w: Workload, num_pairs=None, threads=None) -> HostPairs:
    n = w.num_pairs if num_pairs is None else num_pairs
    threads = threads or min(16, os.cpu_count() or 1)
    cfg = N.SynthCfg(w.seed, w.len_mode, w.len_lo, w.len_hi, w.div_lo, w.div_hi, w.indel_mean,
                     w.run_frac, w.run_mean)
    S = N.synth_lib()
    plen = np.zeros(n, dtype=np.int32)
    tlen = np.zeros(n, dtype=np.int32)
    S.wfa_synth_lengths(ctypes.byref(cfg), n, plen.ctypes.data, tlen.ctypes.data, threads)
    poff, toff, total = layout(plen, tlen, w.seq_format)
    seq = np.zeros(max(total, 1), dtype=np.uint8)
    S.wfa_synth_fill(ctypes.byref(cfg), n, plen.ctypes.data, tlen.ctypes.data, w.seq_format,
                     seq.ctypes.data, poff.ctypes.data, toff.ctypes.data, threads)
    return HostPairs(seq, poff, plen, toff, tlen, w.seq_format)


def from_ascii(patterns, texts, fmt=N.SEQ_ASCII) -> HostPairs:
    """Pack explicit pairs (bytes) into a HostPairs buffer."""
    plen = np.array([len(p) for p in patterns], dtype=np.int32)
    tlen = np.array([len(t) for t in texts], dtype=np.int32)
    poff, toff, total = layout(plen, tlen, fmt)
    seq = np.zeros(max(total, 1), dtype=np.uint8)
    code = np.zeros(256, dtype=np.uint8)
    for c, v in zip(b"ACGTacgt", [0, 1, 2, 3, 0, 1, 2, 3]):
        code[c] = v
    for i, (p, t) in enumerate(zip(patterns, texts)):
        for s, off in ((p, poff[i]), (t, toff[i])):
            a = np.frombuffer(bytes(s), dtype=np.uint8)
            if fmt == N.SEQ_ASCII:
                seq[off:off + len(a)] = a
            else:
                sym = code[a]
                nb = (len(a) + 3) // 4
                padded = np.zeros(nb * 4, dtype=np.uint8)
                padded[:len(a)] = sym
                q = padded.reshape(-1, 4)
                seq[off:off + nb] = q[:, 0] | (q[:, 1] << 2) | (q[:, 2] << 4) | (q[:, 3] << 6)
    return HostPairs(seq, poff, plen, toff, tlen, fmt)


# ---- serving mix: Poisson arrivals of mixed lengths, spans and heuristics ----
# heuristics as the reference's tests parameterize them: wf-adaptive
# (10, 50, 1), X-drop (200, 1), banded-adaptive (-50, 50, 1)
SERVING_HEURISTICS = (("none",), ("wfadaptive", 10, 50, 1), ("xdrop", 200, 1),
                      ("banded_adaptive", -50, 50, 1))
# (share, length mode, shortest, longest): 70% 150 bp, 25% 1-10 kbp,
# 5% 10-100 kbp (log-uniform)
SERVING_LENGTHS = ((0.70, 0, 150, 150), (0.25, 2, 1000, 10_000), (0.05, 2, 10_000, 100_000))


@dataclass
class ServingRequest:
    pattern: bytes        # ASCII
    text: bytes
    span: str             # "end2end", "endsfree" (0.1 of each length free) or "extension"
    heuristic: tuple      # one of SERVING_HEURISTICS
    arrival_s: float      # Poisson arrival time from the start of the stream


def serving_mix(n, rate, seed=0, lengths=SERVING_LENGTHS, div=(0.01, 0.10), indel_mean=2.0):
    """`n` requests arriving at `rate` per second: half end-to-end, a quarter
    ends-free (0.1 of each sequence's length free at every end), a quarter
    extension; heuristic uniform over SERVING_HEURISTICS."""
    rng = np.random.default_rng(seed)
    shares = np.array([c[0] for c in lengths], dtype=np.float64)
    cls = rng.choice(len(lengths), size=n, p=shares / shares.sum())
    span = rng.choice(["end2end", "end2end", "endsfree", "extension"], size=n)
    heur = rng.integers(0, len(SERVING_HEURISTICS), size=n)
    arrival = np.cumsum(rng.exponential(1.0 / rate, size=n)) if n else np.zeros(0)
    pats = [None] * n
    txts = [None] * n
    for c, (_, mode, lo, hi) in enumerate(lengths):
        idx = np.nonzero(cls == c)[0]
        if len(idx) == 0:
            continue
        w = Workload(f"serving{c}", len(idx), mode, lo, hi, div[0], div[1], indel_mean,
                     seq_format=N.SEQ_ASCII, seed=seed * 7 + c + 1)
        p = generate(w)
        for j, i in enumerate(idx):
            pats[i] = p.pattern(j)
            txts[i] = p.text(j)
    return [ServingRequest(pats[i], txts[i], str(span[i]), SERVING_HEURISTICS[heur[i]],
                           float(arrival[i])) for i in range(n)]

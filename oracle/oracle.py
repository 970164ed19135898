"""ctypes binding of the CPU restatement (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline
leg, always as the checker or the CPU baseline, never as the engine.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB_PATH


def lib(attr_type):
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = ctypes.CDLL(LIB_PATH)
        P, I32, VP = ctypes.POINTER, ctypes.c_int32, ctypes.c_void_p
        l.wfo_aligner_new.restype = VP
        l.wfo_aligner_new.argtypes = [P(attr_type), P(I32)]
        l.wfo_aligner_free.argtypes = [VP]
        l.wfo_align_end2end.restype = I32
        l.wfo_align_end2end.argtypes = [VP, ctypes.c_char_p, I32, ctypes.c_char_p, I32]
        l.wfo_align_endsfree.restype = I32
        l.wfo_align_endsfree.argtypes = [VP, ctypes.c_char_p, I32, ctypes.c_char_p, I32]
        for name in ("wfo_align_biwfa", "wfo_align_biwfa_score"):
            f = getattr(l, name)
            f.restype = I32
            f.argtypes = [VP, ctypes.c_char_p, I32, ctypes.c_char_p, I32]
        l.wfo_status.restype = I32
        l.wfo_status.argtypes = [VP]
        l.wfo_end_position.restype = None
        l.wfo_end_position.argtypes = [VP, P(I32)]
        l.wfo_cigar.restype = I32
        l.wfo_cigar.argtypes = [VP, P(P(ctypes.c_uint8))]
        l.wfo_cigar_score.restype = I32
        l.wfo_cigar_score.argtypes = [VP]
        l.wfo_cigar_check.restype = I32
        l.wfo_cigar_check.argtypes = [ctypes.c_char_p, I32, ctypes.c_char_p, I32,
                                      ctypes.c_char_p, I32]
        l.wfo_align_batch.restype = I32
        l.wfo_align_batch.argtypes = [P(attr_type), I32, I32, VP, VP, VP, VP, VP, VP, VP, VP, VP,
                                      VP, I32]
        l.dp_affine_score.restype = I32
        l.dp_affine_score.argtypes = [ctypes.c_char_p, I32, ctypes.c_char_p, I32, I32, I32, I32,
                                      I32, I32, I32]
        l.dp_edit_score.restype = I32
        l.dp_edit_score.argtypes = [ctypes.c_char_p, I32, ctypes.c_char_p, I32, I32]
        l.dp_endsfree_penalty.restype = ctypes.c_int64
        l.dp_endsfree_penalty.argtypes = [ctypes.c_char_p, I32, ctypes.c_char_p, I32] + [I32] * 11
        _lib = l
    return _lib


class OracleAligner:
    """One CPU aligner configured by a wfa_attr_t (ctypes Attr)."""

    def __init__(self, attr):
        self.L = lib(type(attr))
        err = ctypes.c_int32(0)
        self.h = self.L.wfo_aligner_new(ctypes.byref(attr), ctypes.byref(err))
        self.endsfree = attr.span == 1
        self.biwfa = attr.memory_mode == 3
        self.score_only = attr.alignment_scope == 0
        if not self.h:
            raise ValueError(f"oracle rejected the configuration ({err.value})")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.wfo_aligner_free(self.h)
            self.h = None

    def align(self, pattern: bytes, text: bytes):
        """align_biwfa / align_biwfa_score in ultralow memory, align_endsfree
        when the attribute's span is ends-free, else end-to-end."""
        if self.biwfa:
            f = self.L.wfo_align_biwfa_score if self.score_only else self.L.wfo_align_biwfa
            return self._finish(f(self.h, pattern, len(pattern), text, len(text)))
        if self.endsfree:
            return self._finish(self.L.wfo_align_endsfree(self.h, pattern, len(pattern), text,
                                                          len(text)))
        return self.align_end2end(pattern, text)

    def align_end2end(self, pattern: bytes, text: bytes):
        return self._finish(self.L.wfo_align_end2end(self.h, pattern, len(pattern), text,
                                                     len(text)))

    def cigar_score(self):
        """The CIGAR's score field (Cigar.score) after the last alignment."""
        return self.L.wfo_cigar_score(self.h)

    def end_position(self):
        """alignment_end_pos (score, k, offset) after the last alignment."""
        out = (ctypes.c_int32 * 3)()
        self.L.wfo_end_position(self.h, out)
        return tuple(out)

    def _finish(self, s):
        st = self.L.wfo_status(self.h)
        ptr = ctypes.POINTER(ctypes.c_uint8)()
        n = self.L.wfo_cigar(self.h, ctypes.byref(ptr))
        ops = ctypes.string_at(ptr, n) if n > 0 else b""
        return s, st, ops


def align_batch(attr, pairs, cigar=False, threads=1):
    """Align HostPairs on the CPU: (scores, status, cigars or None)."""
    L = lib(type(attr))
    n = len(pairs)
    score = np.zeros(max(n, 1), dtype=np.int32)
    status = np.zeros(max(n, 1), dtype=np.int32)
    ops = off = ln = None
    if cigar:
        cap = pairs.plen.astype(np.int64) + pairs.tlen.astype(np.int64)
        off = np.concatenate([[0], np.cumsum(cap)[:-1]]).astype(np.int64) if n else np.zeros(1, np.int64)
        ops = np.zeros(max(int(cap.sum()), 1), dtype=np.uint8)
        ln = np.zeros(max(n, 1), dtype=np.int32)
    ptr = lambda a: a.ctypes.data if a is not None else None
    r = L.wfo_align_batch(ctypes.byref(attr), n, pairs.seq_format, ptr(pairs.seq), ptr(pairs.poff),
                          ptr(pairs.plen), ptr(pairs.toff), ptr(pairs.tlen), ptr(score),
                          ptr(status), ptr(ops), ptr(off), ptr(ln), threads)
    if r != 0:
        raise RuntimeError(f"oracle batch failed ({r})")
    cigars = None
    if cigar:
        cigars = [bytes(ops[off[i]:off[i] + ln[i]]) for i in range(n)]
    return score[:n], status[:n], cigars


def dp_affine(pattern, text, x, o1, e1, o2=0, e2=0, two_piece=False, attr_type=None):
    L = _lib
    return L.dp_affine_score(pattern, len(pattern), text, len(text), x, o1, e1, o2, e2,
                             1 if two_piece else 0)


def dp_edit(pattern, text, indel=False):
    return _lib.dp_edit_score(pattern, len(pattern), text, len(text), 1 if indel else 0)


def dp_endsfree(pattern, text, match, x, o1, e1, o2, e2, two_piece, ends):
    """Optimal ends-free penalty (matches cost `match` <= 0); ends =
    (pattern begin, pattern end, text begin, text end) free lengths."""
    pb, pe, tb, te = ends
    return _lib.dp_endsfree_penalty(pattern, len(pattern), text, len(text), match, x, o1, e1,
                                    o2, e2, 1 if two_piece else 0, pb, pe, tb, te)

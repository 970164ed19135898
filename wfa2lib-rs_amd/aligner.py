"""The reference's aligner API, over the engine's C ABI.

Names, argument meaning and error behaviour follow HPCBio/wfa2lib-rs:
``WavefrontPenalties`` (src/penalties.rs), ``WavefrontAligner`` with its
setters and ``align_*`` methods (src/aligner.rs:146-211, 462-841),
``AlignStatus`` (src/aligner.rs:120-135) and ``Cigar`` (src/cigar.rs). Every
alignment runs on the GPU through ``libwfa_mi355x.so``; nothing here aligns
on the CPU.
"""
import ctypes
from dataclasses import dataclass
from enum import IntEnum

from . import _native as N


class DistanceMetric(IntEnum):
    """src/penalties.rs:12-24"""
    INDEL = N.METRIC_INDEL
    EDIT = N.METRIC_EDIT
    GAP_LINEAR = N.METRIC_GAP_LINEAR
    GAP_AFFINE = N.METRIC_GAP_AFFINE
    GAP_AFFINE_2P = N.METRIC_GAP_AFFINE_2P


class AlignmentScope(IntEnum):
    """src/aligner.rs:26-31"""
    COMPUTE_SCORE = N.SCOPE_SCORE
    COMPUTE_ALIGNMENT = N.SCOPE_ALIGNMENT


class MemoryMode(IntEnum):
    """src/aligner.rs:44-52; ULTRALOW is BiWFA (src/aligner.rs:2563)."""
    HIGH = N.MEMORY_HIGH
    MED = N.MEMORY_MED
    LOW = N.MEMORY_LOW
    ULTRALOW = N.MEMORY_ULTRALOW


class AlignmentSpan(IntEnum):
    """src/aligner.rs:109-115"""
    END2END = N.SPAN_END2END
    ENDS_FREE = N.SPAN_ENDSFREE


class AlignStatus(IntEnum):
    """src/aligner.rs:120-135, same ordinals."""
    OK = N.STATUS_OK
    END_REACHED = N.STATUS_END_REACHED
    END_UNREACHABLE = N.STATUS_END_UNREACHABLE
    MAX_STEPS_REACHED = N.STATUS_MAX_STEPS_REACHED
    OOM = N.STATUS_OOM
    COMPLETED = N.STATUS_COMPLETED
    PARTIAL = N.STATUS_PARTIAL


@dataclass(frozen=True)
class LinearPenalties:
    """src/penalties.rs:27-35"""
    match_: int
    mismatch: int
    indel: int


@dataclass(frozen=True)
class AffinePenalties:
    """src/penalties.rs:38-48"""
    match_: int
    mismatch: int
    gap_opening: int
    gap_extension: int


@dataclass(frozen=True)
class Affine2pPenalties:
    """src/penalties.rs:52-66"""
    match_: int
    mismatch: int
    gap_opening1: int
    gap_extension1: int
    gap_opening2: int
    gap_extension2: int


class WavefrontPenalties:
    """WavefrontPenalties::new_* (src/penalties.rs:122-296). Invalid values
    raise ValueError where the reference panics."""

    def __init__(self, c_penalties, original):
        self._c = c_penalties
        self.original = original
        self.distance_metric = DistanceMetric(c_penalties.distance_metric)
        code = N.lib().wfa_penalties_check(ctypes.byref(c_penalties))
        if code == N.ERR_INVALID_PENALTIES:
            raise ValueError(f"invalid penalties {original}")
        N.check(code, "wfa_penalties_check")

    @staticmethod
    def _make(metric, match=0, x=0, o1=0, e1=0, o2=0, e2=0, original=None):
        return WavefrontPenalties(N.Penalties(metric, match, x, o1, e1, o2, e2), original)

    @classmethod
    def new_indel(cls):
        return cls._make(N.METRIC_INDEL)

    @classmethod
    def new_edit(cls):
        return cls._make(N.METRIC_EDIT)

    @classmethod
    def new_linear(cls, p: LinearPenalties):
        return cls._make(N.METRIC_GAP_LINEAR, p.match_, p.mismatch, p.indel, original=p)

    @classmethod
    def new_affine(cls, p: AffinePenalties):
        return cls._make(N.METRIC_GAP_AFFINE, p.match_, p.mismatch, p.gap_opening,
                         p.gap_extension, original=p)

    @classmethod
    def new_affine2p(cls, p: Affine2pPenalties):
        return cls._make(N.METRIC_GAP_AFFINE_2P, p.match_, p.mismatch, p.gap_opening1,
                         p.gap_extension1, p.gap_opening2, p.gap_extension2, original=p)


class HeuristicStrategy:
    """HeuristicStrategy (src/heuristic.rs); (kind, p0, p1, p2) for the ABI."""

    def __init__(self, kind=N.HEURISTIC_NONE, p0=0, p1=0, p2=0):
        self.kind, self.params = kind, (p0, p1, p2)

    @classmethod
    def none(cls):
        return cls()

    @classmethod
    def wf_adaptive(cls, min_wavefront_length, max_distance_threshold, steps_between_cutoffs):
        return cls(N.HEURISTIC_WFADAPTIVE, min_wavefront_length, max_distance_threshold,
                   steps_between_cutoffs)

    @classmethod
    def xdrop(cls, xdrop, steps_between_cutoffs):
        return cls(N.HEURISTIC_XDROP, xdrop, steps_between_cutoffs)

    @classmethod
    def zdrop(cls, zdrop, steps_between_cutoffs):
        return cls(N.HEURISTIC_ZDROP, zdrop, steps_between_cutoffs)

    @classmethod
    def banded_static(cls, min_k, max_k):
        return cls(N.HEURISTIC_BANDED_STATIC, min_k, max_k)

    @classmethod
    def banded_adaptive(cls, min_k, max_k, steps_between_cutoffs):
        return cls(N.HEURISTIC_BANDED_ADAPTIVE, min_k, max_k, steps_between_cutoffs)


class Cigar:
    """Cigar (src/cigar.rs:16-29): the operations M/X/I/D, one byte each."""

    def __init__(self, operations=b"", score=-(2 ** 31)):
        self.operations = bytes(operations)
        self.score = score

    def operations_slice(self):
        return self.operations

    def is_null(self):
        return len(self.operations) == 0

    def to_string_rle(self, print_matches=True):
        """src/cigar.rs:522-546"""
        ops = self.operations
        if not ops:
            return ""
        out, last, n = [], ops[0], 1
        for op in ops[1:]:
            if op == last:
                n += 1
            else:
                if print_matches or last != ord("M"):
                    out.append(f"{n}{chr(last)}")
                last, n = op, 1
        if print_matches or last != ord("M"):
            out.append(f"{n}{chr(last)}")
        return "".join(out)

    def score_edit(self):
        """src/cigar.rs:118-127"""
        return sum(0 if op == ord("M") else 1 for op in self.operations)

    def score_gap_linear(self, p: LinearPenalties):
        """src/cigar.rs:130-137"""
        s = 0
        for op in self.operations:
            s -= p.match_ if op == ord("M") else p.mismatch if op == ord("X") else p.indel
        return s

    def score_gap_affine(self, p: AffinePenalties):
        """src/cigar.rs:139-168"""
        s, last = 0, 0
        for op in self.operations:
            if op == ord("M"):
                s -= p.match_
            elif op == ord("X"):
                s -= p.mismatch
            else:
                s -= p.gap_extension + (0 if last == op else p.gap_opening)
            last = op
        return s

    def score_gap_affine2p(self, p: Affine2pPenalties):
        """src/cigar.rs:171-202"""
        def op_score(op, n):
            if op == ord("M"):
                return p.match_ * n
            if op == ord("X"):
                return p.mismatch * n
            return min(p.gap_opening1 + p.gap_extension1 * n, p.gap_opening2 + p.gap_extension2 * n)
        s, last, n = 0, 0, 0
        for op in self.operations:
            if op != last and last != 0:
                s -= op_score(last, n)
                n = 0
            last = op
            n += 1
        if last:
            s -= op_score(last, n)
        return s

    def check_alignment(self, pattern: bytes, text: bytes):
        """src/cigar.rs:446-518: None if valid, else a message."""
        pp = tp = 0
        for i, op in enumerate(self.operations):
            if op in (ord("M"), ord("X")):
                if pp >= len(pattern) or tp >= len(text):
                    return f"op {i} out of bounds"
                if (pattern[pp] == text[tp]) != (op == ord("M")):
                    return f"op {i} {chr(op)} does not match the sequences"
                pp += 1
                tp += 1
            elif op == ord("I"):
                tp += 1
            elif op == ord("D"):
                pp += 1
            else:
                return f"unknown op {op}"
        if pp != len(pattern) or tp != len(text):
            return "alignment does not span both sequences"
        return None


def compute_output_score(cigar, wfa_score, penalties, score_only, plen, tlen):
    """The benchmark driver's printed score (src/bin/align_benchmark.rs:466-504)."""
    m = penalties.distance_metric
    if not score_only:
        if cigar.is_null():
            return -1
        if m in (DistanceMetric.INDEL, DistanceMetric.EDIT):
            return cigar.score_edit()
        if m == DistanceMetric.GAP_LINEAR:
            return cigar.score_gap_linear(penalties.original)
        if m == DistanceMetric.GAP_AFFINE:
            return cigar.score_gap_affine(penalties.original)
        return cigar.score_gap_affine2p(penalties.original)
    if m in (DistanceMetric.INDEL, DistanceMetric.EDIT):
        return wfa_score
    match = penalties.original.match_
    if match < 0:
        # Eizenga: (swg_match * (plen + tlen) - wf_score) / 2 with C's int
        # wrapping and truncation toward zero (src/penalties.rs:300-308)
        v = ((-match) * (plen + tlen) - wfa_score)
        v = (v + 2 ** 31) % 2 ** 32 - 2 ** 31
        return int(v / 2)
    return -wfa_score


class WavefrontAligner:
    """WavefrontAligner<N> (src/aligner.rs:146-211) on the GPU."""

    def __init__(self, penalties: WavefrontPenalties, device: int = 0):
        self.penalties = penalties
        attr = N.Attr()
        N.lib().wfa_attr_default(ctypes.byref(attr), int(penalties.distance_metric))
        attr.penalties = penalties._c
        err = ctypes.c_int32(0)
        h = N.lib().wfa_aligner_new(ctypes.byref(attr), device, ctypes.byref(err))
        if not h:
            N.check(err.value, "wfa_aligner_new")
        self._h = h
        self._cigar = Cigar()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().wfa_aligner_free(h)
            self._h = None

    # ---- setters (src/aligner.rs:462-611) ----
    def set_alignment_scope(self, scope: AlignmentScope):
        N.check(N.lib().wfa_aligner_set_alignment_scope(self._h, int(scope)))

    @property
    def alignment_scope(self):
        raise AttributeError("write-only, as used by the reference's tests")

    @alignment_scope.setter
    def alignment_scope(self, scope):
        self.set_alignment_scope(scope)

    def set_memory_mode(self, mode: MemoryMode):
        N.check(N.lib().wfa_aligner_set_memory_mode(self._h, int(mode)))

    def set_compat(self, compat: int):
        """Compat::C (COMPAT_C, the default) or Compat::Corrected
        (COMPAT_CORRECTED; src/aligner.rs:606-608)."""
        N.check(N.lib().wfa_aligner_set_compat(self._h, int(compat)))

    def set_heuristic(self, strategy: HeuristicStrategy):
        N.check(N.lib().wfa_aligner_set_heuristic(self._h, strategy.kind, *strategy.params))

    def set_alignment_free_ends(self, pattern_begin_free, pattern_end_free, text_begin_free,
                                text_end_free):
        N.check(N.lib().wfa_aligner_set_alignment_free_ends(
            self._h, pattern_begin_free, pattern_end_free, text_begin_free, text_end_free))

    def set_alignment_extension(self):
        N.check(N.lib().wfa_aligner_set_alignment_extension(self._h))

    def set_alignment_end_to_end(self):
        N.check(N.lib().wfa_aligner_set_alignment_end_to_end(self._h))

    def set_max_alignment_steps(self, max_steps: int):
        N.check(N.lib().wfa_aligner_set_max_alignment_steps(self._h, max_steps))

    def set_max_memory_compact(self, nbytes: int):
        """src/aligner.rs:469: the reference compacts its medium/low-memory
        backtrace buffer above this size, which changes no result. Stored
        only: the device plans each pair's CIGAR history pool itself."""
        N.check(N.lib().wfa_aligner_set_max_memory_compact(self._h, nbytes))

    @property
    def max_memory_compact(self) -> int:
        return int(N.lib().wfa_aligner_max_memory_compact(self._h))

    # ---- align (src/aligner.rs:620, 2755) ----
    def _finish(self, score):
        L = N.lib()
        if score == -(2 ** 31):
            N.check(L.wfa_aligner_error(self._h), "align")
        ptr = ctypes.POINTER(ctypes.c_uint8)()
        n = L.wfa_aligner_cigar(self._h, ctypes.byref(ptr))
        ops = ctypes.string_at(ptr, n) if n > 0 else b""
        self._cigar = Cigar(ops, L.wfa_aligner_cigar_score(self._h))
        return score

    def align_end2end(self, pattern: bytes, text: bytes) -> int:
        pattern, text = bytes(pattern), bytes(text)
        return self._finish(N.lib().wfa_align_end2end(self._h, pattern, len(pattern), text,
                                                      len(text)))

    def align_endsfree(self, pattern: bytes, text: bytes) -> int:
        """align_endsfree (src/aligner.rs:707): the configured free ends or
        extension."""
        pattern, text = bytes(pattern), bytes(text)
        return self._finish(N.lib().wfa_align_endsfree(self._h, pattern, len(pattern), text,
                                                       len(text)))

    def align_biwfa(self, pattern: bytes, text: bytes) -> int:
        """align_biwfa (src/aligner.rs:2563): BiWFA with its CIGAR."""
        pattern, text = bytes(pattern), bytes(text)
        return self._finish(N.lib().wfa_align_biwfa(self._h, pattern, len(pattern), text,
                                                    len(text)))

    def align_biwfa_score(self, pattern: bytes, text: bytes) -> int:
        """align_biwfa_score (src/aligner.rs:2650)."""
        pattern, text = bytes(pattern), bytes(text)
        return self._finish(N.lib().wfa_align_biwfa_score(self._h, pattern, len(pattern), text,
                                                          len(text)))

    def align_packed2bits(self, pattern: bytes, pattern_length: int, text: bytes,
                          text_length: int) -> int:
        return self._finish(N.lib().wfa_align_packed2bits(
            self._h, bytes(pattern), pattern_length, bytes(text), text_length))

    def align_lambda_classes(self, pattern: bytes, text: bytes, pattern_class=None,
                             text_class=None) -> int:
        """align_lambda (src/aligner.rs:2743-2752) for match functions that
        compare symbols: pattern[i] matches text[j] iff
        pattern_class[pattern[i]] == text_class[text[j]] (256-entry tables;
        identity when omitted, i.e. byte equality).

        A device cannot call a host closure: closures over positions and
        symbol relations that are not equivalences have no counterpart."""
        ident = bytes(range(256))
        pc = bytes(pattern_class) if pattern_class is not None else ident
        tc = bytes(text_class) if text_class is not None else ident
        if len(pc) != 256 or len(tc) != 256:
            raise ValueError("class tables have 256 entries")
        pattern, text = bytes(pattern), bytes(text)
        return self._finish(N.lib().wfa_align_lambda_classes(self._h, pattern, len(pattern),
                                                             text, len(text), pc, tc))

    def align_lambda(self, match_funct, pattern_length: int, text_length: int) -> int:
        """align_lambda with a Python closure over positions: a device cannot
        call it. Use align_lambda_classes for symbol comparisons."""
        raise NotImplementedError(
            "align_lambda with a closure needs a host call per compared position; the device "
            "path takes symbol classes instead (align_lambda_classes)")

    def status(self) -> AlignStatus:
        return AlignStatus(N.lib().wfa_aligner_status(self._h))

    def cigar(self) -> Cigar:
        return self._cigar

    @property
    def alignment_end_pos(self) -> "WavefrontPos":
        """alignment_end_pos (src/aligner.rs:146-204): the end cell, or a
        Z-drop's best cell and the score it was found at; unset
        (-1, INT32_MAX, OFFSET_NULL) otherwise."""
        out = (ctypes.c_int32 * 3)()
        N.lib().wfa_aligner_end_position(self._h, out)
        return WavefrontPos(out[0], out[1], out[2])


@dataclass(frozen=True)
class WavefrontPos:
    """WavefrontPos (src/wavefront.rs:24-29)."""
    score: int
    k: int
    offset: int


EditAligner = WavefrontAligner
AffineAligner = WavefrontAligner
Affine2pAligner = WavefrontAligner


def encode_packed2bits(ascii_seq: bytes) -> bytes:
    """src/sequences.rs:55-74 (A0 C1 G2 T3, LSB first; others map to A)."""
    code = {ord("A"): 0, ord("a"): 0, ord("C"): 1, ord("c"): 1, ord("G"): 2, ord("g"): 2,
            ord("T"): 3, ord("t"): 3}
    out = bytearray((len(ascii_seq) + 3) // 4)
    for i, ch in enumerate(ascii_seq):
        out[i >> 2] |= code.get(ch, 0) << (2 * (i & 3))
    return bytes(out)

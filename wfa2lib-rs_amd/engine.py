"""Batch engine: many independent pairs in one persistent device launch.

Device memory comes from torch (plumbing only): the batch's arrays are
torch tensors on the GPU, and their data pointers go to
``wfa_engine_align_batch`` through the C ABI together with the current
HIP stream.
"""
import ctypes

import numpy as np

from . import _native as N
from .aligner import WavefrontPenalties


def make_attr(penalties: WavefrontPenalties, scope=N.SCOPE_SCORE, memory_mode=N.MEMORY_HIGH,
              max_steps=2 ** 31 - 1):
    attr = N.Attr()
    N.lib().wfa_attr_default(ctypes.byref(attr), int(penalties.distance_metric))
    attr.penalties = penalties._c
    attr.alignment_scope = int(scope)
    attr.memory_mode = int(memory_mode)
    attr.max_alignment_steps = int(max_steps)
    return attr


def pair_config_row(attr):
    """The wfa_pair_config_t row (12 int32) of an aligner configuration:
    its span, extension, free ends and heuristic."""
    return [attr.span, attr.extension, attr.pattern_begin_free, attr.pattern_end_free,
            attr.text_begin_free, attr.text_end_free, attr.heuristic, *attr.heuristic_p, 0, 0]


class DeviceBatch:
    """A batch resident in device memory (HBM)."""

    def __init__(self, pairs, device="cuda", cigar=False, pair_config=None, end_position=False):
        """`pair_config`: optional (n, 12) int32 rows of wfa_pair_config_t
        (span, extension, four free ends, heuristic, three parameters, two
        zeros), see ``pair_config_row``. `end_position`: also return each
        pair's alignment_end_pos (score, k, offset), ``end_positions()``."""
        import torch
        self.n = len(pairs)
        self.seq_format = pairs.seq_format
        self.max_length = pairs.max_length()
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        self.seq = t(pairs.seq)
        self.poff = t(pairs.poff)
        self.plen = t(pairs.plen)
        self.toff = t(pairs.toff)
        self.tlen = t(pairs.tlen)
        self.score = torch.zeros(max(self.n, 1), dtype=torch.int32, device=device)
        self.status = torch.zeros(max(self.n, 1), dtype=torch.int32, device=device)
        self.cigar = cigar
        self.end_pos = (torch.zeros(3 * max(self.n, 1), dtype=torch.int32, device=device)
                        if end_position else None)
        self.pair_config = None
        if pair_config is not None:
            pc = np.ascontiguousarray(np.asarray(pair_config, dtype=np.int32).reshape(self.n, 12))
            self.pair_config = t(pc)
        if cigar:
            cap = pairs.plen.astype(np.int64) + pairs.tlen.astype(np.int64)
            off = np.concatenate([[0], np.cumsum(cap)[:-1]]).astype(np.int64) if self.n else \
                np.zeros(1, np.int64)
            self.cig_off = t(off)
            self.cig = torch.zeros(max(int(cap.sum()), 1), dtype=torch.uint8, device=device)
            self.cig_len = torch.zeros(max(self.n, 1), dtype=torch.int32, device=device)

    def c_batch(self):
        b = N.Batch()
        b.num_pairs = self.n
        b.seq_format = self.seq_format
        b.max_length = self.max_length
        b.seq_data = self.seq.data_ptr()
        b.pattern_offset = self.poff.data_ptr()
        b.pattern_length = self.plen.data_ptr()
        b.text_offset = self.toff.data_ptr()
        b.text_length = self.tlen.data_ptr()
        b.score = self.score.data_ptr()
        b.status = self.status.data_ptr()
        if self.cigar:
            b.cigar_ops = self.cig.data_ptr()
            b.cigar_offset = self.cig_off.data_ptr()
            b.cigar_length = self.cig_len.data_ptr()
        if self.pair_config is not None:
            b.pair_config = self.pair_config.data_ptr()
        if self.end_pos is not None:
            b.end_position = self.end_pos.data_ptr()
        return b

    def end_positions(self):
        """(n, 3) int32 host array of alignment_end_pos {score, k, offset}."""
        return self.end_pos.cpu().numpy()[:3 * self.n].reshape(self.n, 3)

    def results(self):
        """(scores, status, list of CIGAR bytes or None) on the host."""
        score = self.score.cpu().numpy()[:self.n]
        status = self.status.cpu().numpy()[:self.n]
        cigars = None
        if self.cigar:
            ops = self.cig.cpu().numpy()
            off = self.cig_off.cpu().numpy()
            ln = self.cig_len.cpu().numpy()
            cigars = [bytes(ops[off[i]:off[i] + ln[i]]) for i in range(self.n)]
        return score, status, cigars


class BatchEngine:
    def __init__(self, device=0):
        err = ctypes.c_int32(0)
        self._h = N.lib().wfa_engine_new(device, ctypes.byref(err))
        if not self._h:
            N.check(err.value, "wfa_engine_new")
        self.device = device

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().wfa_engine_free(h)
            self._h = None

    def align(self, attr, batch: DeviceBatch, stream=None, per_pair=False):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        b = batch.c_batch()
        fn = N.lib().wfa_engine_align_batch_per_pair if per_pair else N.lib().wfa_engine_align_batch
        N.check(fn(self._h, ctypes.byref(attr), ctypes.byref(b), ctypes.c_void_p(stream)),
                "wfa_engine_align_batch")

    def autotune(self, attr, batch: DeviceBatch, sample_pairs=0, stream=None):
        """Time the candidate launch shapes of a long-read batch on a
        sample and keep the fastest for its key (wfa_engine_autotune).
        Returns {"waves_per_pair", "groups_per_cu", "default", "trials",
        "sample_pairs"}; waves_per_pair 0 when there is nothing to tune."""
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        rep = N.TuneReport()
        b = batch.c_batch()
        N.check(N.lib().wfa_engine_autotune(self._h, ctypes.byref(attr), ctypes.byref(b),
                                            ctypes.c_void_p(stream), int(sample_pairs),
                                            ctypes.byref(rep)), "wfa_engine_autotune")
        return {
            "waves_per_pair": rep.waves_per_pair, "groups_per_cu": rep.groups_per_cu,
            "default": (rep.default_waves_per_pair, rep.default_groups_per_cu),
            "sample_pairs": rep.sample_pairs,
            "trials": [(t.waves_per_pair, t.groups_per_cu, round(t.ms, 3))
                       for t in rep.trials[:rep.num_trials]],
        }

    def set_launch_shape(self, attr, max_length, waves_per_pair, groups_per_cu):
        """Fix the long-read launch shape for attr's key and max_length's
        bucket (0, 0: back to the fixed plan)."""
        N.check(N.lib().wfa_engine_set_launch_shape(self._h, ctypes.byref(attr), int(max_length),
                                                    int(waves_per_pair), int(groups_per_cu)),
                "wfa_engine_set_launch_shape")

    def launch_shape(self, attr, max_length):
        w, g = ctypes.c_int32(0), ctypes.c_int32(0)
        N.check(N.lib().wfa_engine_get_launch_shape(self._h, ctypes.byref(attr), int(max_length),
                                                    ctypes.byref(w), ctypes.byref(g)),
                "wfa_engine_get_launch_shape")
        return w.value, g.value

    def set_scratch_budget(self, nbytes):
        """Cap one launch's device scratch (pairs that do not fit: OOM)."""
        N.check(N.lib().wfa_engine_set_scratch_budget(self._h, int(nbytes)),
                "wfa_engine_set_scratch_budget")

    def last_retries(self):
        return N.lib().wfa_engine_last_retries(self._h)

    def last_kernel_ms(self):
        """(first persistent launch, all launches) of the last batch, HIP events."""
        a, b = ctypes.c_float(0), ctypes.c_float(0)
        N.check(N.lib().wfa_engine_last_kernel_ms(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

"""ctypes view of the engine's C ABI (include/wfa_mi355x.h).

The shared library is built in-tree (``lib/libwfa_mi355x.so``) by
``__graft_entry__.build()``. Loading fails loudly when it is missing: there is
no CPU fallback anywhere in this package.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
# WFA_MI355X_LIB names another engine build in lib/ (libwfa_mi355x_checks.so:
# the debug build with device index checks, build.py --checks)
LIB_PATH = os.path.join(LIB_DIR, os.path.basename(os.environ.get("WFA_MI355X_LIB",
                                                                 "libwfa_mi355x.so")))
SYNTH_PATH = os.path.join(LIB_DIR, "libwfa_synth.so")

# ---- constants (same values as include/wfa_mi355x.h) ----
METRIC_INDEL, METRIC_EDIT, METRIC_GAP_LINEAR, METRIC_GAP_AFFINE, METRIC_GAP_AFFINE_2P = range(5)
SCOPE_SCORE, SCOPE_ALIGNMENT = 0, 1
MEMORY_HIGH, MEMORY_MED, MEMORY_LOW, MEMORY_ULTRALOW = range(4)
SPAN_END2END, SPAN_ENDSFREE = 0, 1
COMPAT_C, COMPAT_CORRECTED = 0, 1
# the offset of an empty cell (WFA_NULL: INT32_MIN / 2)
OFFSET_NULL = -(2 ** 30)
(HEURISTIC_NONE, HEURISTIC_WFADAPTIVE, HEURISTIC_XDROP, HEURISTIC_ZDROP,
 HEURISTIC_BANDED_STATIC, HEURISTIC_BANDED_ADAPTIVE) = range(6)
(STATUS_OK, STATUS_END_REACHED, STATUS_END_UNREACHABLE, STATUS_MAX_STEPS_REACHED,
 STATUS_OOM, STATUS_COMPLETED, STATUS_PARTIAL) = range(7)
SEQ_ASCII, SEQ_PACKED2BITS = 0, 1
ERRORS = {
    0: "ok", -1: "invalid penalties", -2: "invalid argument", -3: "unsupported on the device path",
    -4: "HIP runtime error", -5: "no HIP device", -6: "out of device memory",
    -7: "rejected combination", -8: "beyond a fixed limit of the no-allocation entry point",
}
ERR_INVALID_PENALTIES = -1
ERR_INVALID_ARGUMENT = -2
ERR_UNSUPPORTED = -3
ERR_LIMIT = -8
# limits of wfa_align_batch_noalloc (include/wfa_mi355x.h)
NOALLOC_SCRATCH_BYTES = 256 << 20
NOALLOC_MAX_PAIRS = 1 << 20
NOALLOC_MAX_LENGTH = 16384


class Penalties(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "distance_metric", "match", "mismatch", "gap_opening1", "gap_extension1",
        "gap_opening2", "gap_extension2")]


class Attr(ctypes.Structure):
    _fields_ = [("penalties", Penalties)] + [(n, ctypes.c_int32) for n in (
        "alignment_scope", "memory_mode", "span", "extension", "pattern_begin_free",
        "pattern_end_free", "text_begin_free", "text_end_free", "heuristic")] + [
        ("heuristic_p", ctypes.c_int32 * 3), ("max_alignment_steps", ctypes.c_int32),
        ("compat", ctypes.c_int32)]


class Batch(ctypes.Structure):
    _fields_ = [
        ("num_pairs", ctypes.c_int32), ("seq_format", ctypes.c_int32),
        ("max_length", ctypes.c_int32),
        ("seq_data", ctypes.c_void_p), ("pattern_offset", ctypes.c_void_p),
        ("pattern_length", ctypes.c_void_p), ("text_offset", ctypes.c_void_p),
        ("text_length", ctypes.c_void_p), ("score", ctypes.c_void_p),
        ("status", ctypes.c_void_p), ("cigar_ops", ctypes.c_void_p),
        ("cigar_offset", ctypes.c_void_p), ("cigar_length", ctypes.c_void_p),
        ("pair_config", ctypes.c_void_p), ("end_position", ctypes.c_void_p),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("num_engines", ctypes.c_int32),
                ("max_batch_pairs", ctypes.c_int32), ("memory_budget", ctypes.c_int64)]


class ServerRequest(ctypes.Structure):
    _fields_ = [("attr", Attr), ("seq_format", ctypes.c_int32),
                ("pattern_length", ctypes.c_int32), ("text_length", ctypes.c_int32),
                ("pattern", ctypes.c_char_p), ("text", ctypes.c_char_p),
                ("tag", ctypes.c_uint64)]


class ServerResult(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("error", ctypes.c_int32), ("score", ctypes.c_int32),
                ("status", ctypes.c_int32), ("cigar_length", ctypes.c_int32),
                ("cigar_ops", ctypes.POINTER(ctypes.c_uint8)), ("latency_us", ctypes.c_double)]


class SynthCfg(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("len_mode", ctypes.c_int32),
                ("len_lo", ctypes.c_int32), ("len_hi", ctypes.c_int32),
                ("div_lo", ctypes.c_double), ("div_hi", ctypes.c_double),
                ("indel_mean", ctypes.c_double), ("run_frac", ctypes.c_double),
                ("run_mean", ctypes.c_double)]


TUNE_MAX_TRIALS = 16


class TuneTrial(ctypes.Structure):
    _fields_ = [("waves_per_pair", ctypes.c_int32), ("groups_per_cu", ctypes.c_int32),
                ("ms", ctypes.c_float)]


class TuneReport(ctypes.Structure):
    _fields_ = [("sample_pairs", ctypes.c_int32), ("num_trials", ctypes.c_int32),
                ("trials", TuneTrial * TUNE_MAX_TRIALS), ("waves_per_pair", ctypes.c_int32),
                ("groups_per_cu", ctypes.c_int32), ("default_waves_per_pair", ctypes.c_int32),
                ("default_groups_per_cu", ctypes.c_int32)]


# Every entry point the header declares, with its ctypes signature.
_P, _I32, _VP = ctypes.POINTER, ctypes.c_int32, ctypes.c_void_p
SIGNATURES = {
    "wfa_attr_default": (None, [_P(Attr), _I32]),
    "wfa_penalties_check": (_I32, [_P(Penalties)]),
    "wfa_aligner_new": (_VP, [_P(Attr), _I32, _P(_I32)]),
    "wfa_aligner_free": (None, [_VP]),
    "wfa_aligner_set_alignment_scope": (_I32, [_VP, _I32]),
    "wfa_aligner_set_memory_mode": (_I32, [_VP, _I32]),
    "wfa_aligner_set_compat": (_I32, [_VP, _I32]),
    "wfa_aligner_set_heuristic": (_I32, [_VP, _I32, _I32, _I32, _I32]),
    "wfa_aligner_set_alignment_free_ends": (_I32, [_VP, _I32, _I32, _I32, _I32]),
    "wfa_aligner_set_alignment_extension": (_I32, [_VP]),
    "wfa_aligner_set_alignment_end_to_end": (_I32, [_VP]),
    "wfa_aligner_set_max_alignment_steps": (_I32, [_VP, _I32]),
    "wfa_aligner_set_max_memory_compact": (_I32, [_VP, ctypes.c_uint64]),
    "wfa_aligner_max_memory_compact": (ctypes.c_uint64, [_VP]),
    "wfa_align_end2end": (_I32, [_VP, ctypes.c_char_p, _I32, ctypes.c_char_p, _I32]),
    "wfa_align_packed2bits": (_I32, [_VP, ctypes.c_char_p, _I32, ctypes.c_char_p, _I32]),
    "wfa_align_lambda_classes": (_I32, [_VP, ctypes.c_char_p, _I32, ctypes.c_char_p, _I32,
                                        ctypes.c_char_p, ctypes.c_char_p]),
    "wfa_align_endsfree": (_I32, [_VP, ctypes.c_char_p, _I32, ctypes.c_char_p, _I32]),
    "wfa_align_biwfa": (_I32, [_VP, ctypes.c_char_p, _I32, ctypes.c_char_p, _I32]),
    "wfa_align_biwfa_score": (_I32, [_VP, ctypes.c_char_p, _I32, ctypes.c_char_p, _I32]),
    "wfa_aligner_status": (_I32, [_VP]),
    "wfa_aligner_cigar": (_I32, [_VP, _P(_P(ctypes.c_uint8))]),
    "wfa_aligner_cigar_score": (_I32, [_VP]),
    "wfa_aligner_end_position": (None, [_VP, _VP]),
    "wfa_aligner_error": (_I32, [_VP]),
    "wfa_engine_new": (_VP, [_I32, _P(_I32)]),
    "wfa_engine_free": (None, [_VP]),
    "wfa_engine_align_batch": (_I32, [_VP, _P(Attr), _P(Batch), _VP]),
    "wfa_engine_align_batch_per_pair": (_I32, [_VP, _P(Attr), _P(Batch), _VP]),
    "wfa_engine_set_scratch_budget": (_I32, [_VP, ctypes.c_int64]),
    "wfa_engine_last_phase_cycles": (_I32, [_VP, _P(ctypes.c_uint64)]),
    "wfa_engine_autotune": (_I32, [_VP, _P(Attr), _P(Batch), _VP, _I32, _P(TuneReport)]),
    "wfa_engine_set_launch_shape": (_I32, [_VP, _P(Attr), _I32, _I32, _I32]),
    "wfa_engine_get_launch_shape": (_I32, [_VP, _P(Attr), _I32, _P(_I32), _P(_I32)]),
    "wfa_align_batch_noalloc": (_I32, [_P(Attr), _P(Batch), _VP]),
    "wfa_engine_last_retries": (_I32, [_VP]),
    "wfa_engine_last_kernel_ms": (_I32, [_VP, _P(ctypes.c_float), _P(ctypes.c_float)]),
    "wfa_server_config_default": (None, [_P(ServerConfig)]),
    "wfa_server_new": (_VP, [_P(ServerConfig), _P(_I32)]),
    "wfa_server_submit": (_I32, [_VP, _P(ServerRequest)]),
    "wfa_server_poll": (_I32, [_VP, _P(ServerResult), _I32, _I32]),
    "wfa_server_stats": (_I32, [_VP] + [_P(ctypes.c_int64)] * 4),
    "wfa_server_free": (None, [_VP]),
    "wfa_version": (ctypes.c_char_p, []),
}

_lib = None
_synth = None


def lib():
    """The engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(the engine has no CPU fallback)")
        # PyTorch-ROCm ships its own HIP runtime beside the system one this
        # library links; in one process torch's must start first, or torch
        # finds no device once this library has initialised the GPU. Device
        # batches are torch tensors, so start it here when it is installed.
        try:
            import torch
            torch.cuda.is_available()
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def synth_lib():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise RuntimeError(f"{SYNTH_PATH} is missing: run __graft_entry__.build()")
        s = ctypes.CDLL(SYNTH_PATH)
        s.wfa_synth_lengths.argtypes = [_P(SynthCfg), ctypes.c_int64, _VP, _VP, _I32]
        s.wfa_synth_lengths.restype = None
        s.wfa_synth_fill.argtypes = [_P(SynthCfg), ctypes.c_int64, _VP, _VP, _I32, _VP, _VP,
                                     _VP, _I32]
        s.wfa_synth_fill.restype = None
        _synth = s
    return _synth


def check(code, what="call"):
    if code != 0:
        raise RuntimeError(f"{what} failed: {ERRORS.get(code, code)} ({code})")

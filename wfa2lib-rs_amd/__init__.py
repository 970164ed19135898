"""wfa2lib-rs_amd: MI355X-native pairwise wavefront alignment.

A drop-in for the aligner of HPCBio/wfa2lib-rs on its hot path: the
reference's ``WavefrontAligner`` API (``aligner``) and a batch engine
(``engine``) over hand-written gfx950 HIP kernels exposed through a C ABI
(``include/wfa_mi355x.h``, ``lib/libwfa_mi355x.so``).
"""
from . import _native
from .aligner import (Affine2pAligner, Affine2pPenalties, AffineAligner, AffinePenalties,
                      AlignmentScope, AlignmentSpan, AlignStatus, Cigar, DistanceMetric,
                      EditAligner, HeuristicStrategy, LinearPenalties, MemoryMode,
                      WavefrontAligner, WavefrontPenalties, WavefrontPos,
                      compute_output_score, encode_packed2bits)

__all__ = [
    "Affine2pAligner", "Affine2pPenalties", "AffineAligner", "AffinePenalties", "AlignmentScope",
    "AlignmentSpan", "AlignStatus", "Cigar", "DistanceMetric", "EditAligner", "HeuristicStrategy",
    "LinearPenalties", "MemoryMode", "WavefrontAligner", "WavefrontPenalties", "WavefrontPos",
    "compute_output_score", "encode_packed2bits",
]

"""align_benchmark on the GPU: the reference's benchmark driver
(src/bin/align_benchmark.rs:20-706) with the same options, input format and
output lines, aligning the input's pairs in batches on the device instead of
one aligner call per pair.

    python wfa2lib-rs_amd/align_benchmark.py -i pairs.seq [-o out.alg]
        [-a gap-affine-wfa|gap-affine2p-wfa|gap-linear-wfa|edit-wfa|indel-wfa]
        [-p M,X,I] [-g M,X,O,E] [--affine2p-penalties M,X,O1,E1,O2,E2]
        [--wfa-score-only] [--wfa-memory high|med|low|ultralow]
        [--wfa-span global|extension|ends-free,P0,Pf,T0,Tf]
        [--wfa-heuristic NAME --wfa-heuristic-parameters P,...]
        [--wfa-max-steps N] [-c] [--batch-size N]

Input: lines alternating `>pattern` and `<text` (prefixes optional; an
unpaired last line is ignored). Output: one `score<TAB>CIGAR` line per pair
(`-` for no CIGAR), exactly as the reference prints them; `Time.Alignment`
(device time of the batches) and, with -c, `Alignments.Correct` on stderr.
Configurations the device path does not run (Z-drop with a medium/low-memory
CIGAR in C-compatible mode) exit with an error instead of falling back to a
CPU; --no-c-compat selects Compat::Corrected.
"""
import argparse
import ctypes
import os
import sys
import time

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import __graft_entry__ as _entry  # noqa: E402
    _entry.load_package()
    from wfa2lib_rs_amd import _native as N  # noqa: E402
    from wfa2lib_rs_amd.aligner import (Affine2pPenalties, AffinePenalties, Cigar,  # noqa: E402
                                        DistanceMetric, LinearPenalties,
                                        WavefrontPenalties, compute_output_score)
else:
    from . import _native as N
    from .aligner import (Affine2pPenalties, AffinePenalties, Cigar, DistanceMetric,
                          LinearPenalties, WavefrontPenalties, compute_output_score)

INT32_MIN = -2 ** 31


def fail(msg):
    print(msg, file=sys.stderr, flush=True)
    sys.exit(1)


# options whose values may start with '-' (negative penalties or band limits)
VALUE_OPTS = ("-p", "--linear-penalties", "-g", "--affine-penalties", "--affine2p-penalties",
              "--wfa-heuristic-parameters", "--wfa-span")


def parse_args(argv):
    argv = list(argv)
    joined = []
    i = 0
    while i < len(argv):
        if argv[i] in VALUE_OPTS and i + 1 < len(argv):
            joined.append(f"{argv[i]}={argv[i + 1]}")
            i += 2
        else:
            joined.append(argv[i])
            i += 1
    argv = joined
    ap = argparse.ArgumentParser(prog="align_benchmark", add_help=True)
    ap.add_argument("-i", "--input", required=True)
    ap.add_argument("-o", "--output")
    ap.add_argument("-a", "--algorithm", default="gap-affine-wfa")
    ap.add_argument("--wfa-score-only", action="store_true")
    ap.add_argument("--wfa-span")
    ap.add_argument("--wfa-memory")
    ap.add_argument("-p", "--linear-penalties")
    ap.add_argument("-g", "--affine-penalties")
    ap.add_argument("--affine2p-penalties")
    ap.add_argument("--wfa-heuristic")
    ap.add_argument("--wfa-heuristic-parameters")
    ap.add_argument("--wfa-max-steps", type=int)
    ap.add_argument("-c", "--check", nargs="?", const="correct")
    ap.add_argument("-v", "--verbose", type=int)
    ap.add_argument("--no-c-compat", action="store_true")
    ap.add_argument("--batch-size", type=int, default=1 << 16,
                    help="pairs per device launch (this driver only)")
    ap.add_argument("--device", type=int, default=0)
    return ap.parse_args(argv)


def read_pairs(path):
    """PairReader (src/bin/align_benchmark.rs:86-140): (pattern, text) bytes."""
    try:
        f = open(path, "rb")
    except OSError as e:
        fail(f"Error reading input file {path!r}: {e}")
    with f:
        while True:
            p = f.readline()
            if not p:
                return
            t = f.readline()
            if not t:
                return
            p = p.rstrip(b"\n").rstrip(b"\r")
            t = t.rstrip(b"\n").rstrip(b"\r")
            if p.startswith(b">"):
                p = p[1:]
            if t.startswith(b"<"):
                t = t[1:]
            yield p, t


def build_penalties(a):
    """build_penalties (src/bin/align_benchmark.rs:148-209)."""
    vals = lambda s: [int(v.strip()) for v in s.split(",")]
    try:
        if a.algorithm == "indel-wfa":
            return WavefrontPenalties.new_indel()
        if a.algorithm == "edit-wfa":
            return WavefrontPenalties.new_edit()
        if a.algorithm == "gap-linear-wfa":
            v = vals(a.linear_penalties) if a.linear_penalties else [0, 4, 2]
            return WavefrontPenalties.new_linear(LinearPenalties(*v[:3]))
        if a.algorithm == "gap-affine-wfa":
            v = vals(a.affine_penalties) if a.affine_penalties else [0, 4, 6, 2]
            return WavefrontPenalties.new_affine(AffinePenalties(*v[:4]))
        if a.algorithm == "gap-affine2p-wfa":
            v = vals(a.affine2p_penalties) if a.affine2p_penalties else [0, 4, 6, 2, 24, 1]
            return WavefrontPenalties.new_affine2p(Affine2pPenalties(*v[:6]))
    except ValueError as e:
        fail(str(e))
    fail(f"Unsupported algorithm: {a.algorithm}")


def parse_span(arg):
    """parse_span (src/bin/align_benchmark.rs:225-253): ("global",),
    ("extension",) or ("ends-free", [p0, pf, t0, tf])."""
    if arg is None or arg == "global":
        return ("global",)
    if arg in ("extension", "extend"):
        return ("extension",)
    if arg.startswith("ends-free") or arg.startswith("endsfree"):
        try:
            v = [float(x.strip()) for x in arg.split(",")[1:]]
        except ValueError:
            fail("Invalid --wfa-span")
        if len(v) != 4:
            fail("Invalid --wfa-span")
        return ("ends-free", v)
    fail("Option '--wfa-span' must be in {'global','extension','ends-free'}")


def nominal_prop(base, factor):
    """C's nominal_prop_u32 (src/bin/align_benchmark.rs:255-263)."""
    if 0.0 <= factor <= 1.0:
        return int(base * factor) & 0xFFFFFFFF
    return int(factor) & 0xFFFFFFFF


HEURISTICS = {"none": N.HEURISTIC_NONE, "wfa-adaptive": N.HEURISTIC_WFADAPTIVE,
              "xdrop": N.HEURISTIC_XDROP, "zdrop": N.HEURISTIC_ZDROP,
              "banded-static": N.HEURISTIC_BANDED_STATIC,
              "banded-adaptive": N.HEURISTIC_BANDED_ADAPTIVE}


def parse_heuristic(name, params):
    """parse_heuristic (src/bin/align_benchmark.rs:265-310): (kind, [p...])."""
    p = [int(x.strip()) for x in params.split(",")] if params else []
    if name not in HEURISTICS:
        fail(f"Unknown heuristic: {name}")
    kind = HEURISTICS[name]
    if kind in (N.HEURISTIC_XDROP, N.HEURISTIC_ZDROP) and len(p) < 2:
        fail("Invalid --wfa-heuristic-parameters")
    need = {N.HEURISTIC_NONE: 0, N.HEURISTIC_WFADAPTIVE: 3, N.HEURISTIC_XDROP: 2,
            N.HEURISTIC_ZDROP: 2, N.HEURISTIC_BANDED_STATIC: 2,
            N.HEURISTIC_BANDED_ADAPTIVE: 3}[kind]
    if len(p) < need:
        fail("Invalid --wfa-heuristic-parameters")
    return kind, (p + [0, 0, 0])[:3]


def main(argv=None, out=None):
    a = parse_args(sys.argv[1:] if argv is None else argv)
    import numpy as np
    import torch
    from wfa2lib_rs_amd.engine import BatchEngine, DeviceBatch
    from wfa2lib_rs_amd.synth import from_ascii

    penalties = build_penalties(a)
    score_only = a.wfa_score_only
    use_biwfa = a.wfa_memory == "ultralow"
    memory = {None: N.MEMORY_HIGH, "high": N.MEMORY_HIGH, "ultralow": N.MEMORY_ULTRALOW,
              "med": N.MEMORY_MED, "low": N.MEMORY_LOW}.get(a.wfa_memory, -1)
    if memory < 0:
        fail("Option '--wfa-memory' must be in {'high','med','low','ultralow'}")
    do_check = a.check is not None and not score_only
    # C's wavefront_align_presets__checks, in its order
    span = parse_span(a.wfa_span)
    if use_biwfa and span[0] != "global":
        fail("[WFA] BiWFA ends-free and extension are not supported")
    heur = parse_heuristic(a.wfa_heuristic, a.wfa_heuristic_parameters) \
        if a.wfa_heuristic else None
    is_drop = heur is not None and heur[0] in (N.HEURISTIC_XDROP, N.HEURISTIC_ZDROP)
    if use_biwfa and is_drop:
        fail("[WFA] BiWFA is not compatible with X-drop/Z-drop heuristics")
    if is_drop and penalties.distance_metric in (DistanceMetric.EDIT, DistanceMetric.INDEL):
        fail("[WFA] Heuristics drops are not compatible with 'edit'/'indel' distance metrics")

    attr = N.Attr()
    N.lib().wfa_attr_default(ctypes.byref(attr), int(penalties.distance_metric))
    attr.penalties = penalties._c
    attr.alignment_scope = N.SCOPE_SCORE if score_only else N.SCOPE_ALIGNMENT
    attr.memory_mode = memory
    if a.no_c_compat:
        attr.compat = N.COMPAT_CORRECTED
    if heur is not None:
        attr.heuristic = heur[0]
        for i, v in enumerate(heur[1]):
            attr.heuristic_p[i] = v
    if a.wfa_max_steps is not None:
        attr.max_alignment_steps = a.wfa_max_steps
    if span[0] == "extension":
        attr.span, attr.extension = N.SPAN_ENDSFREE, 1
    elif span[0] == "ends-free":
        attr.span = N.SPAN_ENDSFREE  # free ends per pair

    close_out = False
    if out is None:
        if a.output:
            try:
                out = open(a.output, "w")
            except OSError as e:
                fail(f"Error creating output file {a.output!r}: {e}")
            close_out = True
        else:
            out = sys.stdout
    eng = BatchEngine(a.device)
    dev = f"cuda:{a.device}"
    total = correct = 0
    align_ns = 0
    pending_error = None

    def run(batch_pairs):
        nonlocal total, correct, align_ns
        n = len(batch_pairs)
        if n == 0:
            return
        pats = [p for p, _ in batch_pairs]
        txts = [t for _, t in batch_pairs]
        hp = from_ascii(pats, txts, N.SEQ_ASCII)
        cfg = None
        if span[0] == "ends-free":
            cfg = np.zeros((n, 12), dtype=np.int32)
            cfg[:, 0] = N.SPAN_ENDSFREE
            for i, (p, t) in enumerate(batch_pairs):
                cfg[i, 2:6] = [nominal_prop(len(p), span[1][0]), nominal_prop(len(p), span[1][1]),
                               nominal_prop(len(t), span[1][2]), nominal_prop(len(t), span[1][3])]
            cfg[:, 6] = attr.heuristic
            cfg[:, 7:10] = list(attr.heuristic_p)
        b = DeviceBatch(hp, device=dev, cigar=not score_only, pair_config=cfg,
                        end_position=not use_biwfa)
        torch.cuda.synchronize()
        t0 = time.perf_counter_ns()
        eng.align(attr, b)
        torch.cuda.synchronize()
        align_ns += time.perf_counter_ns() - t0
        scores, status, cigars = b.results()
        ends = b.end_positions() if not use_biwfa else None
        lines = []
        for i in range(n):
            p, t = batch_pairs[i]
            st = int(status[i])
            wfa_score = int(scores[i])
            cig = Cigar(cigars[i] if cigars is not None else b"")
            biwfa_failed = use_biwfa and score_only and st != N.STATUS_COMPLETED
            # the end position the driver scores a score-only result at
            # (src/bin/align_benchmark.rs:376-391, 630-641): (|P|, |T|) when
            # the end was reached; else the aligner's end position (a
            # Z-drop's best cell, at the score it was found at, or the unset
            # position, whose v = offset - k wraps)
            no_alignment = False
            if not use_biwfa and st in (N.STATUS_PARTIAL, N.STATUS_END_UNREACHABLE):
                ps, pk, poff = (int(x) for x in ends[i])
                no_alignment = poff == N.OFFSET_NULL
                end_v = (poff - pk + 2 ** 31) % 2 ** 32 - 2 ** 31
                end_h = poff
                if ps >= 0:
                    wfa_score = ps
            else:
                end_v, end_h = len(p), len(t)
            # Corrected mode prints -1 for a score-only result with no
            # alignment, as CIGAR mode does (src/bin/align_benchmark.rs:642-651)
            if score_only and a.no_c_compat and (biwfa_failed or no_alignment):
                score = -1
            elif biwfa_failed:
                score = INT32_MIN
            else:
                score = compute_output_score(cig, wfa_score, penalties, score_only, end_v, end_h)
            rle = "-" if score_only or cig.is_null() else cig.to_string_rle(True)
            lines.append(f"{score}\t{rle}\n")
            if do_check and cig.check_alignment(p, t) is None:
                correct += 1
        out.write("".join(lines))
        total += n

    batch = []
    for p, t in read_pairs(a.input):
        if span[0] == "ends-free":
            ends = [nominal_prop(len(p), span[1][0]), nominal_prop(len(p), span[1][1]),
                    nominal_prop(len(t), span[1][2]), nominal_prop(len(t), span[1][3])]
            if ends[0] > len(p) or ends[1] > len(p) or ends[2] > len(t) or ends[3] > len(t):
                pending_error = (
                    "[WFA] Ends-free parameters must be not larger than the sequences "
                    f"(P0={ends[0]},Pf={ends[1]},T0={ends[2]},Tf={ends[3]}). Must be "
                    "(P0<=|P|,Pf<=|P|,T0<=|T|,Tf<=|T|) where (|P|,|T|)="
                    f"({len(p)},{len(t)})")
                break
        batch.append((p, t))
        if len(batch) >= a.batch_size:
            run(batch)
            batch = []
    run(batch)
    out.flush()
    if pending_error is not None:
        fail(pending_error)
    align_ms = align_ns / 1e6
    per_call = align_ms * 1000.0 / total if total else float("nan")
    print(f"Time.Alignment {align_ms:.2f} ms ({total} calls, {per_call:.2f} us/call)",
          file=sys.stderr)
    if do_check:
        print(f"Alignments.Correct {correct}/{total}", file=sys.stderr)
    if close_out:
        out.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

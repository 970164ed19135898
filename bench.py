#!/usr/bin/env python3
"""Benchmark: batched wavefront alignment with bit-exact CIGARs on one MI355X.

The default workload (one "step" = one pass over the whole batch) is
reads100_300: 1,000,000 pairs, gap-affine (match 0, mismatch 4, open 6,
extend 2), pattern lengths uniform 100-300 bp of random ACGT, per-pair
divergence uniform 0-10% split equally among substitution, insertion and
deletion (indel lengths geometric, mean 2), 2-bit packed input, end-to-end,
full CIGAR in high-memory mode. Inputs are resident in HBM before timing
starts. The same run reports the score-only pass, the plain
one-pair-per-launch GPU path on a sample, the CPU oracle on the host cores,
the roofline figures of the dominant kernel and a parity check of a sample
against the oracle (all 1M pairs by default).

Beside the headline the default run times two legs of the metric's other
configurations, reported under "config": long1k_10k (all 100k pairs of 1-10
kbp, gap-affine 2-piece, medium-memory CIGAR, one step; oracle parity on a
sample) and ultralong50k_100k (1k pairs of 50-100 kbp, edit distance, BiWFA,
one step; parity against the committed oracle digests of its first pairs).
--no-legs skips them.

--workload selects the other configurations: short150 (10k x 150 bp ASCII,
gap-affine, CIGAR high), long1k_10k (100k x 1-10 kbp, gap-affine 2-piece
(0,4,6,2,24,1), CIGAR medium memory) and ultralong50k_100k (1k x 50-100 kbp,
BiWFA, edit or --metric gap_affine).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--pairs N]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The serving leg runs one HIP stream per server engine (SERVING_ENGINES)
# beside this process's other streams. HIP maps streams onto at most
# GPU_MAX_HW_QUEUES hardware queues (4 unless set), and streams that share a
# queue run their kernels in order: a 150 bp launch queued behind a 1-10 kbp
# launch waits for it (with 4 engines, p99 of the 150 bp requests 11-18 ms
# on 4 queues, 0.47 ms on 8). At least 16, set before HIP starts.
SERVING_ENGINES = 8
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

METRIC = "aligned pairs/sec (gap-affine, 150bp + 1-10kbp, bit-exact CIGAR) + p99 latency"
HBM_PEAK_GBS = 8000.0

# memory modes: 0 high, 1 medium, 3 ultralow (BiWFA)
WORKLOADS = {
    "reads100_300": dict(metric="gap_affine", pen=(0, 4, 6, 2, 0, 0), memory=0, pairs=1_000_000,
                         steps=5, cpu_sample=100_000, parity=1_000_000, plain=2000,
                         latency=200,
                         desc="{n} pairs gap-affine (0,4,6,2), 100-300 bp, divergence 0-10%, "
                              "2-bit packed, end-to-end, CIGAR high memory, one persistent "
                              "launch per step"),
    "short150": dict(metric="gap_affine", pen=(0, 4, 6, 2, 0, 0), memory=0, pairs=10_000,
                     steps=10, cpu_sample=10_000, parity=10_000, plain=1000, latency=200,
                     desc="{n} pairs gap-affine (0,4,6,2), 150 bp, divergence 5%, ASCII, "
                          "end-to-end, CIGAR high memory"),
    "long1k_10k": dict(metric="gap_affine_2p", pen=(0, 4, 6, 2, 24, 1), memory=1,
                       pairs=100_000, steps=2, cpu_sample=256, parity=256, plain=20, latency=20,
                       autotune=True,
                       desc="{n} pairs gap-affine 2-piece (0,4,6,2,24,1), 1-10 kbp log-uniform, "
                            "divergence 1-10%, 2-bit packed, end-to-end, CIGAR medium memory"),
    "ultralong50k_100k": dict(metric="edit", pen=(0, 0, 0, 0, 0, 0), memory=3, pairs=1000,
                              steps=1, warmup=0, cpu_sample=2, parity=2, plain=0, latency=0,
                              desc="{n} pairs {metric}, 50-100 kbp, divergence 5-15% with 20% "
                                   "indel runs, 2-bit packed, BiWFA (ultralow memory) CIGAR"),
}
METRIC_PEN = {"edit": (0, 0, 0, 0, 0, 0), "gap_affine": (0, 4, 6, 2, 0, 0),
              "gap_affine_2p": (0, 4, 6, 2, 24, 1)}


# the serving mix without its 10-100 kbp share, for the default run
SHORT_MIX = ((0.70, 0, 150, 150), (0.25, 2, 1000, 10_000))


def serving_main(args, log):
    """--workload serving: the full serving mix (150 bp, 1-10 kbp, 10-100 kbp;
    spans and heuristics mixed), one JSON line with requests/s and p50/p99."""
    from wfa2lib_rs_amd.synth import SERVING_LENGTHS
    n = args.serving_requests or 2000
    rate = args.serving_rate or 500.0
    out = serving_leg(n, rate, SERVING_LENGTHS, 0, 300, log)
    line = {
        "metric": METRIC, "value": out["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": round(n / out["pairs_per_s"] * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"serving: {n} requests, Poisson at {rate}/s, 70% 150 bp, 25% "
                               "1-10 kbp, 5% 10-100 kbp, half end-to-end, a quarter ends-free "
                               "0.1, a quarter extension, heuristic none/wf-adaptive/X-drop/"
                               "banded-adaptive, gap-affine (0,4,6,2) CIGAR high memory",
                   **out},
        "roofline": None,
        "cpu_baseline": {"value": out["oracle_sample_pairs_per_s"], "unit": "pairs/s",
                         "cores": 16, "kind": "port",
                         "sample": "the parity sample of the same stream (requests of at most "
                                   "10 kbp), grouped by configuration, 16 threads"},
    }
    print(json.dumps(line), flush=True)


def serving_leg(n, rate, lengths, seed, oracle_sample, log):
    """Replay a Poisson stream of `n` mixed requests at `rate`/s through the
    serving loop; latency = result received - scheduled arrival."""
    import threading

    import numpy as np

    import __graft_entry__ as entry
    from wfa2lib_rs_amd import _native as N
    from wfa2lib_rs_amd.server import AlignmentServer, request_attr
    from wfa2lib_rs_amd.synth import from_ascii, serving_mix

    mix = serving_mix(n, rate, seed=seed, lengths=lengths)
    attrs = [request_attr(r.span, r.heuristic, len(r.pattern), len(r.text)) for r in mix]
    # 8 engines: up to 4 long-read launches at once (half the engines), so
    # 1-10 kbp requests queue less (3000-request mix p99 49 ms with 4
    # engines, 21 ms with 8; 150 bp p99 0.48 / 0.60 ms)
    srv = AlignmentServer(device=0, num_engines=SERVING_ENGINES)
    # warm the engines (device buffers, code objects) on requests of the
    # same mix (another seed)
    warm = serving_mix(min(400, n), 1e9, seed=seed + 1000, lengths=lengths)
    for i, r in enumerate(warm):
        srv.submit(request_attr(r.span, r.heuristic, len(r.pattern), len(r.text)), r.pattern,
                   r.text, tag=10 ** 9 + i)
    srv.drain(len(warm), timeout_s=600)
    got, recv = {}, {}
    done = threading.Event()

    def poller():
        while len(got) < n:
            for r in srv.poll(timeout_us=2000):
                got[r.tag] = r
                recv[r.tag] = time.perf_counter()
        done.set()

    th = threading.Thread(target=poller, daemon=True)
    # The client's own pauses are not the server's latency: Python's cyclic
    # collector walking this process's million-pair results, and the
    # interpreter's 5 ms thread switch between the submitting thread and the
    # poller, would otherwise land in the measured tail.
    import gc
    gc.collect()
    gc.freeze()
    gc.disable()
    switch = sys.getswitchinterval()
    sys.setswitchinterval(1e-4)
    t0 = time.perf_counter()
    th.start()
    for i, r in enumerate(mix):
        wait = t0 + r.arrival_s - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        srv.submit(attrs[i], r.pattern, r.text, tag=i)
    done.wait(timeout=1200)
    sys.setswitchinterval(switch)
    gc.enable()
    gc.unfreeze()
    t_end = max(recv.values())
    st = srv.stats()
    srv.close()
    assert len(got) == n, f"{len(got)}/{n} results"
    lat = np.array([(recv[i] - (t0 + mix[i].arrival_s)) * 1e3 for i in range(n)])
    # the server's own clock: submission to result ready (wfa_server_poll)
    slat = np.array([got[i].latency_us * 1e-3 for i in range(n)])
    short = np.array([len(r.pattern) <= 256 for r in mix])
    errors = sum(1 for r in got.values() if r.error != 0)
    oom = sum(1 for r in got.values() if r.status == N.STATUS_OOM)
    # parity: the oracle on a sample (every configuration appears)
    o = entry.load_oracle()
    idx = [i for i in range(n) if len(mix[i].pattern) <= 10_000][:oracle_sample]
    groups = {}
    for i in idx:
        groups.setdefault(bytes(attrs[i]), []).append(i)
    same = 0
    t = time.perf_counter()
    for g in groups.values():
        pairs = from_ascii([mix[i].pattern for i in g], [mix[i].text for i in g])
        sc, sts, cg = o.align_batch(attrs[g[0]], pairs, cigar=True, threads=16)
        for j, i in enumerate(g):
            r = got[i]
            same += int(r.score == sc[j] and r.status == sts[j] and r.cigar == cg[j])
    cpu_sec = time.perf_counter() - t
    out = {
        "requests": n, "rate_per_s": rate, "engines": SERVING_ENGINES,
        "pairs_per_s": round(n / (t_end - t0), 3),
        "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
        "latency_ms_p50_150bp": round(float(np.percentile(lat[short], 50)), 3) if short.any() else None,
        "latency_ms_p99_150bp": round(float(np.percentile(lat[short], 99)), 3) if short.any() else None,
        "latency_ms_p99_long": round(float(np.percentile(lat[~short], 99)), 3) if (~short).any() else None,
        "server_latency_ms_p50": round(float(np.percentile(slat, 50)), 3),
        "server_latency_ms_p99": round(float(np.percentile(slat, 99)), 3),
        "server_latency_ms_p99_150bp": round(float(np.percentile(slat[short], 99)), 3) if short.any() else None,
        "launches": int(st["launches"]), "errors": errors, "oom": oom,
        "parity_sample": f"{same}/{len(idx)}",
        "lengths": [list(x) for x in lengths],
        "oracle_sample_pairs_per_s": round(len(idx) / cpu_sec, 3) if idx else None,
    }
    log(f"serving: {out['pairs_per_s']:,.1f} requests/s at {rate}/s offered, p50 "
        f"{out['latency_ms_p50']:.2f} ms, p99 {out['latency_ms_p99']:.2f} ms, "
        f"{out['launches']} launches, parity {out['parity_sample']}")
    return out


def prefix(pairs, n):
    """The first n pairs of a HostPairs buffer, without copying."""
    from wfa2lib_rs_amd import _native as N
    from wfa2lib_rs_amd.synth import HostPairs
    if n >= len(pairs):
        return pairs
    last = int(pairs.tlen[n - 1])
    end = int(pairs.toff[n - 1] + ((last + 3) // 4 if pairs.seq_format == N.SEQ_PACKED2BITS
                                   else last))
    return HostPairs(pairs.seq[:end], pairs.poff[:n], pairs.plen[:n], pairs.toff[:n],
                     pairs.tlen[:n], pairs.seq_format)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def extend_phase(workload):
    """The extend phase of the dominant kernel, measured by its phase-clock
    build (profiles/extend_phase_<workload>.json): the HBM bytes it moves
    (the sequences staged into LDS once per pair) and their rate over the
    phase's share of the launch."""
    p = os.path.join(ROOT, "profiles", f"extend_phase_{workload}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def issue_limit(workload):
    """The limit that binds the dominant kernel, as measured by rocprofv3 SQ
    counters (profiles/issue_<workload>.json): the kernel moves a few bytes
    per pair and is bound by instruction issue, not by HBM."""
    p = os.path.join(ROOT, "profiles", f"issue_{workload}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def run_leg(name, npairs, parity, log):
    """One timed CIGAR step over `npairs` pairs of workload `name` (after a
    warm-up launch on a few pairs), with parity against the oracle on the
    first `parity` pairs, or for ultralong50k_100k against the oracle's
    committed digests of its first pairs."""
    import hashlib

    import numpy as np
    import torch

    import __graft_entry__ as entry
    from wfa2lib_rs_amd import _native as N
    from wfa2lib_rs_amd.engine import BatchEngine, DeviceBatch
    from wfa2lib_rs_amd.synth import WORKLOADS as GEN, generate

    cfg = WORKLOADS[name]
    pairs = generate(GEN[name], num_pairs=npairs)
    a = N.Attr()
    N.lib().wfa_attr_default(ctypes.byref(a),
                             {"edit": 1, "gap_affine": 3, "gap_affine_2p": 4}[cfg["metric"]])
    m, x, o1, e1, o2, e2 = cfg["pen"]
    if cfg["metric"] != "edit":
        a.penalties.match, a.penalties.mismatch = m, x
        a.penalties.gap_opening1, a.penalties.gap_extension1 = o1, e1
        a.penalties.gap_opening2, a.penalties.gap_extension2 = o2, e2
    a.alignment_scope = N.SCOPE_ALIGNMENT
    a.memory_mode = cfg["memory"]
    eng = BatchEngine(0)
    warm = DeviceBatch(pairs.subset(range(min(8, npairs))), device="cuda:0", cigar=True)
    eng.align(a, warm)
    batch = DeviceBatch(pairs, device="cuda:0", cigar=True)
    torch.cuda.synchronize()
    tune, fixed_sec = None, None
    if cfg.get("autotune"):
        # the fixed launch plan first, then the autotuner's choice (the
        # tuning trials run on a sample before the timed step)
        t = time.perf_counter()
        eng.align(a, batch)
        torch.cuda.synchronize()
        fixed_sec = time.perf_counter() - t
        t = time.perf_counter()
        tune = eng.autotune(a, batch)
        torch.cuda.synchronize()
        tune["seconds"] = round(time.perf_counter() - t, 3)
    t = time.perf_counter()
    eng.align(a, batch)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t
    kms = eng.last_kernel_ms()
    score, status, cigs = batch.results()
    same, checked = 0, 0
    if name == "ultralong50k_100k":
        with open(os.path.join(ROOT, "tests", "golden", "ultralong50k_100k_biwfa.json")) as f:
            rows = [r for r in json.load(f)["rows"] if r["metric"] == cfg["metric"]]
        for r in rows:
            i = r["pair"]
            checked += 1
            same += int(int(score[i]) == r["score"] and int(status[i]) == r["status"] and
                        hashlib.sha256(bytes(cigs[i])).hexdigest() == r["sha256"])
        how = "oracle digests (tests/golden/ultralong50k_100k_biwfa.json)"
    else:
        o = entry.load_oracle()
        os_, ost, ocig = o.align_batch(a, prefix(pairs, parity), cigar=True, threads=16)
        for i in range(parity):
            checked += 1
            same += int(int(score[i]) == os_[i] and int(status[i]) == ost[i] and
                        bytes(cigs[i]) == ocig[i])
        how = "oracle"
    out = {"workload": f"{name}: " + cfg["desc"].format(n=npairs, metric=cfg["metric"]),
           "pairs": npairs, "pairs_per_s": round(npairs / sec, 3),
           "ms_per_step": round(sec * 1e3, 3),
           "kernel_ms_first_pass": round(kms[0], 3), "kernel_ms_all_passes": round(kms[1], 3),
           "retried_pairs": int(eng.last_retries()),
           "completed": int((np.asarray(status) == N.STATUS_COMPLETED).sum()),
           "parity": f"{same}/{checked} vs the {how}"}
    if tune is not None:
        out["fixed_plan_pairs_per_s"] = round(npairs / fixed_sec, 3)
        out["autotune"] = tune
    log(f"{name}: {npairs / sec:,.1f} pairs/s ({sec * 1e3:.0f} ms for {npairs} pairs), "
        f"parity {same}/{checked}" +
        (f"; fixed plan {npairs / fixed_sec:,.1f} pairs/s, tuned shape "
         f"{tune['waves_per_pair']} waves x {tune['groups_per_cu']} per CU "
         f"(tuning {tune['seconds']} s)" if tune is not None else ""))
    del batch, warm, eng
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="reads100_300", choices=sorted(WORKLOADS) + ["serving"])
    ap.add_argument("--metric", default=None, choices=sorted(METRIC_PEN),
                    help="override the workload's distance metric")
    ap.add_argument("--pairs", type=int, default=None)
    ap.add_argument("--plain-sample", type=int, default=None)
    ap.add_argument("--cpu-sample", type=int, default=None)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--parity-sample", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=None)
    ap.add_argument("--serving-rate", type=float, default=None, help="requests per second")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the fixed launch plan for long reads (profiling runs)")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the long1k_10k and ultralong50k_100k legs")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (calibrates FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: every box here has one MI355X, and pairs are "
              f"independent, so N={args.gpus} would only shard the batch (see DESIGN.md).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import __graft_entry__ as entry
    pkg = entry.load_package()
    from wfa2lib_rs_amd import _native as N
    from wfa2lib_rs_amd.engine import BatchEngine, DeviceBatch
    from wfa2lib_rs_amd.synth import HostPairs, WORKLOADS as GEN, generate

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    if args.workload == "serving":
        serving_main(args, log)
        return
    cfg = dict(WORKLOADS[args.workload])
    if args.metric:
        cfg["metric"] = args.metric
        cfg["pen"] = METRIC_PEN[args.metric]
    steps = args.steps if args.steps is not None else cfg["steps"]
    warmup = args.warmup if args.warmup is not None else cfg.get("warmup", 1)
    npairs = args.pairs if args.pairs is not None else cfg["pairs"]
    t0 = time.time()
    pairs = generate(GEN[args.workload], num_pairs=npairs)
    log(f"{args.workload}: generated {len(pairs)} pairs in {time.time() - t0:.1f}s")

    def attr_of(scope):
        a = N.Attr()
        N.lib().wfa_attr_default(ctypes.byref(a),
                                 {"edit": 1, "gap_affine": 3, "gap_affine_2p": 4}[cfg["metric"]])
        m, x, o1, e1, o2, e2 = cfg["pen"]
        if cfg["metric"] != "edit":
            a.penalties.match, a.penalties.mismatch = m, x
            a.penalties.gap_opening1, a.penalties.gap_extension1 = o1, e1
            a.penalties.gap_opening2, a.penalties.gap_extension2 = o2, e2
        a.alignment_scope = scope
        a.memory_mode = cfg["memory"]
        return a

    attr_cig = attr_of(N.SCOPE_ALIGNMENT)
    attr_score = attr_of(N.SCOPE_SCORE)
    eng = BatchEngine(0)
    batch = DeviceBatch(pairs, device=dev, cigar=True)

    def timed(attr, n_steps, warmup):
        for _ in range(warmup):
            eng.align(attr, batch)
        torch.cuda.synchronize()
        kms = []
        t = time.perf_counter()
        for _ in range(n_steps):
            eng.align(attr, batch)
            kms.append(eng.last_kernel_ms())
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / n_steps, kms

    # ---- long reads: the fixed launch plan, then the autotuner's shape ----
    tune = None
    if cfg.get("autotune") and not args.no_autotune:
        eng.align(attr_cig, batch)
        torch.cuda.synchronize()
        t = time.perf_counter()
        eng.align(attr_cig, batch)
        torch.cuda.synchronize()
        fixed_sec = time.perf_counter() - t
        t = time.perf_counter()
        tune = {"cigar": eng.autotune(attr_cig, batch), "score_only": eng.autotune(attr_score, batch)}
        torch.cuda.synchronize()
        tune["seconds"] = round(time.perf_counter() - t, 3)
        tune["fixed_plan_pairs_per_s"] = round(len(pairs) / fixed_sec, 3)
        log(f"fixed plan: {len(pairs) / fixed_sec:,.1f} pairs/s; tuned shape "
            f"{tune['cigar']['waves_per_pair']} waves x {tune['cigar']['groups_per_cu']} per CU "
            f"(tuning {tune['seconds']} s)")

    # ---- CIGAR mode: the headline ----
    sec, kms = timed(attr_cig, steps, warmup)
    retries = eng.last_retries()
    pps = len(pairs) / sec
    log(f"CIGAR: {sec * 1e3:.1f} ms/step, {pps:,.1f} pairs/s, retries {retries}")
    score = batch.score.cpu().numpy()[:len(pairs)]
    status = batch.status.cpu().numpy()[:len(pairs)]
    cig_len = batch.cig_len.cpu().numpy()[:len(pairs)].astype(np.int64)
    assert (status == N.STATUS_COMPLETED).all(), "not every pair completed"

    # algorithmic bytes of one launch: input sequences + lengths + score/status + CIGAR ops
    if pairs.seq_format == N.SEQ_PACKED2BITS:
        in_bytes = ((pairs.plen.astype(np.int64) + 3) // 4 + (pairs.tlen.astype(np.int64) + 3) // 4).sum()
    else:
        in_bytes = (pairs.plen.astype(np.int64) + pairs.tlen.astype(np.int64)).sum()
    alg_bytes = int(in_bytes + 8 * len(pairs) + 5 * len(pairs) + cig_len.sum())
    first_ms = float(np.mean([k[0] for k in kms]))
    all_ms = float(np.mean([k[1] for k in kms]))
    achieved = alg_bytes / (first_ms * 1e-3) / 1e9

    # ---- score-only ----
    sec_s, _ = timed(attr_score, max(1, steps // 2), min(warmup, 1))
    pps_score = len(pairs) / sec_s
    log(f"score-only: {sec_s * 1e3:.1f} ms/step, {pps_score:,.1f} pairs/s")

    # ---- plain per-pair path on a sample, and single-pair latency ----
    n_plain = min(args.plain_sample if args.plain_sample is not None else cfg["plain"], len(pairs))
    lat, plain_pps = [], None
    if n_plain > 0:
        sub_batch = DeviceBatch(pairs.subset(range(n_plain)), device=dev, cigar=True)
        eng.align(attr_cig, DeviceBatch(pairs.subset([0]), device=dev, cigar=True), per_pair=True)
        torch.cuda.synchronize()
        for i in range(min(cfg["latency"], n_plain)):
            one = DeviceBatch(pairs.subset([i]), device=dev, cigar=True)
            torch.cuda.synchronize()
            a = time.perf_counter()
            eng.align(attr_cig, one, per_pair=True)
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - a) * 1e3)
        t = time.perf_counter()
        eng.align(attr_cig, sub_batch, per_pair=True)
        torch.cuda.synchronize()
        plain_pps = n_plain / (time.perf_counter() - t)
        log(f"plain per-pair path: {plain_pps:,.1f} pairs/s; batched/plain = {pps / plain_pps:.1f}x")

    # ---- parity on a sample vs the oracle ----
    o = entry.load_oracle()
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    ps = min(args.parity_sample or cfg["parity"], len(pairs))
    os_, ost, ocig = o.align_batch(attr_cig, prefix(pairs, ps), cigar=True, threads=threads)
    ops = batch.cig.cpu().numpy()
    off = batch.cig_off.cpu().numpy()
    same = 0
    for i in range(ps):
        g = bytes(ops[off[i]:off[i] + cig_len[i]])
        same += int(g == ocig[i] and score[i] == os_[i] and status[i] == ost[i])
    log(f"parity: {same}/{ps} pairs identical to the oracle (score, status, CIGAR)")

    # ---- serving: a short Poisson stream of mixed requests (reads up to
    # 10 kbp; the full mix with 10-100 kbp reads is --workload serving) ----
    serving = None
    if not args.no_serving:
        serving = serving_leg(args.serving_requests or 3000, args.serving_rate or 2000.0,
                              SHORT_MIX, 0, 400, log)

    # ---- CPU baseline: the oracle on the host cores, bounded sample ----
    cpu = None
    if not args.no_cpu:
        def first(cs):
            return prefix(pairs, cs)

        cs = min(args.cpu_sample or cfg["cpu_sample"], len(pairs))
        csub = first(cs)
        th = min(threads, cs)
        best = None
        for _ in range(3):
            t = time.perf_counter()
            o.align_batch(attr_cig, csub, cigar=True, threads=th)
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        c1 = max(1, cs // 10)
        t = time.perf_counter()
        o.align_batch(attr_cig, first(c1), cigar=True, threads=1)
        one = c1 / (time.perf_counter() - t)
        cpu = {"value": round(cs / best, 3), "unit": "pairs/s", "cores": th, "kind": "port",
               "sample": f"first {cs} pairs of the same {len(pairs)}-pair workload, CIGAR, "
                         f"{th} threads (one oracle aligner per thread, contiguous shards), "
                         f"best of 3: {best:.2f} s wall",
               "one_thread_pairs_per_s": round(one, 3),
               "one_thread_sample": f"first {c1} pairs, 1 thread",
               "cpu_model": cpu_model(),
               "note": "the oracle: an unoptimised C restatement of the reference's "
                       "algorithm (scalar, no SIMD), not the reference's own Rust build"}
        log(f"cpu oracle: {cpu['value']:,.3f} pairs/s on {th} threads (best of 3), "
            f"{one:,.1f} pairs/s on 1 thread")

    legs = None
    if args.workload == "reads100_300" and not args.no_legs:
        legs = {"long1k_10k": run_leg("long1k_10k", 100_000, 48, log),
                "ultralong50k_100k": run_leg("ultralong50k_100k", 1000, 0, log)}

    traffic = None
    tp = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_cigar.json")
    if os.path.exists(tp):
        with open(tp) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    dtype = "int32"  # offsets and scores; short-read rows are stored in 16 bits
    out = {
        "metric": METRIC,
        "value": round(pps, 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: " + cfg["desc"].format(n=len(pairs), metric=cfg["metric"]),
            "pairs": len(pairs),
            "score_only_pairs_per_s": round(pps_score, 3),
            "plain_per_pair_pairs_per_s": round(plain_pps, 3) if plain_pps else None,
            "batched_over_plain": round(pps / plain_pps, 2) if plain_pps else None,
            "plain_latency_ms_p50": round(float(np.percentile(lat, 50)), 3) if lat else None,
            "plain_latency_ms_p99": round(float(np.percentile(lat, 99)), 3) if lat else None,
            "batch_latency_ms": round(sec * 1e3, 3),
            "retried_pairs": int(retries),
            "parity_sample": f"{same}/{ps}",
            "kernel_ms_first_pass": round(first_ms, 3),
            "kernel_ms_all_passes": round(all_ms, 3),
            "serving": serving,
            "legs": legs,
            **({"autotune": tune} if tune else {}),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": traffic,
            # the measured HBM bytes of one launch over this run's kernel time
            "traffic_GBps": round(traffic / (first_ms * 1e-3) / 1e9, 1) if traffic else None,
            "binding": issue_limit(args.workload),
            "extend_phase": extend_phase(args.workload),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

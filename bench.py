#!/usr/bin/env python3
"""Benchmark: batched gap-affine WFA with bit-exact CIGARs on one MI355X.

Workload (one "step" = one pass over the whole batch): 1,000,000 pairs,
gap-affine (match 0, mismatch 4, open 6, extend 2), pattern lengths uniform
100-300 bp of random ACGT, per-pair divergence uniform 0-10% split equally
among substitution, insertion and deletion (indel lengths geometric, mean
2), 2-bit packed input, end-to-end, full CIGAR in high-memory mode. Inputs
are resident in HBM before timing starts. The same run reports the
score-only pass, the plain one-pair-per-launch GPU path on a sample, the
CPU oracle on the host cores, the roofline figures of the dominant kernel
and a parity check of a sample against the oracle.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--pairs N]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "aligned pairs/sec (gap-affine, 150bp + 1-10kbp, bit-exact CIGAR) + p99 latency"
HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--plain-sample", type=int, default=2000)
    ap.add_argument("--cpu-sample", type=int, default=500_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--parity-sample", type=int, default=4000)
    ap.add_argument("--no-cpu", action="store_true")
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
    from wfa2lib_rs_amd.engine import BatchEngine, DeviceBatch, make_attr
    from wfa2lib_rs_amd.synth import WORKLOADS, generate

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    w = WORKLOADS["reads100_300"]
    t0 = time.time()
    pairs = generate(w, num_pairs=args.pairs)
    log(f"generated {len(pairs)} pairs in {time.time() - t0:.1f}s")
    pen = pkg.WavefrontPenalties.new_affine(pkg.AffinePenalties(0, 4, 6, 2))
    attr_cig = make_attr(pen, scope=N.SCOPE_ALIGNMENT)
    attr_score = make_attr(pen, scope=N.SCOPE_SCORE)
    eng = BatchEngine(0)
    batch = DeviceBatch(pairs, device=dev, cigar=True)
    stream = torch.cuda.current_stream()

    def timed(attr, steps, warmup):
        for _ in range(warmup):
            eng.align(attr, batch)
        torch.cuda.synchronize()
        kms = []
        t = time.perf_counter()
        for _ in range(steps):
            eng.align(attr, batch)
            kms.append(eng.last_kernel_ms())
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps, kms

    # ---- CIGAR mode: the headline ----
    sec, kms = timed(attr_cig, args.steps, args.warmup)
    retries = eng.last_retries()
    pps = len(pairs) / sec
    log(f"CIGAR: {sec * 1e3:.1f} ms/step, {pps:,.0f} pairs/s, retries {retries}")
    score, status, _ = batch.score.cpu().numpy()[:len(pairs)], batch.status.cpu().numpy()[:len(pairs)], None
    cig_len = batch.cig_len.cpu().numpy()[:len(pairs)].astype(np.int64)
    assert (status == N.STATUS_COMPLETED).all(), "not every pair completed"

    # algorithmic bytes of one launch: packed input + lengths + score/status + CIGAR ops
    in_bytes = ((pairs.plen.astype(np.int64) + 3) // 4 + (pairs.tlen.astype(np.int64) + 3) // 4).sum()
    alg_bytes = int(in_bytes + 8 * len(pairs) + 5 * len(pairs) + cig_len.sum())
    first_ms = float(np.mean([k[0] for k in kms]))
    all_ms = float(np.mean([k[1] for k in kms]))
    achieved = alg_bytes / (first_ms * 1e-3) / 1e9

    # ---- score-only ----
    sec_s, kms_s = timed(attr_score, max(1, args.steps // 2), 1)
    pps_score = len(pairs) / sec_s
    log(f"score-only: {sec_s * 1e3:.1f} ms/step, {pps_score:,.0f} pairs/s")

    # ---- plain per-pair path on a sample ----
    sub = pairs.subset(range(args.plain_sample))
    sub_batch = DeviceBatch(sub, device=dev, cigar=True)
    eng.align(attr_cig, sub_batch, per_pair=True)  # warm
    torch.cuda.synchronize()
    lat = []
    t = time.perf_counter()
    from wfa2lib_rs_amd.synth import from_ascii
    for i in range(min(200, args.plain_sample)):
        one = DeviceBatch(pairs.subset([i]), device=dev, cigar=True)
        torch.cuda.synchronize()
        a = time.perf_counter()
        eng.align(attr_cig, one, per_pair=True)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    t = time.perf_counter()
    eng.align(attr_cig, sub_batch, per_pair=True)
    torch.cuda.synchronize()
    plain_pps = args.plain_sample / (time.perf_counter() - t)
    log(f"plain per-pair path: {plain_pps:,.0f} pairs/s; batched/plain = {pps / plain_pps:.1f}x")

    # ---- parity on a sample vs the oracle ----
    o = entry.load_oracle()
    ps = min(args.parity_sample, len(pairs))
    psub = pairs.subset(range(ps))
    os_, ost, ocig = o.align_batch(attr_cig, psub, cigar=True, threads=16)
    ops = batch.cig.cpu().numpy()
    off = batch.cig_off.cpu().numpy()
    same = 0
    for i in range(ps):
        g = bytes(ops[off[i]:off[i] + cig_len[i]])
        same += int(g == ocig[i] and score[i] == os_[i] and status[i] == ost[i])
    log(f"parity: {same}/{ps} pairs identical to the oracle (score, status, CIGAR)")

    # ---- CPU baseline: the oracle on the host cores, bounded sample ----
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cs = min(args.cpu_sample, len(pairs))
        csub_idx = range(cs)
        import numpy as _np
        # contiguous prefix of the same workload
        from wfa2lib_rs_amd.synth import HostPairs
        end = int(pairs.toff[cs - 1] + (pairs.tlen[cs - 1] + 3) // 4)
        csub = HostPairs(pairs.seq[:end], pairs.poff[:cs], pairs.plen[:cs], pairs.toff[:cs],
                         pairs.tlen[:cs], pairs.seq_format)
        t = time.perf_counter()
        o.align_batch(attr_cig, csub, cigar=True, threads=threads)
        cpu_sec = time.perf_counter() - t
        cpu = {"value": round(cs / cpu_sec, 1), "unit": "pairs/s", "cores": threads,
               "kind": "port",
               "sample": f"first {cs} pairs of the same 1M-pair workload, CIGAR, "
                         f"{threads} threads (one oracle aligner per thread, contiguous shards), "
                         f"{cpu_sec:.2f} s wall"}
        log(f"cpu oracle: {cpu['value']:,.0f} pairs/s on {threads} threads")

    traffic = None
    tp = os.path.join(ROOT, "profiles", "traffic_reads100_300_cigar.json")
    if os.path.exists(tp):
        with open(tp) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": round(pps, 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": f"reads100_300: {len(pairs)} pairs gap-affine (0,4,6,2), 100-300 bp, "
                        "divergence 0-10%, 2-bit packed, end-to-end, CIGAR high memory, "
                        "one persistent launch per step",
            "pairs": len(pairs),
            "score_only_pairs_per_s": round(pps_score, 1),
            "plain_per_pair_pairs_per_s": round(plain_pps, 1),
            "batched_over_plain": round(pps / plain_pps, 2),
            "plain_latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
            "plain_latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
            "batch_latency_ms": round(sec * 1e3, 3),
            "retried_pairs": int(retries),
            "parity_sample": f"{same}/{ps}",
            "kernel_ms_first_pass": round(first_ms, 3),
            "kernel_ms_all_passes": round(all_ms, 3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Serving loop on the GPU (continuous batching): every request's result is
bit-exact with the oracle and identical to what the same request gets when
it is aligned alone, whatever launch it shared; requests submitted while
others run are admitted and finish."""
import time

import numpy as np
import pytest

from conftest import attr_for, have_gpu

pytestmark = pytest.mark.gpu

# a small serving mix: 150 bp reads and 1-3 kbp reads
SMALL_MIX = ((0.85, 0, 150, 150), (0.15, 2, 1000, 3000))


@pytest.fixture(scope="module")
def server(pkg):
    if not have_gpu():
        pytest.skip("no GPU")
    from wfa2lib_rs_amd.server import AlignmentServer
    s = AlignmentServer(device=0, num_engines=4, memory_budget=8 << 30)
    yield s
    s.close()


def _oracle(oracle, reqs, attrs):
    """Oracle results per request, grouped by configuration."""
    from wfa2lib_rs_amd.synth import from_ascii
    groups = {}
    for i, a in enumerate(attrs):
        groups.setdefault(bytes(a), []).append(i)
    out = {}
    for idx in groups.values():
        a = attrs[idx[0]]
        pairs = from_ascii([reqs[i][0] for i in idx], [reqs[i][1] for i in idx])
        sc, st, cg = oracle.align_batch(a, pairs, cigar=a.alignment_scope == 1, threads=8)
        for j, i in enumerate(idx):
            out[i] = (int(sc[j]), int(st[j]), cg[j] if cg is not None else b"")
    return out


def _alone(engine, reqs, attrs, which):
    """The batch engine on a one-pair launch per request."""
    from wfa2lib_rs_amd.engine import DeviceBatch
    from wfa2lib_rs_amd.synth import from_ascii
    import torch
    out = {}
    for i in which:
        pairs = from_ascii([reqs[i][0]], [reqs[i][1]])
        cig = attrs[i].alignment_scope == 1
        b = DeviceBatch(pairs, device="cuda:0", cigar=cig)
        engine.align(attrs[i], b)
        torch.cuda.synchronize()
        s, st, c = b.results()
        out[i] = (int(s[0]), int(st[0]), c[0] if c else b"")
    return out


def _check(results, expect, label):
    bad = [(i, results[i], expect[i]) for i in expect if results[i] != expect[i]]
    assert not bad, f"{label}: {len(bad)} mismatches, first {bad[0]}"


def test_serving_mix_matches_oracle_and_alone(pkg, N, server, oracle):
    from wfa2lib_rs_amd.engine import BatchEngine
    from wfa2lib_rs_amd.server import request_attr
    from wfa2lib_rs_amd.synth import serving_mix
    mix = serving_mix(240, 1e9, seed=3, lengths=SMALL_MIX)
    reqs = [(r.pattern, r.text) for r in mix]
    attrs = [request_attr(r.span, r.heuristic, len(r.pattern), len(r.text)) for r in mix]
    # score-only, medium-memory and BiWFA requests join the mix
    for i in range(0, len(attrs), 7):
        attrs[i].alignment_scope = N.SCOPE_SCORE
    for i in range(3, len(attrs), 11):
        attrs[i].memory_mode = N.MEMORY_MED
    # Z-drop requests (high memory or score only) share the heuristic launches
    for i in range(2, len(attrs), 17):
        if attrs[i].memory_mode == N.MEMORY_HIGH:
            attrs[i].heuristic = N.HEURISTIC_ZDROP
            attrs[i].heuristic_p[0], attrs[i].heuristic_p[1] = 10, 1
    for i in range(5, len(attrs), 13):
        if attrs[i].heuristic not in (N.HEURISTIC_XDROP, N.HEURISTIC_ZDROP):
            attrs[i].memory_mode = N.MEMORY_ULTRALOW
            attrs[i].span, attrs[i].extension = N.SPAN_END2END, 0
            attrs[i].pattern_begin_free = attrs[i].pattern_end_free = 0
            attrs[i].text_begin_free = attrs[i].text_end_free = 0
    # first half at once, then the rest while the first half runs
    half = len(reqs) // 2
    for i in range(half):
        server.submit(attrs[i], reqs[i][0], reqs[i][1], tag=i)
    got = server.poll(timeout_us=0)
    for i in range(half, len(reqs)):
        server.submit(attrs[i], reqs[i][0], reqs[i][1], tag=i)
    got += server.drain(len(reqs) - len(got), timeout_s=300)
    assert sorted(r.tag for r in got) == list(range(len(reqs)))
    assert all(r.error == 0 for r in got), [r for r in got if r.error][:3]
    assert all(r.latency_us > 0 for r in got)
    results = {r.tag: (r.score, r.status, r.cigar) for r in got}
    _check(results, _oracle(oracle, reqs, attrs), "server vs oracle")
    eng = BatchEngine(0)
    _check(results, _alone(eng, reqs, attrs, range(0, len(reqs), 5)), "server vs alone")
    st = server.stats()
    assert st["queued"] == 0 and st["running"] == 0
    assert st["launched_pairs"] >= len(reqs)


def test_serving_same_request_in_every_mix(pkg, N, server):
    """One request, submitted alone and then inside two different mixes,
    gets the same result each time."""
    from wfa2lib_rs_amd.server import request_attr
    from wfa2lib_rs_amd.synth import serving_mix
    probe = serving_mix(1, 1.0, seed=11, lengths=((1.0, 0, 150, 150),))[0]
    pa = request_attr("end2end", ("xdrop", 200, 1), 150, 150)
    seen = []
    for mix_seed in (None, 4, 5):
        n = 0
        if mix_seed is not None:
            for j, r in enumerate(serving_mix(80, 1e9, seed=mix_seed, lengths=SMALL_MIX)):
                server.submit(request_attr(r.span, r.heuristic, len(r.pattern), len(r.text)),
                              r.pattern, r.text, tag=1000 + j)
                n += 1
        server.submit(pa, probe.pattern, probe.text, tag=7)
        got = server.drain(n + 1, timeout_s=300)
        mine = [r for r in got if r.tag == 7]
        assert len(mine) == 1 and mine[0].error == 0
        seen.append((mine[0].score, mine[0].status, mine[0].cigar))
    assert seen[0] == seen[1] == seen[2]


def test_serving_rejects_invalid_requests(pkg, N, server):
    a = attr_for(N, "gap_affine", (0, 4, 6, 2, 0, 0), "alignment")
    bad = attr_for(N, "gap_affine", (1, 4, 6, 2, 0, 0), "alignment")  # match > 0
    with pytest.raises(RuntimeError, match="invalid penalties"):
        server.submit(bad, b"ACGT", b"ACGT", tag=1)
    # BiWFA takes no Z-drop (src/bin/align_benchmark.rs:529-560)
    z = attr_for(N, "gap_affine", (0, 4, 6, 2, 0, 0), "alignment")
    z.heuristic = N.HEURISTIC_ZDROP
    z.memory_mode = N.MEMORY_ULTRALOW
    with pytest.raises(RuntimeError, match="unsupported"):
        server.submit(z, b"ACGT", b"ACGT", tag=2)
    # an empty pair is served
    server.submit(a, b"", b"", tag=3)
    got = server.drain(1, timeout_s=60)
    assert got[0].tag == 3 and got[0].error == 0 and got[0].score == 0


def test_serving_poisson_stream(pkg, N, server, oracle):
    """Requests arriving over time (Poisson) are admitted while earlier ones
    run; all finish and match the oracle."""
    from wfa2lib_rs_amd.server import request_attr
    from wfa2lib_rs_amd.synth import serving_mix
    mix = serving_mix(300, 3000.0, seed=9, lengths=SMALL_MIX)
    attrs = [request_attr(r.span, r.heuristic, len(r.pattern), len(r.text)) for r in mix]
    t0 = time.perf_counter()
    got = []
    for i, r in enumerate(mix):
        while time.perf_counter() - t0 < r.arrival_s:
            got += server.poll(timeout_us=0)
        server.submit(attrs[i], r.pattern, r.text, tag=i)
    got += server.drain(len(mix) - len(got), timeout_s=300)
    assert sorted(x.tag for x in got) == list(range(len(mix)))
    results = {x.tag: (x.score, x.status, x.cigar) for x in got}
    _check(results, _oracle(oracle, [(r.pattern, r.text) for r in mix], attrs), "poisson")
    lat = np.array([x.latency_us for x in got])
    assert np.isfinite(lat).all() and (lat > 0).all()


def _endsfree_penalty(ops, ends, x=4, o=6, e=2):
    """A gap-affine (0,x,o,e) CIGAR's ends-free penalty: a leading or
    trailing gap run is free up to its end's limit (tests/corrected_endsfree.rs:
    109-146); ends = (pattern begin, pattern end, text begin, text end)."""
    runs = []
    for op in ops:
        if runs and runs[-1][0] == op:
            runs[-1][1] += 1
        else:
            runs.append([op, 1])
    pb, pe, tb, te = ends
    total = 0
    for i, (op, n) in enumerate(runs):
        if op == ord("X"):
            total += x * n
        elif op != ord("M"):
            ch = n
            if i == 0:
                ch -= min(ch, pb if op == ord("D") else tb)
            if i == len(runs) - 1:
                ch -= min(ch, pe if op == ord("D") else te)
            if ch > 0:
                total += o + e * ch
    return total


def test_serving_full_mix(pkg, N, server, oracle):
    """The full serving mix (70% 150 bp, 25% 1-10 kbp, 5% 10-100 kbp; half
    end-to-end, a quarter ends-free, a quarter extension; heuristics none /
    WF-adaptive / X-drop / banded-adaptive), half submitted while the other
    half runs. Requests up to 10 kbp equal the oracle; the 10-100 kbp ones
    are checked at full size by properties: a valid CIGAR (spanning both
    sequences unless extension or cut short), its penalty equal to the score
    for completed end-to-end and ends-free requests, and the same result
    when the request is aligned alone."""
    from test_oracle import valid_prefix
    from wfa2lib_rs_amd.engine import BatchEngine
    from wfa2lib_rs_amd.server import request_attr
    from wfa2lib_rs_amd.synth import SERVING_LENGTHS, serving_mix
    mix = serving_mix(120, 1e9, seed=21, lengths=SERVING_LENGTHS)
    reqs = [(r.pattern, r.text) for r in mix]
    attrs = [request_attr(r.span, r.heuristic, len(r.pattern), len(r.text)) for r in mix]
    long_idx = [i for i, r in enumerate(mix) if max(len(r.pattern), len(r.text)) > 10_000]
    assert len(long_idx) >= 4, "the mix must hold several 10-100 kbp requests"
    half = len(reqs) // 2
    for i in range(half):
        server.submit(attrs[i], reqs[i][0], reqs[i][1], tag=i)
    got = server.poll(timeout_us=0)
    for i in range(half, len(reqs)):
        server.submit(attrs[i], reqs[i][0], reqs[i][1], tag=i)
    got += server.drain(len(reqs) - len(got), timeout_s=600)
    assert sorted(r.tag for r in got) == list(range(len(reqs)))
    assert all(r.error == 0 for r in got), [r for r in got if r.error][:3]
    results = {r.tag: (r.score, r.status, r.cigar) for r in got}
    short = [i for i in range(len(reqs)) if i not in long_idx]
    want = _oracle(oracle, [reqs[i] for i in short], [attrs[i] for i in short])
    _check({i: results[i] for i in short}, {i: want[j] for j, i in enumerate(short)},
           "full mix vs oracle")
    for i in long_idx:
        p, t = reqs[i]
        s, st, c = results[i]
        assert st != N.STATUS_OOM, i
        assert valid_prefix(c, p, t), i
        a = attrs[i]
        whole = pkg.Cigar(c).check_alignment(p, t) is None
        if mix[i].span != "extension" and st == N.STATUS_COMPLETED:
            assert whole, i
            ends = (a.pattern_begin_free, a.pattern_end_free, a.text_begin_free,
                    a.text_end_free)
            assert _endsfree_penalty(c, ends) == s, i
    _check(results, _alone(BatchEngine(0), reqs, attrs, long_idx), "long requests alone")

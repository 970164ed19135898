"""The align_benchmark driver (src/bin/align_benchmark.rs): every literal
output line of the reference's CLI tests (tests/cli_output.rs, transcribed
in tests/golden/) reproduced through the GPU driver; argument rejections
checked on the CPU (they happen before any device work)."""
import io
import os

import pytest

from conftest import expand, have_gpu, load_vectors

HEUR_NAMES = {"wfadaptive": "wfa-adaptive", "xdrop": "xdrop", "zdrop": "zdrop",
              "banded_static": "banded-static", "banded_adaptive": "banded-adaptive"}


def _args(vec):
    m, x, o1, e1, o2, e2 = vec["penalties"]
    if vec["metric"] == "gap_affine":
        a = ["-a", "gap-affine-wfa", "-g", f"{m},{x},{o1},{e1}"]
    elif vec["metric"] == "gap_affine_2p":
        a = ["-a", "gap-affine2p-wfa", "--affine2p-penalties", f"{m},{x},{o1},{e1},{o2},{e2}"]
    elif vec["metric"] == "gap_linear":
        a = ["-a", "gap-linear-wfa", "-p", f"{m},{x},{o1}"]
    else:
        a = ["-a", f"{vec['metric']}-wfa"]
    if vec["scope"] == "score":
        a.append("--wfa-score-only")
    span = vec.get("span")
    if span == "extension":
        a += ["--wfa-span", "extension"]
    elif span is not None:
        a += ["--wfa-span", "ends-free," + ",".join(str(v) for v in span)]
    if "heuristic" in vec:
        h = vec["heuristic"]
        a += ["--wfa-heuristic", HEUR_NAMES[h[0]],
              "--wfa-heuristic-parameters", ",".join(str(v) for v in h[1:])]
    if "memory" in vec:
        a += ["--wfa-memory", vec["memory"]]
    return a


def _cli_groups():
    groups = {}
    for v in load_vectors():
        if "cli" in v:
            groups.setdefault(tuple(_args(v)), []).append(v)
    return sorted(groups.items())


@pytest.mark.gpu
@pytest.mark.parametrize("args,vecs", _cli_groups(), ids=lambda x: x[0]["id"] if isinstance(x, list) else None)
def test_cli_output_lines(args, vecs, pkg, tmp_path):
    if not have_gpu():
        pytest.skip("no GPU")
    from wfa2lib_rs_amd import align_benchmark
    seq = tmp_path / "in.seq"
    with open(seq, "wb") as f:
        for v in vecs:
            f.write(b">" + expand(v["pattern"]) + b"\n<" + expand(v["text"]) + b"\n")
    out = io.StringIO()
    assert align_benchmark.main(["-i", str(seq), "--batch-size", "3"] + list(args), out=out) == 0
    got = out.getvalue().splitlines()
    assert got == [v["cli"] for v in vecs], (args, got)


@pytest.mark.parametrize("args,msg", [
    (["--wfa-memory", "ultralow", "--wfa-span", "extension"], "BiWFA ends-free and extension"),
    (["--wfa-memory", "ultralow", "--wfa-heuristic", "xdrop", "--wfa-heuristic-parameters",
      "20,1"], "BiWFA is not compatible"),
    (["-a", "edit-wfa", "--wfa-heuristic", "xdrop", "--wfa-heuristic-parameters", "20,1"],
     "not compatible with 'edit'/'indel'"),
    (["--wfa-heuristic", "xdrop", "--wfa-heuristic-parameters", "20"],
     "Invalid --wfa-heuristic-parameters"),
    (["--wfa-span", "ends-free,1,2"], "Invalid --wfa-span"),
    (["--wfa-memory", "huge"], "--wfa-memory"),
    (["-g", "1,4,6,2"], "invalid penalties"),
    (["-a", "gap-linear-wfa", "-p", "0,4,0"], "invalid penalties"),
])
def test_cli_rejections(args, msg, pkg, tmp_path, capsys):
    """The reference driver's rejections (src/bin/align_benchmark.rs:
    529-560) exit 1 before any alignment; so do configurations the device
    path does not run."""
    from wfa2lib_rs_amd import align_benchmark
    seq = tmp_path / "in.seq"
    seq.write_bytes(b">ACGT\n<ACGT\n")
    with pytest.raises(SystemExit) as e:
        align_benchmark.main(["-i", str(seq)] + args, out=io.StringIO())
    assert e.value.code == 1
    assert msg in capsys.readouterr().err


@pytest.mark.gpu
def test_cli_corrected_unreachable_prints_minus_one(pkg, tmp_path):
    """tests/cli_output.rs:391-437 (HPCBio/wfa2lib-rs#23): a score-only
    result with no alignment prints C's converted give-up score (-113 with
    a match reward, the negated score with match == 0, INT32_MIN for a
    failed BiWFA score); --no-c-compat prints -1 for each."""
    if not have_gpu():
        pytest.skip("no GPU")
    from wfa2lib_rs_amd import align_benchmark
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

    def run(path, args):
        out = io.StringIO()
        assert align_benchmark.main(["-i", path] + args, out=out) == 0
        return out.getvalue().splitlines()

    issue23 = os.path.join(golden, "issue23_unreachable.seq")
    xdrop = ["-a", "gap-affine-wfa", "-g", "-3,5,3,2", "--wfa-heuristic", "xdrop",
             "--wfa-heuristic-parameters", "20,1", "--wfa-score-only"]
    assert run(issue23, xdrop) == ["-113\t-"]
    assert run(issue23, xdrop + ["--no-c-compat"]) == ["-1\t-"]
    pattern = b"ACGTTGCAAGCT" * 10
    seq = tmp_path / "u.seq"
    seq.write_bytes(b">" + pattern + b"\n<" + pattern[:90] + b"\n")
    band = ["-a", "gap-affine-wfa", "--wfa-heuristic", "banded-static",
            "--wfa-heuristic-parameters", "-5,5", "--wfa-score-only"]
    assert run(str(seq), band) != ["-1\t-"]
    assert run(str(seq), band + ["--no-c-compat"]) == ["-1\t-"]
    bi = band + ["-g", "-3,5,3,2", "--wfa-memory", "ultralow"]
    assert run(str(seq), bi) == ["-2147483648\t-"]
    assert run(str(seq), bi + ["--no-c-compat"]) == ["-1\t-"]


def test_cli_reader_and_nominal_prop(pkg, tmp_path):
    """PairReader strips '>'/'<' and line ends and drops an unpaired last
    line; nominal_prop takes a fraction in [0, 1] of the length, else the
    value itself (src/bin/align_benchmark.rs:86-140, 255-263)."""
    from wfa2lib_rs_amd.align_benchmark import nominal_prop, read_pairs
    seq = tmp_path / "in.seq"
    seq.write_bytes(b">ACGT\r\n<ACGA\nTTT\nGG\n>C\n")
    assert list(read_pairs(str(seq))) == [(b"ACGT", b"ACGA"), (b"TTT", b"GG")]
    assert nominal_prop(20, 0.5) == 10 and nominal_prop(20, 1.0) == 20
    assert nominal_prop(20, 7.0) == 7 and nominal_prop(15, 0.1) == 1


def test_cli_negative_values_parse(pkg):
    """Penalties and band limits may be negative (match rewards, bands
    below the main diagonal)."""
    from wfa2lib_rs_amd.align_benchmark import parse_args
    a = parse_args(["-i", "x", "-g", "-3,5,3,2", "--wfa-heuristic", "banded-static",
                    "--wfa-heuristic-parameters", "-5,5"])
    assert a.affine_penalties == "-3,5,3,2" and a.wfa_heuristic_parameters == "-5,5"

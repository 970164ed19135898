"""Build the engine (gfx950) and the synthetic-workload generator in-tree.

    python wfa2lib-rs_amd/build.py            # the release engine and the generator
    python wfa2lib-rs_amd/build.py --checks   # the debug engine (device index checks)

``libwfa_mi355x.so`` is built with hipcc for gfx950 from csrc/*.hip and
csrc/*.cpp; ``libwfa_synth.so`` with gcc. Outputs go to ``lib/``.
``--phase-clocks`` builds ``libwfa_mi355x_phases.so`` with WFA_PHASE_CLOCKS:
the short-read kernel sums its waves' cycles for sequence staging, the
extension, the CIGAR backtrace and whole pairs (wfa_engine_last_phase_cycles).
``--checks`` builds ``libwfa_mi355x_checks.so`` with WFA_DEVICE_CHECKS: every
ring and history index is tested on the device and a call that found one
out of range fails with WFA_ERR_DEVICE; set WFA_MI355X_LIB to its file name
to load it in place of the release library.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed-FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_engine(checks=False, phases=False):
    """Compile each .hip translation unit to an object in parallel (the
    per-kind kernel sets are the slow part), then link with the host API."""
    os.makedirs(LIB, exist_ok=True)
    variant = "checks" if checks else "phases" if phases else ""
    obj_dir = os.path.join(LIB, "obj_" + variant if variant else "obj")
    os.makedirs(obj_dir, exist_ok=True)
    hips = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    flags = HIP_FLAGS + (["-DWFA_DEVICE_CHECKS"] if checks else []) + (
        ["-DWFA_PHASE_CLOCKS"] if phases else [])
    compile_flags = [f for f in flags if f != "-shared"]
    procs = []
    for f in hips:
        obj = os.path.join(obj_dir, f[:-4] + ".o")
        cmd = [HIPCC] + compile_flags + ["-c", "-o", obj, os.path.join(CSRC, f)]
        print("+", " ".join(cmd), flush=True)
        procs.append((cmd, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                 stderr=subprocess.STDOUT, text=True)))
    objs, failed = [], []
    for cmd, obj, pr in procs:
        out, _ = pr.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
        if out.strip():
            print(out, flush=True)
        if pr.returncode != 0:
            failed.append(" ".join(cmd))
        objs.append(obj)
    if failed:
        raise RuntimeError("build step failed: " + "; ".join(failed))
    out = os.path.join(LIB, f"libwfa_mi355x_{variant}.so" if variant else "libwfa_mi355x.so")
    tmp = out + ".tmp"
    _run([HIPCC] + flags + ["-o", tmp] + objs +
         [os.path.join(CSRC, f) for f in ("wfa_api.cpp", "wfa_biwfa.cpp", "wfa_server.cpp")])
    os.replace(tmp, out)
    return out


def build_synth():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libwfa_synth.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O2", "-Wall", "-fPIC", "-shared", "-o", tmp, os.path.join(CSRC, "synth.c"),
          "-lm", "-lpthread"])
    os.replace(tmp, out)
    return out


def build_all():
    return [build_engine(), build_synth()]


if __name__ == "__main__":
    if "--checks" in sys.argv[1:]:
        build_engine(checks=True)
    elif "--phase-clocks" in sys.argv[1:]:
        build_engine(phases=True)
    else:
        build_all()
    sys.exit(0)

"""Build the engine (gfx950) and the synthetic-workload generator in-tree.

    python wfa2lib-rs_amd/build.py

``libwfa_mi355x.so`` is built with hipcc for gfx950 from csrc/*.hip and
csrc/*.cpp; ``libwfa_synth.so`` with gcc. Outputs go to ``lib/``.
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


def build_engine():
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(CSRC, f) for f in ("wfa_kernels.hip", "wfa_api.cpp")]
    out = os.path.join(LIB, "libwfa_mi355x.so")
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-o", tmp] + srcs)
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
    build_all()
    sys.exit(0)

"""The oracle and the workload generator (csrc/synth.c) under
AddressSanitizer and UndefinedBehaviorSanitizer: `make -C oracle san` runs
every configuration the oracle accepts on small synthetic and empty pairs
and checks each completed CIGAR (oracle/san_check.c). CPU only."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_and_generator_under_asan_ubsan():
    probe = subprocess.run(["gcc", "-fsanitize=address,undefined", "-x", "c", "-", "-o",
                            "/dev/null"], input="int main(void){return 0;}", text=True,
                           capture_output=True)
    if probe.returncode != 0:
        pytest.skip("gcc without sanitizer runtimes")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "0 failures" in r.stdout

"""The two-division receiver rule used by the kernels (csrc/fs_receiver.h)
picks exactly the direction of the reference's eight-division loop
(fastscape_RB.cpp:160-182) on random, tied, near-tied (1-4 ulp) and special
inputs (NaN, infinities, signed zeros, subnormals, huge magnitudes).  The
same header is compiled by hipcc's host pass and run on the CPU."""
import os
import subprocess

from conftest import ROOT

CSRC = os.path.join(ROOT, "barnes2019-landscape_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_fast_receiver_rule_equals_reference(tmp_path):
    exe = tmp_path / "rule"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-I" + CSRC,
                    os.path.join(ROOT, "tests", "native", "receiver_rule_check.cpp"), "-o",
                    str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe), "4000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    cases, bad = map(int, r.stdout.split()[-2:])
    assert cases == 4000000 and bad == 0

"""Accuracy of the FP64 Laguerre weight exp(-x/2) (pe_exp64.h) on the CPU:
against the host's extended-precision expl, the error stays below one unit
in the last place of the correctly rounded double (measured: 0.87) over
x in [0, 1400] and pseudo-random moneyness in [0, 4].  The device runs the
same source; its effect on the induction is pinned by the FP64 LSM parity
tests (test_gpu_lsm.py)."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _checker(tmp, plant=None):
    exe = os.path.join(tmp, "exp_neg_half_check")
    extra = [f"-DPE_PLANT_REFERENCE_SCALE={plant}L"] if plant else []
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off"] + extra +
                          ["-I", os.path.join(ROOT, "rust-pricing-library_amd", "csrc"),
                           os.path.join(ROOT, "tests", "exp_neg_half_check.cpp"), "-o", exe])
    return exe


def test_within_one_ulp():
    with tempfile.TemporaryDirectory() as tmp:
        exe = _checker(tmp)
        for hi in ("1400", "4"):
            err, x = subprocess.check_output([exe, hi, "4000000"], text=True).split()
            assert float(err) < 1.0, (hi, err, x)


def test_checker_catches_a_planted_reference_error():
    """The same check against a reference scaled by 1 + 1e-15 (a mistake
    planted in the reference on the CPU, about 4.5 units in the last place)
    reports an error above the bound."""
    with tempfile.TemporaryDirectory() as tmp:
        err, _ = subprocess.check_output([_checker(tmp, plant="1.000000000000001"), "4", "100000"],
                                         text=True).split()
    assert float(err) > 2.0

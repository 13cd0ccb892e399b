"""Accuracy of the mixed mode's FP32 inverse normal (pe_normal_mixed.h) on
the CPU over every 32-bit word: the relative error of z against the oracle's
FP64 AS241 (Wichura 1988) of the same midpoint probability stays below
7e-7 (measured: 6.3e-7, in the tail form near |z| = 2.9).  The device's
native log2 is checked the same way on the GPU
(test_gpu_parity.py::test_normal_mixed_device_accuracy)."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _checker(tmp, plant=None):
    exe = os.path.join(tmp, "normal_mixed_check")
    extra = [f"-DPE_PLANT_REFERENCE_SCALE={plant}"] if plant else []
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-pthread", "-ffp-contract=off"] + extra +
                          ["-I", os.path.join(ROOT, "rust-pricing-library_amd", "csrc"),
                           "-I", os.path.join(ROOT, "oracle"),
                           os.path.join(ROOT, "tests", "normal_mixed_check.cpp"), "-o", exe])
    return exe


def test_every_word_within_bound():
    with tempfile.TemporaryDirectory() as tmp:
        out = subprocess.check_output([_checker(tmp), "0", str(1 << 32), str(min(8, os.cpu_count() or 1))],
                                      text=True).split()
    err, word = float(out[0]), int(out[1])
    assert err < 7e-7, (err, word)


def test_checker_catches_a_planted_reference_error():
    """The same check against a reference scaled by 1 + 2e-6 (a mistake
    planted in the oracle on the CPU) reports an error above the bound."""
    with tempfile.TemporaryDirectory() as tmp:
        out = subprocess.check_output([_checker(tmp, plant="1.000002"), str((1 << 31) - 4096),
                                       str((1 << 31) + 4096), "2"], text=True).split()
    assert float(out[0]) > 1e-6

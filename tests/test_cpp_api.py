"""The C++ operation API (AFR / FRHIP, AFP / FPHIP, AFFT / FFTHIP over the C ABI), tested by the
C++ program tests/cpp/fr_fp_tests.cpp, which restates the reference's own gtest cases
(tests/operations/fourier_reconstruction/fr_tests.hpp: 12 plane cases with traverse-space known
answers and the zero / non-zero pattern of the volume; tests/operations/fourier_processing/
fp_tests.hpp: 9 crop / normalise / centre / filter / max-frequency / shift cases;
tests/operations/fourier_transformation/fft_tests.hpp: impulses forward / inverse / shifted and
FFT -> IFFT round trips, batched ones included)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "fr_fp_tests")


def run(filter_=None, timeout=120):
    cmd = [EXE] + ([filter_] if filter_ else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    return r


def test_host_cases():
    """Traverse-space known answers and payload bookkeeping: host code only."""
    r = run("Host")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "4 tests, 0 failures" in r.stdout


@pytest.mark.gpu
def test_all_cases_on_gpu():
    """All reference FR / FP / FFT cases through FRHIP / FPHIP / FFTHIP on the device."""
    r = run()
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failures" in r.stdout
    assert r.stdout.count("[ OK ]") >= 34

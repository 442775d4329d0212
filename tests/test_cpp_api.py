"""The C++ operation API (AFR / FRHIP, AFP / FPHIP, AFFT / FFTHIP over the C ABI), tested by the
C++ program tests/cpp/fr_fp_tests.cpp, which restates the reference's own gtest cases
(tests/operations/fourier_reconstruction/fr_tests.hpp: 12 plane cases with traverse-space known
answers and the zero / non-zero pattern of the volume; tests/operations/fourier_processing/
fp_tests.hpp: 9 crop / normalise / centre / filter / max-frequency / shift cases;
tests/operations/fourier_transformation/fft_tests.hpp: impulses forward / inverse / shifted and
FFT -> IFFT round trips, batched ones included), and by tests/cpp/fr_harness.cpp, the
reference's image-to-volume harness (examples/fourier_reconstruction/fr.cpp) on the same
classes, checked against numpy and the oracle."""
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


HARNESS = os.path.join(ROOT, "tests", "cpp", "_build", "fr_harness")


@pytest.mark.gpu
@pytest.mark.parametrize("args", [("32", "32", "20", "8", "C2"), ("48", "48", "9", "4", "D2")])
def test_reference_harness_on_drop_in_classes(args, tmp_path, oracle):
    """tests/cpp/fr_harness.cpp runs the reference harness's chunk loop (fr.cpp:9-76) on
    FFTHIP -> FPHIP -> FRHIP and dumps what it used and produced.  Its crops match numpy's FP64
    FFT put through the oracle's FP within 1e-6; its volume and weights equal, bit for bit, the
    oracle's insertion of those crops with its traverse spaces (exact strategy, the default)."""
    import numpy as np
    import pyoracle
    import frtools as T
    from umpalumpa_amd import fr as F
    r = subprocess.run([HARNESS, str(tmp_path), *args], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    meta = (tmp_path / "meta.txt").read_text().split("\n")
    X, Y, N, B, K = (int(v) for v in meta[0].split()[:5])
    chunks = [int(v) for v in meta[1:] if v.strip()]
    assert sum(chunks) == N and chunks[0] == min(B, N)
    imgs = np.fromfile(tmp_path / "images.f32", dtype=np.float32).reshape(N, Y, X)
    crops = np.fromfile(tmp_path / "crops.c64", dtype=np.complex64).reshape(N, Y, X // 2)
    spaces = np.fromfile(tmp_path / "spaces.bin", dtype=np.uint8).view(pyoracle.TS_DTYPE)
    assert len(spaces) == N * K
    vol = np.fromfile(tmp_path / "volume.c64", dtype=np.complex64).reshape(Y + 1, Y + 1, X + 1)
    wts = np.fromfile(tmp_path / "weights.f32", dtype=np.float32).reshape(Y + 1, Y + 1, X + 1)
    # crops: FFT + centre / normalise / shift
    spec = np.fft.rfft2(imgs.astype(np.float64)).astype(np.complex64)
    want = oracle.fp_execute(spec, [X // 2 + 1, Y, 1, N], [X, Y, 1, N], [X // 2, Y, 1, N],
                             norm_factor=np.float32(1.0 / (X * Y)), normalize=True,
                             center=True, shift=True).reshape(N, Y, X // 2)
    assert T.rel_l2(crops, want) <= 1e-6
    # insertion, batch by batch
    s = T.settings()
    os_ = T.oracle_settings(s)
    table = F.fill_blob_table(10000, F.fr_desc(X // 2, Y, B, X + 1, B * K, 10000), s)
    c = oracle.create_constants(X + 1, Y + 1, table.size, os_)
    rv = np.zeros_like(vol)
    rw = np.zeros_like(wts)
    i = 0
    for m in chunks:
        oracle.fr_execute(os_, rv, rw, np.ascontiguousarray(crops[i:i + m]),
                          spaces[i * K:(i + m) * K], table, c)
        i += m
    assert rw.max() > 0
    assert vol.tobytes() == rv.tobytes() and wts.tobytes() == rw.tobytes()

"""Finalization (umpa_fr_finalize): Hermitian completion and the FP64 weight division.  No
reference counterpart exists; it is pinned by its closed form, computed here in numpy FP64
(IEEE sums and a correctly rounded division, like the kernel's), and by the symmetry the
completed volume must have."""
import numpy as np
import pytest

import frtools as T
from umpalumpa_amd import fr as F

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def closed_form(v, w, hermitian):
    re, im, ww = v.real.astype(np.float64), v.imag.astype(np.float64), w.astype(np.float64)
    if hermitian:
        m = (slice(None, None, -1),) * 3          # voxel mirrored through the centre
        re = re + v.real[m].astype(np.float64)
        im = im - v.imag[m].astype(np.float64)
        ww = ww + w[m].astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        out = np.where(ww > 0, re / ww, 0.0).astype(np.float32) + \
            1j * np.where(ww > 0, im / ww, 0.0).astype(np.float32)
    return out.astype(np.complex64)


@pytest.mark.parametrize("hermitian", [True, False])
def test_finalize_matches_closed_form_on_inserted_volume(hermitian):
    s = T.settings()
    box, n = 48, 40
    fft, spaces, desc, table = T.problem(box, n, s, seed=77)
    v, w, _ = T.run_hip(s, fft, spaces, desc, table, box)
    dv = torch.from_numpy(v).cuda()
    dw = torch.from_numpy(w).cuda()
    out = F.finalize(dv, dw, hermitian=hermitian)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = closed_form(v, w, hermitian)
    assert got.tobytes() == want.tobytes()


def test_finalize_random_volume_and_zero_weights():
    rng = np.random.default_rng(5)
    side = 33
    v = (rng.standard_normal((side,) * 3) + 1j * rng.standard_normal((side,) * 3)).astype(
        np.complex64)
    w = rng.uniform(0, 2, (side,) * 3).astype(np.float32)
    w[rng.uniform(size=w.shape) < 0.3] = 0.0
    for herm in (True, False):
        got = F.finalize(torch.from_numpy(v).cuda(), torch.from_numpy(w).cuda(),
                         hermitian=herm)
        torch.cuda.synchronize()
        assert got.cpu().numpy().tobytes() == closed_form(v, w, herm).tobytes()


def test_completed_volume_is_hermitian():
    """out(-k) = conj(out(k)) for every voxel, and the centre is real."""
    s = T.settings()
    box, n = 40, 30
    fft, spaces, desc, table = T.problem(box, n, s, seed=8)
    v, w, _ = T.run_hip(s, fft, spaces, desc, table, box)
    out = F.finalize(torch.from_numpy(v).cuda(), torch.from_numpy(w).cuda()).cpu().numpy()
    mirrored = out[::-1, ::-1, ::-1]
    assert np.array_equal(out.real, mirrored.real)
    assert np.array_equal(out.imag, -mirrored.imag)
    c = box // 2
    assert out[c, c, c].imag == 0.0
    # the completion fills the skipped half-space: voxels with q.x < -r got no direct weight
    assert (w > 0).sum() < ((w + w[::-1, ::-1, ::-1]) > 0).sum()


def test_finalize_errors():
    v = torch.zeros((8, 8, 8), dtype=torch.complex64, device="cuda")
    w = torch.zeros((8, 8, 8), dtype=torch.float32, device="cuda")
    with pytest.raises(F._lib.UmpaError):
        F.finalize(v, w)          # even side: no centre voxel

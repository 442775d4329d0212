"""The harness's image-to-volume chunk (examples/fourier_reconstruction/fr.cpp:46-62) on the
device: FFT -> crop/centre/normalise/shift -> insertion, checked step by step.

  FFT     against numpy's double-precision rfft2 of the same images (relative L2 <= 1e-6)
  crop    the oracle's FP on the device spectrum, bit-exact (two operations), or within 1e-6
          of the oracle's FP on numpy's spectrum (fused FFT + crop)
  insert  the oracle's FR on the same crop with the same traverse spaces, bit-exact
          (exact strategy), over two chunks accumulated into one volume
"""
import numpy as np
import pytest

import frtools as T
import pyoracle
from umpalumpa_amd import fr as F
from umpalumpa_amd.pipeline import ImagePipeline

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def oracle_crop(oracle, spec, box):
    n = spec.shape[0]
    fin = [box // 2 + 1, box, 1, n]
    fout = [box // 2, box, 1, n]
    out = oracle.fp_execute(spec, fin, [box, box, 1, n], fout,
                            norm_factor=np.float32(1.0 / (box * box)),
                            normalize=True, center=True, shift=True)
    return out.reshape(n, box, box // 2)


@pytest.mark.parametrize("box,batch,group", [(64, 12, None), (48, 8, "C3")])
def test_pipeline_matches_oracle(box, batch, group, oracle):
    """The reference's two operations (hipFFT, then the FP crop): FFT within 1e-6 of numpy,
    crop and insertion bit-exact against the oracle."""
    s = T.settings()
    sym = None if group is None else F.symmetry_operators(group)
    pipe = ImagePipeline(box, batch, s, strategy=F.FR_EXACT, symmetries=sym, fused=False)
    rng = np.random.default_rng(box)
    vol = np.zeros((box + 1,) * 3, np.complex64)
    wts = np.zeros((box + 1,) * 3, np.float32)
    os_ = T.oracle_settings(s)
    table = pipe.table.cpu().numpy()
    c = oracle.create_constants(box + 1, box + 1, table.size, os_)
    for chunk, m in enumerate((batch, batch - 3)):        # a full chunk, then a short one
        imgs = rng.standard_normal((m, box, box)).astype(np.float32)
        rots = pyoracle.uniform_rotations(m, 100 + chunk)
        pipe.insert(torch.from_numpy(imgs).cuda(), rots)
        pipe.synchronize()
        spec = pipe._spec[:m].cpu().numpy()
        want = np.fft.rfft2(imgs.astype(np.float64))
        assert T.rel_l2(spec, want) <= 1e-6
        crop = pipe._crop[:m].cpu().numpy()
        ref_crop = oracle_crop(oracle, spec, box)
        assert crop.tobytes() == ref_crop.tobytes()
        spaces = F.generate_traverse_spaces(rots, box, s, sym)
        oracle.fr_execute(os_, vol, wts, ref_crop, spaces.view(pyoracle.TS_DTYPE), table, c)
    v = pipe.volume.cpu().numpy()
    w = pipe.weights.cpu().numpy()
    assert w.max() > 0
    assert v.tobytes() == vol.tobytes() and w.tobytes() == wts.tobytes()


@pytest.mark.parametrize("box,batch,group", [(64, 12, None), (128, 6, "D2")])
def test_fused_pipeline_matches_oracle(box, batch, group, oracle):
    """Fused FFT + crop: the crop within 1e-6 of the oracle's FP on numpy's FFT; the insertion
    of that crop bit-exact against the oracle's."""
    s = T.settings()
    sym = None if group is None else F.symmetry_operators(group)
    pipe = ImagePipeline(box, batch, s, strategy=F.FR_EXACT, symmetries=sym)
    assert pipe.fused
    rng = np.random.default_rng(box + 1)
    vol = np.zeros((box + 1,) * 3, np.complex64)
    wts = np.zeros((box + 1,) * 3, np.float32)
    os_ = T.oracle_settings(s)
    table = pipe.table.cpu().numpy()
    c = oracle.create_constants(box + 1, box + 1, table.size, os_)
    for chunk, m in enumerate((batch, batch - 2)):
        imgs = rng.standard_normal((m, box, box)).astype(np.float32)
        rots = pyoracle.uniform_rotations(m, 200 + chunk)
        pipe.insert(torch.from_numpy(imgs).cuda(), rots)
        pipe.synchronize()
        crop = pipe._crop[:m].cpu().numpy()
        spec = np.fft.rfft2(imgs.astype(np.float64)).astype(np.complex64)
        want = oracle_crop(oracle, spec, box)
        assert T.rel_l2(crop, want) <= 1e-6
        spaces = F.generate_traverse_spaces(rots, box, s, sym)
        oracle.fr_execute(os_, vol, wts, crop, spaces.view(pyoracle.TS_DTYPE), table, c)
    v = pipe.volume.cpu().numpy()
    w = pipe.weights.cpu().numpy()
    assert v.tobytes() == vol.tobytes() and w.tobytes() == wts.tobytes()


def test_pipeline_rejects_odd_box_and_oversized_batch():
    with pytest.raises(ValueError):
        ImagePipeline(33, 4, T.settings())
    pipe = ImagePipeline(32, 2, T.settings())
    with pytest.raises(ValueError):
        pipe.insert(torch.zeros((3, 32, 32), device="cuda"), np.tile(np.eye(3), (3, 1, 1)))

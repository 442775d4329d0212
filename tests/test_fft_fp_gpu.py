"""The fused FFT + FP operation (umpa_fft_fp_*): the forward R2C transform of 2-D images followed
by the FP step (crop, filter, normalise, centre, max frequency, shift) in one operation.

Reference: numpy's FP64 rfft2 of the same float32 images, rounded to complex64, then the
oracle's FP step (the pinned restatement of ScaleFFT2DCPU, fp_cpu_kernels.hpp:8-41) on it.
The transform differs from numpy's by float32 rounding only, so the bar is relative L2
<= 1e-6 of the whole output, with the zero pattern of the max-frequency crop exact.  The
fused result is also held against the two operations in sequence (umpa_fft then umpa_fp)."""
import itertools

import numpy as np
import pytest

from umpalumpa_amd import fr as F

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def reference(oracle, imgs, spatial_out, flags, filt, max_freq):
    n, ny, nx = imgs.shape
    spec = np.fft.rfft2(imgs.astype(np.float64)).astype(np.complex64)
    fin = [nx // 2 + 1, ny, 1, n]
    fout = [spatial_out[0] // 2 + 1, spatial_out[1], 1, n]
    apply_filter, normalize, center, crop, shift = flags
    out = oracle.fp_execute(spec, fin, [nx, ny, 1, n], fout, filt=filt,
                            norm_factor=np.float32(1.0 / (nx * ny)), max_freq=max_freq,
                            apply_filter=apply_filter, normalize=normalize, center=center,
                            crop_freq=crop, shift=shift)
    return out.reshape(n, fout[1], fout[0])


def run_fused(imgs, spatial_out, flags, filt, max_freq):
    n, ny, nx = imgs.shape
    apply_filter, normalize, center, crop, shift = flags
    op = F.FourierTransformProcessing(0)
    st = F.FPSettings(center, normalize, apply_filter, shift, crop, max_freq)
    assert op.init(F.fp_desc((nx, ny), spatial_out, n), st), op.last_status
    ox, oy = spatial_out[0] // 2 + 1, spatial_out[1]
    d_img = torch.from_numpy(imgs).cuda()
    d_out = torch.full((n, oy, ox), complex(7, 7), dtype=torch.complex64, device="cuda")
    d_f = torch.from_numpy(filt).cuda() if filt is not None else None
    assert op.execute(d_img, n, d_out, d_f), op.last_status
    op.synchronize()
    return d_out.cpu().numpy()


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("flags", list(itertools.product([0, 1], repeat=5)))
def test_flags_vs_numpy_and_oracle(flags, oracle):
    rng = np.random.default_rng(sum(f << i for i, f in enumerate(flags)))
    n, ny, nx = 5, 128, 128
    spatial_out = (126, 128)                     # the harness's crop (fr.cpp:176-178)
    imgs = rng.standard_normal((n, ny, nx)).astype(np.float32)
    filt = rng.standard_normal((spatial_out[1], spatial_out[0] // 2 + 1)).astype(np.float32)
    got = run_fused(imgs, spatial_out, flags, filt, 0.2)
    want = reference(oracle, imgs, spatial_out, flags, filt.ravel(), 0.2)
    assert np.array_equal(got == 0, want == 0)
    assert rel(got, want) <= 1e-6


@pytest.mark.parametrize("shape,spatial_out", [
    ((3, 4, 6), (6, 4)), ((2, 8, 9), (8, 8)), ((4, 64, 64), (62, 64)), ((3, 256, 256), (254, 256)),
    ((2, 256, 100), (60, 128)), ((2, 512, 512), (510, 512)), ((1, 1024, 1024), (1022, 1024)),
    ((3, 32, 48), (20, 16))])
def test_shapes_vs_numpy_and_oracle(shape, spatial_out, oracle):
    """Heights 4 .. 1024 (radix-4 only and with a radix-2 pass), odd and even widths, column and
    row crops, every FP step on."""
    rng = np.random.default_rng(shape[1] + shape[2])
    imgs = rng.standard_normal(shape).astype(np.float32)
    flags = (1, 1, 1, 1, 1)
    filt = rng.standard_normal((spatial_out[1], spatial_out[0] // 2 + 1)).astype(np.float32)
    got = run_fused(imgs, spatial_out, flags, filt, 0.3)
    want = reference(oracle, imgs, spatial_out, flags, filt.ravel(), 0.3)
    assert np.array_equal(got == 0, want == 0)
    assert rel(got, want) <= 1e-6


def test_fused_equals_separate_operations_within_rounding():
    rng = np.random.default_rng(4)
    n, box = 20, 256
    imgs = torch.from_numpy(rng.standard_normal((n, box, box)).astype(np.float32)).cuda()
    desc = F.fp_desc((box, box), (box - 2, box), n)
    st = F.FPSettings(1, 1, 0, 1, 0, 0.0)
    fused = F.FourierTransformProcessing(0)
    assert fused.init(desc, st)
    a = torch.zeros((n, box, box // 2), dtype=torch.complex64, device="cuda")
    assert fused.execute(imgs, n, a)
    fft, fp = F.FourierTransform(0), F.FourierProcessing(0)
    assert fft.init((box, box, 1), n) and fp.init(desc, st)
    spec = torch.zeros((n, box, box // 2 + 1), dtype=torch.complex64, device="cuda")
    b = torch.zeros_like(a)
    assert fft.execute(imgs, spec, n) and fp.execute(spec, n, b, n)
    torch.cuda.synchronize()
    assert rel(a.cpu().numpy(), b.cpu().numpy()) <= 1e-6
    # fewer images than at Init (a plan per count), twice: the same bits each time
    prev = None
    for m in (7, 7):
        c = torch.zeros_like(a)
        assert fused.execute(imgs[:m], m, c)
        torch.cuda.synchronize()
        assert not c[m:].any()
        assert rel(c[:m].cpu().numpy(), a[:m].cpu().numpy()) <= 1e-6
        if prev is not None:
            assert torch.equal(c, prev)
        prev = c


def test_fused_errors():
    from umpalumpa_amd import _lib
    op = F.FourierTransformProcessing(0)
    img = torch.zeros((2, 64, 64), device="cuda")
    out = torch.zeros((2, 64, 32), dtype=torch.complex64, device="cuda")
    assert not op.execute(img, 2, out)
    assert op.last_status == _lib.UMPA_ERR_NOT_INITIALIZED
    st = F.FPSettings(1, 1, 0, 1, 0, 0.0)
    assert not op.init(F.fp_desc((64, 96), (62, 96), 2), st)     # height not a power of two
    assert op.last_status == _lib.UMPA_ERR_UNSUPPORTED
    assert not op.init(F.fp_desc((64, 2048), (62, 2048), 2), st)  # taller than 1024
    assert op.init(F.fp_desc((64, 64), (62, 64), 2), st)
    assert not op.execute(img, 3, out)                            # more images than at Init
    assert op.last_status == _lib.UMPA_ERR_NOT_EQUIVALENT
    assert not op.execute(None, 2, out)
    assert op.last_status == _lib.UMPA_ERR_REJECTED
    filt_st = F.FPSettings(0, 0, 1, 0, 0, 0.0)
    assert op.init(F.fp_desc((64, 64), (62, 64), 2), filt_st)
    assert not op.execute(img, 2, out)                            # filter missing
    assert op.last_status == _lib.UMPA_ERR_INVALID_ARG

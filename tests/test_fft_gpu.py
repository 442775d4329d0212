"""The AFFT operation on the device (umpa_fft_*): the reference's own FFT test cases
(tests/operations/fourier_transformation/fft_tests.hpp + afft_common.hpp: impulse at the
origin forward and inverse, shifted impulse, FFT -> IFFT round trips in batches), and
agreement with numpy's FFT in float and double precision."""
import numpy as np
import pytest

from umpalumpa_amd import fr as F

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def freq_shape(shape):
    x, y, z = shape
    return (z, y, x // 2 + 1) if z > 1 else (y, x // 2 + 1)


def spatial_tensor(shape, n, dtype=torch.float32):
    x, y, z = shape
    return torch.zeros((n, z, y, x) if z > 1 else (n, y, x), dtype=dtype, device="cuda")


def complex_tensor(shape, n, dtype=torch.complex64):
    return torch.zeros((n,) + freq_shape(shape), dtype=dtype, device="cuda")


def test_impulse_origin_forward():
    shape, n = (5, 5, 1), 1
    sp = spatial_tensor(shape, n)
    sp.view(n, -1)[:, 0] = 1.0
    fr = complex_tensor(shape, n)
    op = F.FourierTransform(0)
    assert op.init(shape, n, F.FFT_FORWARD)
    assert op.execute(sp, fr, n)
    op.synchronize()
    out = fr.cpu().numpy()
    assert np.allclose(out.real, 1.0, atol=1e-5) and np.allclose(out.imag, 0.0, atol=1e-5)


def test_impulse_origin_inverse():
    shape, n = (5, 5, 1), 1
    fr = complex_tensor(shape, n)
    fr.fill_(1.0)
    sp = spatial_tensor(shape, n)
    op = F.FourierTransform(0)
    assert op.init(shape, n, F.FFT_INVERSE)
    assert op.execute(fr, sp, n)
    op.synchronize()
    out = sp.cpu().numpy().reshape(-1) / 25.0     # unnormalised inverse
    want = np.zeros_like(out)
    want[0] = 1.0
    assert np.allclose(out, want, atol=1e-5)


def test_impulse_shifted_forward():
    shape, n = (5, 5, 1), 1
    sp = spatial_tensor(shape, n)
    sp.view(n, -1)[:, 1] = 1.0
    fr = complex_tensor(shape, n)
    op = F.FourierTransform(0)
    assert op.init(shape, n, F.FFT_FORWARD)
    assert op.execute(sp, fr, n)
    op.synchronize()
    assert np.allclose(np.abs(fr.cpu().numpy()), 1.0, atol=1e-5)


@pytest.mark.parametrize("shape,n,batch", [((5, 5, 1), 15, 0), ((5, 5, 1), 50, 10),
                                           ((300, 200, 32), 5, 0), ((64, 1, 1), 7, 3)])
def test_fft_ifft_round_trip(shape, n, batch):
    """FFT then inverse FFT gives the input times the image size (afft_common.hpp testFFTIFFT),
    also when Execute runs over fewer images than Init planned (batches)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    sp = torch.randn(spatial_tensor(shape, n).shape, generator=g, device="cuda")
    fr = complex_tensor(shape, n)
    back = spatial_tensor(shape, n)
    fwd, inv = F.FourierTransform(0), F.FourierTransform(0)
    assert fwd.init(shape, n, F.FFT_FORWARD) and inv.init(shape, n, F.FFT_INVERSE)
    step = batch or n
    for a in range(0, n, step):
        b = min(n, a + step)
        assert fwd.execute(sp[a:b], fr[a:b], b - a)
        assert inv.execute(fr[a:b], back[a:b], b - a)
    inv.synchronize()
    size = shape[0] * shape[1] * shape[2]
    assert torch.allclose(back / size, sp, atol=1e-4)


@pytest.mark.parametrize("precision", [F.FFT_F32, F.FFT_F64])
def test_forward_matches_numpy(precision):
    rng = np.random.default_rng(3)
    shape, n = (128, 96, 1), 6
    dt = np.float32 if precision == F.FFT_F32 else np.float64
    img = rng.standard_normal((n, 96, 128)).astype(dt)
    sp = torch.from_numpy(img).cuda()
    fr = torch.zeros((n, 96, 65), dtype=torch.complex64 if precision == F.FFT_F32
                     else torch.complex128, device="cuda")
    op = F.FourierTransform(0)
    assert op.init(shape, n, F.FFT_FORWARD, precision)
    assert op.execute(sp, fr, n)
    op.synchronize()
    want = np.fft.rfft2(img.astype(np.float64))
    got = fr.cpu().numpy()
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < (1e-6 if precision == F.FFT_F32 else 1e-13), err


def test_fft_errors():
    op = F.FourierTransform(0)
    sp = spatial_tensor((8, 8, 1), 2)
    fr = complex_tensor((8, 8, 1), 2)
    assert not op.execute(sp, fr, 2)                       # before Init
    assert op.last_status == F._lib.UMPA_ERR_NOT_INITIALIZED
    assert op.init((8, 8, 1), 2)
    assert not op.execute(sp, fr, 3)                       # more images than at Init
    assert not op.execute(sp, sp, 2)                       # in place
    assert op.last_status == F._lib.UMPA_ERR_UNSUPPORTED


def test_plans_for_many_counts_stay_bounded_and_correct():
    """An Execute over each of 1 .. 13 images (more distinct counts than the eight plans an
    operation keeps), then the Init count again: every result matches numpy."""
    shape, n = (24, 16, 1), 13
    rng = np.random.default_rng(7)
    imgs = rng.standard_normal((n, 16, 24)).astype(np.float32)
    d_img = torch.from_numpy(imgs).cuda()
    want = np.fft.rfft2(imgs.astype(np.float64))
    op = F.FourierTransform(0)
    assert op.init(shape, n, F.FFT_FORWARD)
    for m in list(range(1, n + 1)) + [n, 3]:
        out = complex_tensor(shape, n)
        assert op.execute(d_img, out, m)
        op.synchronize()
        got = out.cpu().numpy()
        assert not got[m:].any()
        err = np.linalg.norm(got[:m] - want[:m]) / np.linalg.norm(want[:m])
        assert err <= 1e-6, (m, err)
    fused = F.FourierTransformProcessing(0)
    st = F.FPSettings(0, 0, 0, 0, 0, 0.0)
    assert fused.init(F.fp_desc((24, 16), (24, 16), n), st)
    for m in list(range(1, n + 1)) + [n]:
        out = torch.zeros((n, 16, 13), dtype=torch.complex64, device="cuda")
        assert fused.execute(d_img, m, out)
        fused.synchronize()
        got = out.cpu().numpy()
        err = np.linalg.norm(got[:m] - want[:m]) / np.linalg.norm(want[:m])
        assert err <= 1e-6, (m, err)

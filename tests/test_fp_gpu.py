"""HIP FP crop/filter/normalize/center/shift vs the reference's golden values and the oracle."""
import itertools

import numpy as np
import pytest

from umpalumpa_amd import fr as F

pytestmark = pytest.mark.gpu

FP_CASES = ["ImageCropping", "ImageNoCropping", "ImageNormalize", "ImageCentering",
            "ImageCropNormalizeCenter", "Filtering", "MaxFreq", "ShiftEven", "ShiftOdd"]


def run_fp(x, sizes, flags, scalars, filt):
    import torch
    dev = torch.device("cuda:0")
    fin, spat, fout = [[int(v) for v in row] for row in sizes]
    apply_filter, normalize, center, crop, shift = [int(v) for v in flags]
    norm, max_freq = [float(v) for v in scalars]
    desc = F.FPDesc(F.Size(*fin), F.Size(*spat), F.Size(*fout),
                    int(round(1.0 / norm)) if norm > 0 else spat[0] * spat[1] * spat[2])
    st = F.FPSettings(center, normalize, apply_filter, shift, crop, max_freq)
    op = F.FourierProcessing(0)
    assert op.init(desc, st), op.last_status
    d_in = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    d_out = torch.zeros(int(np.prod(fout)), dtype=torch.complex64, device=dev)
    d_f = torch.from_numpy(np.ascontiguousarray(filt)).to(dev) if filt is not None else None
    assert op.execute(d_in, fin[3], d_out, fout[3], d_f), op.last_status
    op.synchronize()
    return d_out.cpu().numpy()


@pytest.mark.parametrize("case", FP_CASES)
def test_reference_fp_cases_bit_exact(case, golden):
    g = golden("fp_ref_cases.npz")
    y = run_fp(g[case + "/in"], g[case + "/sizes"], g[case + "/flags"], g[case + "/scalars"],
               g[case + "/filter"])
    assert y.tobytes() == g[case + "/out"].tobytes()


@pytest.mark.parametrize("flags", list(itertools.product([0, 1], repeat=5)))
def test_random_flags_vs_oracle(flags, oracle):
    rng = np.random.default_rng(sum(f << i for i, f in enumerate(flags)))
    spatial_in, spatial_out, n = (128, 128), (126, 128), 7   # the caller's crop (fr.cpp:176-178)
    fin = [spatial_in[0] // 2 + 1, spatial_in[1], 1, n]
    fout = [spatial_out[0] // 2 + 1, spatial_out[1], 1, n]
    x = (rng.standard_normal(int(np.prod(fin))) +
         1j * rng.standard_normal(int(np.prod(fin)))).astype(np.complex64)
    filt = rng.standard_normal(fout[0] * fout[1]).astype(np.float32)
    norm = 1.0 / (spatial_in[0] * spatial_in[1])
    sizes = np.array([fin, [spatial_in[0], spatial_in[1], 1, n], fout], np.uint64)
    y = run_fp(x, sizes, flags, np.array([norm, 0.2], np.float32), filt)
    ref = oracle.fp_execute(x, fin, sizes[1], fout, filt=filt, norm_factor=np.float32(norm),
                            max_freq=0.2, apply_filter=flags[0], normalize=flags[1],
                            center=flags[2], crop_freq=flags[3], shift=flags[4])
    assert y.tobytes() == ref.tobytes()


def test_fp_errors():
    from umpalumpa_amd import _lib
    op = F.FourierProcessing(0)
    assert not op.execute(None, 1, None, 1)
    assert op.last_status == _lib.UMPA_ERR_NOT_INITIALIZED
    st = F.FPSettings(0, 0, 0, 0, 1, 0.7)   # maxFreq above 1/2 (settings.hpp:38-43)
    assert not op.init(F.fp_desc((10, 10), (10, 10), 1), st)


def test_fp_python_mirror_refuses_short_outputs():
    """An output or filter shorter than the descriptor says is refused before the C ABI."""
    import torch
    op = F.FourierProcessing(0)
    st = F.FPSettings(0, 0, 1, 0, 0, 0.0)
    assert op.init(F.fp_desc((16, 16), (14, 16), 3), st)
    inp = torch.zeros((3, 16, 9), dtype=torch.complex64, device="cuda")
    out = torch.zeros((3, 16, 8), dtype=torch.complex64, device="cuda")
    filt = torch.ones((16, 8), dtype=torch.float32, device="cuda")
    assert op.execute(inp, 3, out, 3, filt)
    with pytest.raises(ValueError, match="output"):
        op.execute(inp, 3, out[:2], 3, filt)
    with pytest.raises(ValueError, match="filter"):
        op.execute(inp, 3, out, 3, filt[:8])
    with pytest.raises(ValueError, match="input"):
        op.execute(inp[:, :, :4].contiguous(), 3, out, 3, filt)

"""The allocation-free insertion entry point (umpa_fr_execute_noalloc, C ABI) against the
operation handle's umpa_fr_execute / umpa_fr_execute_f16 on the same inputs: bit-identical
under the exact strategy (and so equal to the oracle), within 1e-6 relative L2 under fast math
(bricks split over wavefronts merge with float atomics, in either run's order).  Calls come
in pairs, the second on a smaller volume with fewer slices: the static scratch the calls
share keeps the first call's work items and counter, so the second must re-fill the items and
clear the counter."""
import ctypes as C

import numpy as np
import pytest

import frtools as T
from umpalumpa_amd import _lib
from umpalumpa_amd import fr as F

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _prepare(box, n, half, seed):
    s = T.settings()
    fft, spaces, desc, table = T.problem(box, n, s, seed=seed)
    hp = None
    if half:
        h, sc, fft = T.quantize_f16(fft)
        hp = (h, sc)
    return s, fft, spaces, desc, table, hp


def _noalloc(s, desc, strategy, fft, spaces, table, hp, box):
    """One umpa_fr_execute_noalloc call on cuda:0's current stream; returns (status, v, w)."""
    dev = torch.device("cuda:0")
    if hp is None:
        d_fft = torch.from_numpy(np.ascontiguousarray(fft)).to(dev)
        d_sc = None
    else:
        d_fft = torch.from_numpy(np.ascontiguousarray(hp[0])).to(dev)
        d_sc = torch.from_numpy(np.ascontiguousarray(hp[1])).to(dev)
    d_sp = torch.from_numpy(np.ascontiguousarray(spaces).view(np.uint8)).to(dev)
    d_tab = torch.from_numpy(table).to(dev)
    vol = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=dev)
    wts = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    st = _lib.load().umpa_fr_execute_noalloc(
        0, C.c_void_p(stream.cuda_stream), C.byref(desc), C.byref(s), strategy,
        C.c_void_p(d_fft.data_ptr()), None if d_sc is None else C.c_void_p(d_sc.data_ptr()),
        d_fft.shape[0], C.c_void_p(vol.data_ptr()), C.c_void_p(wts.data_ptr()),
        C.c_void_p(d_sp.data_ptr()), len(spaces), C.c_void_p(d_tab.data_ptr()))
    torch.cuda.synchronize()
    return st, vol.cpu().numpy(), wts.cpu().numpy()


@pytest.mark.parametrize("strategy,half", [(F.FR_EXACT, False), (F.FR_FAST_MATH, False),
                                           (F.FR_FAST_MATH, True)])
def test_noalloc_matches_operation_twice(strategy, half):
    for box, n, seed in ((128, 48, 21), (48, 12, 22)):
        s, fft, spaces, desc, table, hp = _prepare(box, n, half, seed)
        st, v, w = _noalloc(s, desc, strategy, fft, spaces, table, hp, box)
        assert st == _lib.UMPA_OK, st
        rv, rw, _ = T.run_hip(s, fft, spaces, desc, table, box, strategy=strategy, half=hp)
        assert (w != 0).sum() > 1000
        if strategy == F.FR_EXACT:
            assert v.tobytes() == rv.tobytes() and w.tobytes() == rw.tobytes(), box
            ov, ow, _ = T.run_oracle(s, fft, spaces, table, box)
            assert v.tobytes() == ov.tobytes() and w.tobytes() == ow.tobytes(), box
        else:
            assert T.rel_l2(v, rv) <= 1e-6 and T.rel_l2(w, rw) <= 1e-6, box
            ov, ow, _ = T.run_oracle(s, fft, spaces, table, box)
            assert T.rel_l2(v, ov) <= 1e-5 and T.rel_l2(w, ow) <= 1e-5, box


def test_noalloc_refuses_volumes_beyond_its_scratch():
    """Volumes above 513^3 do not fit the 8 MiB scratch: UMPA_ERR_UNSUPPORTED, outputs
    untouched; 513^3 itself is accepted (checked on a call with no spaces)."""
    s = T.settings()
    lib = _lib.load()
    dev = torch.device("cuda:0")
    fft = torch.zeros((1, 640, 321), dtype=torch.complex64, device=dev)
    tab = torch.zeros(10000, dtype=torch.float32, device=dev)
    sp = torch.zeros(112, dtype=torch.uint8, device=dev)
    for side, want in ((641, _lib.UMPA_ERR_UNSUPPORTED), (513, _lib.UMPA_OK)):
        desc = F.fr_desc(321, 640, 1, side, 0, 10000)
        vol = torch.zeros(8, dtype=torch.complex64, device=dev)   # never written
        wts = torch.full((8,), 7.0, dtype=torch.float32, device=dev)
        st = lib.umpa_fr_execute_noalloc(
            0, None, C.byref(desc), C.byref(s), F.FR_EXACT, C.c_void_p(fft.data_ptr()), None, 1,
            C.c_void_p(vol.data_ptr()), C.c_void_p(wts.data_ptr()), C.c_void_p(sp.data_ptr()),
            0, C.c_void_p(tab.data_ptr()))
        assert st == want, (side, st)
        torch.cuda.synchronize()
        assert float(wts.sum()) == 56.0

"""Shared helpers of the FR tests: synthetic problems, a HIP run through the C ABI, an oracle
run on the same inputs, and error measures."""
from __future__ import annotations

import numpy as np

import pyoracle
from umpalumpa_amd import fr as F


def settings(interpolation=F.FR_LOOKUP, typ=F.FR_PRECISE, alpha=15.0, radius=1.9, order=0):
    return F.FRSettings(interpolation, order, typ, alpha, radius)


def oracle_settings(s):
    return pyoracle.Settings(s.interpolation, s.order, s.type, s.alpha, s.blobRadius)


def synthetic_slices(box, n, seed=42, envelope=True):
    """Hermitian half-plane slices (box/2+1) x box: complex Gaussian, E|z|^2 = 1, times a
    radial 1/(1+k^2) envelope with k measured from the DC row box/2 (shifted layout)."""
    rng = np.random.default_rng(seed)
    hx, hy = box // 2 + 1, box
    z = (rng.standard_normal((n, hy, hx)) + 1j * rng.standard_normal((n, hy, hx))) / np.sqrt(2)
    if envelope:
        kx = np.arange(hx)[None, :]
        ky = np.arange(hy)[:, None] - box // 2
        z = z / (1.0 + kx ** 2 + ky ** 2)
    return z.astype(np.complex64)


def problem(box, n, s, seed=42, symmetries=None, envelope=True):
    fft = synthetic_slices(box, n, seed, envelope)
    rots = pyoracle.uniform_rotations(n, seed + 1)
    spaces = F.traverse_spaces(rots, box, s, symmetries)
    desc = F.fr_desc(box // 2 + 1, box, n, box + 1, len(spaces),
                     10000 if s.interpolation == F.FR_LOOKUP else 0)
    table = F.fill_blob_table(10000, desc, s) if s.interpolation == F.FR_LOOKUP else None
    return fft, spaces, desc, table


def run_oracle(s, fft, spaces, table, box, volume=None, weights=None):
    O = pyoracle.Oracle()
    os_ = oracle_settings(s)
    c = O.create_constants(box + 1, box + 1, 0 if table is None else table.size, os_)
    v = np.zeros((box + 1,) * 3, np.complex64) if volume is None else volume.copy()
    w = np.zeros((box + 1,) * 3, np.float32) if weights is None else weights.copy()
    upd = O.fr_execute(os_, v, w, fft, spaces.view(pyoracle.TS_DTYPE), table, c)
    return v, w, upd


def run_oracle_f64(s, fft, spaces, table, box):
    O = pyoracle.Oracle()
    os_ = oracle_settings(s)
    c = O.create_constants(box + 1, box + 1, table.size, os_)
    v = np.zeros((box + 1,) * 3, np.complex128)
    w = np.zeros((box + 1,) * 3, np.float64)
    O.fr_execute_f64(os_, v, w, fft, spaces.view(pyoracle.TS_DTYPE), table, c)
    return v, w


def run_hip(s, fft, spaces, desc, table, box, strategy=F.FR_EXACT, chunks=1, count=False,
            half=None, volume=None, weights=None, sched=None):
    """Insert on cuda:0; `chunks` splits the spaces over several executes.  `half` is a pair
    (float16 [n, hy, hx, 2] array, float32 scales) for the FP16 path; `sched` forces the
    schedule knobs (a SchedParams) of the strategy."""
    import torch
    dev = torch.device("cuda:0")
    op = F.FourierReconstruction(0)
    assert op.init(desc, s), op.last_status
    op.set_strategy(strategy)
    if sched is not None:
        op.set_sched_params(strategy, sched)
    op.set_count_updates(count)
    if volume is None:
        vol = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=dev)
        wts = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=dev)
    else:
        vol = torch.from_numpy(volume).to(dev)
        wts = torch.from_numpy(weights).to(dev)
    if half is None:
        d_fft = torch.from_numpy(np.ascontiguousarray(fft)).to(dev)
        d_scales = None
    else:
        d_fft = torch.from_numpy(np.ascontiguousarray(half[0])).to(dev)
        d_scales = torch.from_numpy(np.ascontiguousarray(half[1])).to(dev)
    d_tab = None if table is None else torch.from_numpy(table).to(dev)
    updates = 0
    bounds = np.linspace(0, len(spaces), chunks + 1).astype(int)
    for a, b in zip(bounds[:-1], bounds[1:]):
        d_sp = torch.from_numpy(np.ascontiguousarray(spaces[a:b]).view(np.uint8)).to(dev)
        ok = op.execute(d_fft, vol, wts, d_sp, d_tab, n_spaces=b - a, scales=d_scales)
        assert ok, op.last_status
        if count:
            updates += op.last_updates()
    op.synchronize()
    return vol.cpu().numpy(), wts.cpu().numpy(), updates


def quantize_f16(fft):
    """Per-slice max-abs scaling to FP16: store half(v / s), s = max(|re|, |im|)."""
    n = fft.shape[0]
    ri = np.stack([fft.real, fft.imag], axis=-1).astype(np.float32)
    s = np.abs(ri).reshape(n, -1).max(axis=1)
    s[s == 0] = 1.0
    h = (ri / s[:, None, None, None]).astype(np.float16)
    deq = (h.astype(np.float32) * s[:, None, None, None]).astype(np.float32)
    return h, s.astype(np.float32), (deq[..., 0] + 1j * deq[..., 1]).astype(np.complex64)


def rel_l2_big(a, b, rows=16):
    """rel_l2 over large volumes, a few z planes at a time (no full-size temporaries)."""
    num = den = 0.0
    for z in range(0, a.shape[0], rows):
        x = np.asarray(a[z:z + rows], dtype=np.complex128)
        y = np.asarray(b[z:z + rows], dtype=np.complex128)
        num += float(np.sum(np.abs(x - y) ** 2))
        den += float(np.sum(np.abs(y) ** 2))
    return (num / den) ** 0.5 if den > 0 else num ** 0.5


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))

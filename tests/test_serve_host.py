"""Host-side bookkeeping of the serving loop (no GPU): a job's ring of chunk slots, its
footprint, the window of spaces one step may cover and the fast-lane classification."""
import types

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from umpalumpa_amd import serve as S


def job(box=64, n=1300, sym="C4", half=True):
    hy, hx = box, box // 2 + 1
    if half:
        data = (torch.zeros((512, hy, hx, 2), dtype=torch.float16), torch.ones(512))
    else:
        data = torch.zeros((512, hy, hx), dtype=torch.complex64)
    return S.Job(0, box, n, 1.9, sym, np.zeros((n, 3, 3)), lambda c: data)


def test_ring_slots_and_footprint():
    j = job()
    slice_bytes = 64 * 33 * 2 * 2 + 4
    assert j.slice_bytes() == slice_bytes
    assert j.n_spaces() == 4 * 1300
    # as many 512-slice slots as fit the ring budget, 1 to 8, no more than the job's chunks
    assert j.ring_slots(512, 512e6) == 3                 # 1,300 slices: 3 chunks
    assert j.ring_slots(512, 2 * 512 * slice_bytes) == 2
    assert j.ring_slots(512, 1) == 1
    big = job(n=50000)
    assert big.ring_slots(512, 512e6) == 8
    f = j.footprint(512, 2 * 512 * slice_bytes)
    assert f == 12 * 65 ** 3 + 112 * 4 * 1300 + 40000 + 2 * 512 * slice_bytes
    assert job(half=False).slice_bytes() == 64 * 33 * 8


def test_window_covers_the_ring_from_the_cursor():
    srv = types.SimpleNamespace(chunk=512)
    j = job(n=3000, sym="C1")          # 6 chunks of 512 spaces, the last of 440
    j._slots = 2
    win = lambda: S.Server._window_end(srv, j)
    assert win() == 1024               # chunks 0 and 1
    j._cursor = 700                    # inside chunk 1: chunks 1 and 2
    assert win() == 1536
    j._cursor = 2600                   # chunk 5, the last: the job's end
    assert win() == 3000
    assert j.remaining_cost() == (3000 - 2600) * 65 ** 2


def test_fast_lane_classification():
    srv = types.SimpleNamespace(_fast=object(), _rate=None, tick_work=1e8, fast_seconds=0.02,
                                chunk=512)
    small, large = job(n=1000, sym="C1"), job(box=256, n=20000)
    assert S.Server._is_small(srv, small) and not S.Server._is_small(srv, large)
    srv._rate = 1e9                    # work units per second
    assert S.Server._is_small(srv, small)             # 1000 x 65^2 = 4.2e6: 4 ms
    assert not S.Server._is_small(srv, job(n=8000, sym="D2"))   # 32,000 spaces: 135 ms
    srv._fast = None
    assert not S.Server._is_small(srv, small)

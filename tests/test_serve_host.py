"""Host-side bookkeeping of the serving loop (no GPU): a job's ring of chunk slots, its
footprint, the window of spaces one step may cover, the fast-lane classification and the
retirement of a lane's ticks."""
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


class _Event:
    def __init__(self, done):
        self.done = done

    def query(self):
        return self.done


def test_lane_ticks_retire_in_order_and_count_updates():
    """A lane's ticks leave its queue in submission order once their thread has enqueued them
    and their event has completed; their voxel-updates are added once."""
    import concurrent.futures
    srv = types.SimpleNamespace(total_updates=0, _done=S.Server._done)
    lane = types.SimpleNamespace(inflight=__import__("collections").deque())

    def tick(done, n):
        f = concurrent.futures.Future()
        if done is not None:
            f.set_result((_Event(done), n))
        return f

    first, second, third = tick(True, 5), tick(False, 7), tick(True, 11)
    for f, w in ((first, 1.0), (second, 2.0), (third, 4.0)):
        lane.inflight.append((f, w))
    # the second tick's launches have not run: the third waits behind it
    assert S.Server._retire(srv, lane) == 1.0
    assert srv.total_updates == 5 and len(lane.inflight) == 2
    second.result()[0].done = True
    assert S.Server._retire(srv, lane) == 6.0
    assert srv.total_updates == 23 and not lane.inflight
    # a tick whose thread has not enqueued it yet is not done
    pending = tick(None, 0)
    assert not S.Server._done(pending)
    # a failure on the lane's thread is raised in the serving loop
    bad = concurrent.futures.Future()
    bad.set_exception(RuntimeError("umpa_fr_batch_step failed"))
    lane.inflight.append((bad, 1.0))
    with pytest.raises(RuntimeError, match="batch_step"):
        S.Server._retire(srv, lane)

"""The autotuner (umpa_fr_autotune, umpa_fr_batch_autotune): it times candidate schedules on
scratch volumes, keeps the fastest, persists it, and never changes results — the exact
strategy stays bit-identical to the oracle under every schedule it can pick."""
import os

import numpy as np
import pytest

import frtools as T
from umpalumpa_amd import fr as F

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture
def tuning_dir(tmp_path, monkeypatch):
    monkeypatch.setenv("UMPA_TUNING_DIR", str(tmp_path))
    return tmp_path


def _inputs(box, n, s, seed=3):
    dev = torch.device("cuda:0")
    fft, spaces, desc, table = T.problem(box, n, s, seed=seed)
    return (fft, spaces, desc, table, torch.from_numpy(fft).to(dev),
            torch.from_numpy(spaces.view(np.uint8)).to(dev), torch.from_numpy(table).to(dev))


def test_op_autotune_keeps_results_and_persists(tuning_dir):
    s = T.settings()
    box, n = 48, 60
    fft, spaces, desc, table, d_fft, d_sp, d_tab = _inputs(box, n, s)
    dev = torch.device("cuda:0")
    op = F.FourierReconstruction(0)
    assert op.init(desc, s)
    vol = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=dev)
    wts = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=dev)
    p = op.autotune(d_fft, d_sp, d_tab, use_cache=False)
    assert p.depthScale in (0.5, 1.0, 2.0) and p.maxParts == 1   # exact never splits
    assert op.sched_params().as_tuple() == p.as_tuple()
    # the caller's volumes were not touched by the tuning runs
    assert float(wts.abs().sum()) == 0.0
    assert op.execute(d_fft, vol, wts, d_sp, d_tab)
    op.synchronize()
    rv, rw, _ = T.run_oracle(s, fft, spaces, table, box)
    assert vol.cpu().numpy().tobytes() == rv.tobytes()
    assert wts.cpu().numpy().tobytes() == rw.tobytes()
    # persisted, and reused from the cache
    lines = (tuning_dir / "fr_tuning.txt").read_text().splitlines()
    assert len(lines) == 1 and lines[0].startswith("k2:op:")
    q = op.autotune(d_fft, d_sp, d_tab, use_cache=True)
    assert q.as_tuple() == p.as_tuple()
    assert len((tuning_dir / "fr_tuning.txt").read_text().splitlines()) == 1


def test_damaged_cache_entries_are_ignored(tuning_dir):
    """The last valid line of a key wins; lines with knobs outside their ranges (a damaged or
    hand-edited file) are skipped."""
    s = T.settings()
    box, n = 48, 30
    _, _, desc, _, d_fft, d_sp, d_tab = _inputs(box, n, s, seed=4)
    op = F.FourierReconstruction(0)
    assert op.init(desc, s)
    p = op.autotune(d_fft, d_sp, d_tab, use_cache=False)
    path = tuning_dir / "fr_tuning.txt"
    key = path.read_text().split()[0]
    with open(path, "a") as f:
        for bad in ("nan 1 1 0.5 0 0 0", "-2 1 1 0.5 0 0 0", "1 1 999 0.5 0 0 0",
                    "1 1 1 0 0 0 0", "1 1 1 inf 0 0 0", "1 1 1 0.5 7 0 0", "1 1 1 0.5 0 2 0",
                    "1 1 1 0.5 0 0 -1", "1 2 1 0.5 0 0 0", "x y",
                    "2 0 1 0.5"):   # the format without kernel knobs
            f.write(f"{key} {bad} 1.0\n")
    assert op.autotune(d_fft, d_sp, d_tab, use_cache=True).as_tuple() == p.as_tuple()
    with open(path, "a") as f:
        f.write(f"{key} 2 0 1 0.5 1 1 0 0.1\n")
    assert op.autotune(d_fft, d_sp, d_tab, use_cache=True).as_tuple() == \
        (2.0, 0, 1, 0.5, F.FR_KERNEL_PAIR, 1, 0)


def test_op_autotune_fast_math_within_tolerance(tuning_dir):
    s = T.settings()
    box, n = 64, 80
    fft, spaces, desc, table, d_fft, d_sp, d_tab = _inputs(box, n, s, seed=9)
    dev = torch.device("cuda:0")
    op = F.FourierReconstruction(0)
    assert op.init(desc, s)
    op.set_strategy(F.FR_FAST_MATH)
    op.autotune(d_fft, d_sp, d_tab, use_cache=False)
    vol = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=dev)
    wts = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=dev)
    assert op.execute(d_fft, vol, wts, d_sp, d_tab)
    op.synchronize()
    rv, rw, _ = T.run_oracle(s, fft, spaces, table, box)
    assert T.rel_l2(vol.cpu().numpy(), rv) <= 1e-5
    assert T.rel_l2(wts.cpu().numpy(), rw) <= 1e-5


@pytest.mark.parametrize("params", [(0.5, 0, 1, 0.5), (2.0, 1, 16, 0.25)])
def test_every_exact_schedule_is_bit_exact(params):
    """Schedules the tuner may pick, forced on a batch: the exact result does not move."""
    s = T.settings()
    dev = torch.device("cuda:0")
    jobs, ref = [], []
    for k, (box, n) in enumerate([(40, 30), (64, 20)]):
        fft, spaces, desc, table, d_fft, d_sp, d_tab = _inputs(box, n, s, seed=20 + k)
        vol = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=dev)
        wts = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=dev)
        jobs.append(F.BatchJob(d_fft, vol, wts, d_sp, s, d_tab))
        ref.append(T.run_oracle(s, fft, spaces, table, box)[:2])
    b = F.FRBatch(0)
    b.set_sched_params(F.FR_EXACT, F.SchedParams(*params))
    b.plan(jobs, 7, F.FR_EXACT, graph=True)
    b.execute()
    b.synchronize()
    for j, (rv, rw) in zip(jobs, ref):
        assert j.volume.cpu().numpy().tobytes() == rv.tobytes()
        assert j.weights.cpu().numpy().tobytes() == rw.tobytes()


def test_batch_autotune(tuning_dir):
    s = T.settings()
    dev = torch.device("cuda:0")
    jobs, ref = [], []
    for k, (box, n) in enumerate([(32, 50), (48, 20), (64, 35)]):
        fft, spaces, desc, table, d_fft, d_sp, d_tab = _inputs(box, n, s, seed=30 + k)
        h, sc, deq = T.quantize_f16(fft)
        vol = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=dev)
        wts = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=dev)
        jobs.append(F.BatchJob(torch.from_numpy(h).to(dev), vol, wts, d_sp, s, d_tab,
                               scales=torch.from_numpy(sc).to(dev)))
        ref.append(T.run_oracle(s, deq, spaces, table, box)[:2])
    b = F.FRBatch(0)
    p = b.autotune(jobs, 16, F.FR_FAST_MATH, use_cache=False)
    assert p.maxParts >= 1
    for j in jobs:
        assert float(j.weights.abs().sum()) == 0.0      # tuning used scratch volumes
    b.plan(jobs, 16, F.FR_FAST_MATH, graph=True)
    b.execute()
    b.synchronize()
    for j, (rv, rw) in zip(jobs, ref):
        assert T.rel_l2(j.volume.cpu().numpy(), rv) <= 1e-5
        assert T.rel_l2(j.weights.cpu().numpy(), rw) <= 1e-5
    assert (tuning_dir / "fr_tuning.txt").read_text().startswith("k2:batch:")


@pytest.mark.parametrize("radius", [1.2, 1.9])
def test_every_exact_tuning_candidate_is_bit_exact(radius):
    """Every knob set the tuner times under the exact strategy (schedules x staged / per-pair
    kernel x table in LDS / global memory), forced on a handle: bit-identical to the oracle,
    at footprints 3 and 4 (both kernels' instantiations)."""
    s = T.settings(radius=radius)
    box, n = 48, 40
    fft, spaces, desc, table, d_fft, d_sp, d_tab = _inputs(box, n, s, seed=41)
    rv, rw, _ = T.run_oracle(s, fft, spaces, table, box)
    dev = torch.device("cuda:0")
    op = F.FourierReconstruction(0)
    assert op.init(desc, s)
    cands = F.tuning_candidates(F.FR_EXACT)
    assert len(cands) == 24
    for p in cands:
        op.set_sched_params(F.FR_EXACT, p)
        vol = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=dev)
        wts = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=dev)
        assert op.execute(d_fft, vol, wts, d_sp, d_tab)
        op.synchronize()
        assert vol.cpu().numpy().tobytes() == rv.tobytes(), p.as_tuple()
        assert wts.cpu().numpy().tobytes() == rw.tobytes(), p.as_tuple()


@pytest.mark.parametrize("half", [False, True])
def test_every_fast_math_tuning_candidate_within_tolerance(half):
    """Every knob set the tuner times under fast math (schedules, then per-pair / carried
    kernel x sliding / fixed boxes x table placement), FP32 and FP16 slices: within 1e-5 of
    the oracle (of the dequantised slices for FP16)."""
    s = T.settings()
    box, n = 64, 60
    fft, spaces, desc, table, d_fft, d_sp, d_tab = _inputs(box, n, s, seed=43)
    dev = torch.device("cuda:0")
    scales = None
    if half:
        h, sc, fft = T.quantize_f16(fft)
        d_fft, scales = torch.from_numpy(h).to(dev), torch.from_numpy(sc).to(dev)
    rv, rw, _ = T.run_oracle(s, fft, spaces, table, box)
    op = F.FourierReconstruction(0)
    assert op.init(desc, s)
    op.set_strategy(F.FR_FAST_MATH)
    cands = F.tuning_candidates(F.FR_FAST_MATH)
    assert len(cands) == 32
    for p in cands:
        op.set_sched_params(F.FR_FAST_MATH, p)
        vol = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=dev)
        wts = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=dev)
        assert op.execute(d_fft, vol, wts, d_sp, d_tab, scales=scales)
        op.synchronize()
        assert T.rel_l2(vol.cpu().numpy(), rv) <= 1e-5, p.as_tuple()
        assert T.rel_l2(wts.cpu().numpy(), rw) <= 1e-5, p.as_tuple()

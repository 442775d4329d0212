"""Batched FR (umpa_fr_batch_*): several reconstructions of different box sizes and slice
counts through one persistent launch per step.  Each job's result must equal running the job
alone: bit-identical to the oracle (FRCPU's arithmetic) under the exact strategy, for every
chunking and with or without the HIP graph; within fast-math tolerance for FP16 slices."""
import numpy as np
import pytest

import frtools as T
from umpalumpa_amd import fr as F

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def make_jobs(specs, s, half=False, seed=5):
    """specs: list of (box, n_slices).  Returns host problems and device BatchJobs."""
    dev = torch.device("cuda:0")
    host, jobs = [], []
    for k, (box, n) in enumerate(specs):
        fft, spaces, desc, table = T.problem(box, n, s, seed=seed + 7 * k)
        vol = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=dev)
        wts = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=dev)
        d_sp = torch.from_numpy(spaces.view(np.uint8)).to(dev)
        d_tab = torch.from_numpy(table).to(dev)
        if half:
            h, sc, deq = T.quantize_f16(fft)
            d_fft = torch.from_numpy(h).to(dev)
            d_sc = torch.from_numpy(sc).to(dev)
            host.append((box, deq, spaces, desc, table, (h, sc)))
            jobs.append(F.BatchJob(d_fft, vol, wts, d_sp, s, d_tab, scales=d_sc))
        else:
            d_fft = torch.from_numpy(fft).to(dev)
            host.append((box, fft, spaces, desc, table, None))
            jobs.append(F.BatchJob(d_fft, vol, wts, d_sp, s, d_tab))
    return host, jobs


SPECS = [(32, 40), (48, 13), (20, 70), (64, 25)]


@pytest.mark.parametrize("chunk,graph", [(1000, True), (7, True), (7, False), (1, False)])
def test_exact_batch_matches_oracle(chunk, graph):
    s = T.settings()
    host, jobs = make_jobs(SPECS, s)
    b = F.FRBatch(0)
    b.plan(jobs, chunk, F.FR_EXACT, graph=graph)
    assert b.num_steps() == -(-max(j.n_spaces for j in jobs) // chunk)
    b.execute()
    b.synchronize()
    for (box, fft, spaces, desc, table, _), j in zip(host, jobs):
        rv, rw, _ = T.run_oracle(s, fft, spaces, table, box)
        gv, gw = j.volume.cpu().numpy(), j.weights.cpu().numpy()
        assert gv.tobytes() == rv.tobytes(), (box, chunk, graph)
        assert gw.tobytes() == rw.tobytes(), (box, chunk, graph)


@pytest.mark.parametrize("items,graph", [(1, False), (3, True), (64, False)])
def test_bounded_launches_match_oracle(items, graph):
    """Bounded launches (every wavefront leaves after `items` work items, the grid covers the
    items): the same bits as the persistent launch, with or without the HIP graph."""
    s = T.settings()
    host, jobs = make_jobs(SPECS, s, seed=9)
    b = F.FRBatch(0)
    b.set_wave_items(items)
    b.plan(jobs, 11, F.FR_EXACT, graph=graph)
    b.execute()
    b.synchronize()
    for (box, fft, spaces, desc, table, _), j in zip(host, jobs):
        rv, rw, _ = T.run_oracle(s, fft, spaces, table, box)
        assert j.volume.cpu().numpy().tobytes() == rv.tobytes(), (box, items)
        assert j.weights.cpu().numpy().tobytes() == rw.tobytes(), (box, items)


@pytest.mark.parametrize("workgroups,graph", [(1, False), (7, False), (200, False), (7, True)])
def test_capped_persistent_launches_match_oracle(workgroups, graph):
    """Persistent launches on at most N workgroups (umpa_fr_batch_set_max_workgroups), also
    captured into the plan's HIP graph: the same bits as on every CU."""
    s = T.settings()
    host, jobs = make_jobs(SPECS, s, seed=31)
    b = F.FRBatch(0)
    b.set_max_workgroups(workgroups)
    b.plan(jobs, 11, F.FR_EXACT, graph=graph)
    b.execute()
    b.synchronize()
    for (box, fft, spaces, desc, table, _), j in zip(host, jobs):
        rv, rw, _ = T.run_oracle(s, fft, spaces, table, box)
        assert j.volume.cpu().numpy().tobytes() == rv.tobytes(), (box, workgroups)
        assert j.weights.cpu().numpy().tobytes() == rw.tobytes(), (box, workgroups)


def test_two_batches_on_their_streams_overlap_without_changing_results():
    """Two batches (two streams) with bounded launches enqueued back to back, as the server's
    two lanes run: each job's result is the oracle's, bit for bit."""
    s = T.settings()
    host_a, jobs_a = make_jobs(SPECS[:2], s, seed=13)
    host_b, jobs_b = make_jobs(SPECS[2:], s, seed=17)
    ba, bb = F.FRBatch(0), F.FRBatch(0)
    for b in (ba, bb):
        b.set_wave_items(2)
    for c in range(0, 80, 9):
        for b, jobs in ((ba, jobs_a), (bb, jobs_b)):
            b.step(jobs, [min(c, j.n_spaces) for j in jobs],
                   [min(c + 9, j.n_spaces) for j in jobs], F.FR_EXACT)
    ba.synchronize()
    bb.synchronize()
    for host, jobs in ((host_a, jobs_a), (host_b, jobs_b)):
        for (box, fft, spaces, desc, table, _), j in zip(host, jobs):
            rv, rw, _ = T.run_oracle(s, fft, spaces, table, box)
            assert j.volume.cpu().numpy().tobytes() == rv.tobytes(), box
            assert j.weights.cpu().numpy().tobytes() == rw.tobytes(), box


@pytest.mark.parametrize("half,fixed", [(False, 0), (True, 0), (False, 1)])
def test_batch_carried_queues_within_tolerance(half, fixed):
    """Batches on the carried-queue kernel (the knob the batch tuner tries for fast math), FP32
    and FP16 slices, sliding or fixed boxes, split heavy bricks: within 1e-5 of the oracle (on
    the dequantized slices for FP16)."""
    s = T.settings()
    host, jobs = make_jobs(SPECS, s, half=half, seed=21)
    b = F.FRBatch(0)
    b.set_sched_params(F.FR_FAST_MATH,
                       F.SchedParams(0.5, 1, 16, 0.25, F.FR_KERNEL_CARRY, 0, fixed))
    b.plan(jobs, 13, F.FR_FAST_MATH, graph=True)
    b.execute()
    b.synchronize()
    for (box, fft, spaces, desc, table, _), j in zip(host, jobs):
        rv, rw, _ = T.run_oracle(s, fft, spaces, table, box)
        assert T.rel_l2(j.volume.cpu().numpy(), rv) <= 1e-5, box
        assert T.rel_l2(j.weights.cpu().numpy(), rw) <= 1e-5, box


def test_graph_replays_accumulate():
    """Each execute replays the whole chunk loop onto the volumes: two replays = the job's
    spaces inserted twice, still in order."""
    s = T.settings()
    host, jobs = make_jobs(SPECS[:2], s)
    b = F.FRBatch(0)
    b.plan(jobs, 9, F.FR_EXACT, graph=True)
    for _ in range(3):
        b.execute()
        b.synchronize()
    for (box, fft, spaces, desc, table, _), j in zip(host, jobs):
        # the oracle inserts the same records three times (projection indices unchanged)
        rv, rw = T.run_oracle(s, fft, np.concatenate([spaces] * 3), table, box)[:2]
        assert j.volume.cpu().numpy().tobytes() == rv.tobytes()
        assert j.weights.cpu().numpy().tobytes() == rw.tobytes()


def test_step_ranges_advance_independently():
    """Continuous batching: jobs enter at different steps and advance by different ranges;
    every job still equals its solo result bit for bit (exact strategy)."""
    s = T.settings()
    host, jobs = make_jobs(SPECS, s)
    b = F.FRBatch(0)
    pos = [0] * len(jobs)
    start = [0, 2, 3, 5]         # step at which each job is admitted
    rng = np.random.default_rng(3)
    step = 0
    while any(p < j.n_spaces for p, j in zip(pos, jobs)):
        active = [i for i in range(len(jobs)) if step >= start[i] and pos[i] < jobs[i].n_spaces]
        if active:
            ends = [min(jobs[i].n_spaces, pos[i] + int(rng.integers(1, 12))) for i in active]
            b.step([jobs[i] for i in active], [pos[i] for i in active], ends, F.FR_EXACT)
            for i, e in zip(active, ends):
                pos[i] = e
        step += 1
    b.synchronize()
    for (box, fft, spaces, desc, table, _), j in zip(host, jobs):
        rv, rw, _ = T.run_oracle(s, fft, spaces, table, box)
        assert j.volume.cpu().numpy().tobytes() == rv.tobytes()
        assert j.weights.cpu().numpy().tobytes() == rw.tobytes()


def test_batch_update_count_is_sum_of_jobs():
    s = T.settings()
    host, jobs = make_jobs(SPECS, s)
    b = F.FRBatch(0)
    b.plan(jobs, 11, F.FR_EXACT, graph=True)
    b.set_count_updates(True)
    b.execute()
    n = b.last_updates()
    want = sum(T.run_oracle(s, fft, spaces, table, box)[2]
               for box, fft, spaces, desc, table, _ in host)
    assert n == want


def test_f16_batch_matches_solo_f16():
    """FP16 slices (fast-math, split central bricks): each job within 1e-6 rel-L2 of the same
    job run alone through umpa_fr_execute_f16, and within 1e-5 of the FP32 oracle on the
    dequantized slices."""
    s = T.settings()
    host, jobs = make_jobs(SPECS, s, half=True)
    b = F.FRBatch(0)
    b.plan(jobs, 16, F.FR_FAST_MATH, graph=True)
    b.execute()
    b.synchronize()
    for (box, deq, spaces, desc, table, half), j in zip(host, jobs):
        sv, sw, _ = T.run_hip(s, deq, spaces, desc, table, box, half=half)
        gv, gw = j.volume.cpu().numpy(), j.weights.cpu().numpy()
        assert T.rel_l2(gv, sv) <= 1e-6 and T.rel_l2(gw, sw) <= 1e-6
        rv, rw, _ = T.run_oracle(s, deq, spaces, table, box)
        assert T.rel_l2(gv, rv) <= 1e-5 and T.rel_l2(gw, rw) <= 1e-5


def test_mixed_settings_form_separate_launch_groups():
    """Jobs with different blob radii (and so tables and footprints) in one plan."""
    dev = torch.device("cuda:0")
    out = []
    jobs = []
    for k, (box, n, r) in enumerate([(32, 30, 1.5), (32, 30, 1.9), (40, 20, 2.5)]):
        s = T.settings(radius=r)
        fft, spaces, desc, table = T.problem(box, n, s, seed=40 + k)
        vol = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=dev)
        wts = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=dev)
        jobs.append(F.BatchJob(torch.from_numpy(fft).to(dev), vol, wts,
                               torch.from_numpy(spaces.view(np.uint8)).to(dev), s,
                               torch.from_numpy(table).to(dev)))
        out.append((s, box, fft, spaces, table))
    b = F.FRBatch(0)
    b.plan(jobs, 8, F.FR_EXACT, graph=True)
    b.execute()
    b.synchronize()
    for (s, box, fft, spaces, table), j in zip(out, jobs):
        rv, rw, _ = T.run_oracle(s, fft, spaces, table, box)
        assert j.volume.cpu().numpy().tobytes() == rv.tobytes()
        assert j.weights.cpu().numpy().tobytes() == rw.tobytes()


def test_batch_errors():
    s = T.settings()
    host, jobs = make_jobs(SPECS[:1], s)
    b = F.FRBatch(0)
    with pytest.raises(F._lib.UmpaError):
        b.execute()                       # no plan yet
    with pytest.raises(F._lib.UmpaError):
        b.plan(jobs, 0, F.FR_EXACT)       # chunk 0
    bad = F.BatchJob(jobs[0].fft, jobs[0].volume, jobs[0].weights, jobs[0].spaces, s, None)
    with pytest.raises(F._lib.UmpaError) as e:
        b.plan([bad], 8, F.FR_EXACT)      # kLookup without a table: FRCPU returns false
    assert e.value.status == F._lib.UMPA_ERR_REJECTED
    with pytest.raises(F._lib.UmpaError):
        b.step(jobs, [0], [jobs[0].n_spaces + 1], F.FR_EXACT)   # range past the spaces


def test_batch_refuses_tables_beyond_lds():
    """Batches keep the blob table in LDS: 11,776 samples plan, 11,777 are unsupported."""
    s = T.settings()
    host, jobs = make_jobs(SPECS[:1], s)
    dev = torch.device("cuda:0")
    b = F.FRBatch(0)
    for n_table, ok in ((11776, True), (11777, False)):
        tab = torch.zeros(n_table, dtype=torch.float32, device=dev)
        j = F.BatchJob(jobs[0].fft, jobs[0].volume, jobs[0].weights, jobs[0].spaces, s, tab)
        if ok:
            b.plan([j], 8, F.FR_EXACT)
        else:
            with pytest.raises(F._lib.UmpaError) as e:
                b.plan([j], 8, F.FR_EXACT)
            assert e.value.status == F._lib.UMPA_ERR_UNSUPPORTED

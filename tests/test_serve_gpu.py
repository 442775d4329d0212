"""The continuous-batching server: jobs of different boxes, radii and symmetry groups that
overlap on the GPU give the same result as each job served alone (bit for bit under the
exact strategy, and equal to the oracle), admission respects the HBM budget, and FP16 jobs
agree alone vs in the mix within fast-math tolerance."""
import numpy as np
import pytest

import frtools as T
import pyoracle
from umpalumpa_amd import fr as F
from umpalumpa_amd import serve as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SPECS = [  # box, slices, radius, symmetry, arrival (s)
    (32, 150, 1.9, "C1", 0.0),
    (48, 90, 1.5, "C4", 0.0),
    (32, 260, 2.5, "D2", 0.001),
    (40, 70, 1.9, "C1", 0.002),
]


def make_jobs(half=False, seed=11, chunk=64):
    jobs = []
    for k, (box, n, r, sym, t) in enumerate(SPECS):
        fft = T.synthetic_slices(box, n, seed + k)
        rots = pyoracle.uniform_rotations(n, seed + 100 + k)
        if half:
            h, sc, deq = T.quantize_f16(fft)
            dh = torch.from_numpy(h).cuda()
            ds = torch.from_numpy(sc).cuda()
            src = (lambda c, dh=dh, ds=ds: (dh[c * chunk:(c + 1) * chunk],
                                            ds[c * chunk:(c + 1) * chunk]))
            host = deq
        else:
            d = torch.from_numpy(fft).cuda()
            src = (lambda c, d=d: d[c * chunk:(c + 1) * chunk])
            host = fft
        jobs.append((S.Job(k, box, n, r, sym, rots, src, arrival=t), host))
    return jobs


def solo(job_host, strategy, chunk=64):
    job, _ = job_host
    with S.Server(chunk=chunk, strategy=strategy) as srv:   # close() ends its threads
        srv.run([job], realtime=False)
    return job.volume.cpu().numpy(), job.weights.cpu().numpy()


@pytest.mark.parametrize("policy,lanes", [("srpt", 2), ("aged", 2), ("oldest", 2), ("fair", 2),
                                          ("aged", 1)])
def test_exact_alone_equals_mix_and_oracle(policy, lanes):
    """Jobs served together (the smallest on the fast lane unless one lane, the others on the
    main lane, chunks copied into each job's ring) give the bits each gives served alone, and
    FRCPU's."""
    chunk = 64
    mix = make_jobs(chunk=chunk)
    srv = S.Server(chunk=chunk, strategy=F.FR_EXACT, depth=3, policy=policy,
                   tick_work=2 * 64 * 33 ** 2, fast_seconds=2e-2 if lanes != 1 else None,
                   ring_bytes=2 * chunk * 17 * 32 * 8)   # rings of 1-2 slots, reused
    done = srv.run([j for j, _ in mix], realtime=True)
    assert len(done) == len(SPECS)
    alone = make_jobs(chunk=chunk)
    for (jm, host), ja in zip(mix, alone):
        av, aw = solo(ja, F.FR_EXACT, chunk)
        assert jm.volume.cpu().numpy().tobytes() == av.tobytes(), jm.job_id
        assert jm.weights.cpu().numpy().tobytes() == aw.tobytes(), jm.job_id
    # and each equals FRCPU on the job's spaces (chunk-relative projection indices)
    for (jm, host) in mix:
        s = T.settings(radius=jm.radius)
        sym = F.symmetry_operators(jm.symmetry)
        spaces = np.concatenate([
            F.generate_traverse_spaces(jm.rotations[c:c + chunk], jm.box, s, sym, first_index=c)
            for c in range(0, jm.n_slices, chunk)])
        desc = F.fr_desc(jm.box // 2 + 1, jm.box, jm.n_slices, jm.box + 1, len(spaces), 10000)
        table = F.fill_blob_table(10000, desc, s)
        rv, rw, _ = T.run_oracle(s, host, spaces, table, jm.box)
        assert jm.volume.cpu().numpy().tobytes() == rv.tobytes(), jm.job_id
        assert jm.weights.cpu().numpy().tobytes() == rw.tobytes(), jm.job_id


def test_f16_alone_vs_mix_within_tolerance():
    chunk = 64
    mix = make_jobs(half=True, chunk=chunk)
    srv = S.Server(chunk=chunk, strategy=F.FR_FAST_MATH)
    srv.run([j for j, _ in mix], realtime=False)
    alone = make_jobs(half=True, chunk=chunk)
    for (jm, _), ja in zip(mix, alone):
        av, aw = solo(ja, F.FR_FAST_MATH, chunk)
        assert T.rel_l2(jm.volume.cpu().numpy(), av) <= 1e-6
        assert T.rel_l2(jm.weights.cpu().numpy(), aw) <= 1e-6


def test_budget_admission_respects_the_budget():
    """A budget that holds the largest job but not all jobs: at every admission the admitted
    footprints fit the budget, all jobs complete, and results are unchanged."""
    chunk = 64
    jobs = make_jobs(chunk=chunk)
    foot = {j.job_id: j.footprint(chunk) for j, _ in jobs}
    budget = max(foot.values()) * 1.01
    assert sum(foot.values()) > budget
    srv = S.Server(chunk=chunk, strategy=F.FR_EXACT, hbm_budget=budget)
    done = srv.run([j for j, _ in jobs], realtime=False)
    assert len(done) == len(jobs)
    for j in done:
        # jobs holding memory when j was admitted (the server releases a job's budget at the
        # moment it records the job's completion)
        live = [k for k in done if k.admitted <= j.admitted < k.done]
        assert sum(foot[k.job_id] for k in live) <= budget
    alone = make_jobs(chunk=chunk)
    for (jm, _), ja in zip(jobs, alone):
        av, aw = solo(ja, F.FR_EXACT, chunk)
        assert jm.volume.cpu().numpy().tobytes() == av.tobytes()


def test_oversized_job_is_refused():
    chunk = 64
    jobs = make_jobs(chunk=chunk)[:1]
    srv = S.Server(chunk=chunk, hbm_budget=1000)
    with pytest.raises(ValueError):
        srv.run([jobs[0][0]], realtime=False)

"""Parity of each benchmarked workload's own path at that workload's shapes, against the oracle
(FRCPU's arithmetic) and the FP64 oracle, on slice counts the oracle finishes in seconds:

  plain box 256, fast math (the benched strategy) with the schedule forced to split the central
      bricks over wavefronts (their partial sums merge with float atomics), FP32 and FP16
  batched: 16 jobs of the batched workload's boxes (64-320), 8-24 FP16 slices each, one
      persistent launch per step, autotuned and forced-split schedules; FP32 exact bit for bit
  serving: jobs at boxes 96-384, radii 1.5 / 1.9 / 2.5, C1 / C4 / D2, chunks of 512 FP16
      slices streaming through the continuous-batching server
  large symmetric: box 1024 (1025^3 volume), 2 FP16 slices x D7 (14 operators), fast math

Bars: fast math and FP16 within 1e-5 relative L2 of the oracle on the same (dequantised)
slices and within 1e-4 of the FP64 oracle on them.  Each FP16 test also prints the error
against the FP64 oracle of the original FP32 slices (the storage format's own rounding,
2^-11 relative per element, is about 2e-4 to 3e-4 of the volume on these random slices)."""
import numpy as np
import pytest

import frtools as T
from umpalumpa_amd import fr as F
from umpalumpa_amd import serve as S
from umpalumpa_amd._lib import SchedParams

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FORCED_SPLIT = (1.0, 1, 16, 0.25)   # depthScale, priority, maxParts, splitAt


def _split_items(box, params):
    items = F.brick_schedule(box + 1, box + 1, box + 1, True, SchedParams(*params))
    return int((((items[:, 3] >> 8) & 255) > 1).sum())


@pytest.mark.parametrize("half", [False, True])
def test_box256_fast_math_split_vs_frcpu(half):
    """The benched strategy at box 256 (120 slices 129 x 256 into 257^3) with central bricks
    split over wavefronts: within 1e-5 of FRCPU and 1e-4 of FP64 on the same slices."""
    s = T.settings()
    assert _split_items(256, FORCED_SPLIT) > 100
    fft, spaces, desc, table = T.problem(256, 120, s, seed=91)
    hp = None
    if half:
        h, sc, fft_in = T.quantize_f16(fft)
        hp = (h, sc)
    else:
        fft_in = fft
    v, w, _ = T.run_hip(s, fft_in, spaces, desc, table, 256, strategy=F.FR_FAST_MATH, half=hp,
                        sched=SchedParams(*FORCED_SPLIT))
    ov, ow, _ = T.run_oracle(s, fft_in, spaces, table, 256)
    ev, ew = T.rel_l2(v, ov), T.rel_l2(w, ow)
    dv, dw = T.run_oracle_f64(s, fft_in, spaces, table, 256)
    fv, fw = T.rel_l2(v, dv), T.rel_l2(w, dw)
    msg = f"box 256 fast math split {'FP16' if half else 'FP32'}: vs FRCPU {ev:.2e} / {ew:.2e}, " \
          f"vs FP64 {fv:.2e} / {fw:.2e}"
    if half:
        ov64, _ = T.run_oracle_f64(s, fft, spaces, table, 256)
        msg += f", vs FP64 of the FP32 slices {T.rel_l2(v, ov64):.2e}"
    print(msg)
    assert ev <= 1e-5 and ew <= 1e-5
    assert fv <= 1e-4 and fw <= 1e-4


def _batched_specs(n_jobs=16, seed=2024):
    """The batched workload's boxes (its draw from {64, 96, 128, 192, 256, 320}), with 8-24
    slices per job instead of 500-20,000 so the oracle can check every job."""
    rng = np.random.default_rng(seed)
    boxes = rng.choice([64, 96, 128, 192, 256, 320], size=n_jobs)
    counts = rng.integers(8, 25, size=n_jobs)
    return [(int(b), int(c)) for b, c in zip(boxes, counts)]


def _make_batch(specs, s, half, seed=300):
    dev = torch.device("cuda:0")
    host, jobs = [], []
    for k, (box, n) in enumerate(specs):
        fft, spaces, desc, table = T.problem(box, n, s, seed=seed + 11 * k)
        vol = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=dev)
        wts = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=dev)
        d_sp = torch.from_numpy(spaces.view(np.uint8)).to(dev)
        d_tab = torch.from_numpy(table).to(dev)
        if half:
            h, sc, deq = T.quantize_f16(fft)
            jobs.append(F.BatchJob(torch.from_numpy(h).to(dev), vol, wts, d_sp, s, d_tab,
                                   scales=torch.from_numpy(sc).to(dev)))
            host.append((box, fft, deq, spaces, table))
        else:
            jobs.append(F.BatchJob(torch.from_numpy(fft).to(dev), vol, wts, d_sp, s, d_tab))
            host.append((box, fft, fft, spaces, table))
    return host, jobs


@pytest.mark.parametrize("schedule", ["autotuned", "forced_split"])
def test_batched_workload_fp16_vs_oracle(schedule):
    s = T.settings()
    specs = _batched_specs()
    assert {b for b, _ in specs} >= {64, 320}
    host, jobs = _make_batch(specs, s, half=True)
    b = F.FRBatch(0)
    if schedule == "autotuned":
        b.autotune(jobs, 10, F.FR_FAST_MATH, use_cache=False)
    else:
        b.set_sched_params(F.FR_FAST_MATH, SchedParams(*FORCED_SPLIT))
    b.plan(jobs, 10, F.FR_FAST_MATH, graph=True)
    b.execute()
    b.synchronize()
    worst = [0.0, 0.0, 0.0]
    for (box, fft, deq, spaces, table), j in zip(host, jobs):
        gv, gw = j.volume.cpu().numpy(), j.weights.cpu().numpy()
        ov, ow, _ = T.run_oracle(s, deq, spaces, table, box)
        dv, dw = T.run_oracle_f64(s, deq, spaces, table, box)
        e = (max(T.rel_l2(gv, ov), T.rel_l2(gw, ow)), max(T.rel_l2(gv, dv), T.rel_l2(gw, dw)))
        assert e[0] <= 1e-5, (box, e)
        assert e[1] <= 1e-4, (box, e)
        e64, _ = T.run_oracle_f64(s, fft, spaces, table, box)
        worst = [max(worst[0], e[0]), max(worst[1], e[1]), max(worst[2], T.rel_l2(gv, e64))]
    print(f"batched FP16 ({schedule}), worst job: vs FRCPU {worst[0]:.2e}, vs FP64 "
          f"{worst[1]:.2e}, vs FP64 of the FP32 slices {worst[2]:.2e}")


def test_batched_workload_fp32_exact_bit_identical():
    s = T.settings()
    specs = _batched_specs()
    host, jobs = _make_batch(specs, s, half=False)
    b = F.FRBatch(0)
    b.plan(jobs, 10, F.FR_EXACT, graph=True)
    b.execute()
    b.synchronize()
    for (box, fft, _, spaces, table), j in zip(host, jobs):
        ov, ow, _ = T.run_oracle(s, fft, spaces, table, box)
        assert j.volume.cpu().numpy().tobytes() == ov.tobytes(), box
        assert j.weights.cpu().numpy().tobytes() == ow.tobytes(), box


SERVING = [  # box, slices, radius, symmetry, arrival (s)
    (96, 600, 1.5, "C1", 0.0),
    (128, 80, 2.5, "D2", 0.0),
    (192, 30, 1.9, "C4", 0.001),
    (256, 40, 1.5, "D2", 0.002),
    (384, 16, 1.9, "C4", 0.002),
    (384, 12, 2.5, "C1", 0.003),
]


def test_serving_workload_fp16_vs_oracle():
    """Jobs of the serving workload's kinds streamed in chunks of 512 FP16 slices through the
    continuous-batching server, overlapping on the GPU: each job's volume and weights within
    1e-5 of the oracle and 1e-4 of the FP64 oracle on its dequantised slices."""
    chunk = 512
    jobs, hosts = [], []
    for k, (box, n, r, sym, t) in enumerate(SERVING):
        fft = T.synthetic_slices(box, n, 500 + k)
        h, sc, deq = T.quantize_f16(fft)
        dh, ds = torch.from_numpy(h).cuda(), torch.from_numpy(sc).cuda()
        rots = T.pyoracle.uniform_rotations(n, 600 + k)
        src = (lambda c, dh=dh, ds=ds: (dh[c * chunk:(c + 1) * chunk],
                                        ds[c * chunk:(c + 1) * chunk]))
        jobs.append(S.Job(k, box, n, r, sym, rots, src, arrival=t))
        hosts.append((fft, deq))
    srv = S.Server(chunk=chunk, strategy=F.FR_FAST_MATH, depth=3)
    done = srv.run(jobs, realtime=True)
    assert len(done) == len(SERVING)
    assert any(j.n_slices > chunk for j in jobs)
    for j, (fft, deq) in zip(jobs, hosts):
        s = T.settings(radius=j.radius)
        sym = F.symmetry_operators(j.symmetry)
        spaces = np.concatenate([
            F.generate_traverse_spaces(j.rotations[c:c + chunk], j.box, s, sym, first_index=c)
            for c in range(0, j.n_slices, chunk)])
        desc = F.fr_desc(j.box // 2 + 1, j.box, j.n_slices, j.box + 1, len(spaces), 10000)
        table = F.fill_blob_table(10000, desc, s)
        gv, gw = j.volume.cpu().numpy(), j.weights.cpu().numpy()
        ov, ow, _ = T.run_oracle(s, deq, spaces, table, j.box)
        dv, dw = T.run_oracle_f64(s, deq, spaces, table, j.box)
        e = (T.rel_l2(gv, ov), T.rel_l2(gw, ow), T.rel_l2(gv, dv), T.rel_l2(gw, dw))
        e64 = T.rel_l2(gv, T.run_oracle_f64(s, fft, spaces, table, j.box)[0])
        print(f"serving box {j.box} r {j.radius} {j.symmetry}: vs FRCPU {e[0]:.2e} / "
              f"{e[1]:.2e}, vs FP64 {e[2]:.2e} / {e[3]:.2e}, vs FP64 of the FP32 slices "
              f"{e64:.2e}")
        assert e[0] <= 1e-5 and e[1] <= 1e-5, j.job_id
        assert e[2] <= 1e-4 and e[3] <= 1e-4, j.job_id


def test_large_d7_fp16_box1024_vs_oracle():
    """The large symmetric workload's path: box 1024 (513 x 1024 FP16 slices into a 1025^3
    volume, 8.6 GB + 4.3 GB), D7 symmetry (14 operators per slice, FP64 products on the host),
    fast math: 2 slices against the oracle and the FP64 oracle."""
    s = T.settings()
    fft, spaces, desc, table = T.problem(1024, 2, s, seed=77,
                                         symmetries=F.symmetry_operators("D7"))
    assert len(spaces) == 28
    h, sc, deq = T.quantize_f16(fft)
    v, w, _ = T.run_hip(s, deq, spaces, desc, table, 1024, strategy=F.FR_FAST_MATH,
                        half=(h, sc))
    ov, ow, _ = T.run_oracle(s, deq, spaces, table, 1024)
    ev, ew = T.rel_l2_big(v, ov), T.rel_l2_big(w, ow)
    assert (ow > 0).sum() > 1e6
    del ov, ow
    dv, dw = T.run_oracle_f64(s, deq, spaces, table, 1024)
    fv, fw = T.rel_l2_big(v, dv), T.rel_l2_big(w, dw)
    print(f"box 1024 D7 FP16: vs FRCPU {ev:.2e} / {ew:.2e}, vs FP64 {fv:.2e} / {fw:.2e}")
    assert ev <= 1e-5 and ew <= 1e-5
    assert fv <= 1e-4 and fw <= 1e-4

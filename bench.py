"""Benchmark of the Fourier-reconstruction insertion on one MI355X.

Workload (default): 20,000 Hermitian half-plane slices of 129 x 256 complex float (padded box
P = 256), uniform SO(3) orientations, Kaiser-Bessel blob r = 1.9, alpha = 15, order 0, 10,000-
sample lookup table, inserted in chunks of 1,000 slices into a 257^3 complex-float volume plus a
float weight volume.  Slice values: complex Gaussian (E|z|^2 = 1) times 1/(1+k^2).

A step is one AFR::Execute over one chunk of 1,000 slices, all inputs resident in HBM.  The
line reports GVoxel-updates/s (voxel-updates = (voxel, traverse space) pairs that reach the
accumulation step, counted by the kernel's counting variant in an untimed pass), the
algorithmic-byte roofline of the insertion kernel measured with HIP events on the stream it runs
on, and the reference's own CPU insertion (FRCPU compiled from the reference sources, or the
oracle's restatement where that is absent) timed on one host core over one chunk.

    python bench.py [--gpus 1] [--steps 20] [--warmup 2] [--strategy exact|fast|f16]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--box", type=int, default=256)
    p.add_argument("--chunk", type=int, default=1000)
    p.add_argument("--slices", type=int, default=20000)
    p.add_argument("--strategy", default="exact", choices=["exact", "fast", "f16"])
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-slices", type=int, default=1000)
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a plain 256 MiB device-to-device copy, 5 times: the known-size "
                        "kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE to bytes")
    return p.parse_args()


def copy_calibration(torch):
    n = 256 << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n, "launches": 5}))


def make_slices(torch, box, n, seed, dev):
    """Complex Gaussian (E|z|^2 = 1) times 1/(1+k^2), generated on the GPU."""
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    hx, hy = box // 2 + 1, box
    kx = torch.arange(hx, device=dev, dtype=torch.float32)[None, :]
    ky = torch.arange(hy, device=dev, dtype=torch.float32)[:, None] - box // 2
    env = 1.0 / (1.0 + kx * kx + ky * ky)
    out = torch.empty((n, hy, hx), dtype=torch.complex64, device=dev)
    step = 1000
    for a in range(0, n, step):
        b = min(n, a + step)
        re = torch.randn((b - a, hy, hx), generator=g, device=dev) * (0.5 ** 0.5)
        im = torch.randn((b - a, hy, hx), generator=g, device=dev) * (0.5 ** 0.5)
        out[a:b] = torch.complex(re * env, im * env)
    return out


def uniform_rotations(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    x, y, z, w = q.T
    return np.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
        2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
        2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=1).reshape(n, 3, 3)


def cpu_baseline(box, fft_chunk, spaces_chunk, table):
    """FRCPU on one host core over one chunk: the reference's compiled code (oracle/_ref) when
    present, else the oracle's restatement."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    s = pyoracle.Settings(1, 0, 1, 15.0, 1.9)
    O = pyoracle.Oracle()
    c = O.create_constants(box + 1, box + 1, table.size, s)
    sp = spaces_chunk.view(pyoracle.TS_DTYPE)
    v = np.zeros((box + 1,) * 3, np.complex64)
    w = np.zeros((box + 1,) * 3, np.float32)
    kind = "port"
    runner = O
    if pyoracle.Reference.available():
        runner = pyoracle.Reference()
        kind = "reference"
    t0 = time.perf_counter()
    runner.fr_execute(s, v, w, fft_chunk, sp, table, c)
    dt = time.perf_counter() - t0
    # updates of this chunk, from the oracle's instrumented count
    v2 = np.zeros((box + 1,) * 3, np.complex64)
    w2 = np.zeros((box + 1,) * 3, np.float32)
    upd = O.fr_execute(s, v2, w2, fft_chunk, sp, table, c)
    return dt, upd, kind


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    if args.copy_calibration:
        copy_calibration(torch)
        return
    from umpalumpa_amd import fr as F

    dev = torch.device("cuda:0")
    box, chunk = args.box, args.chunk
    n_total = max(args.slices, chunk * (args.steps + args.warmup)) if args.steps else args.slices
    s = F.FRSettings(F.FR_LOOKUP, 0, F.FR_PRECISE, 15.0, 1.9)
    n_chunks = n_total // chunk
    desc = F.fr_desc(box // 2 + 1, box, chunk, box + 1, chunk, 10000)
    table = F.fill_blob_table(10000, desc, s)

    t0 = time.time()
    fft = make_slices(torch, box, n_chunks * chunk, 42, dev)
    rots = uniform_rotations(n_chunks * chunk, 43)
    spaces = []
    for k in range(n_chunks):
        sp = F.traverse_spaces(rots[k * chunk:(k + 1) * chunk], box, s)
        spaces.append(sp)
    d_spaces = [torch.from_numpy(sp.view(np.uint8)).to(dev) for sp in spaces]
    d_table = torch.from_numpy(table).to(dev)
    half = None
    if args.strategy == "f16":
        ri = torch.view_as_real(fft)
        scales = ri.abs().amax(dim=(1, 2, 3)).clamp_min(1e-30)
        half = ((ri / scales[:, None, None, None]).to(torch.float16).contiguous(),
                scales.float().contiguous())
    vol = torch.zeros((box + 1,) * 3, dtype=torch.complex64, device=dev)
    wts = torch.zeros((box + 1,) * 3, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    print(f"# setup {setup_s:.1f}s: {n_chunks} chunks of {chunk} slices", file=sys.stderr,
          flush=True)

    op = F.FourierReconstruction(0)
    assert op.init(desc, s), op.last_status
    op.set_strategy(F.FR_FAST_MATH if args.strategy == "fast" else F.FR_EXACT)
    stream = torch.cuda.current_stream()

    def chunk_inputs(k):
        kk = k % n_chunks
        if half is None:
            return fft[kk * chunk:(kk + 1) * chunk], None, d_spaces[kk]
        return half[0][kk * chunk:(kk + 1) * chunk], half[1][kk * chunk:(kk + 1) * chunk], \
            d_spaces[kk]

    def run(k):
        f, sc, sp = chunk_inputs(k)
        ok = op.execute(f, vol, wts, sp, d_table, scales=sc)
        assert ok, op.last_status

    # voxel-updates per chunk (counting variant, untimed)
    op.set_count_updates(True)
    updates = []
    for k in range(args.steps):
        run(k)
        updates.append(op.last_updates())
    op.set_count_updates(False)
    total_updates = int(sum(updates))

    for k in range(args.warmup):
        run(args.steps + k)
    torch.cuda.synchronize()
    vol.zero_()
    wts.zero_()
    torch.cuda.synchronize()

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t_wall = time.perf_counter()
    ev[0].record(stream)
    for k in range(args.steps):
        run(k)
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    per = np.array([ev[k].elapsed_time(ev[k + 1]) for k in range(args.steps)])  # ms
    total_ms = float(ev[0].elapsed_time(ev[-1]))

    slice_bytes = (box // 2 + 1) * box * (4 if half is not None else 8)
    b_alg = chunk * slice_bytes + (chunk * 4 if half is not None else 0) \
        + 2 * (box + 1) ** 3 * 12 + chunk * 112 + 4 * 10000
    achieved = b_alg / (per.mean() * 1e-3) / 1e9

    result = {
        "metric": "GVoxel-updates/s + HBM GB/s (box 256); p50/p99 job latency, batched vs plain",
        "value": total_updates / (total_ms * 1e-3) / 1e9,
        "unit": "GVoxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16-in/f32-acc" if half is not None else "f32",
        "data": "synthetic",
        "config": {"workload": f"plain FR, box {box}: {args.steps * chunk} slices "
                               f"{box // 2 + 1}x{box} into {box + 1}^3, chunks of {chunk}",
                   "strategy": args.strategy, "chunk": chunk, "blob_radius": 1.9,
                   "alpha": 15.0, "table": 10000, "parallelism": "single GPU"},
        "voxel_updates": total_updates,
        "chunk_ms_p50": float(np.percentile(per, 50)),
        "chunk_ms_p99": float(np.percentile(per, 99)),
        "wall_s": wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "fr_insert_kernel", "alg_bytes_per_launch": b_alg},
    }
    if not args.no_cpu:
        nc = min(args.cpu_slices, chunk)
        sp0 = spaces[0][:nc]
        f0 = fft[:nc].cpu().numpy()
        dt, upd, kind = cpu_baseline(box, f0, sp0, table)
        result["cpu_baseline"] = {"value": upd / dt / 1e9, "unit": "GVoxel-updates/s",
                                  "cores": 1, "kind": kind,
                                  "sample": f"one chunk of {nc} slices {box // 2 + 1}x{box} "
                                            f"into {box + 1}^3 ({upd} voxel-updates, {dt:.1f} s)"}
    print(json.dumps(result))


if __name__ == "__main__":
    main()

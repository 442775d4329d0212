"""Error of the HIP insertion against the FP64 host oracle (the reference's algorithm evaluated
in double), stated per workload on sub-problems the oracle finishes in seconds:

  plain FP32, box 256 (exact and fast math)          slices 129x256, first 40 slices
  FP16 slices, box 256 (batched / serving precision)  same slices, per-slice max-abs FP16
  FP16, serving-like: r 1.5 / 2.5, C4 / D2 symmetry   box 96
  FP16, large symmetric: D7 (14 operators)            box 256, 12 slices

FP16 errors are measured against the FP64 oracle on the dequantized slices (the precision of
the accumulation); the quantization itself is stated separately in the FP16 test of
test_fr_gpu.py.  Bounds: 1e-4 relative L2 for every FP16 path (the stated FP16 tolerance
against FP64) and for FP32."""
import numpy as np
import pytest

import frtools as T
from umpalumpa_amd import fr as F

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def errors(s, fft_host, spaces, desc, table, box, half=None, strategy=F.FR_EXACT):
    v, w, _ = T.run_hip(s, fft_host, spaces, desc, table, box, strategy=strategy, half=half)
    rv, rw = T.run_oracle_f64(s, fft_host, spaces, table, box)
    return T.rel_l2(v, rv), T.rel_l2(w, rw)


@pytest.mark.parametrize("strategy", [F.FR_EXACT, F.FR_FAST_MATH])
def test_box256_fp32_vs_fp64(strategy):
    s = T.settings()
    fft, spaces, desc, table = T.problem(256, 40, s, seed=61)
    ev, ew = errors(s, fft, spaces, desc, table, 256, strategy=strategy)
    print(f"box 256 FP32 strategy {strategy}: volume {ev:.2e}, weights {ew:.2e}")
    assert ev <= 1e-4 and ew <= 1e-4


def test_box256_fp16_vs_fp64():
    s = T.settings()
    fft, spaces, desc, table = T.problem(256, 40, s, seed=61)
    h, sc, deq = T.quantize_f16(fft)
    ev, ew = errors(s, deq, spaces, desc, table, 256, half=(h, sc))
    print(f"box 256 FP16: volume {ev:.2e}, weights {ew:.2e}")
    assert ev <= 1e-4 and ew <= 1e-4


@pytest.mark.parametrize("radius,group", [(1.5, "C4"), (2.5, "D2")])
def test_serving_like_fp16_vs_fp64(radius, group):
    s = T.settings(radius=radius)
    box, n = 96, 30
    fft, spaces, desc, table = T.problem(box, n, s, seed=71,
                                         symmetries=F.symmetry_operators(group))
    h, sc, deq = T.quantize_f16(fft)
    ev, ew = errors(s, deq, spaces, desc, table, box, half=(h, sc))
    print(f"box {box} r {radius} {group} FP16: volume {ev:.2e}, weights {ew:.2e}")
    assert ev <= 1e-4 and ew <= 1e-4


def test_d7_fp16_vs_fp64():
    s = T.settings()
    box, n = 256, 12
    fft, spaces, desc, table = T.problem(box, n, s, seed=81,
                                         symmetries=F.symmetry_operators("D7"))
    assert len(spaces) == 14 * n
    h, sc, deq = T.quantize_f16(fft)
    ev, ew = errors(s, deq, spaces, desc, table, box, half=(h, sc))
    print(f"box {box} D7 FP16: volume {ev:.2e}, weights {ew:.2e}")
    assert ev <= 1e-4 and ew <= 1e-4

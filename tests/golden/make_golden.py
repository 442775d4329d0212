"""Write the golden vectors under tests/golden/ from the reference's own compiled CPU code.

Run in a container that holds the reference tree (it is compiled into oracle/_ref by
`make -C oracle ref`):

    python tests/golden/make_golden.py

Each fixture stores inputs and the reference outputs (no reference source):
  fr_ref_cases.npz   the 12 FR cases of the reference tests (tests/operations/
                     fourier_reconstruction/fr_tests.hpp, afr_common.hpp:316-427)
  fr_random.npz      random slices / orientations, every interpolation mode
  fr_box32.npz       config-1 value distribution at a small box: 40 slices 17x32, 33^3 volume
  fp_ref_cases.npz   the 9 FP cases of fourier_processing/fp_tests.hpp
  ts_random.npz      traverse spaces of random rotations
  fr_orders.npz      blob orders 1-4 (and 0) of every blob mode, dispatched on the order by
                     the reference's ExpandBools (fr_cpu.cpp:34-43); the order-0-only
                     KaiserFourierValue (kaiser.hpp:66-81) makes cIw0 NaN for orders 1-4
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as P  # noqa: E402


def fr_case(R, s, P_box, fft, rots, fast, table):
    """Run the reference FR insertion on fresh zero volumes."""
    n = len(rots)
    sp = np.zeros(n, dtype=P.TS_DTYPE)
    for i in range(n):
        sp[i] = R.compute_traverse_space(P_box // 2, P_box, rots[i], P_box, P_box, fast,
                                         s.blobRadius, 1.0)
        sp[i]["projectionIndex"] = i
    O = P.Oracle()
    c = O.create_constants(P_box + 1, P_box + 1, 0 if table is None else table.size, s)
    vol = np.zeros((P_box + 1,) * 3, np.complex64)
    w = np.zeros((P_box + 1,) * 3, np.float32)
    ok = R.fr_execute(s, vol, w, fft, sp, table, c)
    assert ok
    return sp, vol, w


def reference_fr_cases(R):
    """The reference's 12 tests: 5x6 image (3x6 Hermitian FFT), 7^3 volume."""
    out = {}
    O = P.Oracle()
    planes = {
        "XY": np.eye(3, dtype=np.float32),
        "YZ": np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], np.float32),
        "XZ": np.array([[0, -1, 0], [0, 0, 1], [-1, 0, 0]], np.float32),
    }
    for plane, t in planes.items():
        for typ, tname in ((P.K_PRECISE, "Precise"), (P.K_FAST, "Fast")):
            for interp, iname in ((P.K_DYNAMIC, "Dynamic"), (P.K_LOOKUP, "Lookup")):
                s = P.Settings(interp, 0, typ, 15.0, 1.9)
                # FFT payload: FourierDescriptor(Size(5,6,1,1)) -> 3 x 6 complex
                n_el = 3 * 6
                if plane == "XY":
                    flat = np.ones(2 * n_el, np.float32)
                else:
                    flat = np.arange(1, 2 * n_el + 1, dtype=np.float32)
                fft = flat.view(np.complex64).reshape(1, 6, 3)
                table = O.fill_blob_table(10000, 7, 7, s) if interp == P.K_LOOKUP else None
                sp = np.zeros(1, dtype=P.TS_DTYPE)
                # FillTraverseSpace: imgSizeX = 6/2, imgSizeY = 6, max index 6
                sp[0] = R.compute_traverse_space(3, 6, t, 6, 6, typ == P.K_FAST, 1.9, 1.0)
                c = O.create_constants(7, 7, 0 if table is None else table.size, s)
                vol = np.zeros((7, 7, 7), np.complex64)
                w = np.zeros((7, 7, 7), np.float32)
                assert R.fr_execute(s, vol, w, fft, sp, table, c)
                key = f"{plane}Plane{tname}{iname}"
                out[key + "/fft"] = fft
                out[key + "/space"] = sp
                out[key + "/table"] = np.zeros(0, np.float32) if table is None else table
                out[key + "/volume"] = vol
                out[key + "/weights"] = w
                out[key + "/settings"] = np.array([interp, 0, typ, 15.0, 1.9], np.float64)
    return out


def random_fr(R):
    out = {}
    O = P.Oracle()
    modes = [("lookup", P.Settings(P.K_LOOKUP, 0, P.K_PRECISE, 15.0, 1.9)),
             ("fastkaiser", P.Settings(P.K_DYNAMIC, 0, P.K_PRECISE, 15.0, 1.9)),
             ("precisekaiser", P.Settings(P.K_DYNAMIC, 0, P.K_PRECISE, 16.0, 1.9)),
             ("nearest", P.Settings(P.K_LOOKUP, 0, P.K_FAST, 15.0, 1.9)),
             ("lookup_r25", P.Settings(P.K_LOOKUP, 0, P.K_PRECISE, 15.0, 2.5)),
             ("lookup_r15", P.Settings(P.K_LOOKUP, 0, P.K_PRECISE, 15.0, 1.5))]
    for k, (name, s) in enumerate(modes):
        Pb = 16
        rng = np.random.default_rng(100 + k)
        n = 12
        fft = (rng.standard_normal((n, Pb, Pb // 2 + 1))
               + 1j * rng.standard_normal((n, Pb, Pb // 2 + 1))).astype(np.complex64)
        rots = P.uniform_rotations(n, 200 + k).astype(np.float32)
        table = O.fill_blob_table(10000, Pb + 1, Pb + 1, s) if s.interpolation == P.K_LOOKUP \
            else None
        sp, vol, w = fr_case(R, s, Pb, fft, rots, s.type == P.K_FAST, table)
        out[name + "/fft"] = fft
        out[name + "/spaces"] = sp
        out[name + "/table"] = np.zeros(0, np.float32) if table is None else table
        out[name + "/volume"] = vol
        out[name + "/weights"] = w
        out[name + "/settings"] = np.array([s.interpolation, s.order, s.type, s.alpha,
                                            s.blobRadius], np.float64)
    return out


def orders_fr(R):
    """Blob orders 0-4 in every blob mode: kDynamic with alpha 16 (precise Kaiser,
    fr_common_kernels.hpp:129-161), kDynamic with alpha 15 (kaiserValueFast), kLookup (table
    from FillBlobTable's arithmetic), and the nearest-pixel traversal (order ignored)."""
    out = {}
    O = P.Oracle()
    modes = [("precise", P.K_DYNAMIC, P.K_PRECISE, 16.0),
             ("fast", P.K_DYNAMIC, P.K_PRECISE, 15.0),
             ("lookup", P.K_LOOKUP, P.K_PRECISE, 15.0),
             ("nearest", P.K_DYNAMIC, P.K_FAST, 16.0)]
    for k, (mname, interp, typ, alpha) in enumerate(modes):
        for order in range(5):
            s = P.Settings(interp, order, typ, alpha, 1.9)
            Pb, n = 16, 8
            rng = np.random.default_rng(300 + 10 * k + order)
            fft = (rng.standard_normal((n, Pb, Pb // 2 + 1))
                   + 1j * rng.standard_normal((n, Pb, Pb // 2 + 1))).astype(np.complex64)
            rots = P.uniform_rotations(n, 400 + 10 * k + order).astype(np.float32)
            table = O.fill_blob_table(10000, Pb + 1, Pb + 1, s) if interp == P.K_LOOKUP \
                else None
            sp, vol, w = fr_case(R, s, Pb, fft, rots, typ == P.K_FAST, table)
            name = f"{mname}_o{order}"
            out[name + "/fft"] = fft
            out[name + "/spaces"] = sp
            out[name + "/table"] = np.zeros(0, np.float32) if table is None else table
            out[name + "/volume"] = vol
            out[name + "/weights"] = w
            out[name + "/settings"] = np.array([s.interpolation, s.order, s.type, s.alpha,
                                                s.blobRadius], np.float64)
    return out


def box32(R):
    """Config-1 value distribution (complex Gaussian times 1/(1+k^2)) at P = 32."""
    O = P.Oracle()
    s = P.Settings(P.K_LOOKUP, 0, P.K_PRECISE, 15.0, 1.9)
    Pb, n = 32, 40
    rng = np.random.default_rng(42)
    hx, hy = Pb // 2 + 1, Pb
    z = (rng.standard_normal((n, hy, hx)) + 1j * rng.standard_normal((n, hy, hx))) / np.sqrt(2)
    kx = np.arange(hx)[None, :]
    ky = np.arange(hy)[:, None] - Pb // 2
    fft = (z / (1.0 + kx ** 2 + ky ** 2)).astype(np.complex64)
    rots = P.uniform_rotations(n, 43).astype(np.float32)
    table = O.fill_blob_table(10000, Pb + 1, Pb + 1, s)
    sp, vol, w = fr_case(R, s, Pb, fft, rots, False, table)
    return {"fft": fft, "spaces": sp, "table": table, "volume": vol, "weights": w}


def reference_fp_cases(R):
    """fp_tests.hpp: input (i, i), filter alternating -1 / 0.5, out-of-place."""
    cases = [
        ("ImageCropping", (10, 10, 1, 1), (5, 5, 1, 1), {}),
        ("ImageNoCropping", (10, 10, 1, 1), (10, 10, 1, 1), {}),
        ("ImageNormalize", (10, 10, 1, 1), (10, 10, 1, 1), {"normalize": True}),
        ("ImageCentering", (12, 12, 1, 1), (12, 12, 1, 1), {"center": True}),
        ("ImageCropNormalizeCenter", (40, 40, 1, 5), (20, 20, 1, 5),
         {"normalize": True, "center": True}),
        ("Filtering", (12, 12, 1, 1), (12, 12, 1, 1), {"apply_filter": True}),
        ("MaxFreq", (12, 12, 1, 1), (12, 12, 1, 1), {"crop_freq": True, "max_freq": 0.25}),
        ("ShiftEven", (6, 4, 1, 1), (6, 4, 1, 1), {"shift": True}),
        ("ShiftOdd", (4, 3, 1, 1), (4, 3, 1, 1), {"shift": True}),
    ]
    out = {}
    for name, ins, outs, kw in cases:
        fin = [ins[0] // 2 + 1, ins[1], ins[2], ins[3]]
        fout = [outs[0] // 2 + 1, outs[1], outs[2], outs[3]]
        total = int(np.prod(fin))
        x = (np.arange(total, dtype=np.float32) * (1 + 1j)).astype(np.complex64)
        filt = np.zeros(int(np.prod(fout[:3])), np.float32)
        filt[0::2] = -1.0
        filt[1::2] = 0.5
        norm = 1.0 / (ins[0] * ins[1] * ins[2])
        y = R.fp_execute(x, fin, list(ins), fout, filt=filt, norm_factor=norm, **kw)
        out[name + "/in"] = x
        out[name + "/filter"] = filt
        out[name + "/out"] = y
        out[name + "/sizes"] = np.array([fin, list(ins), fout], np.uint64)
        out[name + "/flags"] = np.array([kw.get("apply_filter", 0), kw.get("normalize", 0),
                                         kw.get("center", 0), kw.get("crop_freq", 0),
                                         kw.get("shift", 0)], np.int32)
        out[name + "/scalars"] = np.array([norm, kw.get("max_freq", 0.0)], np.float32)
    return out


def random_ts(R):
    rots = P.uniform_rotations(64, 5).astype(np.float32)
    recs = []
    params = []
    for k, m in enumerate(rots):
        Pb = [6, 64, 128, 256][k % 4]
        fast = k % 3 == 0
        r = [1.9, 1.5, 2.5][k % 3]
        recs.append(R.compute_traverse_space(Pb // 2, Pb, m, Pb, Pb, fast, r, 1.0))
        params.append([Pb, int(fast), r])
    return {"rotations": rots, "params": np.array(params, np.float64),
            "spaces": np.array(recs, dtype=P.TS_DTYPE)}


def main():
    P.build(with_ref=True)
    R = P.Reference()
    np.savez_compressed(os.path.join(HERE, "fr_ref_cases.npz"), **reference_fr_cases(R))
    np.savez_compressed(os.path.join(HERE, "fr_random.npz"), **random_fr(R))
    np.savez_compressed(os.path.join(HERE, "fr_box32.npz"), **box32(R))
    np.savez_compressed(os.path.join(HERE, "fp_ref_cases.npz"), **reference_fp_cases(R))
    np.savez_compressed(os.path.join(HERE, "ts_random.npz"), **random_ts(R))
    np.savez_compressed(os.path.join(HERE, "fr_orders.npz"), **orders_fr(R))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()

"""Pinning the oracle (CPU, no GPU): reference golden vectors, the reference's own known-answer
tests, closed forms, and brute force on tiny inputs."""
import itertools
import math

import numpy as np
import pytest

import pyoracle as P


def ulps(a, b):
    a = np.float32(a).view(np.int32).astype(np.int64)
    b = np.float32(b).view(np.int32).astype(np.int64)
    return abs(int(a) - int(b))


def float_eq(a, b):
    """gtest ASSERT_FLOAT_EQ: within 4 ulps."""
    return ulps(a, b) <= 4


# ---- traverse spaces: the reference KATs (afr_common.hpp:136-254) ---------------------------
PLANES = {
    "XY": np.eye(3, dtype=np.float32),
    "YZ": np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], np.float32),
    "XZ": np.array([[0, -1, 0], [0, 0, 1], [-1, 0, 0]], np.float32),
}
KATS = {  # (minY, minX, minZ, maxY, maxX, maxZ), maxDistSqr, dir, normal, top, bottom
    ("XY", True): ((0, 3, 3, 6, 6, 3), 9.0, 0, (0, 0, 1), (6, 0, 3), (6, 0, 3)),
    ("XY", False): ((0, 1, 1, 6, 6, 5), 24.0100002, 0, (0, 0, 1), (7.9, -1.9, 1.1),
                    (7.9, -1.9, 4.9)),
    ("YZ", True): ((0, 3, 0, 6, 3, 3), 9.0, 2, (1, 0, 0), (3, 0, 0), (3, 0, 0)),
    ("YZ", False): ((0, 1, 0, 6, 5, 5), 24.0100002, 2, (1, 0, 0), (1.1, -1.9, -1.9),
                    (4.9, -1.9, -1.9)),
    ("XZ", True): ((3, 0, 0, 3, 6, 3), 9.0, 1, (0, 1, 0), (6, 3, 0), (6, 3, 0)),
    ("XZ", False): ((1, 0, 0, 5, 6, 5), 24.0100002, 1, (0, 1, 0), (7.9, 1.1, -1.9),
                    (7.9, 4.9, -1.9)),
}


def check_kat(s, kat):
    box, dist, d, normal, top, bottom = kat
    assert (s["minY"], s["minX"], s["minZ"], s["maxY"], s["maxX"], s["maxZ"]) == box
    assert float_eq(s["maxDistanceSqr"], dist)
    assert s["dir"] == d
    for got, want in ((s["unitNormal"], normal), (s["topOrigin"], top),
                      (s["bottomOrigin"], bottom)):
        assert all(float_eq(g, w) for g, w in zip(got, want)), (got, want)
    assert float_eq(s["weight"], 1.0) and s["projectionIndex"] == 0


@pytest.mark.parametrize("plane,fast", list(itertools.product(PLANES, [True, False])))
def test_traverse_space_reference_kats(plane, fast, oracle):
    s = oracle.compute_traverse_space(3, 6, PLANES[plane], 6, 6, fast, 1.9, 1.0)
    check_kat(s, KATS[(plane, fast)])


def test_traverse_space_closed_forms(oracle):
    """transformInv = T^T, normal = T z, maxDistanceSqr = (P/2 + r)^2, AABB covers the slab."""
    rots = P.uniform_rotations(200, 11).astype(np.float32)
    for m in rots:
        s = oracle.compute_traverse_space(64, 128, m, 128, 128, False, 1.9, 1.0)
        assert np.array_equal(s["transformInv"], m.T)
        assert np.allclose(s["unitNormal"], m[:, 2], atol=1e-6)
        assert s["maxDistanceSqr"] == np.float32(64 + np.float32(1.9)) ** 2
        n = np.abs(m[:, 2])
        assert s["dir"] == (2 if n[0] >= max(n[1], n[2]) else 1 if n[1] >= n[2] else 0)
        # the slab corners, rotated and centred, lie inside the integer AABB
        corners = np.array([[x, y, z] for x in (-1.9, 65.9) for y in (-65.9, 65.9)
                            for z in (-1.9, 1.9)]) @ m.T.astype(np.float64) + 64.0
        lo = np.clip(corners.min(0), 0, 128)
        hi = np.clip(corners.max(0), 0, 128)
        assert s["minX"] <= lo[0] + 1e-4 and s["maxX"] >= hi[0] - 1e-4
        assert s["minY"] <= lo[1] + 1e-4 and s["maxY"] >= hi[1] - 1e-4
        assert s["minZ"] <= lo[2] + 1e-4 and s["maxZ"] >= hi[2] - 1e-4


def test_traverse_space_golden(oracle, golden):
    g = golden("ts_random.npz")
    for m, (box, fast, r), want in zip(g["rotations"], g["params"], g["spaces"]):
        box = int(box)
        got = oracle.compute_traverse_space(box // 2, box, m, box, box, bool(fast), float(r))
        assert got.tobytes() == want.tobytes()


# ---- constants and blob table: closed forms ---------------------------------------------
def test_constants_closed_form(oracle):
    import mpmath as mp
    s = P.Settings(P.K_LOOKUP, 0, P.K_PRECISE, 15.0, 1.9)
    c = oracle.create_constants(129, 129, 10000, s)
    r, a = mp.mpf(float(np.float32(1.9))), mp.mpf(15)
    # Fourier transform of the 3-D Kaiser-Bessel blob at 0 (Lewitt 1990, order 0):
    # (2 pi)^{3/2} r^3 I_{3/2}(alpha) / (I0(alpha) alpha^{3/2})
    fv = (2 * mp.pi) ** 1.5 * r ** 3 * mp.besseli(1.5, a) / (mp.besseli(0, a) * a ** 1.5)
    assert abs(float(c.cIw0) - float(1 / fv)) / float(1 / fv) < 2e-6
    assert abs(c.cIw0 - 0.5811907) < 1e-6
    assert c.cIDeltaSqrt == np.float32(np.float32(9999) / (np.float32(1.9) * np.float32(1.9)))
    assert abs(c.cOneOverBessiOrderAlpha * float(mp.besseli(0, a)) - 1) < 1e-6
    assert c.cMaxVolumeIndexX == 128 and c.cMaxVolumeIndexYZ == 128
    assert c.cOneOverBlobRadiusSqr == np.float32(1) / (np.float32(1.9) * np.float32(1.9))


def test_blob_table_closed_form(oracle):
    import mpmath as mp
    s = P.Settings(P.K_LOOKUP, 0, P.K_PRECISE, 15.0, 1.9)
    t = oracle.fill_blob_table(10000, 129, 129, s)
    c = oracle.create_constants(129, 129, 10000, s)
    assert t[0] == c.cIw0                      # KaiserValue(0) = 1 exactly
    assert np.all(np.diff(t) <= 0)             # monotone in the squared distance
    r, a = 1.9, 15.0
    for i in (1, 17, 2500, 5000, 7777, 9998, 9999):
        d = r * math.sqrt(i / 9999)
        want = float(mp.besseli(0, a * mp.sqrt(1 - (d / r) ** 2)) / mp.besseli(0, a)) * c.cIw0
        # near the table end 1 - (d/r)^2 cancels in float: the reference's own precision
        tol = 2e-6 if i < 9000 else 2e-5
        assert abs(t[i] - want) / want < tol, (i, t[i], want)
    assert abs(t[5000] - 0.0085697) < 1e-6
    assert abs(t[9999] - c.cIw0 / float(mp.besseli(0, a))) < 1e-12


def test_bessel_against_mpmath(oracle):
    import mpmath as mp
    for x in (0.1, 1.0, 3.74, 3.76, 7.5, 15.0):
        for k, f in enumerate((oracle.lib.orc_bessi0, oracle.lib.orc_bessi1)):
            want = float(mp.besseli(k, x))
            assert abs(f(x) - want) / want < 3e-7   # A&S 9.8.1-9.8.4 accuracy


def test_oracle_bessel_bit_exact_vs_reference(oracle, reference):
    xs = np.concatenate([np.linspace(-20, 20, 401), [0.0, 3.75, -3.75]])
    for k in range(5):
        f = getattr(oracle.lib, f"orc_bessi{k}")
        g = getattr(reference.lib, f"ref_bessi{k}")
        for x in xs:
            assert np.float64(f(x)).tobytes() == np.float64(g(x)).tobytes()


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("alpha", [15.0, 16.0])
def test_kaiser_value_orders_closed_form(oracle, order, alpha):
    """KaiserValue for blob orders 1-4 (kaiser.hpp:30-48) against its closed form
    s^k I_k(alpha s) / I_k(alpha), s = sqrt(1 - (d/r)^2), with mpmath's Bessel I_k.  The
    reference evaluates I_2..I_4 by the upward recurrence from polynomial I_0, I_1
    (bessier.hpp:77-84), which loses relative precision for small arguments: the relative
    bound holds up to d = 0.95 r, an absolute bound (against a peak value of 1) everywhere."""
    import mpmath as mp
    r = 1.9
    for i in list(range(0, 10000, 41)) + [9990, 9998, 9999]:
        d = r * math.sqrt(i / 9999)
        sv = mp.sqrt(1 - (mp.mpf(d) / r) ** 2)
        want = float(sv ** order * mp.besseli(order, alpha * sv) / mp.besseli(order, alpha))
        got = oracle.lib.orc_kaiser_value(d, r, alpha, order)
        assert abs(got - want) <= 1e-6, (order, i, got, want)
        if d <= 0.95 * r:
            assert abs(got - want) <= 1e-5 * want, (order, i, got, want)
    assert oracle.lib.orc_kaiser_value(r * 1.001, r, alpha, order) == 0.0


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_orders_make_constants_and_table_nan(oracle, order):
    """KaiserFourierValue exists for order 0 only and returns NaN otherwise (kaiser.hpp:
    73-80), so cIw0 = 1 / NaN and every table entry KaiserValue * cIw0 is NaN (afr.cpp:14,
    22-34); 1 / I_order(alpha) stays finite."""
    s = P.Settings(P.K_LOOKUP, order, P.K_PRECISE, 15.0, 1.9)
    c = oracle.create_constants(33, 33, 10000, s)
    assert math.isnan(c.cIw0)
    assert math.isfinite(c.cOneOverBessiOrderAlpha) and c.cOneOverBessiOrderAlpha > 0
    t = oracle.fill_blob_table(10000, 33, 33, s)
    assert np.isnan(t).all()


# ---- FR insertion: golden vectors of the compiled reference --------------------------------
def _fr_golden_case(oracle, g, key, fft, spaces, table, box):
    interp, order, typ, alpha, radius = g[key + "/settings"] if key else (1, 0, 1, 15.0, 1.9)
    s = P.Settings(int(interp), int(order), int(typ), float(alpha), float(radius))
    c = oracle.create_constants(box + 1, box + 1, 0 if table is None else table.size, s)
    v = np.zeros((box + 1,) * 3, np.complex64)
    w = np.zeros((box + 1,) * 3, np.float32)
    oracle.fr_execute(s, v, w, fft, spaces, table, c)
    return v, w


def test_fr_golden_reference_cases(oracle, golden):
    g = golden("fr_ref_cases.npz")
    keys = sorted({k.split("/")[0] for k in g.files})
    assert len(keys) == 12
    for key in keys:
        table = g[key + "/table"]
        v, w = _fr_golden_case(oracle, g, key, g[key + "/fft"], g[key + "/space"],
                               table if table.size else None, 6)
        assert v.tobytes() == g[key + "/volume"].tobytes(), key
        assert w.tobytes() == g[key + "/weights"].tobytes(), key


def test_fr_golden_random_modes(oracle, golden):
    g = golden("fr_random.npz")
    for key in sorted({k.split("/")[0] for k in g.files}):
        table = g[key + "/table"]
        fft = g[key + "/fft"]
        v, w = _fr_golden_case(oracle, g, key, fft, g[key + "/spaces"],
                               table if table.size else None, fft.shape[1])
        assert v.tobytes() == g[key + "/volume"].tobytes(), key
        assert w.tobytes() == g[key + "/weights"].tobytes(), key


def test_fr_golden_orders(oracle, golden):
    """Blob orders 0-4 in every blob mode and the nearest-pixel traversal, bit for bit (NaN
    bits included) against the compiled reference: for orders 1-4 every voxel a blob reaches
    is NaN (cIw0 is NaN), the nearest-pixel traversal ignores the order."""
    g = golden("fr_orders.npz")
    keys = sorted({k.split("/")[0] for k in g.files})
    assert len(keys) == 20
    for key in keys:
        table = g[key + "/table"]
        fft = g[key + "/fft"]
        v, w = _fr_golden_case(oracle, g, key, fft, g[key + "/spaces"],
                               table if table.size else None, fft.shape[1])
        assert v.tobytes() == g[key + "/volume"].tobytes(), key
        assert w.tobytes() == g[key + "/weights"].tobytes(), key
        gw = g[key + "/weights"]
        order = int(g[key + "/settings"][1])
        if key.startswith("nearest") or order == 0:
            assert not np.isnan(gw).any() and (gw != 0).any(), key
        else:
            assert (np.isnan(gw) == (gw != 0)).all() and np.isnan(gw).any(), key


def test_fr_golden_box32(oracle, golden):
    g = golden("fr_box32.npz")
    v, w = _fr_golden_case(oracle, g, None, g["fft"], g["spaces"], g["table"], 32)
    assert v.tobytes() == g["volume"].tobytes() and w.tobytes() == g["weights"].tobytes()


def test_fr_oracle_vs_compiled_reference(oracle, reference):
    """Fresh random problems in every mode, bit for bit against the compiled reference."""
    for k, (interp, typ, alpha, r) in enumerate(itertools.product(
            (0, 1), (0, 1), (15.0, 16.0), (1.5, 1.9, 2.5))):
        s = P.Settings(interp, 0, typ, alpha, r)
        box, n = 24, 8
        rng = np.random.default_rng(1000 + k)
        fft = (rng.standard_normal((n, box, box // 2 + 1)) +
               1j * rng.standard_normal((n, box, box // 2 + 1))).astype(np.complex64)
        sp = np.zeros(n, dtype=P.TS_DTYPE)
        for i, m in enumerate(P.uniform_rotations(n, 2000 + k).astype(np.float32)):
            sp[i] = reference.compute_traverse_space(box // 2, box, m, box, box, typ == 0, r)
            sp[i]["projectionIndex"] = i
        table = oracle.fill_blob_table(10000, box + 1, box + 1, s) if interp else None
        c = oracle.create_constants(box + 1, box + 1, 10000 if interp else 0, s)
        v1 = np.zeros((box + 1,) * 3, np.complex64)
        w1 = np.zeros((box + 1,) * 3, np.float32)
        v2, w2 = v1.copy(), w1.copy()
        oracle.fr_execute(s, v1, w1, fft, sp, table, c)
        assert reference.fr_execute(s, v2, w2, fft, sp, table, c)
        assert v1.tobytes() == v2.tobytes() and w1.tobytes() == w2.tobytes()


# ---- FR insertion: the reference's zero/non-zero pattern check (afr_common.hpp:256-314) ----
def test_fr_zero_nonzero_pattern(oracle, golden):
    g = golden("fr_ref_cases.npz")
    for key in sorted({k.split("/")[0] for k in g.files}):
        interp, order, typ, alpha, radius = g[key + "/settings"]
        sp = g[key + "/space"][0]
        table = g[key + "/table"]
        v, w = _fr_golden_case(oracle, g, key, g[key + "/fft"], g[key + "/space"],
                               table if table.size else None, 6)
        max_dist = 0.0 if typ == P.K_FAST else np.float32(radius)
        n = sp["unitNormal"]
        for z, y, x in itertools.product(range(7), repeat=3):
            p = np.array([x - 3, y - 3, z - 3], np.float32)
            if float(p @ p) > sp["maxDistanceSqr"]:
                assert v[z, y, x] == 0 and w[z, y, x] == 0
                continue
            q = sp["transformInv"].astype(np.float32) @ p
            off_plane = abs(float(n @ p)) > max_dist
            if q[0] < -max_dist or off_plane:
                assert v[z, y, x] == 0 and w[z, y, x] == 0, (key, x, y, z)
            else:
                assert v[z, y, x].real != 0 and v[z, y, x].imag != 0 and w[z, y, x] != 0, \
                    (key, x, y, z)


# ---- FR insertion: brute force on tiny inputs ------------------------------------------------
def test_fr_brute_force_gather(oracle):
    """Per voxel, a direct sum over *every* slice pixel of the ProcessVoxelBlob predicate in
    float32 (numpy scalars, row-major pixel order) reproduces the oracle bit for bit: the
    traversal, slab culling and blob bounding box only select which pixels are looked at."""
    f32 = np.float32
    box, n, r = 8, 3, f32(1.9)
    s = P.Settings(P.K_LOOKUP, 0, P.K_PRECISE, 15.0, 1.9)
    rng = np.random.default_rng(5)
    hx, hy = box // 2 + 1, box
    fft = (rng.standard_normal((n, hy, hx)) + 1j * rng.standard_normal((n, hy, hx))).astype(
        np.complex64)
    sp = np.zeros(n, dtype=P.TS_DTYPE)
    for i, m in enumerate(P.uniform_rotations(n, 6).astype(np.float32)):
        sp[i] = oracle.compute_traverse_space(box // 2, box, m, box, box, False, 1.9)
        sp[i]["projectionIndex"] = i
    table = oracle.fill_blob_table(10000, box + 1, box + 1, s)
    c = oracle.create_constants(box + 1, box + 1, 10000, s)
    v, w = np.zeros((box + 1,) * 3, np.complex64), np.zeros((box + 1,) * 3, np.float32)
    oracle.fr_execute(s, v, w, fft, sp, table, c)

    bv = np.zeros_like(v)
    bw = np.zeros_like(w)
    r2 = r * r
    half = box // 2
    for k in range(n):
        t = sp[k]["transformInv"].astype(np.float32)
        for z, y, x in itertools.product(range(box + 1), repeat=3):
            px, py, pz = f32(x - half), f32(y - half), f32(z - half)
            if px * px + py * py + pz * pz > sp[k]["maxDistanceSqr"]:
                continue
            qx = t[0, 0] * px + t[0, 1] * py + t[0, 2] * pz
            qy = t[1, 0] * px + t[1, 1] * py + t[1, 2] * pz
            qz = t[2, 0] * px + t[2, 1] * py + t[2, 2] * pz
            if qx < -r:
                continue
            qy = qy + f32(half)
            zs = qz * qz
            if zs > r2:
                continue
            acc_w, acc_r, acc_i = f32(0), f32(0), f32(0)
            for i in range(hy):
                yz = (qy - f32(i)) * (qy - f32(i)) + zs
                if yz > r2:
                    continue
                for j in range(hx):
                    xd = qx - f32(j)
                    d2 = xd * xd + yz
                    if d2 > r2:
                        continue
                    wt = table[int(d2 * c.cIDeltaSqrt + f32(0.5))] * f32(1.0)
                    acc_w = acc_w + wt
                    acc_r = acc_r + fft[k, i, j].real * wt
                    acc_i = acc_i + fft[k, i, j].imag * wt
            bv[z, y, x] = np.complex64(complex(bv[z, y, x].real + acc_r,
                                               bv[z, y, x].imag + acc_i))
            bw[z, y, x] = bw[z, y, x] + acc_w
    assert bv.tobytes() == v.tobytes() and bw.tobytes() == w.tobytes()


def test_fr_linearity_and_f64(oracle):
    """Volume is linear in slice values; weights ignore them.  FP32 vs FP64: the table index
    is computed from a float squared distance, so neighbouring table entries get picked near
    index boundaries; the resulting FP32 error is a few 1e-5 (stated, bounded by 1e-4)."""
    s = P.Settings(P.K_LOOKUP, 0, P.K_PRECISE, 15.0, 1.9)
    box, n = 32, 30
    rng = np.random.default_rng(3)
    a = (rng.standard_normal((n, box, 17)) + 1j * rng.standard_normal((n, box, 17))).astype(
        np.complex64)
    b = (rng.standard_normal((n, box, 17)) + 1j * rng.standard_normal((n, box, 17))).astype(
        np.complex64)
    sp = np.zeros(n, dtype=P.TS_DTYPE)
    for i, m in enumerate(P.uniform_rotations(n, 4).astype(np.float32)):
        sp[i] = oracle.compute_traverse_space(16, 32, m, 32, 32, False, 1.9)
        sp[i]["projectionIndex"] = i
    table = oracle.fill_blob_table(10000, 33, 33, s)
    c = oracle.create_constants(33, 33, 10000, s)

    def run(f):
        v, w = np.zeros((33,) * 3, np.complex64), np.zeros((33,) * 3, np.float32)
        oracle.fr_execute(s, v, w, f, sp, table, c)
        return v, w
    va, wa = run(a)
    vb, wb = run(b)
    vab, wab = run((a + b).astype(np.complex64))
    assert wa.tobytes() == wb.tobytes() == wab.tobytes()
    err = np.linalg.norm(vab - (va + vb)) / np.linalg.norm(vab)
    assert err < 1e-6
    dv = np.zeros((33,) * 3, np.complex128)
    dw = np.zeros((33,) * 3, np.float64)
    oracle.fr_execute_f64(s, dv, dw, a, sp, table, c)
    assert np.linalg.norm(va - dv) / np.linalg.norm(dv) < 1e-4
    assert np.linalg.norm(wa - dw) / np.linalg.norm(dw) < 1e-4


def _axis_rotations():
    """Rotations whose entries are 0 and +-1: with them q = T^T p is exact in any precision."""
    mats = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3), np.float32)
            for i, j in enumerate(perm):
                m[i, j] = signs[i]
            if round(float(np.linalg.det(m))) == 1:
                mats.append(m)
    return mats


def test_fr_f64_exact_on_integer_inputs(oracle):
    """On inputs where FP32 arithmetic is exact -- axis-aligned rotations (q = T^T p exactly),
    integer slice values, an integer-valued table and integer space weights -- the FP64
    oracle must equal the FP32 one exactly.  The weights 1, 2 and 3 of the spaces make a
    dropped or misplaced space weight visible, the non-constant table a wrong table index,
    and the 24 rotations a transposed operand."""
    s = P.Settings(P.K_LOOKUP, 0, P.K_PRECISE, 15.0, 1.9)
    box = 16
    rots = _axis_rotations()
    assert len(rots) == 24
    n = len(rots)
    rng = np.random.default_rng(8)
    hx, hy = box // 2 + 1, box
    fft = (rng.integers(-4, 5, (n, hy, hx)) + 1j * rng.integers(-4, 5, (n, hy, hx))).astype(
        np.complex64)
    sp = np.zeros(n, dtype=P.TS_DTYPE)
    for i, m in enumerate(rots):
        sp[i] = oracle.compute_traverse_space(box // 2, box, m, box, box, False, 1.9,
                                              float(1 + i % 3))
        sp[i]["projectionIndex"] = i
    table = (np.arange(10000) % 5 + 1).astype(np.float32)
    c = oracle.create_constants(box + 1, box + 1, 10000, s)
    v = np.zeros((box + 1,) * 3, np.complex64)
    w = np.zeros((box + 1,) * 3, np.float32)
    oracle.fr_execute(s, v, w, fft, sp, table, c)
    dv = np.zeros((box + 1,) * 3, np.complex128)
    dw = np.zeros((box + 1,) * 3, np.float64)
    oracle.fr_execute_f64(s, dv, dw, fft, sp, table, c)
    assert (w != 0).sum() > 500
    assert np.array_equal(dv, v.astype(np.complex128))
    assert np.array_equal(dw, w.astype(np.float64))
    # the same records with all weights 1 give a different volume: the weight is used
    sp1 = sp.copy()
    sp1["weight"] = 1.0
    dw1 = np.zeros_like(dw)
    oracle.fr_execute_f64(s, np.zeros_like(dv), dw1, fft, sp1, table, c)
    assert not np.array_equal(dw1, dw)


def test_fr_f64_non_unit_weights(oracle):
    """Random orientations with space weights drawn from [0.25, 4]: the FP64 oracle agrees
    with the FP32 one to FP32 accumulation error, and scaling every weight by 2 scales the
    FP64 volume and weights by exactly 2 (a power of two changes no rounding)."""
    s = P.Settings(P.K_LOOKUP, 0, P.K_PRECISE, 15.0, 1.9)
    box, n = 24, 20
    rng = np.random.default_rng(12)
    fft = (rng.standard_normal((n, box, box // 2 + 1)) +
           1j * rng.standard_normal((n, box, box // 2 + 1))).astype(np.complex64)
    sp = np.zeros(n, dtype=P.TS_DTYPE)
    wts = rng.uniform(0.25, 4.0, n).astype(np.float32)
    for i, m in enumerate(P.uniform_rotations(n, 13).astype(np.float32)):
        sp[i] = oracle.compute_traverse_space(box // 2, box, m, box, box, False, 1.9,
                                              float(wts[i]))
        sp[i]["projectionIndex"] = i
    table = oracle.fill_blob_table(10000, box + 1, box + 1, s)
    c = oracle.create_constants(box + 1, box + 1, 10000, s)

    def f64(spaces):
        dv = np.zeros((box + 1,) * 3, np.complex128)
        dw = np.zeros((box + 1,) * 3, np.float64)
        oracle.fr_execute_f64(s, dv, dw, fft, spaces, table, c)
        return dv, dw
    v = np.zeros((box + 1,) * 3, np.complex64)
    w = np.zeros((box + 1,) * 3, np.float32)
    oracle.fr_execute(s, v, w, fft, sp, table, c)
    dv, dw = f64(sp)
    assert np.linalg.norm(v - dv) / np.linalg.norm(dv) < 1e-4
    assert np.linalg.norm(w - dw) / np.linalg.norm(dw) < 1e-4
    sp2 = sp.copy()
    sp2["weight"] = sp["weight"] * 2
    dv2, dw2 = f64(sp2)
    assert np.array_equal(dv2, 2 * dv) and np.array_equal(dw2, 2 * dw)


def test_fr_empty_table_rejected(oracle):
    s = P.Settings(P.K_LOOKUP, 0, P.K_PRECISE, 15.0, 1.9)
    c = oracle.create_constants(9, 9, 0, s)
    v, w = np.zeros((9,) * 3, np.complex64), np.zeros((9,) * 3, np.float32)
    sp = np.zeros(1, dtype=P.TS_DTYPE)
    assert oracle.fr_execute(s, v, w, np.zeros((1, 8, 5), np.complex64), sp, None, c) is None


# ---- FP: reference value tests (afp_common.hpp:132-178) and golden vectors -----------------
def reference_check(out, x, fin, spatial, fout, flags, norm, max_freq, filt):
    """The reference test's own restatement of the expected FP output."""
    apply_filter, normalize, center, crop, shift = flags
    ox, oy, _, on = fout
    ix, iy = fin[0], fin[1]
    for n in range(on):
        crop_index = oy // 2 + 1
        for y in range(oy):
            for x_ in range(ox):
                in_y = y if y < crop_index else iy - oy + y
                ii = n * ix * iy + in_y * ix + x_
                sy = (y + oy + oy // 2) % oy
                oi = n * ox * oy + (sy if shift else y) * ox + x_
                re, im = np.float32(x[ii].real), np.float32(x[ii].imag)
                if apply_filter:
                    re, im = re * filt[y * ox + x_], im * filt[y * ox + x_]
                if normalize:
                    re, im = re * np.float32(norm), im * np.float32(norm)
                if center:
                    cc = np.float32(1 - 2 * ((x_ + y) & 1))
                    re, im = re * cc, im * cc
                if crop:
                    def fq(i, m):
                        return ((i if i <= m / 2 else i - m)) / m
                    fx, fy = fq(x_, spatial[0]), fq(y, spatial[1])
                    if fx * fx + fy * fy > max_freq:
                        re = im = np.float32(0)
                if not (float_eq(re, out[oi].real) and float_eq(im, out[oi].imag)):
                    return False
    return True


def test_fp_reference_cases(oracle, golden):
    g = golden("fp_ref_cases.npz")
    names = sorted({k.split("/")[0] for k in g.files})
    assert len(names) == 9
    for name in names:
        fin, spatial, fout = [[int(v) for v in row] for row in g[name + "/sizes"]]
        flags = [int(v) for v in g[name + "/flags"]]
        norm, max_freq = [float(v) for v in g[name + "/scalars"]]
        x, filt = g[name + "/in"], g[name + "/filter"]
        y = oracle.fp_execute(x, fin, spatial, fout, filt=filt, norm_factor=norm,
                              max_freq=max_freq, apply_filter=flags[0], normalize=flags[1],
                              center=flags[2], crop_freq=flags[3], shift=flags[4])
        assert y.tobytes() == g[name + "/out"].tobytes(), name
        assert reference_check(y, x, fin, spatial, fout, flags, norm, max_freq, filt), name


def test_fp_oracle_vs_compiled_reference(oracle, reference):
    rng = np.random.default_rng(9)
    for flags in itertools.product([0, 1], repeat=5):
        fin, spatial, fout = [33, 32, 1, 3], [32, 32, 1, 3], [16, 30, 1, 3]
        x = (rng.standard_normal(int(np.prod(fin))) +
             1j * rng.standard_normal(int(np.prod(fin)))).astype(np.complex64)
        filt = rng.standard_normal(16 * 30).astype(np.float32)
        kw = dict(filt=filt, norm_factor=1 / 1024, max_freq=0.15, apply_filter=flags[0],
                  normalize=flags[1], center=flags[2], crop_freq=flags[3], shift=flags[4])
        assert oracle.fp_execute(x, fin, spatial, fout, **kw).tobytes() == \
            reference.fp_execute(x, fin, spatial, fout, **kw).tobytes()

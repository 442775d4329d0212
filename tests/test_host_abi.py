"""The product library on the CPU side: it loads without a GPU, exports every entry point that
include/umpalumpa_amd.h declares, and its host helpers (AFR::CreateConstants,
AFR::FillBlobTable, ComputeTraverseSpace) give the oracle's bits."""
import os
import re

import numpy as np
import pytest

import pyoracle as P
from umpalumpa_amd import _lib
from umpalumpa_amd import fr as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "umpalumpa_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(umpa_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert {p[0] for p in _lib.PROTOTYPES} == set(names)


def test_create_without_device_fails_cleanly():
    import ctypes as C
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    st = _lib.load().umpa_fr_create(C.byref(h), 0, None)
    assert st == _lib.UMPA_ERR_DEVICE and not h.value


def test_status_strings():
    lib = _lib.load()
    for st in range(8):
        assert lib.umpa_status_string(st)


@pytest.mark.parametrize("radius,alpha,n", [(1.9, 15.0, 10000), (2.5, 15.0, 10000),
                                            (1.5, 10.0, 4096), (1.9, 16.0, 777)])
def test_host_constants_and_table_match_oracle(oracle, radius, alpha, n):
    s = F.FRSettings(F.FR_LOOKUP, 0, F.FR_PRECISE, alpha, radius)
    desc = F.fr_desc(65, 128, 1, 129, 1, n)
    c = F.create_constants(desc, s)
    oc = oracle.create_constants(129, 129, n, P.Settings(1, 0, 1, alpha, radius))
    assert bytes(c) == bytes(oc)
    t = F.fill_blob_table(n, desc, s)
    assert t.tobytes() == oracle.fill_blob_table(n, 129, 129,
                                                 P.Settings(1, 0, 1, alpha, radius)).tobytes()


def test_host_traverse_spaces_match_oracle_and_golden(oracle, golden):
    g = golden("ts_random.npz")
    for m, (box, fast, r), want in zip(g["rotations"], g["params"], g["spaces"]):
        box = int(box)
        got = F.compute_traverse_space(box // 2, box, m, box, box, bool(fast), float(r))
        assert got.tobytes() == want.tobytes()
    for m in P.uniform_rotations(300, 17).astype(np.float32):
        for fast in (False, True):
            a = F.compute_traverse_space(128, 256, m, 256, 256, fast, 1.9)
            b = oracle.compute_traverse_space(128, 256, m, 256, 256, fast, 1.9)
            assert a.tobytes() == b.tobytes()


def test_traverse_spaces_caller_convention():
    """fr.cpp:278-312: spaces ordered slice-major then symmetry, projectionIndex = slice."""
    s = F.FRSettings(F.FR_LOOKUP, 0, F.FR_PRECISE, 15.0, 1.9)
    rots = P.uniform_rotations(3, 1)
    sym = np.stack([np.eye(3), np.diag([-1.0, -1.0, 1.0])])
    sp = F.traverse_spaces(rots, 64, s, sym)
    assert list(sp["projectionIndex"]) == [0, 0, 1, 1, 2, 2]
    assert np.allclose(sp[1]["transformInv"], (rots[0] @ sym[1]).T.astype(np.float32))

"""ORACLE -- test infrastructure only.

ctypes bindings for the oracle's C restatement (oracle/liboracle.so) and, where it was built,
for the reference's own CPU code compiled from the reference tree (oracle/_ref/libumpa_ref.so).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module; the
product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libumpa_ref.so")

# 112-byte TraverseSpace record (fourier_reconstruction/traverse_space.hpp:14-48)
TS_DTYPE = np.dtype([
    ("minY", "<i4"), ("minX", "<i4"), ("minZ", "<i4"),
    ("maxY", "<i4"), ("maxX", "<i4"), ("maxZ", "<i4"),
    ("maxDistanceSqr", "<f4"), ("dir", "<i4"),
    ("unitNormal", "<f4", (3,)), ("topOrigin", "<f4", (3,)), ("bottomOrigin", "<f4", (3,)),
    ("projectionIndex", "<i4"), ("transformInv", "<f4", (3, 3)), ("weight", "<f4"),
])
assert TS_DTYPE.itemsize == 112

# 32-byte Constants record (fourier_reconstruction/constants.hpp:6-16)
CONST_DTYPE = np.dtype([
    ("cMaxVolumeIndexX", "<i4"), ("cMaxVolumeIndexYZ", "<i4"), ("cBlobRadius", "<f4"),
    ("cOneOverBlobRadiusSqr", "<f4"), ("cBlobAlpha", "<f4"), ("cIw0", "<f4"),
    ("cIDeltaSqrt", "<f4"), ("cOneOverBessiOrderAlpha", "<f4"),
])
assert CONST_DTYPE.itemsize == 32

K_DYNAMIC, K_LOOKUP = 0, 1
K_FAST, K_PRECISE = 0, 1


class Settings(C.Structure):
    """fourier_reconstruction/settings.hpp:8-67 defaults: kDynamic, k0, kFast, 15, 1.9."""
    _fields_ = [("interpolation", C.c_int), ("order", C.c_int), ("type", C.c_int),
                ("alpha", C.c_float), ("blobRadius", C.c_float)]

    def __init__(self, interpolation=K_DYNAMIC, order=0, type=K_FAST, alpha=15.0,
                 blobRadius=1.9):
        super().__init__(interpolation, order, type, alpha, blobRadius)


class Constants(C.Structure):
    _fields_ = [(n, C.c_int if t.kind == "i" else C.c_float) for n, (t, _) in
                CONST_DTYPE.fields.items()]

    def as_array(self):
        return np.frombuffer(bytes(self), dtype=CONST_DTYPE)[0]


_P = C.c_void_p
_SZ = C.c_size_t


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def build(with_ref: bool | None = None) -> None:
    """Compile the oracle (and, when the reference tree is present, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/libumpalumpa")
    if with_ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


class Oracle:
    """The C restatement (oracle/umpa_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build(with_ref=False)
        L = self.lib = C.CDLL(path)
        L.orc_bessi0.restype = C.c_double
        L.orc_bessi0.argtypes = [C.c_double]
        for n in ("orc_bessi1", "orc_bessi2", "orc_bessi3", "orc_bessi4"):
            getattr(L, n).restype = C.c_double
            getattr(L, n).argtypes = [C.c_double]
        L.orc_bessi1_5f.restype = C.c_float
        L.orc_bessi1_5f.argtypes = [C.c_float]
        L.orc_get_bessi_order_alpha.restype = C.c_float
        L.orc_get_bessi_order_alpha.argtypes = [C.c_int, C.c_float]
        L.orc_kaiser_value.restype = C.c_float
        L.orc_kaiser_value.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int]
        L.orc_kaiser_fourier_value.restype = C.c_float
        L.orc_kaiser_fourier_value.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int]
        L.orc_create_constants.restype = Constants
        L.orc_create_constants.argtypes = [_SZ, _SZ, _SZ, C.POINTER(Settings)]
        L.orc_fill_blob_table.restype = None
        L.orc_fill_blob_table.argtypes = [_P, _SZ, _SZ, _SZ, C.POINTER(Settings)]
        L.orc_compute_traverse_space.restype = None
        L.orc_compute_traverse_space.argtypes = [C.c_uint32, C.c_uint32, _P, _P, C.c_uint32,
                                                 C.c_uint32, C.c_int, C.c_float, C.c_float]
        L.orc_fr_execute.restype = C.c_int
        L.orc_fr_execute.argtypes = [C.POINTER(Settings), _P, _P, C.c_int, C.c_int, _P, _P,
                                     _SZ, _P, _SZ, C.POINTER(Constants),
                                     C.POINTER(C.c_longlong)]
        L.orc_fr_execute_f64.restype = C.c_int
        L.orc_fr_execute_f64.argtypes = [C.POINTER(Settings), _P, _P, C.c_int, C.c_int, _P, _P,
                                         _SZ, _P, _SZ, C.POINTER(Constants)]
        L.orc_fp_execute.restype = None
        L.orc_fp_execute.argtypes = [_P, _P, _P, _P, _P, _P, C.c_float, C.c_float, C.c_int,
                                     C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_idx2freq.restype = C.c_float
        L.orc_idx2freq.argtypes = [_SZ, _SZ]

    # -- blob table / constants ---------------------------------------------------------------
    def create_constants(self, volume_x, volume_y, table_elems, s: Settings) -> Constants:
        return self.lib.orc_create_constants(volume_x, volume_y, table_elems, C.byref(s))

    def fill_blob_table(self, n, volume_x, volume_y, s: Settings) -> np.ndarray:
        t = np.zeros(n, dtype=np.float32)
        self.lib.orc_fill_blob_table(_ptr(t), n, volume_x, volume_y, C.byref(s))
        return t

    def compute_traverse_space(self, img_x, img_y, transform, max_x, max_yz, use_fast,
                               blob_radius, weight=1.0) -> np.ndarray:
        t = np.ascontiguousarray(transform, dtype=np.float32).reshape(9)
        out = np.zeros(1, dtype=TS_DTYPE)
        self.lib.orc_compute_traverse_space(img_x, img_y, _ptr(t), _ptr(out), max_x, max_yz,
                                            int(use_fast), blob_radius, weight)
        return out[0]

    # -- insertion ----------------------------------------------------------------------------
    def fr_execute(self, s: Settings, volume, weights, fft, spaces, table, c: Constants):
        """volume: complex64 [Z,Y,X] (in/out), weights float32 [Z,Y,X] (in/out),
        fft: complex64 [N, Hy, Hx].  Returns the voxel-update count, or None when the
        reference would return false."""
        assert volume.dtype == np.complex64 and weights.dtype == np.float32
        assert fft.dtype == np.complex64 and spaces.dtype == TS_DTYPE
        assert volume.flags.c_contiguous and weights.flags.c_contiguous
        fft = np.ascontiguousarray(fft)
        spaces = np.ascontiguousarray(spaces)
        n_table = 0 if table is None else table.size
        if table is not None:
            table = np.ascontiguousarray(table, dtype=np.float32)
        upd = C.c_longlong(0)
        rc = self.lib.orc_fr_execute(C.byref(s), _ptr(volume), _ptr(weights), fft.shape[-1],
                                     fft.shape[-2], _ptr(fft), _ptr(spaces), spaces.size,
                                     _ptr(table), n_table, C.byref(c), C.byref(upd))
        return None if rc != 0 else int(upd.value)

    def fr_execute_f64(self, s: Settings, volume, weights, fft, spaces, table, c: Constants):
        assert volume.dtype == np.complex128 and weights.dtype == np.float64
        fft = np.ascontiguousarray(fft, dtype=np.complex64)
        spaces = np.ascontiguousarray(spaces)
        table = np.ascontiguousarray(table, dtype=np.float32)
        rc = self.lib.orc_fr_execute_f64(C.byref(s), _ptr(volume), _ptr(weights), fft.shape[-1],
                                         fft.shape[-2], _ptr(fft), _ptr(spaces), spaces.size,
                                         _ptr(table), table.size, C.byref(c))
        if rc != 0:
            raise ValueError("FP64 oracle supports kPrecise + kLookup only")

    def fp_execute(self, inp, in_size, in_spatial, out_size, filt=None, norm_factor=1.0,
                   max_freq=0.0, apply_filter=False, normalize=False, center=False,
                   crop_freq=False, shift=False):
        inp = np.ascontiguousarray(inp, dtype=np.complex64)
        out = np.zeros(int(np.prod(out_size)), dtype=np.complex64)
        sz = [np.asarray(v, dtype=np.uint64) for v in (in_size, in_spatial, out_size)]
        if filt is not None:
            filt = np.ascontiguousarray(filt, dtype=np.float32)
        self.lib.orc_fp_execute(_ptr(inp), _ptr(out), _ptr(sz[0]), _ptr(sz[1]), _ptr(sz[2]),
                                _ptr(filt), norm_factor, max_freq, int(apply_filter),
                                int(normalize), int(center), int(crop_freq), int(shift))
        return out


class Reference:
    """The reference's own CPU code compiled from the reference tree (oracle/_ref)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        for n in ("ref_bessi0", "ref_bessi1", "ref_bessi2", "ref_bessi3", "ref_bessi4"):
            getattr(L, n).restype = C.c_double
            getattr(L, n).argtypes = [C.c_double]
        L.ref_bessi1_5f.restype = C.c_float
        L.ref_bessi1_5f.argtypes = [C.c_float]
        L.ref_get_bessi_order_alpha.restype = C.c_float
        L.ref_get_bessi_order_alpha.argtypes = [C.c_int, C.c_float]
        L.ref_fr_execute.restype = C.c_int
        L.ref_fr_execute.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, _P, _P, C.c_int,
                                     C.c_int, _P, _P, _SZ, _P, _SZ, _P]
        L.ref_fp_execute.restype = None
        L.ref_fp_execute.argtypes = [_P, _P, _P, _P, _P, _P, C.c_float, C.c_float, C.c_int,
                                     C.c_int, C.c_int, C.c_int, C.c_int]
        L.ref_compute_traverse_space.restype = None
        L.ref_compute_traverse_space.argtypes = [C.c_uint32, C.c_uint32, _P, _P, C.c_uint32,
                                                 C.c_uint32, C.c_int, C.c_float, C.c_float]

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_SO)

    def compute_traverse_space(self, img_x, img_y, transform, max_x, max_yz, use_fast,
                               blob_radius, weight=1.0) -> np.ndarray:
        t = np.ascontiguousarray(transform, dtype=np.float32).reshape(9)
        out = np.zeros(1, dtype=TS_DTYPE)
        self.lib.ref_compute_traverse_space(img_x, img_y, _ptr(t), _ptr(out), max_x, max_yz,
                                            int(use_fast), blob_radius, weight)
        return out[0]

    def fr_execute(self, s: Settings, volume, weights, fft, spaces, table, c: Constants):
        fft = np.ascontiguousarray(fft)
        spaces = np.ascontiguousarray(spaces)
        n_table = 0 if table is None else table.size
        if table is not None:
            table = np.ascontiguousarray(table, dtype=np.float32)
        rc = self.lib.ref_fr_execute(s.type, s.interpolation, s.order, s.alpha, _ptr(volume),
                                     _ptr(weights), fft.shape[-1], fft.shape[-2], _ptr(fft),
                                     _ptr(spaces), spaces.size, _ptr(table), n_table,
                                     C.byref(c))
        return rc == 0

    def fp_execute(self, inp, in_size, in_spatial, out_size, filt=None, norm_factor=1.0,
                   max_freq=0.0, apply_filter=False, normalize=False, center=False,
                   crop_freq=False, shift=False):
        inp = np.ascontiguousarray(inp, dtype=np.complex64)
        out = np.zeros(int(np.prod(out_size)), dtype=np.complex64)
        sz = [np.asarray(v, dtype=np.uint64) for v in (in_size, in_spatial, out_size)]
        if filt is not None:
            filt = np.ascontiguousarray(filt, dtype=np.float32)
        self.lib.ref_fp_execute(_ptr(inp), _ptr(out), _ptr(sz[0]), _ptr(sz[1]), _ptr(sz[2]),
                                _ptr(filt), norm_factor, max_freq, int(apply_filter),
                                int(normalize), int(center), int(crop_freq), int(shift))
        return out


# ---- synthetic inputs shared by tests and the CPU-baseline leg -------------------------------
def uniform_rotations(n: int, seed: int = 43) -> np.ndarray:
    """n rotation matrices uniform on SO(3): unit quaternions from 4 Gaussians, FP64."""
    rng = np.random.default_rng(seed)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    x, y, z, w = q.T
    m = np.empty((n, 3, 3))
    m[:, 0, 0] = 1 - 2 * (y * y + z * z)
    m[:, 0, 1] = 2 * (x * y - z * w)
    m[:, 0, 2] = 2 * (x * z + y * w)
    m[:, 1, 0] = 2 * (x * y + z * w)
    m[:, 1, 1] = 1 - 2 * (x * x + z * z)
    m[:, 1, 2] = 2 * (y * z - x * w)
    m[:, 2, 0] = 2 * (x * z - y * w)
    m[:, 2, 1] = 2 * (y * z + x * w)
    m[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return m

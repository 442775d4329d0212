"""ctypes binding of the C ABI declared in include/umpalumpa_amd.h.

The product library is libumpalumpa_amd.so next to this file, built from umpalumpa_amd/csrc by
``__graft_entry__.build()`` (or ``make -C umpalumpa_amd/csrc``).  There is no fallback: when
the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libumpalumpa_amd.so")

UMPA_OK = 0
UMPA_ERR_INVALID_ARG = 1
UMPA_ERR_NOT_INITIALIZED = 2
UMPA_ERR_NOT_EQUIVALENT = 3
UMPA_ERR_REJECTED = 4
UMPA_ERR_UNSUPPORTED = 5
UMPA_ERR_DEVICE = 6
UMPA_ERR_OUT_OF_MEMORY = 7

FR_DYNAMIC, FR_LOOKUP = 0, 1
FR_FAST, FR_PRECISE = 0, 1
FR_EXACT, FR_FAST_MATH = 0, 1
FFT_FORWARD, FFT_INVERSE = 0, 1
FFT_F32, FFT_F64 = 0, 1

# umpa_traverse_space, 112 B (fourier_reconstruction/traverse_space.hpp:14-48)
TRAVERSE_SPACE_DTYPE = np.dtype([
    ("minY", "<i4"), ("minX", "<i4"), ("minZ", "<i4"),
    ("maxY", "<i4"), ("maxX", "<i4"), ("maxZ", "<i4"),
    ("maxDistanceSqr", "<f4"), ("dir", "<i4"),
    ("unitNormal", "<f4", (3,)), ("topOrigin", "<f4", (3,)), ("bottomOrigin", "<f4", (3,)),
    ("projectionIndex", "<i4"), ("transformInv", "<f4", (3, 3)), ("weight", "<f4"),
])
assert TRAVERSE_SPACE_DTYPE.itemsize == 112


class Size(C.Structure):
    _fields_ = [("x", C.c_uint64), ("y", C.c_uint64), ("z", C.c_uint64), ("n", C.c_uint64)]


class FRConstants(C.Structure):
    _fields_ = [("cMaxVolumeIndexX", C.c_int32), ("cMaxVolumeIndexYZ", C.c_int32),
                ("cBlobRadius", C.c_float), ("cOneOverBlobRadiusSqr", C.c_float),
                ("cBlobAlpha", C.c_float), ("cIw0", C.c_float), ("cIDeltaSqrt", C.c_float),
                ("cOneOverBessiOrderAlpha", C.c_float)]


class FRSettings(C.Structure):
    """fourier_reconstruction::Settings; defaults as settings.hpp:61-66."""
    _fields_ = [("interpolation", C.c_int32), ("order", C.c_int32), ("type", C.c_int32),
                ("alpha", C.c_float), ("blobRadius", C.c_float)]

    def __init__(self, interpolation=FR_DYNAMIC, order=0, type=FR_FAST, alpha=15.0,
                 blobRadius=1.9):
        super().__init__(interpolation, order, type, alpha, blobRadius)


class FRDesc(C.Structure):
    _fields_ = [("fft", Size), ("volume", Size), ("traverseSpaces", C.c_uint64),
                ("blobTable", C.c_uint64)]


class FRJob(C.Structure):
    """umpa_fr_job: one reconstruction of a batch (device pointers)."""
    _fields_ = [("desc", FRDesc), ("settings", FRSettings), ("fft", C.c_void_p),
                ("scales", C.c_void_p), ("volume", C.c_void_p), ("weights", C.c_void_p),
                ("spaces", C.c_void_p), ("table", C.c_void_p)]


class SchedParams(C.Structure):
    """umpa_fr_sched_params: the autotuner's schedule and kernel knobs (0 = default for the
    kernel knobs)."""
    _fields_ = [("depthScale", C.c_float), ("priority", C.c_int32), ("maxParts", C.c_int32),
                ("splitAt", C.c_float), ("kernel", C.c_int32), ("tableGlobal", C.c_int32),
                ("boxFixed", C.c_int32)]

    def as_tuple(self):
        return (self.depthScale, self.priority, self.maxParts, self.splitAt, self.kernel,
                self.tableGlobal, self.boxFixed)


# umpa_fr_sched_params.kernel
FR_KERNEL_AUTO, FR_KERNEL_PAIR, FR_KERNEL_CARRY, FR_KERNEL_STAGED = 0, 1, 2, 3


class FFTDesc(C.Structure):
    _fields_ = [("spatial", Size), ("direction", C.c_int32), ("precision", C.c_int32)]


class FPSettings(C.Structure):
    _fields_ = [("center", C.c_int32), ("normalize", C.c_int32), ("applyFilter", C.c_int32),
                ("shift", C.c_int32), ("hasMaxFreq", C.c_int32), ("maxFreq", C.c_float)]


class FPDesc(C.Structure):
    _fields_ = [("in_", Size), ("inSpatial", Size), ("out", Size),
                ("inPaddedSpatialSingle", C.c_uint64)]


_P = C.c_void_p
_U64 = C.c_uint64

# (name, restype, argtypes) for every entry point of include/umpalumpa_amd.h
PROTOTYPES = [
    ("umpa_version", C.c_int, []),
    ("umpa_status_string", C.c_char_p, [C.c_int]),
    ("umpa_fr_create", C.c_int, [C.POINTER(_P), C.c_int, _P]),
    ("umpa_fr_init", C.c_int, [_P, C.POINTER(FRDesc), C.POINTER(FRSettings)]),
    ("umpa_fr_execute", C.c_int, [_P, _P, _U64, _P, _P, _P, _U64, _P]),
    ("umpa_fr_execute_f16", C.c_int, [_P, _P, _P, _U64, _P, _P, _P, _U64, _P]),
    ("umpa_fr_execute_noalloc", C.c_int, [C.c_int, _P, C.POINTER(FRDesc), C.POINTER(FRSettings),
                                          C.c_int, _P, _P, _U64, _P, _P, _P, _U64, _P]),
    ("umpa_fr_set_strategy", C.c_int, [_P, C.c_int]),
    ("umpa_fr_synchronize", C.c_int, [_P]),
    ("umpa_fr_cleanup", C.c_int, [_P]),
    ("umpa_fr_destroy", None, [_P]),
    ("umpa_fr_set_count_updates", C.c_int, [_P, C.c_int]),
    ("umpa_fr_last_updates", C.c_int, [_P, C.POINTER(_U64)]),
    ("umpa_fr_create_constants", C.c_int, [C.POINTER(FRDesc), C.POINTER(FRSettings),
                                          C.POINTER(FRConstants)]),
    ("umpa_fr_fill_blob_table", C.c_int, [_P, _U64, C.POINTER(FRDesc), C.POINTER(FRSettings)]),
    ("umpa_fr_compute_traverse_space", C.c_int, [C.c_uint32, C.c_uint32, _P, _P, C.c_uint32,
                                                C.c_uint32, C.c_int, C.c_float, C.c_float]),
    ("umpa_fr_generate_traverse_spaces", C.c_int, [_P, _U64, _P, C.c_uint32, C.c_uint32,
                                                   C.POINTER(FRSettings), C.c_float, C.c_int32,
                                                   _P]),
    ("umpa_fr_autotune", C.c_int, [_P, _P, _P, _U64, _P, _U64, _P, C.c_int,
                                   C.POINTER(SchedParams)]),
    ("umpa_fr_get_sched_params", C.c_int, [_P, C.POINTER(SchedParams)]),
    ("umpa_fr_set_sched_params", C.c_int, [_P, C.c_int, C.POINTER(SchedParams)]),
    ("umpa_fr_tuning_candidates", C.c_int, [C.c_int, C.POINTER(SchedParams), _U64,
                                            C.POINTER(_U64)]),
    ("umpa_fr_brick_schedule", C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                         C.POINTER(SchedParams), _P, _U64, C.POINTER(_U64)]),
    ("umpa_fr_batch_autotune", C.c_int, [_P, C.POINTER(FRJob), C.c_uint32, C.c_uint32, C.c_int,
                                         C.c_int, C.POINTER(SchedParams)]),
    ("umpa_fr_batch_set_sched_params", C.c_int, [_P, C.c_int, C.POINTER(SchedParams)]),
    ("umpa_fr_batch_set_wave_items", C.c_int, [_P, C.c_uint32]),
    ("umpa_fr_batch_set_max_workgroups", C.c_int, [_P, C.c_uint32]),
    ("umpa_fr_symmetry_operators", C.c_int, [C.c_char_p, _P, C.c_uint32,
                                             C.POINTER(C.c_uint32)]),
    ("umpa_fr_finalize", C.c_int, [_P, _P, _U64, _P, C.c_int, _P]),
    ("umpa_fr_batch_create", C.c_int, [C.POINTER(_P), C.c_int, _P]),
    ("umpa_fr_batch_plan", C.c_int, [_P, C.POINTER(FRJob), C.c_uint32, C.c_uint32, C.c_int,
                                     C.c_int]),
    ("umpa_fr_batch_execute", C.c_int, [_P]),
    ("umpa_fr_batch_step", C.c_int, [_P, C.POINTER(FRJob), C.POINTER(_U64), C.POINTER(_U64),
                                     C.c_uint32, C.c_int]),
    ("umpa_fr_batch_check_jobs", C.c_int, [C.POINTER(FRJob), C.c_uint32, C.c_int]),
    ("umpa_fr_batch_num_steps", C.c_int, [_P, C.POINTER(C.c_uint32)]),
    ("umpa_fr_batch_set_count_updates", C.c_int, [_P, C.c_int]),
    ("umpa_fr_batch_last_updates", C.c_int, [_P, C.POINTER(_U64)]),
    ("umpa_fr_batch_record", C.c_int, [_P, C.POINTER(_U64)]),
    ("umpa_fr_batch_query", C.c_int, [_P, _U64, C.POINTER(C.c_int)]),
    ("umpa_fr_batch_synchronize", C.c_int, [_P]),
    ("umpa_fr_batch_destroy", None, [_P]),
    ("umpa_fft_create", C.c_int, [C.POINTER(_P), C.c_int, _P]),
    ("umpa_fft_init", C.c_int, [_P, C.POINTER(FFTDesc)]),
    ("umpa_fft_execute", C.c_int, [_P, _P, _P, _U64]),
    ("umpa_fft_synchronize", C.c_int, [_P]),
    ("umpa_fft_cleanup", C.c_int, [_P]),
    ("umpa_fft_destroy", None, [_P]),
    ("umpa_fp_create", C.c_int, [C.POINTER(_P), C.c_int, _P]),
    ("umpa_fp_init", C.c_int, [_P, C.POINTER(FPDesc), C.POINTER(FPSettings)]),
    ("umpa_fp_execute", C.c_int, [_P, _P, _U64, _P, _U64, _P]),
    ("umpa_fp_synchronize", C.c_int, [_P]),
    ("umpa_fp_cleanup", C.c_int, [_P]),
    ("umpa_fp_destroy", None, [_P]),
    ("umpa_fft_fp_create", C.c_int, [C.POINTER(_P), C.c_int, _P]),
    ("umpa_fft_fp_init", C.c_int, [_P, C.POINTER(FPDesc), C.POINTER(FPSettings)]),
    ("umpa_fft_fp_execute", C.c_int, [_P, _P, _U64, _P, _P]),
    ("umpa_fft_fp_synchronize", C.c_int, [_P]),
    ("umpa_fft_fp_cleanup", C.c_int, [_P]),
    ("umpa_fft_fp_destroy", None, [_P]),
]

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the product library (once).  Raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"umpalumpa_amd: HIP library {path} is missing; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (there is no CPU fallback)")
    try:
        # torch ships its own HIP runtime.  Loaded first, it is the one this library binds to
        # (same soname), so the process holds a single runtime and torch's device pointers are
        # valid here; loaded after this library, torch would map a second runtime beside the
        # one under /opt/rocm, and device pointers of one would be invalid in the other.
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, res, args in PROTOTYPES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.umpa_version() < 101:   # 101: umpa_fr_sched_params with kernel knobs
        raise RuntimeError("umpalumpa_amd: library version mismatch")
    _lib = lib
    return lib


class UmpaError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        msg = load().umpa_status_string(status).decode()
        super().__init__(f"{what}: {msg} (status {status})")


def check(status: int, what: str) -> None:
    if status != UMPA_OK:
        raise UmpaError(status, what)

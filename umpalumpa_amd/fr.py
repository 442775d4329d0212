"""Python mirror of the reference's AFR / AFP operations over the C ABI.

``FourierReconstruction`` follows AFR (fourier_reconstruction/afr.hpp:13-55) and
``FourierProcessing`` follows AFP (fourier_processing/afp.hpp:13-53): ``init`` and ``execute``
return ``True``/``False`` like the reference's ``[[nodiscard]] bool Init/Execute``
(basic_operation.hpp:28-66), ``synchronize`` waits for the enqueued work and ``cleanup``
drops the Init state.  Device buffers are torch tensors on the operation's GPU; the work is
enqueued on torch's current stream of that device unless a stream is given.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib
from ._lib import (FFT_F32, FFT_F64, FFT_FORWARD, FFT_INVERSE, FR_DYNAMIC, FR_EXACT,
                   FR_FAST, FR_FAST_MATH, FR_KERNEL_AUTO, FR_KERNEL_CARRY, FR_KERNEL_PAIR,
                   FR_KERNEL_STAGED, FR_LOOKUP, FR_PRECISE, TRAVERSE_SPACE_DTYPE, FFTDesc,
                   FPDesc, FPSettings, FRConstants, FRDesc, FRJob, FRSettings, SchedParams, Size,
                   check)

__all__ = [
    "FR_DYNAMIC", "FR_LOOKUP", "FR_FAST", "FR_PRECISE", "FR_EXACT", "FR_FAST_MATH",
    "FR_KERNEL_AUTO", "FR_KERNEL_PAIR", "FR_KERNEL_CARRY", "FR_KERNEL_STAGED",
    "tuning_candidates",
    "TRAVERSE_SPACE_DTYPE", "FRSettings", "FRDesc", "FPSettings", "Size",
    "FourierReconstruction", "FourierProcessing", "create_constants", "fill_blob_table",
    "compute_traverse_space", "traverse_spaces", "generate_traverse_spaces", "BatchJob",
    "FRBatch", "SchedParams", "symmetry_operators", "finalize", "FourierTransform",
    "FFT_FORWARD", "FFT_INVERSE", "FFT_F32", "FFT_F64", "FourierTransformProcessing",
    "check_batch_jobs", "brick_schedule",
]


def _stream_handle(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def _own(obj, destroy: str) -> None:
    """Destroys obj._h with the given C entry point when obj is collected or the interpreter
    exits (the entry point is bound now, so it outlives the module's globals at shutdown)."""
    obj._finalizer = weakref.finalize(obj, getattr(_lib.load(), destroy), obj._h)


def _ptr(t, what: str = "tensor", nbytes: int = 0):
    """Device pointer of a contiguous CUDA tensor holding at least `nbytes` bytes (None for
    None, which the C entry points turn into their own status).  The C ABI takes plain
    pointers, so a host tensor, a strided view or a buffer shorter than the operation's
    descriptor would reach the kernels as something else: refused here with ValueError."""
    if t is None:
        return None
    if not getattr(t, "is_cuda", False):
        raise ValueError(f"{what}: a CUDA tensor is required")
    if not t.is_contiguous():
        raise ValueError(f"{what}: the tensor must be contiguous")
    have = t.numel() * t.element_size()
    if have < nbytes:
        raise ValueError(f"{what}: {have} bytes, at least {nbytes} needed")
    return C.c_void_p(t.data_ptr())


def _elems(s) -> int:
    return int(s.x) * int(s.y) * int(s.z)


def _in_use_on(stream, tensors) -> None:
    """Marks device tensors as used by work enqueued on `stream` (torch's record_stream): a
    tensor made on another stream and dropped by the caller right after the call is then not
    handed out again by the caching allocator until that work has run."""
    for t in tensors:
        if t is not None:
            t.record_stream(stream)


def fr_desc(slice_x: int, slice_y: int, n_slices: int, volume: int, n_spaces: int = 0,
            n_table: int = 0) -> FRDesc:
    """Descriptor of AFR's payloads: slices Hx x Hy (n of them), volume^3, spaces, table."""
    return FRDesc(Size(slice_x, slice_y, 1, n_slices), Size(volume, volume, volume, 1),
                  n_spaces, n_table)


# ---- host statics (no GPU) ----------------------------------------------------------------
def create_constants(desc: FRDesc, s: FRSettings) -> FRConstants:
    """AFR::CreateConstants (afr.cpp:6-19)."""
    out = FRConstants()
    check(_lib.load().umpa_fr_create_constants(C.byref(desc), C.byref(s), C.byref(out)),
          "create_constants")
    return out


def fill_blob_table(n: int, desc: FRDesc, s: FRSettings) -> np.ndarray:
    """AFR::FillBlobTable (afr.cpp:22-34) into a new host array of n floats."""
    t = np.zeros(n, dtype=np.float32)
    check(_lib.load().umpa_fr_fill_blob_table(t.ctypes.data_as(C.c_void_p), n, C.byref(desc),
                                              C.byref(s)), "fill_blob_table")
    return t


def compute_traverse_space(img_x: int, img_y: int, transform, max_x: int, max_yz: int,
                           use_fast: bool, blob_radius: float, weight: float = 1.0):
    """ComputeTraverseSpace (traverse_space_generator.hpp:94-145) for one 3x3 transform."""
    t = np.ascontiguousarray(transform, dtype=np.float32).reshape(9)
    out = np.zeros(1, dtype=TRAVERSE_SPACE_DTYPE)
    check(_lib.load().umpa_fr_compute_traverse_space(
        img_x, img_y, t.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), max_x,
        max_yz, int(use_fast), blob_radius, weight), "compute_traverse_space")
    return out[0]


def traverse_spaces(rotations, padded_box: int, s: FRSettings, symmetries=None,
                    weight: float = 1.0) -> np.ndarray:
    """Traverse spaces of a chunk the way the reference caller builds them
    (examples/fourier_reconstruction/fr.cpp:278-312): for every slice i and every symmetry
    operator S, transform = R_i @ S (FP64 on the host, cast to float), projectionIndex = i.
    The slice extent along x is padded_box/2 (afr_common.hpp:57-59)."""
    rot = np.asarray(rotations, dtype=np.float64).reshape(-1, 3, 3)
    sym = np.eye(3)[None] if symmetries is None else np.asarray(symmetries, np.float64)
    out = np.zeros(len(rot) * len(sym), dtype=TRAVERSE_SPACE_DTYPE)
    k = 0
    for i, r in enumerate(rot):
        for s_op in sym:
            rec = compute_traverse_space(padded_box // 2, padded_box, (r @ s_op), padded_box,
                                         padded_box, s.type == FR_FAST, s.blobRadius, weight)
            rec["projectionIndex"] = i
            out[k] = rec
            k += 1
    return out


def generate_traverse_spaces(rotations, padded_box: int, s: FRSettings, symmetries=None,
                             weight: float = 1.0, first_index: int = 0) -> np.ndarray:
    """traverse_spaces in native code (umpa_fr_generate_traverse_spaces): same records."""
    rot = np.ascontiguousarray(np.asarray(rotations, dtype=np.float64).reshape(-1, 3, 3))
    sym = None if symmetries is None else np.ascontiguousarray(
        np.asarray(symmetries, dtype=np.float64).reshape(-1, 3, 3))
    k = 1 if sym is None else len(sym)
    out = np.zeros(len(rot) * k, dtype=TRAVERSE_SPACE_DTYPE)
    check(_lib.load().umpa_fr_generate_traverse_spaces(
        rot.ctypes.data_as(C.c_void_p), len(rot),
        None if sym is None else sym.ctypes.data_as(C.c_void_p), k, padded_box, C.byref(s),
        weight, first_index, out.ctypes.data_as(C.c_void_p)), "generate_traverse_spaces")
    return out


def symmetry_operators(group: str) -> np.ndarray:
    """Rotation operators (k, 3, 3) float64 of the point group "C<n>" or "D<n>"."""
    lib = _lib.load()
    n = C.c_uint32(0)
    check(lib.umpa_fr_symmetry_operators(group.encode(), None, 0, C.byref(n)),
          f"symmetry_operators({group})")
    out = np.zeros((n.value, 3, 3), dtype=np.float64)
    check(lib.umpa_fr_symmetry_operators(group.encode(), out.ctypes.data_as(C.c_void_p),
                                         n.value, C.byref(n)), f"symmetry_operators({group})")
    return out


def finalize(volume, weights, out=None, hermitian: bool = True, stream=None):
    """Hermitian completion and FP64 weight normalisation of an inserted volume
    (umpa_fr_finalize) into a new (or the given) complex64 tensor on the same device."""
    import torch
    side = volume.shape[-1]
    if out is None:
        out = torch.empty_like(volume)
    n = side ** 3
    check(_lib.load().umpa_fr_finalize(_ptr(volume, "volume", 8 * n),
                                       _ptr(weights, "weights", 4 * n), side,
                                       _ptr(out, "out", 8 * n), int(hermitian),
                                       _stream_handle(stream)),
          "umpa_fr_finalize")
    return out


# ---- AFR ----------------------------------------------------------------------------------
class FourierReconstruction:
    """AFR on one GPU: hand-written gfx950 voxel-owned gather kernels."""

    def __init__(self, device: int = 0, stream=None):
        import torch
        self.device = device
        self._stream = stream if stream is not None else torch.cuda.current_stream(device)
        self._h = C.c_void_p()
        check(_lib.load().umpa_fr_create(C.byref(self._h), device, _stream_handle(self._stream)),
              "umpa_fr_create")
        _own(self, "umpa_fr_destroy")
        self.desc = None
        self.settings = None
        self.last_status = 0

    def init(self, desc: FRDesc, settings: FRSettings) -> bool:
        self.last_status = _lib.load().umpa_fr_init(self._h, C.byref(desc), C.byref(settings))
        ok = self.last_status == 0
        self.desc, self.settings = (desc, settings) if ok else (None, None)
        return ok

    def set_strategy(self, strategy: int) -> None:
        check(_lib.load().umpa_fr_set_strategy(self._h, strategy), "set_strategy")

    def set_count_updates(self, enable: bool) -> None:
        check(_lib.load().umpa_fr_set_count_updates(self._h, int(enable)), "set_count_updates")

    def last_updates(self) -> int:
        v = C.c_uint64(0)
        check(_lib.load().umpa_fr_last_updates(self._h, C.byref(v)), "last_updates")
        return int(v.value)

    def execute(self, fft, volume, weights, spaces, table=None, n_spaces=None,
                scales=None) -> bool:
        """fft: complex64 [n, Hy, Hx] (or float16 [n, Hy, Hx, 2] with `scales`), volume:
        complex64 [Z, Y, X], weights: float32 [Z, Y, X], spaces: uint8 device tensor of
        112-byte records, table: float32 device tensor."""
        n_sp = spaces.numel() // 112 if n_spaces is None else n_spaces
        p = self._pointers(fft, scales, spaces, n_sp, table, volume, weights)
        import torch
        if self._stream != torch.cuda.current_stream(self.device):
            _in_use_on(self._stream, (fft, scales, spaces, table, volume, weights))
        lib = _lib.load()
        if scales is None:
            self.last_status = lib.umpa_fr_execute(
                self._h, p["fft"], fft.shape[0], p["volume"], p["weights"], p["spaces"], n_sp,
                p["table"])
        else:
            self.last_status = lib.umpa_fr_execute_f16(
                self._h, p["fft"], p["scales"], fft.shape[0], p["volume"], p["weights"],
                p["spaces"], n_sp, p["table"])
        return self.last_status == 0

    def _pointers(self, fft, scales, spaces, n_sp, table, volume=None, weights=None) -> dict:
        """Checked device pointers, each tensor at least as large as the Init descriptor says
        (a count above Init's is left to the C entry point, which refuses it)."""
        d = self.desc
        n = 0 if fft is None else min(fft.shape[0], d.fft.n if d is not None else 0)
        px = 0 if d is None else int(d.fft.x) * int(d.fft.y) * (4 if scales is not None else 8)
        vox = 0 if d is None else _elems(d.volume)
        return {"fft": _ptr(fft, "fft", n * px), "scales": _ptr(scales, "scales", 4 * n),
                "spaces": _ptr(spaces, "spaces", 112 * n_sp),
                "table": _ptr(table, "table", 0 if d is None else 4 * int(d.blobTable)),
                "volume": _ptr(volume, "volume", 8 * vox),
                "weights": _ptr(weights, "weights", 4 * vox)}

    def autotune(self, fft, spaces, table=None, n_spaces=None, scales=None,
                 use_cache: bool = True) -> SchedParams:
        """Times the candidate schedules for inputs like these on scratch volumes and keeps
        the fastest for the current strategy (umpa_fr_autotune)."""
        n_sp = spaces.numel() // 112 if n_spaces is None else n_spaces
        out = SchedParams()
        p = self._pointers(fft, scales, spaces, n_sp, table)
        check(_lib.load().umpa_fr_autotune(self._h, p["fft"], p["scales"], fft.shape[0],
                                           p["spaces"], n_sp, p["table"], int(use_cache),
                                           C.byref(out)), "umpa_fr_autotune")
        return out

    def set_sched_params(self, strategy: int, p: SchedParams) -> None:
        """Forces the schedule knobs of `strategy` (umpa_fr_set_sched_params)."""
        check(_lib.load().umpa_fr_set_sched_params(self._h, strategy, C.byref(p)),
              "set_sched_params")

    def sched_params(self) -> SchedParams:
        out = SchedParams()
        check(_lib.load().umpa_fr_get_sched_params(self._h, C.byref(out)), "sched_params")
        return out

    def synchronize(self) -> None:
        check(_lib.load().umpa_fr_synchronize(self._h), "synchronize")

    def cleanup(self) -> None:
        check(_lib.load().umpa_fr_cleanup(self._h), "cleanup")
        self.desc = self.settings = None


# ---- batched AFR --------------------------------------------------------------------------
class BatchJob:
    """One reconstruction of a batch: device tensors as for FourierReconstruction.execute."""

    def __init__(self, fft, volume, weights, spaces, settings: FRSettings, table=None,
                 scales=None, n_table: int | None = None):
        self.fft, self.volume, self.weights, self.spaces = fft, volume, weights, spaces
        self.settings, self.table, self.scales = settings, table, scales
        self.n_table = (table.numel() if table is not None else 0) if n_table is None else n_table

    @property
    def n_spaces(self) -> int:
        return self.spaces.numel() // 112

    def tensors(self):
        return (self.fft, self.scales, self.volume, self.weights, self.spaces, self.table)

    def to_c(self) -> FRJob:
        n, hy, hx = self.fft.shape[0], self.fft.shape[1], self.fft.shape[2]
        v = self.volume.shape[-1]
        desc = fr_desc(hx, hy, n, v, self.n_spaces, self.n_table)
        px = 4 if self.scales is not None else 8

        def ptr(t, what, nbytes):
            p = _ptr(t, what, nbytes)
            return None if p is None else p.value

        return FRJob(desc, self.settings, ptr(self.fft, "fft", n * hy * hx * px),
                     ptr(self.scales, "scales", 4 * n), ptr(self.volume, "volume", 8 * v ** 3),
                     ptr(self.weights, "weights", 4 * v ** 3),
                     ptr(self.spaces, "spaces", 112 * self.n_spaces),
                     ptr(self.table, "table", 4 * self.n_table))


def brick_schedule(vx: int, vy: int, vz: int, split_heavy: bool,
                   params: SchedParams | None = None) -> np.ndarray:
    """The insertion's work items (umpa_fr_brick_schedule): int32 [n, 4] rows {x, y, z, flags}
    with depth = flags & 15, priority bit 4, parts (flags >> 8) & 255, part (flags >> 16)."""
    lib = _lib.load()
    n = C.c_uint64(0)
    p = C.byref(params) if params is not None else None
    check(lib.umpa_fr_brick_schedule(vx, vy, vz, int(split_heavy), p, None, 0, C.byref(n)),
          "umpa_fr_brick_schedule")
    out = np.zeros((n.value, 4), np.int32)
    check(lib.umpa_fr_brick_schedule(vx, vy, vz, int(split_heavy), p,
                                     out.ctypes.data_as(C.c_void_p), n.value, C.byref(n)),
          "umpa_fr_brick_schedule")
    return out


def tuning_candidates(strategy: int) -> list:
    """The knobs the autotuner times under `strategy` (umpa_fr_tuning_candidates)."""
    lib = _lib.load()
    n = C.c_uint64(0)
    check(lib.umpa_fr_tuning_candidates(strategy, None, 0, C.byref(n)),
          "umpa_fr_tuning_candidates")
    arr = (SchedParams * n.value)()
    check(lib.umpa_fr_tuning_candidates(strategy, arr, n.value, C.byref(n)),
          "umpa_fr_tuning_candidates")
    return list(arr)


def check_batch_jobs(jobs, strategy: int = FR_EXACT) -> int:
    """The checks a batch plan / step applies to `jobs`, on the host only
    (umpa_fr_batch_check_jobs): returns the status code (0 when the jobs are accepted)."""
    arr = (FRJob * len(jobs))(*[j.to_c() for j in jobs])
    return int(_lib.load().umpa_fr_batch_check_jobs(arr, len(jobs), strategy))


class FRBatch:
    """Several reconstructions through one persistent insertion launch per step
    (umpa_fr_batch_*).  Runs on its own stream (a HIP graph cannot capture the null stream);
    every call first makes that stream wait for torch's current stream, so inputs written
    there are visible, and ``wait()`` makes torch's current stream wait for the batch."""

    def __init__(self, device: int = 0):
        import torch
        self.device = device
        self.stream = torch.cuda.Stream(device)
        self._h = C.c_void_p()
        check(_lib.load().umpa_fr_batch_create(C.byref(self._h), device,
                                               C.c_void_p(self.stream.cuda_stream)),
              "umpa_fr_batch_create")
        _own(self, "umpa_fr_batch_destroy")
        self._jobs = None

    def _follow(self):
        import torch
        self.stream.wait_stream(torch.cuda.current_stream(self.device))

    def plan(self, jobs, chunk: int, strategy: int = FR_EXACT, graph: bool = True) -> None:
        self._follow()
        arr = (FRJob * len(jobs))(*[j.to_c() for j in jobs])
        self._jobs = (jobs, arr)
        for j in jobs:
            _in_use_on(self.stream, j.tensors())
        check(_lib.load().umpa_fr_batch_plan(self._h, arr, len(jobs), chunk, strategy,
                                             int(graph)), "umpa_fr_batch_plan")

    def autotune(self, jobs, chunk: int, strategy: int = FR_EXACT,
                 use_cache: bool = True) -> SchedParams:
        """Tunes the schedule of later plans / steps on the first step of these jobs."""
        self._follow()
        arr = (FRJob * len(jobs))(*[j.to_c() for j in jobs])
        out = SchedParams()
        check(_lib.load().umpa_fr_batch_autotune(self._h, arr, len(jobs), chunk, strategy,
                                                 int(use_cache), C.byref(out)),
              "umpa_fr_batch_autotune")
        return out

    def set_sched_params(self, strategy: int, p: SchedParams) -> None:
        check(_lib.load().umpa_fr_batch_set_sched_params(self._h, strategy, C.byref(p)),
              "set_sched_params")

    def set_wave_items(self, items_per_wave: int) -> None:
        """Bounded launches (umpa_fr_batch_set_wave_items): every wavefront leaves after that
        many work items (0: persistent workgroups), so another stream's launches get CUs while
        this batch's launches run."""
        check(_lib.load().umpa_fr_batch_set_wave_items(self._h, items_per_wave),
              "set_wave_items")

    def set_max_workgroups(self, workgroups: int) -> None:
        """Persistent launches use at most this many workgroups (one per CU; 0: every CU),
        leaving CUs to another stream (umpa_fr_batch_set_max_workgroups)."""
        check(_lib.load().umpa_fr_batch_set_max_workgroups(self._h, workgroups),
              "set_max_workgroups")

    def num_steps(self) -> int:
        v = C.c_uint32(0)
        check(_lib.load().umpa_fr_batch_num_steps(self._h, C.byref(v)), "num_steps")
        return int(v.value)

    def execute(self) -> None:
        self._follow()
        if self._jobs is not None:
            for j in self._jobs[0]:
                _in_use_on(self.stream, j.tensors())
        check(_lib.load().umpa_fr_batch_execute(self._h), "umpa_fr_batch_execute")

    def step(self, jobs, begin, end, strategy: int = FR_EXACT) -> None:
        """Enqueues one step; the job tensors are marked as in use on the batch stream, so
        the caller may drop them (e.g. a freshly made slice chunk) as soon as this returns."""
        self._follow()
        for j in jobs:
            _in_use_on(self.stream, j.tensors())
        n = len(jobs)
        arr = (FRJob * n)(*[j.to_c() for j in jobs])
        b = (C.c_uint64 * n)(*[int(x) for x in begin])
        e = (C.c_uint64 * n)(*[int(x) for x in end])
        check(_lib.load().umpa_fr_batch_step(self._h, arr, b, e, n, strategy),
              "umpa_fr_batch_step")

    def set_count_updates(self, enable: bool) -> None:
        check(_lib.load().umpa_fr_batch_set_count_updates(self._h, int(enable)),
              "set_count_updates")

    def last_updates(self) -> int:
        v = C.c_uint64(0)
        check(_lib.load().umpa_fr_batch_last_updates(self._h, C.byref(v)), "last_updates")
        return int(v.value)

    def record(self) -> int:
        t = C.c_uint64(0)
        check(_lib.load().umpa_fr_batch_record(self._h, C.byref(t)), "record")
        return int(t.value)

    def query(self, ticket: int) -> bool:
        d = C.c_int(0)
        check(_lib.load().umpa_fr_batch_query(self._h, ticket, C.byref(d)), "query")
        return bool(d.value)

    def synchronize(self) -> None:
        check(_lib.load().umpa_fr_batch_synchronize(self._h), "synchronize")

    def wait(self) -> None:
        import torch
        torch.cuda.current_stream(self.device).wait_stream(self.stream)


# ---- AFFT ---------------------------------------------------------------------------------
class FourierTransform:
    """AFFT on one GPU (batched out-of-place R2C forward / unnormalised C2R inverse, hipFFT)."""

    def __init__(self, device: int = 0, stream=None):
        import torch
        self.device = device
        self._stream = stream if stream is not None else torch.cuda.current_stream(device)
        self._h = C.c_void_p()
        check(_lib.load().umpa_fft_create(C.byref(self._h), device,
                                          _stream_handle(self._stream)), "umpa_fft_create")
        _own(self, "umpa_fft_destroy")
        self.last_status = 0
        self.desc = None

    def init(self, spatial, n: int, direction: int = FFT_FORWARD,
             precision: int = FFT_F32) -> bool:
        """spatial: (x, y[, z]) image size; n: images."""
        x, y, z = (tuple(spatial) + (1, 1))[:3]
        d = FFTDesc(Size(x, y, z, n), direction, precision)
        self.last_status = _lib.load().umpa_fft_init(self._h, C.byref(d))
        self.desc = d if self.last_status == 0 else None
        return self.last_status == 0

    def execute(self, inp, out, count: int) -> bool:
        d = self.desc
        real = freq = 0
        if d is not None:
            m = min(count, int(d.spatial.n))
            s = d.spatial
            w = 8 if d.precision == FFT_F64 else 4
            real = m * _elems(s) * w
            freq = m * (int(s.x) // 2 + 1) * int(s.y) * int(s.z) * 2 * w
        src, dst = (real, freq) if d is None or d.direction == FFT_FORWARD else (freq, real)
        if inp is not None and out is not None and inp.data_ptr() == out.data_ptr():
            dst = 0  # in place: the C entry point refuses it (UMPA_ERR_UNSUPPORTED)
        self.last_status = _lib.load().umpa_fft_execute(self._h, _ptr(inp, "input", src),
                                                        _ptr(out, "output", dst), count)
        return self.last_status == 0

    def synchronize(self) -> None:
        check(_lib.load().umpa_fft_synchronize(self._h), "synchronize")


# ---- fused FFT + FP --------------------------------------------------------------------------
class FourierTransformProcessing:
    """The forward R2C transform of 2-D images followed by the FP step (crop / filter /
    normalise / centre / shift), fused (umpa_fft_fp_*): hipFFT for the rows, one gfx950
    kernel for the column transforms and the FP arithmetic.  Same descriptor and settings as
    FourierProcessing, with desc.in the R2C size of desc.inSpatial; image heights are powers of
    two from 4 to 1024."""

    def __init__(self, device: int = 0, stream=None):
        import torch
        self.device = device
        self._stream = stream if stream is not None else torch.cuda.current_stream(device)
        self._h = C.c_void_p()
        check(_lib.load().umpa_fft_fp_create(C.byref(self._h), device,
                                             _stream_handle(self._stream)), "umpa_fft_fp_create")
        _own(self, "umpa_fft_fp_destroy")
        self.last_status = 0
        self.desc = None

    def init(self, desc: FPDesc, settings: FPSettings) -> bool:
        self.last_status = _lib.load().umpa_fft_fp_init(self._h, C.byref(desc),
                                                        C.byref(settings))
        self.desc = desc if self.last_status == 0 else None
        return self.last_status == 0

    def execute(self, images, count: int, out, filt=None) -> bool:
        """images: float32 [count, y, x]; out: complex64 [count, out.y, out.x]."""
        d = self.desc
        m = 0 if d is None else min(count, int(d.inSpatial.n))
        img = 0 if d is None else 4 * m * _elems(d.inSpatial)
        res = 0 if d is None else 8 * m * _elems(d.out)
        flt = 0 if d is None else 4 * int(d.out.x) * int(d.out.y)
        self.last_status = _lib.load().umpa_fft_fp_execute(
            self._h, _ptr(images, "images", img), count, _ptr(out, "out", res),
            _ptr(filt, "filter", flt))
        return self.last_status == 0

    def synchronize(self) -> None:
        check(_lib.load().umpa_fft_fp_synchronize(self._h), "synchronize")


# ---- AFP ----------------------------------------------------------------------------------
class FourierProcessing:
    """AFP on one GPU (out-of-place crop / filter / normalize / center / shift)."""

    def __init__(self, device: int = 0, stream=None):
        import torch
        self.device = device
        self._stream = stream if stream is not None else torch.cuda.current_stream(device)
        self._h = C.c_void_p()
        check(_lib.load().umpa_fp_create(C.byref(self._h), device, _stream_handle(self._stream)),
              "umpa_fp_create")
        _own(self, "umpa_fp_destroy")
        self.last_status = 0
        self.desc = None

    def init(self, desc: FPDesc, settings: FPSettings) -> bool:
        self.last_status = _lib.load().umpa_fp_init(self._h, C.byref(desc), C.byref(settings))
        self.desc = desc if self.last_status == 0 else None
        return self.last_status == 0

    def execute(self, inp, in_count, out, out_count, filt=None) -> bool:
        d = self.desc
        src = 0 if d is None else 8 * min(in_count, int(d.in_.n)) * _elems(d.in_)
        # the kernel writes the first in_count output images
        dst = 0 if d is None else 8 * min(in_count, out_count, int(d.out.n)) * _elems(d.out)
        flt = 0 if d is None else 4 * int(d.out.x) * int(d.out.y)
        self.last_status = _lib.load().umpa_fp_execute(
            self._h, _ptr(inp, "input", src), in_count, _ptr(out, "output", dst), out_count,
            _ptr(filt, "filter", flt))
        return self.last_status == 0

    def synchronize(self) -> None:
        check(_lib.load().umpa_fp_synchronize(self._h), "synchronize")

    def cleanup(self) -> None:
        check(_lib.load().umpa_fp_cleanup(self._h), "cleanup")


def fp_desc(in_spatial, out_spatial, n: int, padded_spatial_single: int | None = None) -> FPDesc:
    """Descriptor for a crop of spatial (x, y) images to (x', y'): frequency sizes x/2+1."""
    ix, iy = in_spatial
    ox, oy = out_spatial
    single = ix * iy if padded_spatial_single is None else padded_spatial_single
    return FPDesc(Size(ix // 2 + 1, iy, 1, n), Size(ix, iy, 1, n), Size(ox // 2 + 1, oy, 1, n),
                  single)

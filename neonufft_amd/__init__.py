"""MI355X (gfx950) NeoNUFFT: GPU type 1/2/3 NUFFT plans.

Python mirror of the reference's GPU plan interface
(``neonufft::gpu::Plan<T,DIM>`` / ``neonufft::gpu::PlanT3<T,DIM>``,
include/neonufft/gpu/plan.hpp of the reference): same constructor arguments,
method names, argument meaning and error classes. All arrays are device
tensors (``torch`` CUDA/HIP tensors) or raw device addresses (``int``); every
call goes through the C ABI of ``libneonufft_amd.so`` (include/neonufft_amd.h).
There is no CPU fallback: a missing library raises ``ImportError``.
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass
from typing import Optional, Sequence, Union

from . import _lib as _L

__all__ = [
    "Options", "ModeOrder", "Plan", "PlanT3", "BatchPlan", "autotune", "GenericError", "InputError", "InternalError",
    "MemoryAllocError", "NotImplementedError_", "GPUError", "UnknownError", "version",
]


class ModeOrder(enum.IntEnum):
    """NeonufftModeOrder (include/neonufft/enums.h:6-9)."""
    FFT = 0
    CMCL = 1


@dataclass
class Options:
    """neonufft::Options (include/neonufft/plan.hpp:19-29), same defaults."""
    tol: float = 0.001
    upsampfac: float = 2.0
    recenter_threshold: float = 0.1
    num_threads: int = 0
    sort_input: bool = True
    sort_output: bool = True
    kernel_approximation: bool = False
    order: ModeOrder = ModeOrder.CMCL
    kernel_type: int = 0

    def _c(self) -> _L.NnfOptions:
        o = _L.NnfOptions()
        o.tol = float(self.tol)
        o.upsampfac = float(self.upsampfac)
        o.recenter_threshold = float(self.recenter_threshold)
        o.num_threads = int(self.num_threads)
        o.sort_input = int(bool(self.sort_input))
        o.sort_output = int(bool(self.sort_output))
        o.kernel_approximation = int(bool(self.kernel_approximation))
        o.order = int(self.order)
        o.kernel_type = int(self.kernel_type)
        return o


# Exception classes with the reference's names and NeonufftError codes
# (include/neonufft/exceptions.hpp:16-103, errors.h:5-38).
class GenericError(RuntimeError):
    code = 1


class MemoryAllocError(GenericError):
    code = 2


class UnknownError(GenericError):
    code = 3


class InternalError(GenericError):
    code = 4


class InputError(GenericError):
    code = 5


class GPUError(GenericError):
    code = 6


class NotImplementedError_(GenericError):
    code = 7


_ERRORS = {c.code: c for c in (GenericError, MemoryAllocError, UnknownError, InternalError,
                               InputError, GPUError, NotImplementedError_)}


def _check(rc: int) -> None:
    if rc != 0:
        msg = _L.lib().nnf_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, GenericError)(msg or f"neonufft error {rc}")


def version() -> str:
    return _L.lib().nnf_version().decode()


Pointer = Union[int, "torch.Tensor"]  # noqa: F821


def _ptr(a) -> int:
    if a is None:
        return 0
    if isinstance(a, int):
        return a
    return int(a.data_ptr())


def _precision_of(x) -> int:
    import torch
    if isinstance(x, torch.Tensor):
        if x.dtype in (torch.float32, torch.complex64):
            return 0
        if x.dtype in (torch.float64, torch.complex128):
            return 1
        raise InputError(f"unsupported dtype {x.dtype}")
    raise InputError("pass dtype= when giving raw pointers")


def _stream(stream) -> int:
    if stream is None:
        import torch
        return int(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


def _vpp(ptrs: Sequence[int]):
    arr = (ctypes.c_void_p * max(1, len(ptrs)))(*[ctypes.c_void_p(p) for p in ptrs])
    return ctypes.cast(arr, ctypes.POINTER(ctypes.c_void_p)), arr


def _i64p(vals: Sequence[int]):
    arr = (ctypes.c_int64 * 3)(*([int(v) for v in vals] + [0] * (3 - len(vals))))
    return arr


def _dtype_code(dtype) -> int:
    import torch
    if dtype in ("float32", "f32", torch.float32, torch.complex64):
        return 0
    if dtype in ("float64", "f64", torch.float64, torch.complex128):
        return 1
    if dtype == "mixed":
        # type 3 only: double coordinates, complex128 I/O, fp32 bulk work
        return 2
    raise InputError(f"unsupported dtype {dtype}")


def autotune(opt: "Options", sign: int, num_nu: int, loc: Sequence[Pointer],
             modes: Sequence[int], types: int = 3, reps: int = 3, stream=None, dtype=None,
             use_cache: bool = True) -> dict:
    """Fastest subgrid tile and per-workgroup point cap for a plan with these
    arguments, timed on these device points (extension; no reference
    counterpart). types: 1 = spreading, 2 = interpolation, 3 = both. 2D plans
    spreading with row bands also get the fastest band height (type 1 timed).
    Returns {"bin_size": [...], "max_points": int, "band_rows": int, "ms": float,
    "source": int} (band_rows 0 = default; source 0 = timed by this call, 1 =
    this process's cache, 2 = the cache file of earlier processes)."""
    L = _L.lib()
    dim = len(modes)
    prec = _dtype_code(dtype) if dtype is not None else _precision_of(loc[0])
    vpp, _keep = _vpp([_ptr(p) for p in loc])
    best = _L.NnfTuning()
    ms = ctypes.c_double()
    _check(L.nnf_autotune(prec, dim, ctypes.byref(opt._c()), int(sign), int(num_nu), vpp,
                          _i64p(modes), ctypes.c_void_p(_stream(stream)), int(types), int(reps),
                          int(bool(use_cache)), ctypes.byref(best), ctypes.byref(ms)))
    return {"bin_size": list(best.bin_size)[:dim], "max_points": int(best.max_points),
            "band_rows": int(best.band_rows), "ms": float(ms.value),
            "source": int(best.reserved[0])}


class Plan:
    """Type 1 / type 2 plan: ``gpu::Plan<T,DIM>`` (include/neonufft/gpu/plan.hpp:25-46).

    ``loc`` holds DIM device arrays of num_nu coordinates (float32 or float64,
    which selects T). The constructor folds, rescales and bins the points and
    synchronises the stream; the transforms are asynchronous on the stream.
    """

    def __init__(self, opt: Options, sign: int, num_nu: int, loc: Sequence[Pointer],
                 modes: Sequence[int], stream=None, dtype=None, bin_size=None,
                 max_points: int = 0, spread_mode: str = "auto", band_rows: int = 0):
        """bin_size / max_points: optional tuning (extension; None/0 = defaults).
        bin_size="auto" runs autotune() (both kernels) on these points first.
        spread_mode ("auto" | "tiles" | "bands" | "bands_grid") and band_rows
        select the type-1 spreading of 2D plans (row bands with the fused
        x-FFT by default where the grid suits them)."""
        self._L = _L.lib()
        if isinstance(bin_size, str):
            if bin_size != "auto":
                raise InputError(f"bin_size must be a sequence or 'auto', not {bin_size!r}")
            t = autotune(opt, sign, num_nu, loc, modes, stream=stream, dtype=dtype)
            bin_size, max_points = t["bin_size"], t["max_points"]
            band_rows = band_rows or t["band_rows"]
        self.dim = len(modes)
        self.modes = [int(m) for m in modes]
        self.precision = _dtype_code(dtype) if dtype is not None else _precision_of(loc[0])
        self.num_nu = int(num_nu)
        self._h = ctypes.c_void_p()
        vpp, _keep = _vpp([_ptr(p) for p in loc])
        tune = _tuning(bin_size, max_points, spread_mode, band_rows)
        _check(self._L.nnf_plan_create_tuned(ctypes.byref(self._h), self.precision, self.dim,
                                             ctypes.byref(opt._c()), int(sign), self.num_nu,
                                             vpp, _i64p(self.modes),
                                             ctypes.c_void_p(_stream(stream)), tune))

    def set_points(self, num_nu: int, loc: Sequence[Pointer]) -> None:
        vpp, _keep = _vpp([_ptr(p) for p in loc])
        _check(self._L.nnf_plan_set_points(self._h, int(num_nu), vpp))
        self.num_nu = int(num_nu)

    def default_strides(self):
        s = [1]
        for m in self.modes[:-1]:
            s.append(s[-1] * m)
        return s

    def transform_type_1(self, in_: Pointer, out: Pointer, out_strides=None) -> None:
        st = _i64p(out_strides if out_strides is not None else self.default_strides())
        _check(self._L.nnf_plan_transform_type_1(self._h, ctypes.c_void_p(_ptr(in_)),
                                                 ctypes.c_void_p(_ptr(out)), st))

    def transform_type_2(self, in_: Pointer, in_strides, out: Pointer) -> None:
        st = _i64p(in_strides if in_strides is not None else self.default_strides())
        _check(self._L.nnf_plan_transform_type_2(self._h, ctypes.c_void_p(_ptr(in_)), st,
                                                 ctypes.c_void_p(_ptr(out))))

    # ---- stage-level extension (no reference counterpart) ----
    def spread(self, in_: Pointer, accumulate: bool = False) -> None:
        _check(self._L.nnf_plan_spread(self._h, ctypes.c_void_p(_ptr(in_)), int(accumulate)))

    def spread_tiles(self, in_: Pointer) -> None:
        """The spreading kernel alone (subgrid tiles; see spread())."""
        _check(self._L.nnf_plan_spread_tiles(self._h, ctypes.c_void_p(_ptr(in_))))

    def reduce_tiles(self, accumulate: bool = False) -> None:
        """The tile reduction alone: fine grid (+)= sum of covering tiles."""
        _check(self._L.nnf_plan_reduce_tiles(self._h, int(accumulate)))

    def interp(self, out: Pointer) -> None:
        _check(self._L.nnf_plan_interp(self._h, ctypes.c_void_p(_ptr(out))))

    def fine_grid_ptr(self) -> int:
        p = ctypes.c_void_p()
        _check(self._L.nnf_plan_fine_grid(self._h, ctypes.byref(p)))
        return int(p.value or 0)

    def fseries_ptr(self, d: int) -> int:
        p = ctypes.c_void_p()
        _check(self._L.nnf_plan_fseries(self._h, int(d), ctypes.byref(p)))
        return int(p.value or 0)

    def info(self) -> dict:
        i = _L.NnfPlanInfo()
        _check(self._L.nnf_plan_get_info(self._h, ctypes.byref(i)))
        return _info_dict(i)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _check(self._L.nnf_plan_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlanT3:
    """Type 3 plan: ``gpu::PlanT3<T,DIM>`` (include/neonufft/gpu/plan.hpp:48-82).

    Either ``PlanT3(opt, sign, num_in, input_points, num_out, output_points,
    stream, batch_size)`` or ``PlanT3.from_bounds(...)``.
    """

    def __init__(self, opt: Options, sign: int, num_in: int, input_points: Sequence[Pointer],
                 num_out: int, output_points: Sequence[Pointer], stream=None, batch_size: int = 1,
                 dtype=None, _handle=None, _dim=None, _precision=None):
        self._L = _L.lib()
        if _handle is not None:
            self._h = _handle
            self.dim = _dim
            self.precision = _precision
            return
        self.dim = len(input_points)
        self.precision = _dtype_code(dtype) if dtype is not None else _precision_of(input_points[0])
        self._h = ctypes.c_void_p()
        vin, _k1 = _vpp([_ptr(p) for p in input_points])
        vout, _k2 = _vpp([_ptr(p) for p in output_points])
        _check(self._L.nnf_plan_t3_create(ctypes.byref(self._h), self.precision, self.dim,
                                          ctypes.byref(opt._c()), int(sign), int(num_in), vin,
                                          int(num_out), vout, ctypes.c_void_p(_stream(stream)),
                                          int(batch_size)))

    @staticmethod
    def _bounds(vals):
        return (ctypes.c_double * 3)(*([float(v) for v in vals] + [0.0] * (3 - len(vals))))

    @classmethod
    def from_bounds(cls, opt: Options, sign: int, input_min, input_max, output_min, output_max,
                    stream=None, batch_size: int = 1, dtype="float32") -> "PlanT3":
        L = _L.lib()
        h = ctypes.c_void_p()
        prec = _dtype_code(dtype)
        dim = len(input_min)
        _check(L.nnf_plan_t3_create_bounds(ctypes.byref(h), prec, dim, ctypes.byref(opt._c()),
                                           int(sign), cls._bounds(input_min),
                                           cls._bounds(input_max), cls._bounds(output_min),
                                           cls._bounds(output_max),
                                           ctypes.c_void_p(_stream(stream)), int(batch_size)))
        return cls(opt, sign, 0, [], 0, [], _handle=h, _dim=dim, _precision=prec)

    @classmethod
    def grid_memory_size(cls, opt: Options, input_min, input_max, output_min, output_max,
                         dtype="float32") -> int:
        L = _L.lib()
        out = ctypes.c_uint64()
        _check(L.nnf_plan_t3_grid_memory_size(_dtype_code(dtype), len(input_min),
                                              ctypes.byref(opt._c()), cls._bounds(input_min),
                                              cls._bounds(input_max), cls._bounds(output_min),
                                              cls._bounds(output_max), ctypes.byref(out)))
        return int(out.value)

    def set_input_points(self, num_in: int, points: Sequence[Pointer]) -> None:
        v, _k = _vpp([_ptr(p) for p in points])
        _check(self._L.nnf_plan_t3_set_input_points(self._h, int(num_in), v))

    def set_output_points(self, num_out: int, points: Sequence[Pointer]) -> None:
        v, _k = _vpp([_ptr(p) for p in points])
        _check(self._L.nnf_plan_t3_set_output_points(self._h, int(num_out), v))

    def add_input(self, in_: Pointer, bdist: int = 0) -> None:
        _check(self._L.nnf_plan_t3_add_input(self._h, ctypes.c_void_p(_ptr(in_)), int(bdist)))

    def transform(self, out: Pointer, bdist: int = 0) -> None:
        _check(self._L.nnf_plan_t3_transform(self._h, ctypes.c_void_p(_ptr(out)), int(bdist)))

    def reset(self) -> None:
        _check(self._L.nnf_plan_t3_reset(self._h))

    def info(self) -> dict:
        i = _L.NnfPlanInfo()
        _check(self._L.nnf_plan_t3_get_info(self._h, ctypes.byref(i)))
        return _info_dict(i)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _check(self._L.nnf_plan_t3_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchPlan:
    """Many independent type 1 / type 2 transforms of uneven sizes, issued
    together (extension; no reference counterpart).

    Equivalent to ``count`` plans ``Plan(opt, sign, num_nu[t], loc[t],
    modes[t])`` run one after the other, but each transform call makes one
    spreading or interpolation launch over every transform's subproblems and
    one batched FFT per distinct mode shape; with ``use_graph`` the sequence
    is captured once into a HIP graph and replayed. Mode arrays are contiguous
    with x fastest.
    """

    def __init__(self, opt: Options, sign: int, num_nu: Sequence[int],
                 loc: Sequence[Sequence[Pointer]], modes: Sequence[Sequence[int]],
                 stream=None, dtype=None, use_graph: bool = True, spread_mode: str = "auto",
                 band_rows: int = 0):
        self._L = _L.lib()
        self.count = len(modes)
        if self.count == 0 or len(num_nu) != self.count or len(loc) != self.count:
            raise InputError("num_nu, loc and modes need one entry per transform")
        self.dim = len(modes[0])
        if any(len(m) != self.dim for m in modes) or any(len(l) != self.dim for l in loc):
            raise InputError("every transform needs the same dimension")
        self.modes = [[int(v) for v in m] for m in modes]
        self.precision = _dtype_code(dtype) if dtype is not None else _precision_of(loc[0][0])
        self.num_nu = [int(v) for v in num_nu]
        self._h = ctypes.c_void_p()
        flat_loc = [_ptr(p) for l in loc for p in l]
        vpp, _keep = _vpp(flat_loc)
        nn = (ctypes.c_int64 * self.count)(*self.num_nu)
        md = (ctypes.c_int64 * (self.count * self.dim))(*[v for m in self.modes for v in m])
        _check(self._L.nnf_batch_create_tuned(ctypes.byref(self._h), self.precision, self.dim,
                                              ctypes.byref(opt._c()), int(sign), self.count,
                                              nn, vpp, md, ctypes.c_void_p(_stream(stream)),
                                              int(bool(use_graph)),
                                              _tuning(None, 0, spread_mode, band_rows)))

    def set_points(self, num_nu: Sequence[int], loc: Sequence[Sequence[Pointer]]) -> None:
        if len(num_nu) != self.count or len(loc) != self.count:
            raise InputError("num_nu and loc need one entry per transform")
        vpp, _keep = _vpp([_ptr(p) for l in loc for p in l])
        nn = (ctypes.c_int64 * self.count)(*[int(v) for v in num_nu])
        _check(self._L.nnf_batch_set_points(self._h, nn, vpp))
        self.num_nu = [int(v) for v in num_nu]

    def _tables(self, ins, outs):
        if len(ins) != self.count or len(outs) != self.count:
            raise InputError("one input and one output per transform")
        a, ka = _vpp([_ptr(p) for p in ins])
        b, kb = _vpp([_ptr(p) for p in outs])
        return a, b, (ka, kb)

    def transform_type_1(self, ins: Sequence[Pointer], outs: Sequence[Pointer]) -> None:
        a, b, _keep = self._tables(ins, outs)
        _check(self._L.nnf_batch_transform_type_1(self._h, a, b))

    def transform_type_2(self, ins: Sequence[Pointer], outs: Sequence[Pointer]) -> None:
        a, b, _keep = self._tables(ins, outs)
        _check(self._L.nnf_batch_transform_type_2(self._h, a, b))

    def info(self) -> dict:
        i = _L.NnfBatchInfo()
        _check(self._L.nnf_batch_get_info(self._h, ctypes.byref(i)))
        return {
            "precision": "float32" if i.precision == 0 else "float64",
            "dim": i.dim, "n_spread": i.n_spread, "count": i.count,
            "num_points": i.num_points, "num_subproblems": i.num_subproblems,
            "num_shapes": i.num_shapes, "grid_cells": i.grid_cells, "graph": bool(i.graph),
            "spread_mode": _SPREAD_NAMES.get(i.spread_mode, "tiles"),
        }

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _check(self._L.nnf_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_SPREAD_MODES = {"auto": 0, "tiles": 1, "bands": 2, "bands_grid": 3}
_SPREAD_NAMES = {1: "tiles", 2: "bands", 3: "bands_grid"}


def _tuning(bin_size, max_points, spread_mode, band_rows):
    """nnf_tuning by reference, or None when everything is default."""
    if spread_mode not in _SPREAD_MODES:
        raise InputError(f"spread_mode must be one of {sorted(_SPREAD_MODES)}, not {spread_mode!r}")
    if bin_size is None and not max_points and spread_mode == "auto" and not band_rows:
        return None
    tune = _L.NnfTuning()
    for d, b in enumerate(list(bin_size or [])[:3]):
        tune.bin_size[d] = int(b)
    tune.max_points = int(max_points)
    tune.spread_mode = _SPREAD_MODES[spread_mode]
    tune.band_rows = int(band_rows)
    return ctypes.byref(tune)


def _info_dict(i: _L.NnfPlanInfo) -> dict:
    return {
        "precision": {0: "float32", 1: "float64", 2: "mixed"}.get(i.precision, "?"),
        "dim": i.dim,
        "n_spread": i.n_spread,
        "es_beta": i.es_beta,
        "modes": list(i.modes)[: i.dim],
        "grid": list(i.grid)[: i.dim],
        "spread_grid": list(i.spread_grid)[: i.dim],
        "bin_size": list(i.bin_size)[: i.dim],
        "num_bins": i.num_bins,
        "num_subproblems": i.num_subproblems,
        "num_points": i.num_points,
        "num_out_points": i.num_out_points,
        "horner_degree": i.horner_degree,
        "quad_nodes": i.quad_nodes,
        "spread_mode": _SPREAD_NAMES.get(i.spread_mode, "tiles"),
        "band_rows": i.band_rows,
    }

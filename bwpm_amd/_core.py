"""ctypes binding of libbwpm_hip.so: the module surface of the reference's `bwpm._core`.

Reference: cpp/python/bindings.cpp:122-180 (`transmission_matrix`, `reflection_matrix`,
`longitudinal_wavenumber`, `CudaSolver(...).run`, `HAS_CUDA`, `cuda_device_count`). Here the solver
is `GpuSolver` (alias `CudaSolver`), computing on an MI355X through the C ABI of
include/bwpm_hip.h. There is no CPU fallback: if the library is missing the import fails, and
without a GPU every compute call raises RuntimeError.

Error mapping as in the reference binding: invalid arguments → ValueError; device, allocation,
non-convergence → RuntimeError.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Any

import numpy as np

from bwpm_amd.specs import FluctuationSynthesis, Geometry, Uniform

_LIB_PATH = Path(os.environ.get("BWPM_HIP_LIB", Path(__file__).resolve().parent / "lib" / "libbwpm_hip.so"))
if not _LIB_PATH.exists():
    raise ImportError(
        f"libbwpm_hip.so not found at {_LIB_PATH}; build it with `python -m bwpm_amd.build`"
    )
_lib = ctypes.CDLL(str(_LIB_PATH))

OK, EINVAL, ENOCONV, EDEVICE, ENOMEM, EINTERNAL = 0, 1, 2, 3, 4, 5
PRECISIONS = {"fp64": 64, "mixed": 48, "bfp": 40, "fp32": 32, "fp32_plain": 31}

_d = ctypes.c_double
_i64 = ctypes.c_int64
_p = ctypes.c_void_p
_int = ctypes.c_int


def _sig(name: str, *args: Any) -> Any:
    f = getattr(_lib, name)
    f.argtypes = list(args)
    f.restype = ctypes.c_int
    return f


_lib.bwpm_last_error.restype = ctypes.c_char_p
_lib.bwpm_last_error.argtypes = []
_device_count = _sig("bwpm_device_count", ctypes.POINTER(_int))
_tmat = _sig("bwpm_transmission_matrix", _p, _p, _i64, _d, _d, _d, _d, _d, _p)
_rmat = _sig("bwpm_reflection_matrix", _p, _p, _i64, _d, _d, _d, _d, _d, _p)
_kz = _sig("bwpm_longitudinal_wavenumber", _d, _d, _d, _p, _i64, _p)
_create = _sig("bwpm_solver_create_ex", _i64, _i64, _d, _d, _d, _d, _p, _i64, _p, _i64, _d, _d, _p, _p, _p,
               _int, _int, ctypes.POINTER(_p))
_fresnel = _sig("bwpm_fresnel_coefficients", _p, _p, _i64, _d, _d, _d, _d, _d, _p)
_flux = _sig("bwpm_power_flux", _p, _i64, _i64, _d, _d, _d, _d, _d, _int, _int, ctypes.POINTER(_d))


def power_flux(field: Any, dy: float, dx: float, k0: float, index: complex, direction: int = 1,
               device: int = 0) -> float:
    """Poynting flux of a (ny, nx, 2) field in a homogeneous medium (device FP64 reduction)."""
    f = np.ascontiguousarray(np.asarray(field, dtype=np.complex128))
    if f.ndim != 3 or f.shape[2] != 2:
        raise ValueError("field must have shape (ny, nx, 2)")
    n = complex(index)
    out = _d(0)
    _check(_flux(_ptr(f), f.shape[0], f.shape[1], float(dy), float(dx), float(k0), n.real, n.imag, int(direction),
                 int(device), ctypes.byref(out)))
    return float(out.value)
_destroy = _sig("bwpm_solver_destroy", _p)
_run = _sig("bwpm_solver_run", _p, _p, _int, _p, _i64, ctypes.POINTER(_p))
_res_field = _sig("bwpm_result_field", _p, _int, _p)
_res_iters = _sig("bwpm_result_iterations", _p, ctypes.POINTER(_i64))
_res_iter_field = _sig("bwpm_result_iter_field", _p, _int, _i64, _p)
_res_plane = _sig("bwpm_result_plane", _p, _int, _i64, _p, ctypes.POINTER(_int))
_res_stored = _sig("bwpm_result_stored_values", _p, ctypes.POINTER(_i64))
_res_stats = _sig("bwpm_result_stats", _p, ctypes.POINTER(_d), ctypes.POINTER(_i64), ctypes.POINTER(_i64))
_res_free = _sig("bwpm_result_free", _p)
_upload = _sig("bwpm_solver_upload", _p, _p)
_run_resident = _sig("bwpm_solver_run_resident", _p, _int, _int, ctypes.POINTER(_d), ctypes.POINTER(_d),
                     ctypes.POINTER(_d), ctypes.POINTER(_i64), ctypes.POINTER(_i64))
_dev_bytes = _sig("bwpm_solver_device_bytes", _p, ctypes.POINTER(_i64))
_copy_probe = _sig("bwpm_copy_probe", _i64, _int, _int, ctypes.POINTER(_d))
_fetch = _sig("bwpm_solver_fetch", _p, _int, ctypes.POINTER(_p))
_q_create = _sig("bwpm_queue_create", _int, _d, _int, _int, ctypes.POINTER(_p))
_q_destroy = _sig("bwpm_queue_destroy", _p)
_q_submit = _sig("bwpm_queue_submit", _p, _i64, _i64, _d, _d, _d, _d, _p, _i64, _p, _i64, _d, _d, _p, _p, _p, _int,
                 _p, _int, ctypes.POINTER(_i64))
_q_submit_synth = _sig("bwpm_queue_submit_synth", _p, _i64, _i64, _d, _d, _d, _d, _p, _i64, _p, _int, _i64, _d, _d,
                       _p, _p, _d, _d, ctypes.c_uint64, _int, _p, _int, ctypes.POINTER(_i64))
_create_synth = _sig("bwpm_solver_create_synth", _i64, _i64, _d, _d, _d, _d, _p, _i64, _p, _int, _i64, _d, _d, _p,
                     _p, _d, _d, ctypes.c_uint64, _int, _int, ctypes.POINTER(_p))
_fluct = _sig("bwpm_solver_fluctuation", _p, _p)
_q_state = _sig("bwpm_queue_state", _p, _i64, ctypes.POINTER(_int))
_q_wait = _sig("bwpm_queue_wait", _p, _i64, _d, ctypes.POINTER(_int))
_q_result = _sig("bwpm_queue_result", _p, _i64, ctypes.POINTER(_p))
_q_stats = _sig("bwpm_queue_stats", _p, _p)
_a_create = _sig("bwpm_analysis_create", _int, _i64, _i64, _d, _d, _d, ctypes.POINTER(_p))
_a_destroy = _sig("bwpm_analysis_destroy", _p)
_a_power = _sig("bwpm_plane_power", _p, _p, _int, _i64, _d, _p)
_a_refl = _sig("bwpm_reference_reflectance", _p, _p, _int, _d, _d, ctypes.POINTER(_d))
_a_axial = _sig("bwpm_axial_profile", _p, _p, _int, _d, _p, _i64, _d, _p)
_a_strehl = _sig("bwpm_strehl_profile", _p, _p, _int, _d, _d, _p, _i64, _p)
_a_strehl_modes = _sig("bwpm_strehl_modes", _p, _p, _p, _i64, _d, _p, _i64, _p)
_geom_labels = _sig("bwpm_geometry_labels", _i64, _i64, _d, _d, _d, _i64, _d, _int, _p, _i64, _i64, _p)
_create_geom = _sig("bwpm_solver_create_geometry", _i64, _i64, _d, _d, _d, _d, _p, _i64, _i64, _d, _int, _p, _i64, _d,
                    _d, _p, _p, _d, _d, ctypes.c_uint64, _int, _int, ctypes.POINTER(_p))
_q_submit_geom = _sig("bwpm_queue_submit_geometry", _p, _i64, _i64, _d, _d, _d, _d, _p, _i64, _i64, _d, _int, _p, _i64,
                      _d, _d, _p, _p, _d, _d, ctypes.c_uint64, _int, _p, _int, ctypes.POINTER(_i64))
_graph_stats = _sig("bwpm_lane_graph_stats", ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64))
_fft_plans = _sig("bwpm_fft_plans", ctypes.c_char_p, _i64, ctypes.POINTER(_i64))
_solver_plane = _sig("bwpm_solver_plane", _p, _int, _i64, ctypes.POINTER(_p), ctypes.POINTER(_int))
_batch = _sig("bwpm_batch_run", _p, _i64, _int, _int, _int, ctypes.POINTER(_d), ctypes.POINTER(_d),
               ctypes.POINTER(_d), ctypes.POINTER(_i64), ctypes.POINTER(_i64))

LIBRARY = str(_LIB_PATH)


def _check(status: int) -> None:
    if status == OK:
        return
    msg = (_lib.bwpm_last_error() or b"").decode("utf-8", "replace")
    if status == EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg or f"bwpm_hip error {status}")


def geometry_labels(g: Geometry, ny: int, nx: int, dy: float, dx: float, dz: float, nlabels: int) -> np.ndarray:
    """The (nz, ny, nx) int32 label volume the library makes from a geometry (host code, no GPU)."""
    out = np.empty((int(g.nz), int(ny), int(nx)), dtype=np.int32)
    pk = g.packed()
    _check(_geom_labels(int(ny), int(nx), float(dy), float(dx), float(dz), int(g.nz), float(g.z0), int(g.background),
                        _ptr(pk) if pk.size else None, pk.shape[0], int(nlabels), _ptr(out)))
    return out


def lane_graph_stats() -> dict[str, int]:
    """Process-wide counts of continuous-lane graphs: built, replayed, fallbacks to direct launches."""
    b, r, f = _i64(0), _i64(0), _i64(0)
    _check(_graph_stats(ctypes.byref(b), ctypes.byref(r), ctypes.byref(f)))
    return {"built": int(b.value), "replays": int(r.value), "fallbacks": int(f.value)}


def fft_plans() -> list[dict]:
    """The radix plans the plan tuner chose in this process: per precision code, direction and line
    length, the plan taken (0: default, 1: alternative) and the timed launch of each (µs)."""
    n = _i64(0)
    _check(_fft_plans(None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(int(n.value) + 1)
    _check(_fft_plans(buf, int(n.value) + 1, ctypes.byref(n)))
    out = []
    for line in buf.value.decode().splitlines():
        prec, direction, length, plan, us0, us1 = line.split()
        out.append({"precision": int(prec), "direction": direction, "n": int(length), "plan": int(plan),
                    "us": [float(us0), float(us1)]})
    return out


def gpu_device_count() -> int:
    n = _int(0)
    _check(_device_count(ctypes.byref(n)))
    return int(n.value)


def copy_probe(nbytes: int, reps: int = 20, device: int = 0) -> float:
    """Average duration in ms of the plain copy kernel `k_copy_probe` over `nbytes`."""
    ms = _d(0)
    _check(_copy_probe(int(nbytes), int(reps), int(device), ctypes.byref(ms)))
    return float(ms.value)


HAS_GPU = True
HAS_CUDA = True  # name of the reference's flag: "the compiled core has a GPU backend"
cuda_device_count = gpu_device_count


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _f64(a: Any) -> np.ndarray:
    return np.array(a, dtype=np.float64, order="C", copy=True)


def _matrices(fn: Any, k_x: Any, k_y: Any, index_in: complex, index_out: complex, k0: float) -> np.ndarray:
    kx, ky = _f64(k_x), _f64(k_y)
    if kx.shape != ky.shape:
        raise ValueError("k_x and k_y must have the same shape")
    out = np.empty(kx.shape + (2, 2), dtype=np.complex128)
    a, b = complex(index_in), complex(index_out)
    _check(fn(_ptr(kx), _ptr(ky), kx.size, a.real, a.imag, b.real, b.imag, float(k0), _ptr(out)))
    return out


def transmission_matrix(k_x: Any, k_y: Any, index_in: complex, index_out: complex, k0: float) -> np.ndarray:
    """T of wende2023 eqs. (2), (3) without propagation factor, shape (..., 2, 2), complex128."""
    return _matrices(_tmat, k_x, k_y, index_in, index_out, k0)


def reflection_matrix(k_x: Any, k_y: Any, index_in: complex, index_out: complex, k0: float) -> np.ndarray:
    """R of wende2023 eqs. (8), (9) without propagation factor, shape (..., 2, 2), complex128."""
    return _matrices(_rmat, k_x, k_y, index_in, index_out, k0)


def fresnel_coefficients(k_x: Any, k_y: Any, index_in: complex, index_out: complex, k0: float) -> np.ndarray:
    """(r_s, r_p, t_s, t_p, k_z,in, k_z,out) stacked on the last axis, Convention A."""
    kx, ky = _f64(k_x), _f64(k_y)
    if kx.shape != ky.shape:
        raise ValueError("k_x and k_y must have the same shape")
    out = np.empty(kx.shape + (6,), dtype=np.complex128)
    a, b = complex(index_in), complex(index_out)
    _check(_fresnel(_ptr(kx), _ptr(ky), kx.size, a.real, a.imag, b.real, b.imag, float(k0), _ptr(out)))
    return out


def longitudinal_wavenumber(index: complex, k0: float, k_perp_sq: Any) -> Any:
    """k_z = sqrt(ñ²k₀² − k⊥²) with Im ≥ 0 and the grazing rule; scalar in → complex out."""
    kp2 = _f64(k_perp_sq)
    out = np.empty(kp2.shape, dtype=np.complex128)
    n = complex(index)
    _check(_kz(n.real, n.imag, float(k0), _ptr(kp2), kp2.size, _ptr(out)))
    return complex(out) if out.ndim == 0 else out


def _solver_arrays(ny: int, nx: int, indices: Any, labels: Any, mask: Any, fluctuation: Any, labels_before: Any,
                   precision: str) -> tuple:
    """Checked, contiguous copies of a solve's arrays (shapes as cpp/python/bindings.cpp:100-115)."""
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
    if isinstance(labels, (Uniform, Geometry)):
        if int(labels.nz) < 1:
            raise ValueError("a uniform or geometry volume needs nz >= 1")
        lab = labels
    else:
        lab = np.asarray(labels)
        if lab.ndim != 3:
            raise ValueError("labels must have shape (nz, ny, nx)")
        lab = np.ascontiguousarray(lab, dtype=np.int32)
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.complex128).ravel())
    m = None if mask is None else _f64(mask)
    if m is not None and m.size != int(ny) * int(nx):
        raise ValueError("mask must have shape (ny, nx)")
    f = None
    if isinstance(fluctuation, FluctuationSynthesis):
        if fluctuation.sigma < 0 or fluctuation.corr_mm < 0:
            raise ValueError("sigma and corr_mm must be >= 0")
        f = fluctuation
    elif fluctuation is not None:
        if isinstance(lab, (Uniform, Geometry)):
            raise ValueError("a fluctuation array needs a label array")
        f = np.ascontiguousarray(np.asarray(fluctuation, dtype=np.complex128))
        if f.shape != lab.shape:
            raise ValueError("fluctuation must have the shape of labels")
    if isinstance(lab, Uniform):
        if not isinstance(f, FluctuationSynthesis):
            raise ValueError("a uniform label volume is given with a synthesised fluctuation")
    elif isinstance(lab, Geometry):
        pass
    elif lab.shape[1:] != (int(ny), int(nx)):
        raise ValueError("labels must have shape (nz, ny, nx)")
    lb = None
    if labels_before is not None:
        lb = np.ascontiguousarray(np.asarray(labels_before), dtype=np.int32)
        if lb.size != int(ny) * int(nx):
            raise ValueError("labels_before must have shape (ny, nx)")
    return idx, lab, m, f, lb


def _collect(res: Any, ny: int, nx: int, keep_list: list[int]) -> dict[str, Any]:
    """Result dict of a bwpm_result (freed here): the keys of cpp/python/bindings.cpp:71-85 plus stats."""
    try:
        shape = (ny, nx, 2)

        def field_of(which: int) -> np.ndarray:
            out = np.empty(shape, dtype=np.complex128)
            _check(_res_field(res, which, _ptr(out)))
            return out

        out: dict[str, Any] = {"transmitted": field_of(0), "reflected": field_of(1)}
        n = _i64(0)
        _check(_res_iters(res, ctypes.byref(n)))
        ti, ri = [], []
        for it in range(n.value):
            for which, lst in ((0, ti), (1, ri)):
                a = np.empty(shape, dtype=np.complex128)
                _check(_res_iter_field(res, which, it, _ptr(a)))
                lst.append(a)
        out["transmitted_iter"] = ti
        out["reflected_iter"] = ri
        fw: dict[int, np.ndarray] = {}
        bw: dict[int, np.ndarray] = {}
        present = _int(0)
        for p in sorted(set(keep_list)):
            for which, dct in ((0, fw), (1, bw)):
                a = np.empty(shape, dtype=np.complex128)
                _check(_res_plane(res, which, p, _ptr(a), ctypes.byref(present)))
                if present.value:
                    dct[p] = a
        out["forward"] = fw
        out["backward"] = bw
        sv = _i64(0)
        _check(_res_stored(res, ctypes.byref(sv)))
        out["stored_values"] = int(sv.value)
        ms, steps, gm = _d(0), _i64(0), _i64(0)
        _check(_res_stats(res, ctypes.byref(ms), ctypes.byref(steps), ctypes.byref(gm)))
        out["gpu_ms"] = float(ms.value)
        out["slab_steps"] = int(steps.value)
        out["gmres_iterations"] = int(gm.value)
        return out
    finally:
        _res_free(res)


class GpuSolver:
    """FPWPM/BWPM on one MI355X: `GpuSolver(ny, nx, dy, dx, dz, k0, indices, labels, index_before,
    mask, device=0, fluctuation=None, precision="fp64")` with `.run(field, iterations, keep=[])`.

    Arguments and result dict as the reference's `CudaSolver` (cpp/python/bindings.cpp:158-173);
    `precision` selects complex128 ("fp64") or the compensated complex64 bulk path ("fp32").
    """

    def __init__(self, ny: int, nx: int, dy: float, dx: float, dz: float, k0: float, indices: Any,
                 labels: Any, index_before: complex, mask: Any = None, device: int = 0,
                 fluctuation: Any = None, precision: str = "fp64", labels_before: Any = None) -> None:
        self._h = _p(None)
        idx, lab, m, f, lb = _solver_arrays(ny, nx, indices, labels, mask, fluctuation, labels_before, precision)
        ib = complex(index_before)
        nz = int(lab.nz) if isinstance(lab, (Uniform, Geometry)) else int(lab.shape[0])
        self.ny, self.nx, self.nz = int(ny), int(nx), nz
        self.precision = precision
        self.device = int(device)
        if isinstance(lab, Geometry):
            pk = lab.packed()
            sig = f if isinstance(f, FluctuationSynthesis) else FluctuationSynthesis(0.0, 0.0, 0)
            _check(_create_geom(int(ny), int(nx), float(dy), float(dx), float(dz), float(k0), _ptr(idx), idx.size, nz,
                                float(lab.z0), int(lab.background), _ptr(pk) if pk.size else None, pk.shape[0],
                                ib.real, ib.imag, None if m is None else _ptr(m), None if lb is None else _ptr(lb),
                                float(sig.sigma), float(sig.corr_mm), int(sig.seed), int(device),
                                PRECISIONS[precision], ctypes.byref(self._h)))
            return
        if isinstance(f, FluctuationSynthesis):
            arr = None if isinstance(lab, Uniform) else _ptr(lab)
            ul = int(lab.label) if isinstance(lab, Uniform) else -1
            _check(_create_synth(int(ny), int(nx), float(dy), float(dx), float(dz), float(k0), _ptr(idx), idx.size,
                                 arr, ul, nz, ib.real, ib.imag, None if m is None else _ptr(m),
                                 None if lb is None else _ptr(lb), float(f.sigma), float(f.corr_mm), int(f.seed),
                                 int(device), PRECISIONS[precision], ctypes.byref(self._h)))
            return
        _check(_create(int(ny), int(nx), float(dy), float(dx), float(dz), float(k0), _ptr(idx), idx.size,
                       _ptr(lab), lab.shape[0], ib.real, ib.imag, None if m is None else _ptr(m),
                       None if f is None else _ptr(f), None if lb is None else _ptr(lb), int(device),
                       PRECISIONS[precision], ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h and self._h.value:
            _destroy(self._h)
            self._h = _p(None)

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    def _field(self, field: Any) -> np.ndarray:
        a = np.asarray(field)
        if a.ndim != 3 or a.shape[2] != 2 or a.shape[:2] != (self.ny, self.nx):
            raise ValueError("field must have shape (ny, nx, 2)")
        return np.ascontiguousarray(a, dtype=np.complex128)

    def run(self, field: Any, iterations: int, keep: Any = ()) -> dict[str, Any]:
        """Bidirectional solve; the dict of cpp/python/bindings.cpp:71-85 plus `gpu_ms`."""
        e = self._field(field)
        if int(iterations) < 0:
            raise ValueError("iterations must be >= 0")
        keep_list = [int(k) for k in keep]
        kp = np.ascontiguousarray(np.asarray(keep_list, dtype=np.int64))
        res = _p(None)
        _check(_run(self._h, _ptr(e), int(iterations), _ptr(kp) if kp.size else None, kp.size, ctypes.byref(res)))
        return self._collect(res, keep_list)

    def fetch(self, iterations: int) -> dict[str, Any]:
        """Results of the last resident or batched run (no kept planes)."""
        res = _p(None)
        _check(_fetch(self._h, int(iterations), ctypes.byref(res)))
        return self._collect(res, [])

    def _collect(self, res: Any, keep_list: list[int]) -> dict[str, Any]:
        return _collect(res, self.ny, self.nx, keep_list)

    # device-resident timing path (fields stay in HBM)
    def upload(self, field: Any) -> None:
        e = self._field(field)
        _check(_upload(self._h, _ptr(e)))

    def run_resident(self, iterations: int, time_kernels: bool = False) -> dict[str, float]:
        ms, cms, rms = _d(0), _d(0), _d(0)
        cl, rl = _i64(0), _i64(0)
        _check(_run_resident(self._h, int(iterations), 1 if time_kernels else 0, ctypes.byref(ms),
                             ctypes.byref(cms), ctypes.byref(rms), ctypes.byref(cl), ctypes.byref(rl)))
        return {"gpu_ms": ms.value, "col_ms": cms.value, "row_ms": rms.value, "col_launches": cl.value,
                "row_launches": rl.value}

    def fluctuation(self) -> np.ndarray:
        """The δn the solver holds, (nz, ny, nx) complex128 (generated or given; zeros without)."""
        out = np.empty((self.nz, self.ny, self.nx), dtype=np.complex128)
        _check(_fluct(self._h, _ptr(out)))
        return out

    def device_bytes(self) -> int:
        b = _i64(0)
        _check(_dev_bytes(self._h, ctypes.byref(b)))
        return int(b.value)

    def device_plane(self, which: str, plane: int = 0) -> "DevicePlane":
        """A plane of the last run left on the device: "transmitted", "reflected", or the
        "forward" / "backward" field at a kept plane. Valid until the next run or close; the
        analysis functions of `bwpm_amd.pilot_imc3d` read it without a copy to the host."""
        ptr, prec = _p(None), _int(0)
        _check(_solver_plane(self._h, PLANE_KINDS[which], int(plane), ctypes.byref(ptr), ctypes.byref(prec)))
        return DevicePlane(int(ptr.value), prec.value == 64, self.ny, self.nx, self)


PLANE_KINDS = {"transmitted": 0, "reflected": 1, "forward": 2, "backward": 3}
PLANE_HOST_C128, PLANE_DEVICE_C128, PLANE_DEVICE_C64 = 0, 1, 2


class DevicePlane:
    """Address of a (ny, nx, 2) plane in device memory (complex128 if `fp64`, else complex64),
    with a reference to its owner so that the memory outlives the object."""

    def __init__(self, ptr: int, fp64: bool, ny: int, nx: int, owner: Any = None) -> None:
        self.ptr, self.fp64, self.ny, self.nx, self._owner = ptr, fp64, ny, nx, owner

    @property
    def where(self) -> int:
        return PLANE_DEVICE_C128 if self.fp64 else PLANE_DEVICE_C64


class GpuAnalysis:
    """Analysis reductions of one grid on one device (bwpm_analysis_*): flux, Fresnel reference
    reflectance, axial intensity scan and Strehl profiles, FP64 with deterministic sums."""

    def __init__(self, ny: int, nx: int, dy: float, dx: float, k0: float, device: int = 0) -> None:
        self.ny, self.nx = int(ny), int(nx)
        self._h = _p(None)
        _check(_a_create(int(device), self.ny, self.nx, float(dy), float(dx), float(k0), ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h and self._h.value:
            _a_destroy(self._h)
            self._h = _p(None)

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    def _plane(self, field: Any, count: int | None = None) -> tuple[Any, int, Any]:
        if isinstance(field, DevicePlane):
            if (field.ny, field.nx) != (self.ny, self.nx):
                raise ValueError("the plane's grid differs from the analysis grid")
            return _p(field.ptr), field.where, field
        a = np.ascontiguousarray(np.asarray(field, dtype=np.complex128))
        if a.shape[-3:] != (self.ny, self.nx, 2):
            raise ValueError(f"field must have shape (..., {self.ny}, {self.nx}, 2)")
        return _ptr(a), PLANE_HOST_C128, a

    def plane_power(self, field: Any, index: float) -> np.ndarray:
        """Flux of each plane; field (ny, nx, 2), (count, ny, nx, 2) or a DevicePlane."""
        ptr, where, keep = self._plane(field)
        count = 1 if where != PLANE_HOST_C128 or keep.ndim == 3 else int(np.prod(keep.shape[:-3]))
        out = np.zeros(count, dtype=np.float64)
        _check(_a_power(self._h, ptr, where, count, float(index), _ptr(out)))
        return out

    def reference_reflectance(self, field: Any, n1: float, n2: float) -> float:
        ptr, where, keep = self._plane(field)
        r = _d(0)
        _check(_a_refl(self._h, ptr, where, float(n1), float(n2), ctypes.byref(r)))
        return float(r.value)

    def axial_profile(self, field: Any, index: float, offsets: Any, na: float | None = None) -> np.ndarray:
        ptr, where, keep = self._plane(field)
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.float64).ravel())
        out = np.zeros(off.size, dtype=np.float64)
        _check(_a_axial(self._h, ptr, where, float(index), _ptr(off), off.size, -1.0 if na is None else float(na),
                        _ptr(out)))
        return out

    def strehl_profile(self, spectrum: Any, k: float, na: float, offsets: Any) -> np.ndarray:
        a = np.ascontiguousarray(np.asarray(spectrum, dtype=np.complex128))
        if a.shape != (self.ny, self.nx):
            raise ValueError(f"spectrum must have shape ({self.ny}, {self.nx})")
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.float64).ravel())
        out = np.zeros(off.size, dtype=np.float64)
        _check(_a_strehl(self._h, _ptr(a), PLANE_HOST_C128, float(k), float(na), _ptr(off), off.size, _ptr(out)))
        return out

    def strehl_modes(self, ky: Any, spectrum: Any, k: float, offsets: Any) -> np.ndarray:
        kyv = np.ascontiguousarray(np.asarray(ky, dtype=np.float64).ravel())
        a = np.ascontiguousarray(np.asarray(spectrum, dtype=np.complex128).ravel())
        if a.size != kyv.size:
            raise ValueError("ky and spectrum must have the same length")
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.float64).ravel())
        out = np.zeros(off.size, dtype=np.float64)
        _check(_a_strehl_modes(self._h, _ptr(kyv), _ptr(a), a.size, float(k), _ptr(off), off.size, _ptr(out)))
        return out


CudaSolver = GpuSolver  # the reference's name for its GPU solver class


class Solver(GpuSolver):
    """The reference's `Solver` signature, `Solver(ny, nx, dy, dx, dz, k0, indices, labels,
    index_before, mask, threads=1, fluctuation=None)` with `.run(field, iterations, keep=[])`
    (cpp/python/bindings.cpp:139-153), so that callers choosing the reference's non-CUDA branch
    (src/bwpm/fast.py:86-98) keep working. It computes on GPU 0 in FP64 through the same C ABI as
    `GpuSolver`; `threads` (the reference's OpenMP thread count, any int, clamped to >= 1 there) is
    accepted and has no effect on the result.
    """

    def __init__(self, ny: int, nx: int, dy: float, dx: float, dz: float, k0: float, indices: Any,
                 labels: Any, index_before: complex, mask: Any = None, threads: int = 1,
                 fluctuation: Any = None) -> None:
        self.threads = max(int(threads), 1)
        super().__init__(ny, nx, dy, dx, dz, k0, indices, labels, index_before, mask, device=0,
                         fluctuation=fluctuation, precision="fp64")


def run_batch(solvers: list[GpuSolver], iterations: int, max_lane: int = 64, time_kernels: bool = False) -> Any:
    """Run every solver (fields uploaded with `upload`) as one batch; returns the wall time in ms
    (with `time_kernels`, a dict that adds the summed per-launch HIP-event durations).

    Solvers of one grid and precision share lanes of at most `max_lane` that step in lockstep
    (one column and one row launch per slab step for the whole lane); collect results with
    `fetch(iterations)`. Each result is bit-identical to the solver's run alone."""
    arr = (_p * len(solvers))(*[s._h.value for s in solvers])
    ms, cms, rms = _d(0), _d(0), _d(0)
    cl, rl = _i64(0), _i64(0)
    _check(_batch(arr, len(solvers), int(iterations), int(max_lane), 1 if time_kernels else 0, ctypes.byref(ms),
                  ctypes.byref(cms), ctypes.byref(rms), ctypes.byref(cl), ctypes.byref(rl)))
    if not time_kernels:
        return float(ms.value)
    return {"wall_ms": ms.value, "col_ms": cms.value, "row_ms": rms.value, "col_launches": cl.value,
            "row_launches": rl.value}


TICKET_STATES = ("pending", "admitted", "done", "failed")
QUEUE_STATS = ("p50_ms", "p99_ms", "mean_ms", "updates_per_s", "completed", "failed", "pending", "running",
               "used_bytes", "budget_bytes", "peak_bytes")


class GpuQueue:
    """Request queue with continuous batching on one device (the bwpm_queue_* entry points).

    `submit(...)` takes the arguments of `GpuSolver` plus the input field and iterations and returns
    a ticket at once; admission by HBM footprint, lane scheduling and the result copy run on the
    library's own threads. `result(ticket)` returns the dict of `GpuSolver.run` (no kept planes)."""

    def __init__(self, device: int = 0, hbm_budget: float | None = None, slots: int = 16, chunk: int = 64) -> None:
        self._h = _p(None)
        self.device = int(device)
        _check(_q_create(int(device), float(hbm_budget or 0.0), int(slots), int(chunk), ctypes.byref(self._h)))
        self._shape: dict[int, tuple[int, int]] = {}

    def close(self) -> None:
        if self._h and self._h.value:
            _q_destroy(self._h)
            self._h = _p(None)

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    def submit(self, ny: int, nx: int, dy: float, dx: float, dz: float, k0: float, indices: Any, labels: Any,
               index_before: complex, mask: Any, field: Any, iterations: int, fluctuation: Any = None,
               precision: str = "fp64", labels_before: Any = None) -> int:
        idx, lab, m, f, lb = _solver_arrays(ny, nx, indices, labels, mask, fluctuation, labels_before, precision)
        e = np.asarray(field)
        if e.shape != (int(ny), int(nx), 2):
            raise ValueError("field must have shape (ny, nx, 2)")
        e = np.ascontiguousarray(e, dtype=np.complex128)
        if int(iterations) < 0:
            raise ValueError("iterations must be >= 0")
        ib = complex(index_before)
        t = _i64(0)
        if isinstance(lab, Geometry):
            pk = lab.packed()
            sig = f if isinstance(f, FluctuationSynthesis) else FluctuationSynthesis(0.0, 0.0, 0)
            _check(_q_submit_geom(self._h, int(ny), int(nx), float(dy), float(dx), float(dz), float(k0), _ptr(idx),
                                  idx.size, int(lab.nz), float(lab.z0), int(lab.background),
                                  _ptr(pk) if pk.size else None, pk.shape[0], ib.real, ib.imag,
                                  None if m is None else _ptr(m), None if lb is None else _ptr(lb), float(sig.sigma),
                                  float(sig.corr_mm), int(sig.seed), PRECISIONS[precision], _ptr(e), int(iterations),
                                  ctypes.byref(t)))
        elif isinstance(f, FluctuationSynthesis):
            uni = isinstance(lab, Uniform)
            _check(_q_submit_synth(self._h, int(ny), int(nx), float(dy), float(dx), float(dz), float(k0), _ptr(idx),
                                   idx.size, None if uni else _ptr(lab), int(lab.label) if uni else -1,
                                   int(lab.nz) if uni else lab.shape[0], ib.real, ib.imag,
                                   None if m is None else _ptr(m), None if lb is None else _ptr(lb), float(f.sigma),
                                   float(f.corr_mm), int(f.seed), PRECISIONS[precision], _ptr(e), int(iterations),
                                   ctypes.byref(t)))
        else:
            _check(_q_submit(self._h, int(ny), int(nx), float(dy), float(dx), float(dz), float(k0), _ptr(idx),
                             idx.size, _ptr(lab), lab.shape[0], ib.real, ib.imag, None if m is None else _ptr(m),
                             None if f is None else _ptr(f), None if lb is None else _ptr(lb), PRECISIONS[precision],
                             _ptr(e), int(iterations), ctypes.byref(t)))
        self._shape[int(t.value)] = (int(ny), int(nx))
        return int(t.value)

    def state(self, ticket: int) -> str:
        st = _int(0)
        _check(_q_state(self._h, int(ticket), ctypes.byref(st)))
        return TICKET_STATES[st.value]

    def wait(self, ticket: int, timeout: float | None = None) -> str:
        st = _int(0)
        _check(_q_wait(self._h, int(ticket), -1.0 if timeout is None else float(timeout), ctypes.byref(st)))
        return TICKET_STATES[st.value]

    def result(self, ticket: int) -> dict[str, Any]:
        ny, nx = self._shape.pop(int(ticket))
        res = _p(None)
        _check(_q_result(self._h, int(ticket), ctypes.byref(res)))
        return _collect(res, ny, nx, [])

    def stats(self) -> dict[str, float]:
        out = np.zeros(len(QUEUE_STATS), dtype=np.float64)
        _check(_q_stats(self._h, _ptr(out)))
        return dict(zip(QUEUE_STATS, (float(v) for v in out)))

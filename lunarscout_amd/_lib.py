"""ctypes binding of liblunarscout_hip.so (declared in include/lunarscout_hip.h).

Loading fails loudly: if the library is missing or no HIP device is visible,
``HipSession`` raises ``HipBackendError`` -- there is no CPU fallback in the
product path.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

from .errors import ContractValidationError, HipBackendError, HipExecutionError

LIBRARY_PATH = Path(__file__).resolve().parent / "liblunarscout_hip.so"

LSH_OK = 0
LSH_EINVAL = -1
LSH_ENODEV = -2
LSH_EDEVICE = -3
LSH_ENOMEM = -4
LSH_EIO = -5

TILE_ABSENT = 0
TILE_COMPLETE = 1
TILE_INCOMPLETE = 2

HORIZON_DEGREES = 1
HORIZON_CAP = 2
HORIZON_COUNTERS = 4
HORIZON_F32_VALUES = 8
HORIZON_AZIMUTH_MAJOR = 16

LIGHT_SUN_BYTES = 1
LIGHT_SUN_FRACTION = 2
LIGHT_SUN_MEAN = 4
LIGHT_EARTH_MARGIN = 8
LIGHT_EARTH_VISIBLE = 16
LIGHT_PSR = 32
LIGHT_PSR_MARGIN = 64

REDUCE_LONGEST_DURATION = 0
REDUCE_SAFE_HAVEN = 1
SIGNAL_SUN_FRACTION = 0
SIGNAL_SUN_MARGIN = 1

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_f64 = ctypes.c_double
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t

_lib = None
_lock = threading.Lock()


class DurationSpec(ctypes.Structure):
    """lsh_duration_spec (include/lunarscout_hip.h)."""

    _fields_ = [
        ("reducer", _i32),
        ("sun_signal", _i32),
        ("sun_threshold", _f64),
        ("use_earth", _i32),
        ("earth_threshold_deg", _f64),
        ("time_count", _i32),
        ("sample_hours", _vp),
        ("interval_count", _i32),
        ("interval_first", _vp),
        ("interval_end", _vp),
        ("output_days", _i32),
        ("month_of", _vp),
        ("band_count", _i32),
        ("step_hours", _f64),
    ]

# (name, argtypes) of every exported entry point; tests check the library
# exports each of them.
SIGNATURES = {
    "lsh_last_error": [ctypes.c_char_p, _sz],
    "lsh_version": [],
    "lsh_device_count": [ctypes.POINTER(_i32)],
    "lsh_device_info": [_i32, ctypes.c_char_p, _sz, ctypes.POINTER(_i32)],
    "lsh_session_create": [_i32, ctypes.POINTER(_vp)],
    "lsh_session_destroy": [_vp],
    "lsh_alloc": [_vp, _sz, ctypes.POINTER(_vp)],
    "lsh_free": [_vp, _vp],
    "lsh_copy_to_device": [_vp, _vp, _vp, _sz],
    "lsh_copy_to_host": [_vp, _vp, _vp, _sz],
    "lsh_memset_device": [_vp, _vp, ctypes.c_int, _sz],
    "lsh_synchronize": [_vp],
    "lsh_dem_create": [_vp, _vp, _i32, _i32, _vp, _vp, _vp, ctypes.POINTER(_i32)],
    "lsh_dem_destroy": [_vp, _i32],
    "lsh_pyramid_layout": [_i32, _i32, _vp, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i64)],
    "lsh_dem_pyramid": [_vp, _i32, _vp, _i64],
    "lsh_dem_device_pointers": [_vp, _i32, ctypes.POINTER(_vp), ctypes.POINTER(_vp)],
    "lsh_segments_tile": [_vp, _vp, _i32, _i32, _i32, _i32, _i32, _f64, _f64, _i32, _vp],
    "lsh_horizon_tile": [_vp, _vp, _i32, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _f32, _i32,
                         _vp, _vp],
    "lsh_horizon_tile_full": [_vp, _vp, _i32, _i32, _i32, _i32, _f64, _f32, _i32, _vp, _vp],
    "lsh_lighting_tiles": [_vp, _i32, _vp, _i32, _vp, _i32, _i32, _vp, _i32, _vp, _i32, _vp, _i32,
                           _f64, _f64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "lsh_lighting_tiles_static": [_vp, _i32, _vp, _i32, _vp, _i32, _i32, _vp, _i32, _vp, _i32,
                                  _vp, _i32, _f64, _f64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "lsh_lighting_tiles_host": [_vp, _i32, _i32, _vp, _vp, _vp, _i32, _vp, _i32, _i32, _vp, _i32,
                                _vp, _i32, _vp, _i32, _f64, _f64, _i32, _i32, _vp, _vp, _vp, _vp,
                                _vp, _vp, _vp],
    "lsh_duration_workspace_bytes": [_vp, _i32, ctypes.POINTER(_i64)],
    "lsh_duration_tiles": [_vp, _i32, _vp, _i32, _vp, _i32, _i32, _vp, _vp, _f64, _vp, _vp, _vp],
    "lsh_duration_tiles_host": [_vp, _i32, _i32, _vp, _vp, _vp, _i32, _vp, _i32, _i32, _vp, _vp,
                                _f64, _vp, _i32, _vp, _vp],
    "lsh_psr_reduce": [_vp, _i32, _i32, _vp, _vp, _vp, _i32, _i32, _vp, ctypes.POINTER(_i32)],
    "lsh_sun_disk": [_f64, _vp, _vp],
    "lsh_cbin_encode": [_vp, _i32, _i32, _vp, _i64, ctypes.POINTER(_i64), _i32],
    "lsh_cbin_decode": [_vp, _i64, _i32, _i32, _vp, _i32],
    "lsh_cbin_complete": [_vp, _i64, _i32, _i32, ctypes.POINTER(_i32)],
    "lsh_tile_file_status": [ctypes.c_char_p, _i32, _i32, ctypes.POINTER(_i32)],
    "lsh_tile_file_read": [ctypes.c_char_p, _i32, _i32, _vp, _i32, _vp],
    "lsh_tile_file_write": [ctypes.c_char_p, _vp, _i32, _i32, _i32, _i32],
    "lsh_horizon_forward": [_vp, _vp, _i32, _vp, _i32, _i32, _i32, _f32, _i32, _vp, _vp],
    "lsh_sun_loss": [_vp, _i32, _vp, _i32, _vp, _i32, _vp, _i32, _f64, _f64, _vp, _vp],
    "lsh_horizon_adjoint": [_vp, _vp, _i32, _vp, _i32, _i32, _i32, _f32, _vp, _vp, _vp, _f64, _vp],
    "lsh_batch_shape_count": [ctypes.POINTER(_i32)],
    "lsh_batch_shape": [_i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i32)],
    "lsh_batch_resident_workgroups": [_vp, _i32, ctypes.POINTER(_i32)],
    "lsh_batch_create": [_vp, _vp, _i32, _i32, _i32, _f64, ctypes.POINTER(_vp)],
    "lsh_batch_run": [_vp, _vp, _i32, _i32, _i32, _i32, _i32, _vp],
    "lsh_batch_destroy": [_vp],
    "lsh_probe_copy": [_vp, _vp, _vp, ctypes.c_size_t],
    "lsh_probe_gather": [_vp, _vp, _i32, _i32, _i32, _i32, _vp],
    "lsh_event_create": [_vp, ctypes.POINTER(_vp)],
    "lsh_event_record": [_vp, _vp],
    "lsh_event_elapsed_ms": [_vp, _vp, ctypes.POINTER(_f32)],
    "lsh_event_destroy": [_vp],
}


def load_library() -> ctypes.CDLL:
    """Load the HIP library once; raise HipBackendError if it is absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIBRARY_PATH.exists():
            raise HipBackendError(
                f"{LIBRARY_PATH.name} is not built; run `python -m lunarscout_amd.build`"
            )
        try:
            lib = ctypes.CDLL(str(LIBRARY_PATH))
        except OSError as error:
            raise HipBackendError(f"cannot load {LIBRARY_PATH.name}: {error}") from error
        for name, argtypes in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_char_p if name == "lsh_version" else ctypes.c_int
        _lib = lib
        return lib


def last_error() -> str:
    lib = load_library()
    buf = ctypes.create_string_buffer(4096)
    lib.lsh_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int, what: str) -> None:
    if status == LSH_OK:
        return
    message = f"{what}: {last_error()}"
    if status == LSH_EINVAL:
        raise ContractValidationError(message)
    if status == LSH_ENODEV:
        raise HipBackendError(message)
    if status == LSH_EIO:
        raise OSError(message)
    raise HipExecutionError(message)


def device_count() -> int:
    try:
        lib = load_library()
    except HipBackendError:
        return 0
    n = _i32()
    lib.lsh_device_count(ctypes.byref(n))
    return int(n.value)


def _ptr(array: np.ndarray) -> int:
    return array.ctypes.data


class DeviceBuffer:
    """Device allocation owned by a session (freed on ``free`` or collection)."""

    def __init__(self, session: "Session", nbytes: int, dtype, shape):
        self.session = session
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape)
        ptr = _vp()
        check(load_library().lsh_alloc(session.handle, max(self.nbytes, 1), ctypes.byref(ptr)),
              "lsh_alloc")
        self.ptr = int(ptr.value)

    @classmethod
    def empty(cls, session: "Session", shape, dtype) -> "DeviceBuffer":
        shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        nbytes = int(np.prod(shape, dtype=np.int64)) * np.dtype(dtype).itemsize
        return cls(session, nbytes, dtype, shape)

    @classmethod
    def from_host(cls, session: "Session", array: np.ndarray) -> "DeviceBuffer":
        array = np.ascontiguousarray(array)
        buf = cls(session, array.nbytes, array.dtype, array.shape)
        buf.upload(array)
        return buf

    def upload(self, array: np.ndarray) -> None:
        array = np.ascontiguousarray(array, dtype=self.dtype)
        if array.nbytes != self.nbytes:
            raise ContractValidationError("upload size does not match the device buffer")
        check(load_library().lsh_copy_to_device(self.session.handle, self.ptr, _ptr(array),
                                                array.nbytes), "lsh_copy_to_device")

    def to_host(self, out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.empty(self.shape, dtype=self.dtype)
        if out.nbytes != self.nbytes or not out.flags.c_contiguous:
            raise ContractValidationError("host buffer does not match the device buffer")
        check(load_library().lsh_copy_to_host(self.session.handle, _ptr(out), self.ptr,
                                              self.nbytes), "lsh_copy_to_host")
        return out

    def zero(self) -> None:
        check(load_library().lsh_memset_device(self.session.handle, self.ptr, 0, self.nbytes),
              "lsh_memset_device")

    def slot(self, index: int) -> "DeviceBuffer":
        """A view of sub-array ``index`` along the first axis (the parent owns the memory)."""
        if not 0 <= index < self.shape[0]:
            raise ContractValidationError("slot index out of range")
        view = DeviceBuffer.__new__(DeviceBuffer)
        view.session = self.session
        view.dtype = self.dtype
        view.shape = self.shape[1:]
        view.nbytes = int(np.prod(view.shape, dtype=np.int64)) * view.dtype.itemsize
        view.ptr = self.ptr + index * view.nbytes
        view._parent = self
        view.free = lambda: None
        return view

    def slots(self, start: int, count: int) -> "DeviceBuffer":
        """A view of sub-arrays ``start .. start + count - 1`` along the first axis."""
        if not (0 <= start and 0 < count and start + count <= self.shape[0]):
            raise ContractValidationError("slot range out of range")
        view = DeviceBuffer.__new__(DeviceBuffer)
        view.session = self.session
        view.dtype = self.dtype
        view.shape = (int(count),) + self.shape[1:]
        per = int(np.prod(self.shape[1:], dtype=np.int64)) * self.dtype.itemsize
        view.nbytes = per * int(count)
        view.ptr = self.ptr + int(start) * per
        view._parent = self
        view.free = lambda: None
        return view

    def free(self) -> None:
        if self.ptr and self.session.handle:
            load_library().lsh_free(self.session.handle, self.ptr)
        self.ptr = 0

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.free()
        except Exception:
            pass


class Event:
    """A hipEvent recorded on a session's stream."""

    def __init__(self, session: "Session"):
        self.session = session
        handle = _vp()
        check(load_library().lsh_event_create(session.handle, ctypes.byref(handle)),
              "lsh_event_create")
        self.handle = handle.value

    def record(self) -> "Event":
        check(load_library().lsh_event_record(self.session.handle, self.handle), "lsh_event_record")
        return self

    def elapsed_ms(self, stop: "Event") -> float:
        ms = _f32()
        check(load_library().lsh_event_elapsed_ms(self.handle, stop.handle, ctypes.byref(ms)),
              "lsh_event_elapsed_ms")
        return float(ms.value)

    def __del__(self):  # pragma: no cover
        try:
            load_library().lsh_event_destroy(self.handle)
        except Exception:
            pass


class Session:
    """One HIP device and stream (lsh_session)."""

    def __init__(self, device_id: int = 0):
        lib = load_library()
        handle = _vp()
        check(lib.lsh_session_create(int(device_id), ctypes.byref(handle)), "lsh_session_create")
        self.handle = handle.value
        self.device_id = int(device_id)
        self.lib = lib

    def close(self) -> None:
        if self.handle:
            self.lib.lsh_session_destroy(self.handle)
            self.handle = None

    def synchronize(self) -> None:
        check(self.lib.lsh_synchronize(self.handle), "kernel execution")

    def device_name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        cus = _i32()
        check(self.lib.lsh_device_info(self.device_id, buf, len(buf), ctypes.byref(cus)),
              "lsh_device_info")
        return buf.value.decode()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

"""Device memory, streams and events through the C-ABI (no framework types)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import call, hip, ptr


def device_count() -> int:
    n = C.c_int(0)
    status = hip().rta_hip_device_count(C.byref(n))
    return n.value if status == 0 else 0


def device_name() -> str:
    buf = C.create_string_buffer(256)
    call("rta_hip_device_name", buf, 256)
    return buf.value.decode()


def synchronize() -> None:
    call("rta_hip_device_synchronize")


class DeviceBuffer:
    """An owning hipMalloc allocation."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        call("rta_hip_malloc", C.byref(p), C.c_size_t(max(self.nbytes, 1)))
        self.ptr = p.value or 0

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        arr = np.ascontiguousarray(arr)
        if offset + arr.nbytes > self.nbytes:
            raise ValueError("upload exceeds the buffer")
        call("rta_hip_memcpy_htod", C.c_void_p(self.ptr + offset), ptr(arr), C.c_size_t(arr.nbytes))

    def download(self, dtype, count: int | None = None, offset: int = 0) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes + offset > self.nbytes:
            raise ValueError("download exceeds the buffer")
        call("rta_hip_memcpy_dtoh", ptr(out), C.c_void_p(self.ptr + offset), C.c_size_t(out.nbytes))
        return out

    def fill_u32(self, value: int, count: int | None = None, stream=None) -> None:
        count = self.nbytes // 4 if count is None else count
        call("rta_hip_fill_u32", C.c_void_p(self.ptr), C.c_uint32(value), C.c_size_t(count),
             ptr(stream))

    def free(self) -> None:
        if self.ptr:
            call("rta_hip_free", C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            if self.ptr:
                hip().rta_hip_free(C.c_void_p(self.ptr))
                self.ptr = 0
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream at default priority."""

    def __init__(self):
        s = C.c_void_p()
        call("rta_hip_stream_create", C.byref(s))
        self.ptr = s.value or 0

    def synchronize(self) -> None:
        call("rta_hip_stream_synchronize", C.c_void_p(self.ptr))

    def wait(self, event: "Event") -> None:
        """Work enqueued after this waits for the event's last record."""
        call("rta_hip_stream_wait_event", C.c_void_p(self.ptr), C.c_void_p(event.ptr))

    def __del__(self):
        try:
            if self.ptr:
                hip().rta_hip_stream_destroy(C.c_void_p(self.ptr))
                self.ptr = 0
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        call("rta_hip_event_create", C.byref(e))
        self.ptr = e.value or 0

    def record(self, stream=None) -> None:
        call("rta_hip_event_record", C.c_void_p(self.ptr), ptr(stream))

    def synchronize(self) -> None:
        call("rta_hip_event_synchronize", C.c_void_p(self.ptr))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0)
        call("rta_hip_event_elapsed_ms", C.byref(ms), C.c_void_p(self.ptr), C.c_void_p(end.ptr))
        return float(ms.value)

    def __del__(self):
        try:
            if self.ptr:
                hip().rta_hip_event_destroy(C.c_void_p(self.ptr))
                self.ptr = 0
        except Exception:
            pass

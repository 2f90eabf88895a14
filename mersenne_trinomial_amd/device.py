"""Device memory, streams and events over the C ABI (no torch needed)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._capi import check, lib

_initialised = False


def init(device: int = 0) -> None:
    """Select the GPU and create its context (idempotent)."""
    global _initialised
    if not _initialised:
        check(lib.gf2t_init(device), "gf2t_init")
        _initialised = True


def device_count() -> int:
    n = C.c_int(0)
    check(lib.gf2t_device_count(C.byref(n)), "gf2t_device_count")
    return n.value


def device_name() -> str:
    init()
    buf = C.create_string_buffer(256)
    check(lib.gf2t_device_name(buf, 256), "gf2t_device_name")
    return buf.value.decode()


def mem_info():
    init()
    f, t = C.c_size_t(), C.c_size_t()
    check(lib.gf2t_mem_info(C.byref(f), C.byref(t)))
    return f.value, t.value


def synchronize() -> None:
    check(lib.gf2t_device_sync(), "gf2t_device_sync")


class DeviceBuffer:
    """An owned device allocation of `nbytes` bytes."""

    def __init__(self, nbytes: int):
        init()
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib.gf2t_malloc(C.byref(p), max(8, self.nbytes)), "gf2t_malloc")
        self.ptr = p.value

    @classmethod
    def words(cls, n: int) -> "DeviceBuffer":
        return cls(8 * int(n))

    @classmethod
    def from_array(cls, a: np.ndarray, extra_words: int = 0) -> "DeviceBuffer":
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = cls(8 * (a.size + extra_words))
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset_words: int = 0) -> None:
        a = np.ascontiguousarray(a, dtype=np.uint64)
        assert 8 * (offset_words + a.size) <= self.nbytes
        check(lib.gf2t_memcpy_h2d(self.ptr + 8 * offset_words, a.ctypes.data, a.nbytes, None),
              "gf2t_memcpy_h2d")

    def download(self, nwords: int | None = None, offset_words: int = 0) -> np.ndarray:
        n = self.nbytes // 8 - offset_words if nwords is None else int(nwords)
        out = np.empty(n, dtype=np.uint64)
        check(lib.gf2t_memcpy_d2h(out.ctypes.data, self.ptr + 8 * offset_words, out.nbytes, None),
              "gf2t_memcpy_d2h")
        return out

    def fill(self, byte: int) -> None:
        check(lib.gf2t_memset(self.ptr, byte, self.nbytes, None), "gf2t_memset")

    def word_ptr(self, offset_words: int) -> int:
        return self.ptr + 8 * offset_words

    def free(self) -> None:
        if getattr(self, "ptr", None):
            lib.gf2t_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        init()
        p = C.c_void_p()
        check(lib.gf2t_stream_create(C.byref(p)), "gf2t_stream_create")
        self.handle = p.value

    def sync(self) -> None:
        check(lib.gf2t_stream_sync(self.handle), "gf2t_stream_sync")

    def __del__(self):
        try:
            if self.handle:
                lib.gf2t_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    """HIP event; record() on the stream the measured kernels run on."""

    def __init__(self):
        init()
        p = C.c_void_p()
        check(lib.gf2t_event_create(C.byref(p)), "gf2t_event_create")
        self.handle = p.value

    def record(self, stream=None) -> None:
        check(lib.gf2t_event_record(self.handle, getattr(stream, "handle", stream)),
              "gf2t_event_record")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib.gf2t_event_elapsed_ms(self.handle, end.handle, C.byref(ms)),
              "gf2t_event_elapsed_ms")
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                lib.gf2t_event_destroy(self.handle)
        except Exception:
            pass

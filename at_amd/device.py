"""Device memory, pinned host memory, streams and events of the engine.

Thin Python handles over the C ABI's memory and stream entry points
(include/atx.h: atx_device_alloc ... atx_event_elapsed), so that callers
that keep particles and loss maps resident on the device (batched launches,
the serving loop, the autotuner) need no other GPU runtime binding.

* ``DeviceArray`` -- a typed, C-contiguous device buffer: ``ptr``,
  ``shape``, ``dtype``; ``from_host``/``to_host``/``copy_from``;
* ``HostArray`` -- pinned (page-locked) host memory seen as a numpy array,
  the source or target of asynchronous copies;
* ``Stream`` and ``Event`` -- HIP streams (non-blocking, default priority)
  and events, with record / query / synchronize / elapsed time.

Memory is released when the handle is closed or collected; ``Stream`` and
``Event`` likewise.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import engine


def lib():
    return engine.load_library()


def check(rc):
    """Raise the engine's error for a failing entry-point return code."""
    if rc != engine.ATX_OK:
        engine._raise(rc, lib())


_lib, _check = lib, check


class Stream:
    """A HIP stream of the engine's runtime (``handle`` is the hipStream_t
    passed as ``stream`` to every asynchronous entry point)."""

    def __init__(self):
        h = C.c_void_p()
        _check(_lib().atx_stream_create(C.byref(h)))
        self.handle = h.value

    def synchronize(self) -> None:
        _check(_lib().atx_stream_synchronize(self.handle))

    def wait(self, event: "Event") -> None:
        """Later work on this stream waits for ``event``."""
        _check(_lib().atx_stream_wait_event(self.handle, event.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            _lib().atx_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        h = C.c_void_p()
        _check(_lib().atx_event_create(C.byref(h)))
        self.handle = h.value

    def record(self, stream: Stream | None = None) -> "Event":
        _check(_lib().atx_event_record(self.handle, None if stream is None else stream.handle))
        return self

    def query(self) -> bool:
        done = C.c_int(0)
        _check(_lib().atx_event_query(self.handle, C.byref(done)))
        return bool(done.value)

    def synchronize(self) -> None:
        _check(_lib().atx_event_synchronize(self.handle))

    def elapsed_ms(self, end: "Event") -> float:
        """Milliseconds from this event to ``end`` (both recorded and done)."""
        ms = C.c_float(0.0)
        _check(_lib().atx_event_elapsed(C.byref(ms), self.handle, end.handle))
        return float(ms.value)

    def close(self) -> None:
        if getattr(self, "handle", None):
            _lib().atx_event_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synchronize() -> None:
    """Wait for all work on the null stream."""
    _check(_lib().atx_stream_synchronize(None))


def _stream(stream):
    return None if stream is None else stream.handle


class HostArray:
    """Pinned host memory viewed as a numpy array (``.numpy``)."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(_lib().atx_host_alloc(C.byref(p), self.nbytes))
        self.ptr = p.value or 0
        if self.nbytes:
            buf = (C.c_char * self.nbytes).from_address(self.ptr)
            self.numpy = np.frombuffer(buf, dtype=self.dtype).reshape(self.shape)
        else:
            self.numpy = np.empty(self.shape, self.dtype)

    def close(self) -> None:
        if getattr(self, "ptr", 0):
            self.numpy = None
            _lib().atx_host_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """C-contiguous device buffer of ``shape`` and ``dtype``."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(_lib().atx_device_alloc(C.byref(p), self.nbytes))
        self.ptr = p.value or 0

    @classmethod
    def from_host(cls, a: np.ndarray, stream: Stream | None = None) -> "DeviceArray":
        """Device copy of a host array (C order), complete when it returns."""
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        if stream is None:
            synchronize()
        else:
            stream.synchronize()
        return d

    @classmethod
    def zeros(cls, shape, dtype=np.float64, stream: Stream | None = None) -> "DeviceArray":
        return cls.from_host(np.zeros(shape, dtype), stream)

    def upload(self, a, stream: Stream | None = None) -> None:
        """Copy ``a`` (numpy array or HostArray of the same size) in;
        asynchronous on ``stream`` for a HostArray source."""
        src = a.ptr if isinstance(a, HostArray) else np.ascontiguousarray(a, self.dtype)
        n = a.nbytes
        if n != self.nbytes:
            raise ValueError(f"size mismatch: {n} bytes into {self.nbytes}")
        addr = src if isinstance(a, HostArray) else src.ctypes.data
        _check(_lib().atx_memcpy(self.ptr, addr, self.nbytes, _stream(stream)))
        if not isinstance(a, HostArray):
            # a pageable source may be gone once this returns
            (stream.synchronize() if stream is not None else synchronize())

    def download(self, out, stream: Stream | None = None) -> None:
        """Copy into ``out`` (HostArray: asynchronous on ``stream``; numpy:
        synchronous)."""
        if out.nbytes != self.nbytes:
            raise ValueError(f"size mismatch: {self.nbytes} bytes into {out.nbytes}")
        if isinstance(out, HostArray):
            _check(_lib().atx_memcpy(out.ptr, self.ptr, self.nbytes, _stream(stream)))
            return
        if not out.flags.c_contiguous:
            raise ValueError("target must be C-contiguous")
        _check(_lib().atx_memcpy(out.ctypes.data, self.ptr, self.nbytes, _stream(stream)))
        if stream is None:
            synchronize()
        else:
            stream.synchronize()

    def to_host(self, stream: Stream | None = None) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        self.download(out, stream)
        return out

    def copy_from(self, other: "DeviceArray", stream: Stream | None = None) -> None:
        """Device-to-device copy, asynchronous on ``stream``."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _check(_lib().atx_memcpy(self.ptr, other.ptr, self.nbytes, _stream(stream)))

    def clone(self, stream: Stream | None = None) -> "DeviceArray":
        d = DeviceArray(self.shape, self.dtype)
        d.copy_from(self, stream)
        return d

    def close(self) -> None:
        if getattr(self, "ptr", 0):
            _lib().atx_device_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

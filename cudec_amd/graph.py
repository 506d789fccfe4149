"""HIP graph capture of a batch launch.

A caller that launches the same batch decode again and again (fixed tables,
new compressed bytes in the same arenas) can capture one
`cudec_lz4_decompress_batch` / `cudec_snappy_decompress_batch` call into a HIP
graph and replay it. The tuned launch is capture-safe: its scratch (cost
queue and chunk order) becomes a graph allocation node, and it records no
memset node (the queue header is cleared by a kernel).

The HIP runtime entry points are resolved through the product library's
handle, so they are the runtime that library (and torch, when it is loaded
first) already uses; no second runtime is opened.
"""
from __future__ import annotations

import ctypes

from . import lib

_CAPTURE_RELAXED = 2  # hipStreamCaptureModeRelaxed
_hip = None


def _runtime():
    global _hip
    if _hip is None:
        L = lib()
        vp = ctypes.c_void_p
        L.hipStreamBeginCapture.argtypes = [vp, ctypes.c_int]
        L.hipStreamEndCapture.argtypes = [vp, ctypes.POINTER(vp)]
        L.hipGraphInstantiate.argtypes = [ctypes.POINTER(vp), vp, vp, vp, ctypes.c_size_t]
        L.hipGraphLaunch.argtypes = [vp, vp]
        L.hipGraphExecDestroy.argtypes = [vp]
        L.hipGraphDestroy.argtypes = [vp]
        for name in ("hipStreamBeginCapture", "hipStreamEndCapture", "hipGraphInstantiate",
                     "hipGraphLaunch", "hipGraphExecDestroy", "hipGraphDestroy"):
            getattr(L, name).restype = ctypes.c_int
        _hip = L
    return _hip


def _handle(stream) -> int:
    h = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
    if not h:
        raise ValueError("capture needs a created stream, not the null stream")
    return h


class CapturedLaunch:
    """Captures `launch()` (which must submit on `stream` and return the
    cudec status) into a graph; `replay()` launches the graph on the same
    stream. Raises RuntimeError with the HIP or cudec code on failure."""

    def __init__(self, launch, stream):
        rt = _runtime()
        self._stream = _handle(stream)
        self._graph = ctypes.c_void_p()
        self._exec = ctypes.c_void_p()
        e = rt.hipStreamBeginCapture(self._stream, _CAPTURE_RELAXED)
        if e != 0:
            raise RuntimeError(f"hipStreamBeginCapture failed: {e}")
        st = launch()
        e = rt.hipStreamEndCapture(self._stream, ctypes.byref(self._graph))
        if st != 0 or e != 0:
            if self._graph:
                rt.hipGraphDestroy(self._graph)
            raise RuntimeError(f"capture failed: cudec status {st}, hipStreamEndCapture {e}")
        e = rt.hipGraphInstantiate(ctypes.byref(self._exec), self._graph, None, None, 0)
        if e != 0:
            rt.hipGraphDestroy(self._graph)
            raise RuntimeError(f"hipGraphInstantiate failed: {e}")

    def replay(self) -> None:
        e = _runtime().hipGraphLaunch(self._exec, self._stream)
        if e != 0:
            raise RuntimeError(f"hipGraphLaunch failed: {e}")

    def close(self) -> None:
        rt = _runtime()
        if self._exec:
            rt.hipGraphExecDestroy(self._exec)
            self._exec = ctypes.c_void_p()
        if self._graph:
            rt.hipGraphDestroy(self._graph)
            self._graph = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

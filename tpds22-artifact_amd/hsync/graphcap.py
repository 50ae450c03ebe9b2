"""hipGraph capture of repeated launches through libhs's C ABI
(hs_graph_capture_begin / _end / hs_graph_replay / hs_graph_destroy).

The libhs launchers enqueue on the caller's stream and never call
hipMemsetAsync (device words are cleared by fill kernels), so a sequence of
launches can be captured once and replayed.  Capture runs on a side stream of
default priority; replays go to the caller's current stream.
"""
from __future__ import annotations

import ctypes as C

from . import _lib as L


def _torch():
    import torch
    return torch


class Replay:
    """Capture fn() (which must launch libhs work on the current stream and
    allocate nothing new) into a hipGraph; replay() re-runs it."""

    def __init__(self, fn, warmup: int = 1):
        torch = _torch()
        lib = L.load()
        self._lib = lib
        self.h = None
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):  # first launches may load code objects: not capturable
                fn()
        side.synchronize()
        s = C.c_void_p(side.cuda_stream)
        with torch.cuda.stream(side):
            L.check(lib.hs_graph_capture_begin(s), "hs_graph_capture_begin")
            try:
                fn()
            finally:
                h = C.c_void_p()
                rc = lib.hs_graph_capture_end(s, C.byref(h))
        L.check(rc, "hs_graph_capture_end")
        self.h = h
        self._side = side
        torch.cuda.current_stream().wait_stream(side)

    def replay(self, stream=None) -> None:
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        L.check(self._lib.hs_graph_replay(self.h, C.c_void_p(s)), "hs_graph_replay")

    def close(self) -> None:
        if self.h:
            self._lib.hs_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

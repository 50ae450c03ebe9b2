"""hipGraph capture of repeated launches.

The libhs launchers enqueue on the caller's stream and never call
hipMemsetAsync (device words are cleared by fill kernels), so a sequence of
launches can be captured once and replayed.  Capture runs on a side stream of
default priority.
"""
from __future__ import annotations


def _torch():
    import torch
    return torch


class Replay:
    """Capture fn() (which must launch libhs work on the current stream and
    allocate nothing new) into a hipGraph; replay() re-runs it."""

    def __init__(self, fn, warmup: int = 1):
        torch = _torch()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                fn()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            fn()
        torch.cuda.synchronize()

    def replay(self) -> None:
        self.graph.replay()

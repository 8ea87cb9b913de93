"""Batched pairs of mixed size over the C ABI (rts_batch_*): the rank + LoG +
NCC cost with box aggregation, chunked under a device budget, the box stage
as one persistent work-queue launch per chunk, each chunk captured into a
HIP graph and replayed."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import PipelineDesc, RtsError, call, lib
from .device import DeviceArray


class BatchItem(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("disp", C.c_int),
                ("left_rgb", C.c_void_p), ("right_rgb", C.c_void_p),
                ("dl", C.c_void_p), ("dr", C.c_void_p), ("mask", C.c_void_p)]


class BatchStats(C.Structure):
    _fields_ = [("pairs", C.c_int), ("chunks", C.c_int), ("graphs_built", C.c_int), ("graph_replays", C.c_int),
                ("box_items", C.c_ulonglong), ("arena_bytes", C.c_size_t), ("box_rows", C.c_int),
                ("cost_disparities", C.c_int), ("tune_ms", C.c_float * 6)]


_SIGS = {
    "rts_batch_create": ([C.POINTER(PipelineDesc), C.c_size_t, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "rts_batch_destroy": ([C.c_void_p], C.c_int),
    "rts_batch_run": ([C.c_void_p, C.POINTER(BatchItem), C.c_int, C.c_void_p], C.c_int),
    "rts_batch_get_stats": ([C.c_void_p, C.POINTER(BatchStats)], C.c_int),
    "rts_batch_pair_bytes": ([C.c_int, C.c_int, C.c_int], C.c_size_t),
}
for _n, (_a, _r) in _SIGS.items():
    _f = getattr(lib, _n)
    _f.argtypes = _a
    _f.restype = _r


def pair_bytes(W: int, H: int, D: int) -> int:
    return int(lib.rts_batch_pair_bytes(W, H, D))


class Pair:
    """One pair's device buffers: colour images in, D_L / D_R / mask out."""

    def __init__(self, left: np.ndarray, right: np.ndarray, disp: int):
        H, W = left.shape[:2]
        self.shape = (W, H, disp)
        self.left, self.right = DeviceArray.from_host(left), DeviceArray.from_host(right)
        self.dl, self.dr, self.mask = (DeviceArray((H, W), np.uint8) for _ in range(3))

    def item(self) -> BatchItem:
        W, H, D = self.shape
        return BatchItem(W, H, D, self.left.ptr, self.right.ptr, self.dl.ptr, self.dr.ptr, self.mask.ptr)

    def results(self, stream=None):
        return self.dl.download(stream), self.dr.download(stream), self.mask.download(stream)

    def free(self):
        for a in (self.left, self.right, self.dl, self.dr, self.mask):
            a.free()


class Batch:
    def __init__(self, desc: PipelineDesc, budget_bytes: int, use_graphs: bool = True):
        p = C.c_void_p()
        call("rts_batch_create", C.byref(desc), budget_bytes, int(use_graphs), C.byref(p))
        self.handle = p.value

    def run(self, pairs, stream) -> None:
        arr = (BatchItem * len(pairs))(*[p.item() for p in pairs])
        call("rts_batch_run", self.handle, arr, len(pairs), stream)

    def stats(self) -> BatchStats:
        st = BatchStats()
        call("rts_batch_get_stats", self.handle, C.byref(st))
        return st

    def close(self) -> None:
        if self.handle:
            lib.rts_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

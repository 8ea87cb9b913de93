"""Resident stereo pipelines over the C ABI (rts_pipeline_*).

A pipeline owns its device volumes and per-pixel buffers; ``run`` takes device
images and leaves D_L, D_R and the left-right mask in device buffers.  The
named presets are the measured workloads with the reference driver's stage
order (seq/Main.cpp:180-257).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import PipelineDesc, call, lib
from .device import DeviceArray


def defaults(W: int, H: int, D: int) -> PipelineDesc:
    d = PipelineDesc()
    lib.rts_pipeline_defaults(C.byref(d), W, H, D)
    return d


def preset(name: str, W: int, H: int, D: int, **over) -> PipelineDesc:
    """Named stage configurations.

    plain:        colour AD truncated at 20, 9x9 box, WTA, left-right check tol 1
    census_cross: census 9x7 + colour AD combined (lambda 30 / 10), cross-based
                  aggregation (L 17, tao 20, one H/V pass), WTA, check tol 1
    aw:           colour AD truncated at 40, adaptive weights 33x33 (gc 7, gp 36),
                  WTA, check tol 1
    rank_box / log_box / ncc_box: the remaining cost families with an 11x11 box
    rank_log_ncc: the batched workload: rank 7x7 -> AD, LoG 5x5 -> AD and NCC
                  7x7 combined as three exponential terms (lambda 8 / 10 / 0.5,
                  NCC as 1 - ncc), 11x11 box, WTA, check tol 1
    """
    d = defaults(W, H, D)
    d.lr_check, d.lr_tol = 1, 1
    if name == "plain":
        d.cost1, d.ad_trunc_enable, d.ad_trunc = L.RTS_COST_AD, 1, 20.0
        d.aggr, d.box_r = L.RTS_AGGR_BOX, 4
    elif name == "census_cross":
        d.cost1, d.cost2, d.lambda1, d.lambda2 = L.RTS_COST_CENSUS, L.RTS_COST_AD, 30.0, 10.0
        d.ad_trunc_enable = 0
        d.aggr, d.cross_L, d.cross_tao, d.cross_iters = L.RTS_AGGR_CROSS, 17, 20, 1
    elif name == "aw":
        d.cost1, d.ad_trunc_enable, d.ad_trunc = L.RTS_COST_AD, 1, 40.0
        d.aggr, d.aw_r, d.aw_gamma_c, d.aw_gamma_p = L.RTS_AGGR_AW, 16, 7.0, 36.0
    elif name == "rank_box":
        d.cost1, d.rank_m, d.aggr, d.box_r = L.RTS_COST_RANK, 3, L.RTS_AGGR_BOX, 5
    elif name == "log_box":
        d.cost1, d.aggr, d.box_r = L.RTS_COST_LOG, L.RTS_AGGR_BOX, 5
    elif name == "ncc_box":
        d.cost1, d.ncc_hr, d.ncc_wr, d.aggr, d.box_r = L.RTS_COST_NCC, 3, 3, L.RTS_AGGR_BOX, 5
    elif name == "rank_log_ncc":
        d.cost1, d.cost2, d.cost3 = L.RTS_COST_RANK, L.RTS_COST_LOG, L.RTS_COST_NCC
        d.rank_m, d.ncc_hr, d.ncc_wr = 3, 3, 3
        d.lambda1, d.lambda2, d.lambda3 = 8.0, 10.0, 0.5
        d.aggr, d.box_r = L.RTS_AGGR_BOX, 5
    else:
        raise ValueError(f"unknown preset {name!r}")
    for k, v in over.items():
        setattr(d, k, v)
    return d


class Pipeline:
    def __init__(self, desc: PipelineDesc):
        self.desc = desc
        p = C.c_void_p()
        call("rts_pipeline_create", C.byref(desc), C.byref(p))
        self.handle = p.value
        W, H = desc.width, desc.height
        self.dl = DeviceArray((H, W), np.uint8)
        self.dr = DeviceArray((H, W), np.uint8)
        self.mask = DeviceArray((H, W), np.uint8)

    @property
    def bytes(self) -> int:
        return int(lib.rts_pipeline_bytes(C.byref(self.desc)))

    def run(self, left_rgb: DeviceArray, right_rgb: DeviceArray, left_lab: DeviceArray | None = None,
            right_lab: DeviceArray | None = None, stream=None, left_gray=None, right_gray=None) -> None:
        call("rts_pipeline_run", self.handle, left_rgb.ptr, right_rgb.ptr,
             left_gray.ptr if left_gray is not None else None,
             right_gray.ptr if right_gray is not None else None,
             left_lab.ptr if left_lab is not None else None,
             right_lab.ptr if right_lab is not None else None,
             self.dl.ptr, self.dr.ptr, self.mask.ptr, stream)

    def enable_timing(self, on: bool = True) -> None:
        call("rts_pipeline_enable_timing", self.handle, int(on))

    def enable_graphs(self, on: bool = True) -> None:
        """Replay the run's launch sequence as a HIP graph for repeated runs
        with the same buffers and stream (captured on the second such run)."""
        call("rts_pipeline_enable_graphs", self.handle, int(on))

    def stage_ms(self) -> dict:
        """Stage durations (ms) of the last run, measured with HIP events on
        the run's stream; empty unless timing is enabled."""
        ms = (C.c_float * 8)()
        names = (C.c_char_p * 8)()
        n = C.c_int(0)
        call("rts_pipeline_stage_ms", self.handle, ms, names, 8, C.byref(n))
        return {names[i].decode(): float(ms[i]) for i in range(n.value)}

    def volume(self) -> np.ndarray:
        """Final aggregated volume of the last run, [D][H][W]."""
        d = self.desc
        ptr = lib.rts_pipeline_volume(self.handle)
        out = np.empty((d.disp, d.height, d.width), np.float32)
        call("rts_memcpy_d2h", out.ctypes.data, ptr, out.nbytes, None)
        call("rts_stream_sync", None)
        return out

    def results(self):
        return self.dl.download(), self.dr.download(), self.mask.download()

    def close(self):
        if self.handle:
            lib.rts_pipeline_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_host(desc: PipelineDesc, left_rgb: np.ndarray, right_rgb: np.ndarray, left_lab=None, right_lab=None):
    """One pipeline run from host images; returns (D_L, D_R, mask, volume)."""
    p = Pipeline(desc)
    a, b = DeviceArray.from_host(left_rgb), DeviceArray.from_host(right_rgb)
    la = DeviceArray.from_host(left_lab) if left_lab is not None else None
    lb = DeviceArray.from_host(right_lab) if right_lab is not None else None
    p.run(a, b, la, lb)
    call("rts_device_sync")
    dl, dr, m = p.results()
    vol = p.volume()
    p.close()
    return dl, dr, m, vol

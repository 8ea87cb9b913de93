"""ctypes binding of librtstereo.so (the C ABI in include/rtstereo.h).

The library is loaded from this package's directory.  There is no fallback:
if the HIP library is missing, importing fails with instructions to build it.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librtstereo.so")
SEQ_LIB_PATH = os.path.join(HERE, "librtstereo_seq.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `python -m rtstereo_amd.build` "
                      "(or __graft_entry__.build()); there is no CPU fallback")

lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)

vp, sz, i32, f32, u8 = C.c_void_p, C.c_size_t, C.c_int, C.c_float, C.c_uint8
cp = C.c_char_p

RTS_OK, RTS_ERR_INVALID, RTS_ERR_HIP, RTS_ERR_NOMEM, RTS_ERR_BUDGET, RTS_ERR_CANCELLED, RTS_ERR_NOTREADY = range(7)
RTS_L2R, RTS_R2L = 0, 1
RTS_NCC_SEQ, RTS_NCC_INTEGRAL, RTS_NCC_KEEP_FLAT = 0, 1, 0x100
RTS_AW_FP32, RTS_AW_FP16 = 0, 1
RTS_CROSS_EXACT, RTS_CROSS_FAST, RTS_CROSS_ARMS_LE17 = 0, 1, 0x100
RTS_WTA_MAX, RTS_WTA_LITERAL = 1, 2
RTS_COST_AD, RTS_COST_RANK, RTS_COST_CENSUS, RTS_COST_LOG, RTS_COST_NCC = 1, 2, 3, 4, 5
RTS_AGGR_NONE, RTS_AGGR_BOX, RTS_AGGR_AW, RTS_AGGR_CROSS = 0, 1, 2, 3


class PipelineDesc(C.Structure):
    _fields_ = [("width", i32), ("height", i32), ("disp", i32),
                ("cost1", i32), ("cost2", i32),
                ("lambda1", f32), ("lambda2", f32),
                ("ad_trunc_enable", i32), ("ad_trunc", f32), ("ad_norm", i32),
                ("census_hr", i32), ("census_wr", i32),
                ("rank_m", i32),
                ("ncc_hr", i32), ("ncc_wr", i32),
                ("aggr", i32),
                ("box_r", i32),
                ("cross_L", i32), ("cross_tao", i32), ("cross_iters", i32), ("cross_exact", i32),
                ("aw_r", i32), ("aw_gamma_c", f32), ("aw_gamma_p", f32), ("aw_precision", i32),
                ("wta_literal", i32),
                ("lr_check", i32), ("lr_tol", i32),
                ("cost3", i32), ("lambda3", f32)]


# name -> argtypes (all return rts_status unless listed in _RESTYPES)
_SIGS = {
    "rts_last_error": [], "rts_version": [],
    "rts_device_count": [C.POINTER(i32)], "rts_set_device": [i32],
    "rts_malloc": [C.POINTER(vp), sz], "rts_free": [vp],
    "rts_host_alloc": [C.POINTER(vp), sz], "rts_host_free": [vp],
    "rts_memcpy_h2d": [vp, vp, sz, vp], "rts_memcpy_d2h": [vp, vp, sz, vp], "rts_memcpy_d2d": [vp, vp, sz, vp],
    "rts_fill_u8": [vp, u8, sz, vp],
    "rts_stream_create": [C.POINTER(vp)], "rts_stream_destroy": [vp], "rts_stream_sync": [vp],
    "rts_device_sync": [],
    "rts_event_create": [C.POINTER(vp)], "rts_event_destroy": [vp], "rts_event_record": [vp, vp],
    "rts_event_sync": [vp], "rts_event_elapsed_ms": [vp, vp, C.POINTER(f32)],
    "rts_mem_info": [C.POINTER(sz), C.POINTER(sz)],
    "rts_vol_to_dm": [vp, vp, i32, i32, i32, vp], "rts_vol_to_pm": [vp, vp, i32, i32, i32, vp],
    "rts_copy_kernel": [vp, vp, sz, vp],
    "rts_read_kernel": [vp, sz, vp, vp],
    "rts_write_kernel": [vp, sz, vp],
    "rts_volume_to_half": [vp, vp, sz, vp],
    "rts_ad_gray": [vp, vp, vp, i32, i32, i32, i32, vp],
    "rts_ad_float": [vp, vp, vp, i32, i32, i32, i32, vp],
    "rts_ad_color": [vp, vp, vp, i32, i32, i32, i32, i32, f32, i32, vp],
    "rts_ad_color_half": [vp, vp, vp, i32, i32, i32, i32, i32, f32, i32, vp],
    "rts_census": [vp, vp, i32, i32, i32, i32, vp],
    "rts_census_unpack": [vp, vp, i32, i32, i32, vp], "rts_census_pack": [vp, vp, i32, i32, i32, vp],
    "rts_hamming": [vp, vp, vp, i32, i32, i32, i32, vp],
    "rts_rank": [vp, vp, i32, i32, i32, vp],
    "rts_convolve": [vp, vp, C.POINTER(f32), i32, i32, i32, i32, vp],
    "rts_ncc": [vp, vp, vp, i32, i32, i32, i32, i32, i32, vp],
    "rts_combine": [vp, vp, f32, f32, sz, vp],
    "rts_box": [vp, vp, i32, i32, i32, i32, vp],
    "rts_box_int": [vp, vp, i32, i32, i32, i32, vp],
    "rts_prox_weight": [vp, f32, i32, vp],
    "rts_compute_weights": [vp, vp, vp, f32, i32, i32, i32, vp],
    "rts_aw_weights": [vp, vp, vp, vp, i32, i32, i32, i32, i32, vp],
    "rts_aw": [vp, vp, vp, vp, i32, i32, i32, i32, f32, f32, i32, vp],
    "rts_aw_ws": [vp, vp, vp, vp, i32, i32, i32, i32, f32, f32, i32, vp, sz, vp],
    "rts_aw_half_ws": [vp, vp, vp, vp, i32, i32, i32, i32, f32, f32, vp, sz, vp],
    "rts_find_cross": [vp, vp, i32, i32, i32, i32, vp],
    "rts_arms_unpack": [vp, vp, vp, i32, i32, vp], "rts_arms_pack": [vp, vp, vp, i32, i32, vp],
    "rts_hprefix": [vp, i32, i32, i32, vp],
    "rts_cross_v": [vp, vp, vp, vp, i32, i32, i32, i32, vp],
    "rts_cross_v_wta": [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, vp, sz, vp],
    "rts_cross_fused": [vp, vp, vp, vp, i32, i32, i32, vp],
    "rts_wta": [vp, vp, vp, vp, i32, i32, i32, i32, i32, vp],
    "rts_wta_noalloc": [vp, vp, vp, vp, i32, i32, i32, i32, i32, vp],
    "rts_cross_check": [vp, vp, vp, i32, i32, i32, vp],
    "rts_disp_refine": [vp, vp, vp, i32, i32, i32, i32, vp],
    "rts_pipeline_defaults": [C.POINTER(PipelineDesc), i32, i32, i32],
    "rts_pipeline_bytes": [C.POINTER(PipelineDesc)],
    "rts_aw_workspace_bytes": [i32, i32, i32, i32],
    "rts_aw_half_supported": [i32, i32, i32],
    "rts_cost_rank_log_ncc": [vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, f32, f32, f32, vp],
    "rts_cost_rank_log_ncc_dch": [vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, f32, f32, f32, i32, vp],
    "rts_cost_rank_log_ncc_ws": [vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, f32, f32, f32, i32, vp, sz, vp],
    "rts_cost_rank_log_ncc_ws_bytes": [i32, i32],
    "rts_pipeline_create": [C.POINTER(PipelineDesc), C.POINTER(vp)],
    "rts_pipeline_destroy": [vp],
    "rts_pipeline_run": [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp],
    "rts_pipeline_volume": [vp],
    "rts_pipeline_enable_timing": [vp, i32],
    "rts_pipeline_enable_graphs": [vp, i32],
    "rts_pipeline_stage_ms": [vp, C.POINTER(f32), C.POINTER(cp), i32, C.POINTER(i32)],
    "rts_rgb_to_gray": [vp, vp, i32, i32, vp],
    "rts_rgb_to_lab": [vp, vp, i32, i32, vp],
    "rts_median3": [vp, vp, i32, i32, vp],
    "rts_png_read": [cp, i32, C.POINTER(vp), C.POINTER(i32), C.POINTER(i32)],
    "rts_png_decode": [vp, sz, i32, C.POINTER(vp), C.POINTER(i32), C.POINTER(i32)],
    "rts_png_free": [vp],
    "rts_png_write": [cp, vp, i32, i32, i32],
    "rts_disp_scale": [vp, i32, i32, f32],
    "rts_cross_v_wta_ws_bytes": [i32, i32],
}
_RESTYPES = {"rts_last_error": cp, "rts_version": cp, "rts_pipeline_defaults": None, "rts_png_free": None,
             "rts_pipeline_bytes": sz, "rts_pipeline_volume": vp,
             "rts_aw_workspace_bytes": sz, "rts_cost_rank_log_ncc_ws_bytes": sz, "rts_cross_v_wta_ws_bytes": sz}

for _name, _args in _SIGS.items():
    if not hasattr(lib, _name):
        continue
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _RESTYPES.get(_name, i32)


class RtsError(RuntimeError):
    def __init__(self, status: int, where: str):
        msg = lib.rts_last_error().decode(errors="replace")
        super().__init__(f"{where}: status {status}: {msg}")
        self.status = status


def check(status: int, where: str = "rtstereo") -> None:
    if status != RTS_OK:
        raise RtsError(status, where)


def call(name: str, *args) -> None:
    check(getattr(lib, name)(*args), name)

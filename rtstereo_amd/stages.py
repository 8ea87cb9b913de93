"""Host-array conveniences over the per-stage C ABI.

Each function uploads numpy inputs, runs one HIP stage and returns numpy
outputs.  Volumes are disparity-major ``[D][H][W]`` float32 (the device
layout); ``to_pm`` / ``to_dm`` convert to and from the reference's
pixel-major ``[H][W][D]``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (RTS_L2R, RTS_CROSS_EXACT, RTS_AW_FP32, RTS_NCC_SEQ, call, lib)
from .device import DeviceArray


def to_pm(vol_dm: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.moveaxis(vol_dm, 0, -1))


def to_dm(vol_pm: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.moveaxis(vol_pm, -1, 0))


def _vol(D, H, W):
    return DeviceArray((D, H, W), np.float32)


def ad_gray(ref, tar, D, direction=RTS_L2R):
    H, W = ref.shape
    a, b, v = DeviceArray.from_host(ref), DeviceArray.from_host(tar), _vol(D, H, W)
    call("rts_ad_gray", a.ptr, b.ptr, v.ptr, W, H, D, direction, None)
    return v.download()


def ad_float(ref, tar, D, direction=RTS_L2R):
    H, W = ref.shape
    a, b, v = DeviceArray.from_host(ref.astype(np.float32)), DeviceArray.from_host(tar.astype(np.float32)), _vol(D, H, W)
    call("rts_ad_float", a.ptr, b.ptr, v.ptr, W, H, D, direction, None)
    return v.download()


def ad_color(ref, tar, D, direction=RTS_L2R, trunc=None, norm=False):
    H, W, _ = ref.shape
    a, b, v = DeviceArray.from_host(ref), DeviceArray.from_host(tar), _vol(D, H, W)
    call("rts_ad_color", a.ptr, b.ptr, v.ptr, W, H, D, direction, int(trunc is not None),
         float(trunc or 0.0), int(norm), None)
    return v.download()


def ad_color_half(ref, tar, D, direction=RTS_L2R, trunc=None, norm=False):
    """Colour AD as float16 [D][H][W] (each float cost rounded to nearest)."""
    H, W, _ = ref.shape
    a, b, v = DeviceArray.from_host(ref), DeviceArray.from_host(tar), DeviceArray((D, H, W), np.float16)
    call("rts_ad_color_half", a.ptr, b.ptr, v.ptr, W, H, D, direction, int(trunc is not None),
         float(trunc or 0.0), int(norm), None)
    return v.download()


def volume_to_half(vol):
    a = DeviceArray.from_host(vol.astype(np.float32))
    h = DeviceArray(vol.shape, np.float16)
    call("rts_volume_to_half", a.ptr, h.ptr, a.nbytes // 4, None)
    return h.download()


def census(img, hr, wr):
    H, W = img.shape
    nw = ((2 * hr + 1) * (2 * wr + 1) + 63) // 64
    a, w = DeviceArray.from_host(img), DeviceArray((nw, H, W), np.uint64)
    call("rts_census", a.ptr, w.ptr, W, H, hr, wr, None)
    return w.download()


def census_bools(img, hr, wr):
    H, W = img.shape
    N = (2 * hr + 1) * (2 * wr + 1)
    words = census(img, hr, wr)
    w, b = DeviceArray.from_host(words), DeviceArray((H, W, N), np.uint8)
    call("rts_census_unpack", w.ptr, b.ptr, W, H, N, None)
    return b.download()


def census_pack(bools):
    H, W, N = bools.shape
    nw = (N + 63) // 64
    b, w = DeviceArray.from_host(bools.astype(np.uint8)), DeviceArray((nw, H, W), np.uint64)
    call("rts_census_pack", b.ptr, w.ptr, W, H, N, None)
    return w.download()


def hamming(cref, ctar, D):
    nw, H, W = cref.shape
    a, b, v = DeviceArray.from_host(cref), DeviceArray.from_host(ctar), _vol(D, H, W)
    call("rts_hamming", a.ptr, b.ptr, v.ptr, W, H, D, nw, None)
    return v.download()


def rank(img, m):
    H, W = img.shape
    a, o = DeviceArray.from_host(img), DeviceArray((H, W), np.uint8)
    call("rts_rank", a.ptr, o.ptr, W, H, m, None)
    return o.download()


def convolve(img, mask, hr, wr):
    H, W = img.shape
    m = np.ascontiguousarray(mask, dtype=np.float32)
    a, o = DeviceArray.from_host(img), DeviceArray((H, W), np.float32)
    call("rts_convolve", a.ptr, o.ptr, m.ctypes.data_as(C.POINTER(C.c_float)), W, H, hr, wr, None)
    return o.download()


def ncc(ref, tar, D, hr, wr, mode=RTS_NCC_SEQ, init=None):
    H, W = ref.shape
    a, b, v = DeviceArray.from_host(ref), DeviceArray.from_host(tar), _vol(D, H, W)
    if init is not None:
        v.upload(init)
    call("rts_ncc", a.ptr, b.ptr, v.ptr, W, H, D, hr, wr, mode, None)
    return v.download()


def rank_log_ncc(gref, gtar, rref, rtar, lref, ltar, D, hr, wr, l1, l2, l3, dch=16, ws=False):
    """Three-term cost (rank AD, LoG AD, NCC) as a [D][H][W] volume; with
    ``ws`` through the workspace entry point (per-pixel records)."""
    H, W = gref.shape
    bufs = [DeviceArray.from_host(np.ascontiguousarray(a)) for a in (gref, gtar, rref, rtar)]
    bufs += [DeviceArray.from_host(np.ascontiguousarray(a, np.float32)) for a in (lref, ltar)]
    v = _vol(D, H, W)
    ptrs = [b.ptr for b in bufs]
    if ws:
        n = int(lib.rts_cost_rank_log_ncc_ws_bytes(W, H))
        w = DeviceArray((n,), np.uint8)
        call("rts_cost_rank_log_ncc_ws", *ptrs, v.ptr, W, H, D, hr, wr, l1, l2, l3, dch, w.ptr, n, None)
    else:
        call("rts_cost_rank_log_ncc_dch", *ptrs, v.ptr, W, H, D, hr, wr, l1, l2, l3, dch, None)
    return v.download()


def combine(c1, c2, l1, l2):
    a, b = DeviceArray.from_host(c1.astype(np.float32)), DeviceArray.from_host(c2.astype(np.float32))
    call("rts_combine", a.ptr, b.ptr, float(l1), float(l2), c1.size, None)
    return a.download()


def box(vol, r):
    D, H, W = vol.shape
    a, o = DeviceArray.from_host(vol.astype(np.float32)), _vol(D, H, W)
    call("rts_box", a.ptr, o.ptr, W, H, D, r, None)
    return o.download()


def box_int(vol, r):
    """rts_box_int: running-sum box for integer-valued volumes."""
    D, H, W = vol.shape
    a, o = DeviceArray.from_host(vol.astype(np.float32)), _vol(D, H, W)
    call("rts_box_int", a.ptr, o.ptr, W, H, D, r, None)
    return o.download()


def prox_weight(gamma_p, r):
    p = DeviceArray(((2 * r + 1) ** 2,), np.float32)
    call("rts_prox_weight", p.ptr, float(gamma_p), r, None)
    return p.download()


def compute_weights(lab, prox, gamma_c, r):
    H, W, _ = lab.shape
    A = (2 * r + 1) ** 2
    l, p, w = DeviceArray.from_host(lab), DeviceArray.from_host(prox.astype(np.float32)), DeviceArray((H, W, A), np.float32)
    call("rts_compute_weights", w.ptr, p.ptr, l.ptr, float(gamma_c), W, H, r, None)
    return w.download()


def aw_weights(wref, wtar, vol, r, direction=RTS_L2R):
    D, H, W = vol.shape
    a, b = DeviceArray.from_host(wref), DeviceArray.from_host(wtar)
    v, o = DeviceArray.from_host(vol.astype(np.float32)), _vol(D, H, W)
    call("rts_aw_weights", a.ptr, b.ptr, v.ptr, o.ptr, W, H, D, r, direction, None)
    return o.download()


def aw(lab_ref, lab_tar, vol, r, gamma_c, gamma_p, precision=RTS_AW_FP32):
    D, H, W = vol.shape
    a, b = DeviceArray.from_host(lab_ref), DeviceArray.from_host(lab_tar)
    v, o = DeviceArray.from_host(vol.astype(np.float32)), _vol(D, H, W)
    call("rts_aw", a.ptr, b.ptr, v.ptr, o.ptr, W, H, D, r, float(gamma_c), float(gamma_p), precision, None)
    return o.download()


def aw_half(lab_ref, lab_tar, vol_half, r, gamma_c, gamma_p):
    """FP16 adaptive weights from an FP16 cost volume (rts_aw_half_ws)."""
    D, H, W = vol_half.shape
    a, b = DeviceArray.from_host(lab_ref), DeviceArray.from_host(lab_tar)
    v, o = DeviceArray.from_host(vol_half.astype(np.float16)), _vol(D, H, W)
    nbytes = int(lib.rts_aw_workspace_bytes(W, H, r, 1))
    ws = DeviceArray((nbytes,), np.uint8)
    call("rts_aw_half_ws", a.ptr, b.ptr, v.ptr, o.ptr, W, H, D, r, float(gamma_c), float(gamma_p), ws.ptr, nbytes,
         None)
    return o.download()


def find_cross(rgb, L=17, tao=20):
    """Returns (packed arms uint32 [H][W], horz int32 [H][W][2], vert int32 [H][W][2])."""
    H, W, _ = rgb.shape
    a, arms = DeviceArray.from_host(rgb), DeviceArray((H, W), np.uint32)
    call("rts_find_cross", a.ptr, arms.ptr, W, H, L, tao, None)
    h, v = DeviceArray((H, W, 2), np.int32), DeviceArray((H, W, 2), np.int32)
    call("rts_arms_unpack", arms.ptr, h.ptr, v.ptr, W, H, None)
    return arms.download(), h.download(), v.download()


def pack_arms(horz, vert):
    H, W, _ = horz.shape
    h, v = DeviceArray.from_host(horz.astype(np.int32)), DeviceArray.from_host(vert.astype(np.int32))
    a = DeviceArray((H, W), np.uint32)
    call("rts_arms_pack", h.ptr, v.ptr, a.ptr, W, H, None)
    return a.download()


def hprefix(vol):
    D, H, W = vol.shape
    a = DeviceArray.from_host(vol.astype(np.float32))
    call("rts_hprefix", a.ptr, W, H, D, None)
    return a.download()


def cross_v(P, arms_ref, arms_tar, mode=RTS_CROSS_EXACT):
    D, H, W = P.shape
    a, o = DeviceArray.from_host(P.astype(np.float32)), _vol(D, H, W)
    ar, at = DeviceArray.from_host(arms_ref.astype(np.uint32)), DeviceArray.from_host(arms_tar.astype(np.uint32))
    call("rts_cross_v", a.ptr, o.ptr, ar.ptr, at.ptr, W, H, D, mode, None)
    return o.download()


def cross_v_wta(P, arms_ref, arms_tar, lr=True, tol=0):
    """rts_cross_v_wta: the arms <= 17 sweep fused with the WTA; returns
    (D_L, D_R, mask) (D_R and mask None unless lr)."""
    D, H, W = P.shape
    a = DeviceArray.from_host(P.astype(np.float32))
    ar, at = DeviceArray.from_host(arms_ref.astype(np.uint32)), DeviceArray.from_host(arms_tar.astype(np.uint32))
    dl = DeviceArray((H, W), np.uint8)
    dr = DeviceArray((H, W), np.uint8) if lr else None
    m = DeviceArray((H, W), np.uint8) if lr else None
    nws = int(lib.rts_cross_v_wta_ws_bytes(W, H))
    ws = DeviceArray((nws,), np.uint8) if lr else None
    call("rts_cross_v_wta", a.ptr, ar.ptr, at.ptr, dl.ptr, dr.ptr if lr else None, m.ptr if lr else None, W, H, D,
         tol, ws.ptr if lr else None, nws if lr else 0, None)
    return dl.download(), (dr.download() if lr else None), (m.download() if lr else None)


def cross_fused(cost, arms_ref, arms_tar):
    D, H, W = cost.shape
    a, o = DeviceArray.from_host(cost.astype(np.float32)), _vol(D, H, W)
    ar, at = DeviceArray.from_host(arms_ref.astype(np.uint32)), DeviceArray.from_host(arms_tar.astype(np.uint32))
    call("rts_cross_fused", a.ptr, o.ptr, ar.ptr, at.ptr, W, H, D, None)
    return o.download()


def wta(vol, flags=0, lr=True, tol=0):
    """Returns (D_L, D_R, mask) (D_R and mask None unless lr)."""
    D, H, W = vol.shape
    v = DeviceArray.from_host(vol.astype(np.float32))
    dl = DeviceArray((H, W), np.uint8)
    dr = DeviceArray((H, W), np.uint8) if lr else None
    m = DeviceArray((H, W), np.uint8) if lr else None
    call("rts_wta", v.ptr, dl.ptr, dr.ptr if lr else None, m.ptr if lr else None, W, H, D, flags, tol, None)
    return dl.download(), (dr.download() if lr else None), (m.download() if lr else None)


def cross_check(dl, dr, tol=0):
    H, W = dl.shape
    a, b, m = DeviceArray.from_host(dl), DeviceArray.from_host(dr), DeviceArray((H, W), np.uint8)
    call("rts_cross_check", a.ptr, b.ptr, m.ptr, W, H, tol, None)
    return m.download()


def disp_refine(disp, arms, D, literal=False):
    H, W = disp.shape
    d, a = DeviceArray.from_host(disp), DeviceArray.from_host(arms.astype(np.uint32))
    if literal:
        call("rts_disp_refine", d.ptr, None, a.ptr, W, H, D, 1, None)
        return d.download()
    o = DeviceArray((H, W), np.uint8)
    call("rts_disp_refine", d.ptr, o.ptr, a.ptr, W, H, D, 0, None)
    return o.download()


def rgb_to_gray(rgb):
    H, W, _ = rgb.shape
    a, g = DeviceArray.from_host(rgb), DeviceArray((H, W), np.uint8)
    call("rts_rgb_to_gray", a.ptr, g.ptr, W, H, None)
    return g.download()


def rgb_to_lab(rgb):
    H, W, _ = rgb.shape
    a, o = DeviceArray.from_host(rgb), DeviceArray((H, W, 3), np.uint8)
    call("rts_rgb_to_lab", a.ptr, o.ptr, W, H, None)
    return o.download()


def median3(rgb):
    H, W, _ = rgb.shape
    a, o = DeviceArray.from_host(rgb), DeviceArray((H, W, 3), np.uint8)
    call("rts_median3", a.ptr, o.ptr, W, H, None)
    return o.download()

"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/liboracle.so (the C restatement of the reference's
seq/ stage arithmetic in rts_oracle.c) and a loader for oracle/_ref/libseqref.so
(the reference's own seq/ sources compiled here, when they were available at
build time).  Only the tests, the smoke check and the benchmark's CPU-baseline
leg use this package, always as the checker, never as the measured path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libseqref.so")
REF_SRC = "/root/reference/seq"

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def build_ref() -> bool:
    """Compiles the reference's seq/ stage files when they are present."""
    if not os.path.isdir(REF_SRC):
        return False
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile.ref")], check=True)
    return True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        P, I, F, D_ = C.c_void_p, C.c_int, C.c_float, C.c_double
        sigs = {
            "ora_ad_gray": [P, P, P, I, I, I, I],
            "ora_ad_float": [P, P, P, I, I, I, I],
            "ora_ad_color": [P, P, P, I, I, I, I, I, F, I],
            "ora_census": [P, P, I, I, I, I],
            "ora_hamming": [P, P, P, I, I, I, I, I],
            "ora_rank": [P, P, I, I, I],
            "ora_convolve": [P, P, P, I, I, I, I],
            "ora_convolve_f64": [P, P, P, I, I, I, I],
            "ora_ncc": [P, P, P, I, I, I, I, I, I],
            "ora_ncc_f64": [P, P, P, I, I, I, I, I],
            "ora_combine": [P, P, F, F, I, I, I],
            "ora_combine3": [P, P, P, F, F, F, C.c_size_t],
            "ora_combine_f64": [P, P, D_, D_, P, C.c_size_t],
            "ora_box": [P, P, I, I, I, I, I],
            "ora_prox_weight": [P, F, I],
            "ora_compute_weights": [P, P, P, F, I, I, I],
            "ora_aw_l2r": [P, P, P, P, I, I, I, I],
            "ora_aw_l2r_f64": [P, P, P, P, P, I, P, I, I, I, I, D_, D_],
            "ora_aw_rows_f64": [P, P, P, P, I, P, I, I, I, I, D_, D_],
            "ora_find_cross": [P, P, P, I, I, I, I],
            "ora_hprefix": [P, I, I, I],
            "ora_cross_v": [P, P, P, P, P, P, I, I, I],
            "ora_cross_f64": [P, P, P, P, P, P, I, I, I],
            "ora_disp_refine": [P, P, P, P, I, I, I, I],
            "ora_wta": [P, P, I, I, I, I, I],
            "ora_wta_r2l_diag": [P, P, I, I, I, I],
            "ora_cross_check": [P, P, P, I, I, I],
        }
        for n, a in sigs.items():
            f = getattr(_lib, n)
            f.argtypes = a
            f.restype = None
    return _lib


def has_ref() -> bool:
    return os.path.exists(REF_LIB)


def ref_lib():
    """The reference's own stage functions (SeqLib), or None if not built."""
    if not has_ref():
        return None
    from rtstereo_amd.seqabi import SeqLib
    return SeqLib(REF_LIB)


def _p(a):
    assert a.flags.c_contiguous, "oracle arrays must be C-contiguous"
    return a.ctypes.data


# ---- numpy wrappers (reference layouts: volumes [H][W][D]) ----

def ad_gray(ref, tar, D, direction=1):
    H, W = ref.shape
    out = np.zeros((H, W, D), np.float32)
    lib().ora_ad_gray(_p(ref), _p(tar), _p(out), D, H, W, direction)
    return out


def ad_float(ref, tar, D, direction=1):
    H, W = ref.shape
    out = np.zeros((H, W, D), np.float32)
    ref, tar = np.ascontiguousarray(ref, np.float32), np.ascontiguousarray(tar, np.float32)
    lib().ora_ad_float(_p(ref), _p(tar), _p(out), D, H, W, direction)
    return out


def ad_color(ref, tar, D, direction=1, trunc=None, norm=False):
    H, W, _ = ref.shape
    out = np.zeros((H, W, D), np.float32)
    lib().ora_ad_color(_p(ref), _p(tar), _p(out), D, H, W, direction, int(trunc is not None),
                       float(trunc or 0.0), int(norm))
    return out


def census(img, hr, wr):
    H, W = img.shape
    out = np.zeros((H, W, (2 * hr + 1) * (2 * wr + 1)), np.uint8)
    lib().ora_census(_p(img), _p(out), hr, wr, H, W)
    return out


def hamming(bref, btar, hr, wr, D):
    H, W, _ = bref.shape
    out = np.zeros((H, W, D), np.float32)
    lib().ora_hamming(_p(np.ascontiguousarray(bref, np.uint8)), _p(np.ascontiguousarray(btar, np.uint8)), _p(out),
                      hr, wr, H, W, D)
    return out


def rank(img, m):
    H, W = img.shape
    out = np.zeros((H, W), np.uint8)
    lib().ora_rank(_p(img), _p(out), m, H, W)
    return out


def convolve(img, mask, hr, wr, f64=False):
    H, W = img.shape
    m = np.ascontiguousarray(mask, np.float32)
    if f64:
        out = np.zeros((H, W), np.float64)
        lib().ora_convolve_f64(_p(img), _p(out), _p(m), H, W, hr, wr)
    else:
        out = np.zeros((H, W), np.float32)
        lib().ora_convolve(_p(img), _p(out), _p(m), H, W, hr, wr)
    return out


def ncc(ref, tar, D, hr, wr, literal=False, init=None):
    H, W = ref.shape
    out = np.zeros((H, W, D), np.float32) if init is None else np.ascontiguousarray(init, np.float32).copy()
    lib().ora_ncc(_p(ref), _p(tar), _p(out), D, H, W, hr, wr, int(literal))
    return out


def ncc_f64(ref, tar, D, hr, wr):
    H, W = ref.shape
    out = np.zeros((H, W, D), np.float64)
    lib().ora_ncc_f64(_p(ref), _p(tar), _p(out), D, H, W, hr, wr)
    return out


def combine(c1, c2, l1, l2):
    H, W, D = c1.shape
    out = np.ascontiguousarray(c1, np.float32).copy()
    lib().ora_combine(_p(out), _p(np.ascontiguousarray(c2, np.float32)), l1, l2, H, W, D)
    return out


def combine3(c1, c2, ncc, l1, l2, l3):
    out = np.ascontiguousarray(c1, np.float32).copy()
    lib().ora_combine3(_p(out), _p(np.ascontiguousarray(c2, np.float32)), _p(np.ascontiguousarray(ncc, np.float32)),
                       l1, l2, l3, out.size)
    return out


def combine_f64(c1, c2, l1, l2):
    c1 = np.ascontiguousarray(c1, np.float32)
    out = np.zeros(c1.shape, np.float64)
    lib().ora_combine_f64(_p(c1), _p(np.ascontiguousarray(c2, np.float32)), l1, l2, _p(out), c1.size)
    return out


def box(cost, r, literal=False):
    H, W, D = cost.shape
    src = np.ascontiguousarray(cost, np.float32).copy()
    out = np.zeros_like(src)
    lib().ora_box(_p(src), _p(out), H, W, D, r, int(literal))
    return src if literal else out


def prox_weight(gp, r):
    out = np.zeros(((2 * r + 1) ** 2,), np.float32)
    lib().ora_prox_weight(_p(out), gp, r)
    return out


def compute_weights(lab, prox, gc, r):
    H, W, _ = lab.shape
    out = np.zeros((H, W, (2 * r + 1) ** 2), np.float32)
    lib().ora_compute_weights(_p(out), _p(np.ascontiguousarray(prox, np.float32)), _p(lab), gc, H, W, r)
    return out


def aw_l2r(wref, wtar, cost, r):
    H, W, D = cost.shape
    out = np.zeros((H, W, D), np.float32)
    lib().ora_aw_l2r(_p(wref), _p(wtar), _p(np.ascontiguousarray(cost, np.float32)), _p(out), H, W, D, r)
    return out


def aw_l2r_f64(labL, labR, cost, ys, xs, r, gc, gp):
    H, W, D = cost.shape
    ys = np.ascontiguousarray(ys, np.int32)
    xs = np.ascontiguousarray(xs, np.int32)
    out = np.zeros((len(ys), D), np.float64)
    lib().ora_aw_l2r_f64(_p(labL), _p(labR), _p(np.ascontiguousarray(cost, np.float32)), _p(ys), _p(xs), len(ys),
                         _p(out), H, W, D, r, gc, gp)
    return out


def aw_rows_f64(labL, labR, cost, rows, r, gc, gp):
    """FP64 adaptive-weight aggregation of whole rows: [len(rows)][W][D]."""
    H, W, D = cost.shape
    assert r <= 16
    rows = np.ascontiguousarray(rows, np.int32)
    out = np.zeros((len(rows), W, D), np.float64)
    lib().ora_aw_rows_f64(_p(labL), _p(labR), _p(np.ascontiguousarray(cost, np.float32)), _p(rows), len(rows),
                          _p(out), H, W, D, r, gc, gp)
    return out


def median3(rgb):
    """3x3 median per channel with replicated borders (opencl/Main.cpp:342-345
    applies OpenCV's medianBlur with ksize 3 before the cross arms)."""
    H, W, _ = rgb.shape
    p = np.pad(rgb, ((1, 1), (1, 1), (0, 0)), mode="edge")
    stack = np.stack([p[dy:dy + H, dx:dx + W] for dy in range(3) for dx in range(3)], 0)
    return np.sort(stack, 0)[4].astype(np.uint8)


def find_cross(rgb, L=17, tao=20):
    H, W, _ = rgb.shape
    horz = np.zeros((H, W, 2), np.int32)
    vert = np.zeros((H, W, 2), np.int32)
    lib().ora_find_cross(_p(rgb), _p(horz), _p(vert), H, W, L, tao)
    return horz, vert


def hprefix(cost):
    H, W, D = cost.shape
    out = np.ascontiguousarray(cost, np.float32).copy()
    lib().ora_hprefix(_p(out), H, W, D)
    return out


def cross_v(P_, href, htar, vref, vtar):
    H, W, D = P_.shape
    out = np.zeros((H, W, D), np.float32)
    lib().ora_cross_v(_p(np.ascontiguousarray(P_, np.float32)), _p(out), _p(href), _p(htar), _p(vref), _p(vtar),
                      H, W, D)
    return out


def cross_f64(cost, href, htar, vref, vtar):
    H, W, D = cost.shape
    out = np.zeros((H, W, D), np.float64)
    lib().ora_cross_f64(_p(np.ascontiguousarray(cost, np.float32)), _p(out), _p(href), _p(htar), _p(vref),
                        _p(vtar), H, W, D)
    return out


def disp_refine(disp, horz, vert, D, literal=False):
    H, W = disp.shape
    src = np.ascontiguousarray(disp).copy()
    out = np.zeros_like(src)
    lib().ora_disp_refine(_p(src), _p(out), _p(horz), _p(vert), H, W, D, int(literal))
    return src if literal else out


def wta(cost, maximize=False, literal=False):
    H, W, D = cost.shape
    out = np.zeros((H, W), np.uint8)
    lib().ora_wta(_p(np.ascontiguousarray(cost, np.float32)), _p(out), H, W, D, int(maximize), int(literal))
    return out


def wta_r2l_diag(cost, maximize=False):
    H, W, D = cost.shape
    out = np.zeros((H, W), np.uint8)
    lib().ora_wta_r2l_diag(_p(np.ascontiguousarray(cost, np.float32)), _p(out), H, W, D, int(maximize))
    return out


def cross_check(dl, dr, tol=0):
    H, W = dl.shape
    out = np.zeros((H, W), np.uint8)
    lib().ora_cross_check(_p(dl), _p(dr), _p(out), H, W, tol)
    return out


# ---- pipelines in the reference driver's order (seq/Main.cpp:180-257) ----

LOG5 = np.array([0.0239, 0.046, 0.0499, 0.046, 0.0239,
                 0.046, 0.0061, -0.0923, 0.0061, 0.046,
                 0.0499, -0.0923, -0.3182, -0.0923, 0.0499,
                 0.046, 0.0061, -0.0923, 0.0061, 0.046,
                 0.0239, 0.046, 0.0499, 0.046, 0.0239], np.float32)


def pipeline(desc, left_rgb, right_rgb, left_gray=None, right_gray=None, left_lab=None, right_lab=None):
    """Oracle of rts_pipeline_run for a PipelineDesc-like object (attributes
    as rtstereo_amd._lib.PipelineDesc).  Returns (D_L, D_R, mask, volume)."""
    from rtstereo_amd import gen
    d = desc
    D = d.disp
    if left_gray is None:
        left_gray, right_gray = gen.to_gray(left_rgb), gen.to_gray(right_rgb)

    def cost(which):
        if which == 1:
            return ad_color(left_rgb, right_rgb, D, 1, d.ad_trunc if d.ad_trunc_enable else None, bool(d.ad_norm))
        if which == 2:
            return ad_gray(rank(left_gray, d.rank_m), rank(right_gray, d.rank_m), D)
        if which == 3:
            return hamming(census(left_gray, d.census_hr, d.census_wr), census(right_gray, d.census_hr, d.census_wr),
                           d.census_hr, d.census_wr, D)
        if which == 4:
            return ad_float(convolve(left_gray, LOG5, 2, 2), convolve(right_gray, LOG5, 2, 2), D)
        if which == 5:
            return ncc_f64(left_gray, right_gray, D, d.ncc_hr, d.ncc_wr).astype(np.float32)
        raise ValueError(which)

    vol = cost(d.cost1)
    if getattr(d, "cost3", 0):
        vol = combine3(vol, cost(d.cost2), cost(d.cost3), d.lambda1, d.lambda2, d.lambda3)
    elif d.cost2:
        vol = combine(vol, cost(d.cost2), d.lambda1, d.lambda2)
    if d.aggr == 1:
        vol = box(vol, d.box_r)
    elif d.aggr == 2:
        prox = prox_weight(d.aw_gamma_p, d.aw_r)
        wl = compute_weights(left_lab, prox, d.aw_gamma_c, d.aw_r)
        wr = compute_weights(right_lab, prox, d.aw_gamma_c, d.aw_r)
        vol = aw_l2r(wl, wr, vol, d.aw_r)
    elif d.aggr == 3:
        hl, vl = find_cross(left_rgb, d.cross_L, d.cross_tao)
        hr, vr = find_cross(right_rgb, d.cross_L, d.cross_tao)
        for _ in range(d.cross_iters):
            vol = cross_v(hprefix(vol), hl, hr, vl, vr)
    maximize = d.cost1 == 5
    dl = wta(vol, maximize, bool(d.wta_literal))
    dr = wta_r2l_diag(vol, maximize)
    m = cross_check(dl, dr, d.lr_tol)
    return dl, dr, m, vol

"""Writes the golden fixtures of the reference's seq/ stage functions.

Runs the reference's own sequential stage functions, compiled from its
sources by oracle/Makefile.ref into oracle/_ref/libseqref.so, on small seeded
inputs and stores inputs and outputs as .npz data files next to this script.
The reference ships no test vectors of its own; these fixtures are the
known answers the oracle and the GPU path are checked against.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from rtstereo_amd import gen  # noqa: E402


def small_case(seed: int, W: int, H: int, D: int, ref) -> dict:
    p = gen.make_pair(W, H, D, 4, 2.0, seed)
    L, R = p.left, p.right
    gl, gr = p.gray()
    ll, lr = p.lab()
    out = dict(left=L, right=R, gray_l=gl, gray_r=gr, lab_l=ll, lab_r=lr, D=np.int32(D))
    out["ad_gray_l2r"] = ref.ad("pixelBasedL2R", gl, gr, D)
    out["ad_gray_r2l"] = ref.ad("pixelBasedR2L", gl, gr, D)
    fl = (gl.astype(np.float32) * 0.37).astype(np.float32)
    fr = (gr.astype(np.float32) * 0.37).astype(np.float32)
    out["float_l"], out["float_r"] = fl, fr
    out["ad_float_l2r"] = ref.ad("pixelBasedCostL2R_Float", fl, fr, D)
    out["ad_float_r2l"] = ref.ad("pixelBasedCostR2L_Float", fl, fr, D)
    out["ad_color"] = ref.ad("pixelBasedCostL2R_Color", L, R, D, 0)
    out["ad_color_norm"] = ref.ad("pixelBasedCostL2R_Color", L, R, D, 1)
    out["ad_color_t20"] = ref.ad("pixelBasedCostL2R_Color_trunc", L, R, D, 20.0, 0)
    out["ad_color_t20_norm"] = ref.ad("pixelBasedCostL2R_Color_trunc", L, R, D, 20.0, 1)
    out["ad_color_r2l"] = ref.ad("pixelBasedR2L_Color", L, R, D, 0)
    out["ad_color_r2l_t24"] = ref.ad("pixelBasedR2L_Color_trunc", L, R, D, 24.0, 0)
    cl, cr = ref.census(gl, 4, 3), ref.census(gr, 4, 3)
    out["census_l"], out["census_r"] = cl.astype(np.uint8), cr.astype(np.uint8)
    out["hamming"] = ref.hamming(cl, cr, 4, 3, D)
    out["rank3_l"] = ref.rank(gl, 3)
    out["rank7_l"] = ref.rank(gl, 7)
    out["log_l"] = ref.convolve(gl, oracle.LOG5, 2, 2)
    out["log_r"] = ref.convolve(gr, oracle.LOG5, 2, 2)
    # NCC leaves flat-window entries untouched: start from a marker value.
    init = np.full((H, W, D), -7.0, np.float32)
    out["ncc_r1"] = ref.ncc(gl, gr, D, 1, 1, init=init)
    out["ncc_r3"] = ref.ncc(gl, gr, D, 3, 3, init=init)
    out["combine"] = ref.combine(out["hamming"], out["ad_color"], 30.0, 10.0)
    out["box_r2_inplace"] = ref.box(out["ad_color_t20"], 2)
    out["prox_r3"] = ref.prox(36.0, 3)
    wl = ref.weights(out["prox_r3"], ll, 7.0, 3)
    wr = ref.weights(out["prox_r3"], lr, 7.0, 3)
    out["weights_l_r3_row0"] = wl[0].copy()
    out["aw_r3"] = ref.aw_l2r(wl, wr, out["ad_color_t20"], 3)
    hl, vl = ref.find_cross(L)
    hr, vr = ref.find_cross(R)
    out["horz_l"], out["vert_l"], out["horz_r"], out["vert_r"] = hl, vl, hr, vr
    P = ref.hprefix(out["combine"])
    out["hprefix"] = P
    out["cross"] = ref.cross_v(P, hl, hr, vl, vr)
    out["wta"] = ref.wta(out["cross"])
    out["wta_ncc"] = ref.wta(out["ncc_r1"], ncc=True)
    out["wta_box"] = ref.wta(out["box_r2_inplace"])
    dr = oracle.wta_r2l_diag(out["cross"])
    out["dr_diag"] = dr
    out["cross_check"] = ref.cross_check(out["wta"], dr)
    out["refine"] = ref.refine(out["wta"], hl, vl, D)
    return out


def plain_full(ref) -> dict:
    """The reference-order plain pipeline at 384x288, D=16 (in-place box and
    WTA carry included, as the reference computes them)."""
    p = gen.plain_384x288()
    cost = ref.ad("pixelBasedCostL2R_Color_trunc", p.left, p.right, 16, 20.0, 0)
    agg = ref.box(cost, 4)
    disp = ref.wta(agg)
    return dict(digest=np.array(p.digest()), disp_literal=disp)


def census_cross_small(ref) -> dict:
    """Reference-order census + AD -> combine -> cross -> WTA at 160x96, D=24."""
    p = gen.make_pair(160, 96, 24, 10, 3.0, seed=11)
    gl, gr = p.gray()
    cl, cr = ref.census(gl, 4, 3), ref.census(gr, 4, 3)
    c1 = ref.hamming(cl, cr, 4, 3, 24)
    c2 = ref.ad("pixelBasedCostL2R_Color", p.left, p.right, 24, 0)
    c1 = ref.combine(c1, c2, 30.0, 10.0)
    hl, vl = ref.find_cross(p.left)
    hr, vr = ref.find_cross(p.right)
    A = ref.cross_v(ref.hprefix(c1), hl, hr, vl, vr)
    return dict(digest=np.array(p.digest()), disp_literal=ref.wta(A), vol_sha=np.array(_sha(A)))


def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    oracle.build_ref()
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("the reference's seq/ sources are not available: cannot regenerate fixtures")
    np.savez_compressed(os.path.join(HERE, "seq_small_40x24.npz"), **small_case(7, 40, 24, 8, ref))
    np.savez_compressed(os.path.join(HERE, "seq_small_29x21.npz"), **small_case(8, 29, 21, 5, ref))
    np.savez_compressed(os.path.join(HERE, "seq_plain_384x288.npz"), **plain_full(ref))
    np.savez_compressed(os.path.join(HERE, "seq_census_cross_160x96.npz"), **census_cross_small(ref))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()

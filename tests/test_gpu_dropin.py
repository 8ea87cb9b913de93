"""The drop-in (librtstereo_seq.so, the reference's seq/ signatures on the
GPU) against the reference's own outputs stored in the golden fixtures.

Every stage must reproduce the reference bit-exactly, except where the
drop-in deliberately differs (documented in include/rtstereo_seq.h): the box
filter is a true box (checked against the oracle's out-of-place mode), the
exponentials of combineCost/calcProxWeight/computeWeights come from the
device's expf (checked to 1e-5 relative / a few ulp).
"""
import numpy as np
import pytest

import oracle
from conftest import load_golden, golden_files

pytestmark = pytest.mark.gpu

import rtstereo_amd  # noqa: E402

seq = rtstereo_amd.seq_lib()


def test_all_symbols_exported():
    from rtstereo_amd.seqabi import SYMBOLS
    missing = [n for n in SYMBOLS if not seq.has(n)]
    assert not missing, missing


def test_dropin_matches_reference(golden):
    g = golden
    print(f"[case] drop-in {g['left'].shape} D={int(g['D'])}", flush=True)
    L, R, gl, gr, D = g["left"], g["right"], g["gray_l"], g["gray_r"], int(g["D"])
    H, W = gl.shape
    assert np.array_equal(seq.ad("pixelBasedL2R", gl, gr, D), g["ad_gray_l2r"])
    assert np.array_equal(seq.ad("pixelBasedR2L", gl, gr, D), g["ad_gray_r2l"])
    assert np.array_equal(seq.ad("pixelBasedCostL2R_Float", g["float_l"], g["float_r"], D), g["ad_float_l2r"])
    assert np.array_equal(seq.ad("pixelBasedCostR2L_Float", g["float_l"], g["float_r"], D), g["ad_float_r2l"])
    assert np.array_equal(seq.ad("pixelBasedCostL2R_Color", L, R, D, 0), g["ad_color"])
    assert np.array_equal(seq.ad("pixelBasedCostL2R_Color", L, R, D, 1), g["ad_color_norm"])
    assert np.array_equal(seq.ad("pixelBasedCostL2R_Color_trunc", L, R, D, 20.0, 0), g["ad_color_t20"])
    assert np.array_equal(seq.ad("pixelBasedCostL2R_Color_trunc", L, R, D, 20.0, 1), g["ad_color_t20_norm"])
    assert np.array_equal(seq.ad("pixelBasedR2L_Color", L, R, D, 0), g["ad_color_r2l"])
    assert np.array_equal(seq.ad("pixelBasedCostR2L_Color", L, R, D, 0), g["ad_color_r2l"])
    assert np.array_equal(seq.ad("pixelBasedR2L_Color_trunc", L, R, D, 24.0, 0), g["ad_color_r2l_t24"])
    cl = seq.census(gl, 4, 3)
    assert np.array_equal(cl.astype(np.uint8), g["census_l"])
    cr = seq.census(gr, 4, 3)
    assert np.array_equal(seq.hamming(cl, cr, 4, 3, D), g["hamming"])
    assert np.array_equal(seq.rank(gl, 3), g["rank3_l"])
    assert np.array_equal(seq.rank(gl, 7), g["rank7_l"])
    assert np.array_equal(seq.convolve(gl, oracle.LOG5, 2, 2), g["log_l"])
    init = np.full((H, W, D), -7.0, np.float32)
    assert np.array_equal(seq.ncc(gl, gr, D, 1, 1, init=init), g["ncc_r1"])
    assert np.array_equal(seq.ncc(gl, gr, D, 3, 3, init=init), g["ncc_r3"])
    comb = seq.combine(g["hamming"], g["ad_color"], 30.0, 10.0)
    assert np.max(np.abs(comb - g["combine"]) / np.maximum(g["combine"], 1e-30)) <= 1e-5
    # true box (the reference's in-place update is not reproduced)
    assert np.array_equal(seq.box(g["ad_color_t20"], 2), oracle.box(g["ad_color_t20"], 2))
    prox = seq.prox(36.0, 3)
    assert np.allclose(prox, g["prox_r3"], rtol=3e-7, atol=0)
    wl = seq.weights(g["prox_r3"], g["lab_l"], 7.0, 3)
    assert np.allclose(wl[0], g["weights_l_r3_row0"], rtol=3e-7, atol=1e-30)
    wlo = oracle.compute_weights(g["lab_l"], g["prox_r3"], 7.0, 3)
    wro = oracle.compute_weights(g["lab_r"], g["prox_r3"], 7.0, 3)
    assert np.array_equal(seq.aw_l2r(wlo, wro, g["ad_color_t20"], 3), g["aw_r3"])
    hl, vl = seq.find_cross(L)
    hr, vr = seq.find_cross(R)
    assert np.array_equal(hl, g["horz_l"]) and np.array_equal(vl, g["vert_l"])
    assert np.array_equal(hr, g["horz_r"]) and np.array_equal(vr, g["vert_r"])
    P = seq.hprefix(g["combine"])
    assert np.array_equal(P, g["hprefix"])
    assert np.array_equal(seq.cross_v(P, hl, hr, vl, vr), g["cross"])
    assert np.array_equal(seq.wta(g["cross"]), g["wta"])
    assert np.array_equal(seq.wta(g["ncc_r1"], ncc=True), g["wta_ncc"])
    assert np.array_equal(seq.wta(g["box_r2_inplace"]), g["wta_box"])
    assert np.array_equal(seq.cross_check(g["wta"], g["dr_diag"]), g["cross_check"])
    assert np.array_equal(seq.refine(g["wta"], hl, vl, D), g["refine"])


def test_dropin_scalar_helpers():
    import ctypes as C
    ref = oracle.ref_lib()
    f = seq.fn["calcEuclideanDist3P"]
    assert f(10, 3, 200, 7, 0, 255) == np.float32(np.sqrt(np.float32(49 + 193 ** 2 + 255 ** 2)))
    g = load_golden(golden_files()[0])
    L = g["left"]
    H, W, _ = L.shape
    a, b = C.c_int(), C.c_int()
    for (y, x) in [(0, 0), (3, 5), (H - 1, W - 1)]:
        seq.fn["calc_x_range"](L.ctypes.data, y, x, H, W, C.byref(a), C.byref(b))
        assert (a.value, b.value) == tuple(g["horz_l"][y, x])
        seq.fn["calc_y_range"](L.ctypes.data, y, x, H, W, C.byref(a), C.byref(b))
        assert (a.value, b.value) == tuple(g["vert_l"][y, x])

"""The vertical cross sweep fused with the disparity selection
(rts_cross_v_wta) against the sweep followed by the WTA and the check
(rts_cross_v + rts_wta): D_L, D_R and the mask bit for bit, on integer
costs (many ties: the lowest d must win in both views), on the combined
census + AD cost, at widths and heights that are not multiples of the 64 x 64
block, with the right view's wrap at the left border, at D = 1 and D = 256,
and on a 1080p/192 frame."""
import numpy as np
import pytest

import oracle
from rtstereo_amd import gen

pytestmark = pytest.mark.gpu

from rtstereo_amd import _lib as L  # noqa: E402
from rtstereo_amd import stages as S  # noqa: E402


def _dm(v):
    return np.ascontiguousarray(np.transpose(v, (2, 0, 1)))


def _inputs(W, H, D, seed, kind):
    print(f"[case] fused sweep+wta {W}x{H} D={D} seed={seed} cost={kind}", flush=True)
    p = gen.make_pair(W, H, D, 6, 2.5, seed)
    if kind == "ad":
        c = oracle.ad_color(p.left, p.right, D, 1, 20.0, False)            # integers: many ties
    else:
        gl, gr = p.gray()
        h = oracle.hamming(oracle.census(gl, 4, 3), oracle.census(gr, 4, 3), 4, 3, D)
        c = oracle.combine(h, oracle.ad_color(p.left, p.right, D, 1, None, False), 30.0, 10.0)
    arms_l = S.find_cross(p.left)[0]
    arms_r = S.find_cross(p.right)[0]
    P = S.hprefix(_dm(c))
    return P, arms_l, arms_r


def _check(P, al, ar, tol):
    agg = S.cross_v(P, al, ar, L.RTS_CROSS_EXACT | L.RTS_CROSS_ARMS_LE17)
    want = S.wta(agg, 0, True, tol)
    got = S.cross_v_wta(P, al, ar, True, tol)
    for g, w, name in zip(got, want, ("D_L", "D_R", "mask")):
        assert np.array_equal(g, w), name
    dl_only = S.cross_v_wta(P, al, ar, False)[0]
    assert np.array_equal(dl_only, want[0])


@pytest.mark.parametrize("W,H,D,seed,kind", [
    (64, 64, 16, 1, "ad"), (97, 71, 24, 2, "ad"), (130, 130, 40, 3, "combined"), (200, 45, 64, 4, "combined"),
    (33, 80, 1, 5, "ad"), (300, 96, 256, 6, "ad"), (257, 129, 100, 7, "combined")])
def test_fused_equals_sweep_then_wta(W, H, D, seed, kind):
    P, al, ar = _inputs(W, H, D, seed, kind)
    _check(P, al, ar, 1)


def test_fused_1080p():
    """The benchmark frame's size: census Hamming + colour AD combined,
    prefixed, then both paths on the same prefix volume."""
    W, H, D = 1920, 1080, 192
    print(f"[case] fused sweep+wta {W}x{H} D={D}", flush=True)
    p = gen.census_cross_1080p()
    gl, gr = p.gray()
    c = S.combine(S.hamming(S.census(gl, 4, 3), S.census(gr, 4, 3), D), S.ad_color(p.left, p.right, D), 30.0, 10.0)
    P = S.hprefix(c)
    del c
    _check(P, S.find_cross(p.left)[0], S.find_cross(p.right)[0], 1)

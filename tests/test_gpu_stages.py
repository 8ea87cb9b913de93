"""Per-stage parity of the HIP kernels against the CPU oracle.

Integer stages and the float stages that keep the reference's operation
order are checked bit-exactly; stages that reorder float arithmetic are
checked against the FP64 oracle with the tolerance written in the test.
"""
import sys

import numpy as np
import pytest

import oracle
from rtstereo_amd import gen
from conftest import dm, pm

pytestmark = pytest.mark.gpu

from rtstereo_amd import stages as S  # noqa: E402  (loads the HIP library; fails loudly if missing)
from rtstereo_amd import _lib as L  # noqa: E402
from rtstereo_amd import pipeline as PL  # noqa: E402

SIZES = [(40, 24, 8, 7), (29, 21, 5, 8), (67, 41, 13, 9), (130, 70, 64, 10)]


def _pair(W, H, D, seed):
    print(f"[case] {W}x{H} D={D} seed={seed}", flush=True)
    return gen.make_pair(W, H, D, 4, 2.5, seed)


@pytest.mark.parametrize("W,H,D,seed", SIZES)
def test_ad(W, H, D, seed):
    p = _pair(W, H, D, seed)
    gl, gr = p.gray()
    assert np.array_equal(pm(S.ad_gray(gl, gr, D, L.RTS_L2R)), oracle.ad_gray(gl, gr, D, 1))
    assert np.array_equal(pm(S.ad_gray(gl, gr, D, L.RTS_R2L)), oracle.ad_gray(gl, gr, D, -1))
    fl, fr = (gl * np.float32(0.37)).astype(np.float32), (gr * np.float32(0.41)).astype(np.float32)
    assert np.array_equal(pm(S.ad_float(fl, fr, D)), oracle.ad_float(fl, fr, D, 1))
    assert np.array_equal(pm(S.ad_float(fl, fr, D, L.RTS_R2L)), oracle.ad_float(fl, fr, D, -1))
    for trunc, norm in ((None, False), (None, True), (20.0, False), (40.0, True), (7.5, False)):
        for direction, odir in ((L.RTS_L2R, 1), (L.RTS_R2L, -1)):
            got = pm(S.ad_color(p.left, p.right, D, direction, trunc, norm))
            assert np.array_equal(got, oracle.ad_color(p.left, p.right, D, odir, trunc, norm)), (trunc, norm, direction)


@pytest.mark.parametrize("W,H,D,seed", SIZES)
@pytest.mark.parametrize("hr,wr", [(4, 3), (1, 1), (2, 6), (4, 7)])
def test_census_hamming(W, H, D, seed, hr, wr):
    p = _pair(W, H, D, seed)
    gl, gr = p.gray()
    bl, br = oracle.census(gl, hr, wr), oracle.census(gr, hr, wr)
    assert np.array_equal(S.census_bools(gl, hr, wr), bl)
    wl, wr_ = S.census(gl, hr, wr), S.census(gr, hr, wr)
    assert np.array_equal(S.census_pack(bl), wl)
    assert np.array_equal(pm(S.hamming(wl, wr_, D)), oracle.hamming(bl, br, hr, wr, D))


@pytest.mark.parametrize("W,H", [(1024, 40), (1030, 37), (520, 17)])
def test_census_wide(W, H):
    """Rows wide enough for the census kernel's aligned interior blocks (and
    widths that are not a multiple of 4, which stage every row byte-wise),
    with the pipeline's 7x9 window."""
    p = _pair(W, H, 8, 41)
    gl, _ = p.gray()
    assert np.array_equal(S.census(gl, 4, 3), S.census_pack(oracle.census(gl, 4, 3)))


@pytest.mark.parametrize("W,H,D,seed", SIZES[:3])
def test_rank_log(W, H, D, seed):
    p = _pair(W, H, D, seed)
    gl, _ = p.gray()
    for m in (1, 3, 7):
        if m < min(W, H):
            assert np.array_equal(S.rank(gl, m), oracle.rank(gl, m)), m
    assert np.array_equal(S.convolve(gl, oracle.LOG5, 2, 2), oracle.convolve(gl, oracle.LOG5, 2, 2))
    mask = np.linspace(-1, 1, 21).astype(np.float32)
    assert np.array_equal(S.convolve(gl, mask, 1, 3), oracle.convolve(gl, mask, 1, 3))


@pytest.mark.parametrize("W,H,D,seed", SIZES[:3])
@pytest.mark.parametrize("hr,wr", [(1, 1), (3, 3), (2, 1)])
def test_ncc(W, H, D, seed, hr, wr):
    p = _pair(W, H, D, seed)
    gl, gr = p.gray()
    # Reference order: bit-exact, including the untouched flat-window entries.
    init = np.full((D, H, W), -7.0, np.float32)
    got = pm(S.ncc(gl, gr, D, hr, wr, L.RTS_NCC_SEQ | L.RTS_NCC_KEEP_FLAT, init=init))
    want = oracle.ncc(gl, gr, D, hr, wr, literal=True, init=pm(init))
    assert np.array_equal(got, want)
    assert np.array_equal(pm(S.ncc(gl, gr, D, hr, wr, L.RTS_NCC_SEQ)), oracle.ncc(gl, gr, D, hr, wr))
    # Integral form: exact integer sums, FP64 combine; vs the FP64 oracle.
    f64 = oracle.ncc_f64(gl, gr, D, hr, wr)
    got = pm(S.ncc(gl, gr, D, hr, wr, L.RTS_NCC_INTEGRAL)).astype(np.float64)
    assert np.all(got[f64 == 0] == 0)                     # flat windows decided exactly
    assert np.all(np.abs(got - f64) <= 1e-6 * np.abs(f64) + 1e-9)


@pytest.mark.parametrize("W,H,D,seed", [(96, 64, 24, 13), (67, 41, 13, 14), (9, 21, 5, 15), (200, 70, 40, 16)])
def test_rank_log_ncc_records(W, H, D, seed):
    """The workspace form of the three-term cost (per-pixel window-sum
    records) equals the plain entry point bit for bit, for every disparity
    chunk size; a window of another size ignores the workspace."""
    p = _pair(W, H, D, seed)
    gl, gr = p.gray()
    rl, rr = S.rank(gl, 3), S.rank(gr, 3)
    ll, lr = S.convolve(gl, oracle.LOG5, 2, 2), S.convolve(gr, oracle.LOG5, 2, 2)
    for dch in (8, 16, 32):
        plain = S.rank_log_ncc(gl, gr, rl, rr, ll, lr, D, 3, 3, 8.0, 10.0, 0.5, dch=dch)
        rec = S.rank_log_ncc(gl, gr, rl, rr, ll, lr, D, 3, 3, 8.0, 10.0, 0.5, dch=dch, ws=True)
        assert np.array_equal(plain, rec), f"dch {dch}"
    plain = S.rank_log_ncc(gl, gr, rl, rr, ll, lr, D, 2, 1, 8.0, 10.0, 0.5)
    assert np.array_equal(plain, S.rank_log_ncc(gl, gr, rl, rr, ll, lr, D, 2, 1, 8.0, 10.0, 0.5, ws=True))


def test_combine():
    rng = np.random.default_rng(0)
    c1 = rng.integers(0, 63, size=(5, 17, 33)).astype(np.float32)
    c2 = (rng.random((5, 17, 33)) * 765).astype(np.float32)
    got = S.combine(c1, c2, 30.0, 10.0).astype(np.float64)
    f64 = oracle.combine_f64(c1, c2, 30.0, 10.0)
    # relative to the FP64 oracle: <= 1e-5 (north-star float tolerance)
    assert np.max(np.abs(got - f64) / np.maximum(f64, 1e-30)) <= 1e-5
    ref32 = oracle.combine(c1, c2, 30.0, 10.0)
    assert np.max(np.abs(got - ref32) / np.maximum(ref32, 1e-30)) <= 1e-5


@pytest.mark.parametrize("W,H,D,seed", SIZES)
@pytest.mark.parametrize("r", [0, 1, 4, 5, 8, 12, 17])
def test_box(W, H, D, seed, r):
    if r >= min(W, H):
        pytest.skip("radius larger than the image")
    p = _pair(W, H, D, seed)
    c = oracle.ad_color(p.left, p.right, D, 1, 20.0, False)
    assert np.array_equal(pm(S.box(dm(c), r)), oracle.box(c, r))
    f = oracle.ad_float(oracle.convolve(p.gray()[0], oracle.LOG5, 2, 2), oracle.convolve(p.gray()[1], oracle.LOG5, 2, 2), D)
    assert np.array_equal(pm(S.box(dm(f), r)), oracle.box(f, r))


@pytest.mark.parametrize("r", [2, 3, 6, 7])
def test_box_interior_tiles(r):
    """Tiles whose window lies inside the image (no wrap) next to border tiles."""
    p = _pair(257, 131, 5, 12)
    c = oracle.ad_color(p.left, p.right, 5, 1, 20.0, False)
    assert np.array_equal(pm(S.box(dm(c), r)), oracle.box(c, r))


@pytest.mark.parametrize("W,H,D,seed", SIZES)
@pytest.mark.parametrize("r", [0, 1, 4, 5, 8, 12, 16])
def test_box_int(W, H, D, seed, r):
    """Running-sum box on the integer-valued costs it is for (truncated AD,
    colour AD up to 765, census Hamming): equal to the reference's row-major
    float window sum / area, the oracle's box."""
    if r >= min(W, H):
        pytest.skip("radius larger than the image")
    p = _pair(W, H, D, seed)
    gl, gr = p.gray()
    for c in (oracle.ad_color(p.left, p.right, D, 1, 20.0, False),
              oracle.ad_color(p.left, p.right, D, 1, None, False),
              oracle.hamming(oracle.census(gl, 4, 3), oracle.census(gr, 4, 3), 4, 3, D)):
        assert np.array_equal(pm(S.box_int(dm(c), r)), oracle.box(c, r))


def test_box_int_full_size_equals_box():
    """At 1920x1080, D = 64 (the plain configuration's AD truncated at 20,
    r = 4 and 5): the running-sum box equals the row-major box bit for bit."""
    print("[case] box_int vs box 1920x1080 D=64", flush=True)
    p = gen.make_pair(1920, 1080, 64, 20, 2.5, 77)
    c = dm(oracle.ad_color(p.left, p.right, 64, 1, 20.0, False))
    for r in (4, 5):
        assert np.array_equal(S.box_int(c, r), S.box(c, r))


@pytest.mark.parametrize("W,H,D,seed", SIZES[:3])
def test_aw_weights(W, H, D, seed):
    p = _pair(W, H, D, seed)
    ll, lr = p.lab()
    r = 3
    prox = S.prox_weight(36.0, r)
    assert np.allclose(prox, oracle.prox_weight(36.0, r), rtol=2e-7, atol=0)
    po = oracle.prox_weight(36.0, r)
    wl, wr = S.compute_weights(ll, po, 7.0, r), S.compute_weights(lr, po, 7.0, r)
    wlo, wro = oracle.compute_weights(ll, po, 7.0, r), oracle.compute_weights(lr, po, 7.0, r)
    assert np.allclose(wl, wlo, rtol=3e-7, atol=1e-30)
    c = oracle.ad_color(p.left, p.right, D, 1, 40.0, False)
    got = pm(S.aw_weights(wlo, wro, dm(c), r))
    assert np.array_equal(got, oracle.aw_l2r(wlo, wro, c, r))


@pytest.mark.parametrize("W,H,D,seed,r", [s + (3,) for s in SIZES[:3]] + [s + (16,) for s in SIZES[:2]])
def test_aw_on_the_fly(W, H, D, seed, r):
    p = _pair(W, H, D, seed)
    ll, lr = p.lab()
    c = oracle.ad_color(p.left, p.right, D, 1, 40.0, False)
    ys, xs = np.mgrid[0:H, 0:W]
    f64 = oracle.aw_l2r_f64(ll, lr, c, ys.ravel(), xs.ravel(), r, 7.0, 36.0).reshape(H, W, D)
    got32 = pm(S.aw(ll, lr, dm(c), r, 7.0, 36.0, L.RTS_AW_FP32)).astype(np.float64)
    err32 = np.max(np.abs(got32 - f64) / np.maximum(np.abs(f64), 1e-30))
    assert err32 <= 1e-5, err32
    got16 = pm(S.aw(ll, lr, dm(c), r, 7.0, 36.0, L.RTS_AW_FP16)).astype(np.float64)
    err16 = np.max(np.abs(got16 - f64))
    print(f"AW r={r}: fp32 max rel err {err32:.2e}, fp16 max abs err {err16:.3f}")
    d64 = np.argmin(f64, -1)
    srt = np.sort(f64, -1)
    margin = srt[..., 1] - srt[..., 0]
    differ = np.argmin(got16, -1) != d64
    # FP16 weights move costs by at most err16; only pixels whose FP64 winner
    # leads by more than that can be judged (tiny images here are dominated
    # by truncated, near-tied costs).  The full-size rate is reported by the
    # AW benchmark.
    assert err16 < 0.05
    assert differ[margin > 2 * err16].mean() <= 1e-3 if (margin > 2 * err16).any() else True


@pytest.mark.parametrize("r", [1, 5, 8, 10])
def test_aw_padded_radius(r):
    """Radii below the kernel's compiled window run with zero-padded taps."""
    W, H, D, seed = 67, 41, 13, 9
    p = _pair(W, H, D, seed)
    ll, lr = p.lab()
    c = oracle.ad_color(p.left, p.right, D, 1, 40.0, False)
    ys, xs = np.mgrid[0:H, 0:W]
    f64 = oracle.aw_l2r_f64(ll, lr, c, ys.ravel(), xs.ravel(), r, 7.0, 36.0).reshape(H, W, D)
    got32 = pm(S.aw(ll, lr, dm(c), r, 7.0, 36.0, L.RTS_AW_FP32)).astype(np.float64)
    assert np.max(np.abs(got32 - f64) / np.maximum(np.abs(f64), 1e-30)) <= 1e-5
    got16 = pm(S.aw(ll, lr, dm(c), r, 7.0, 36.0, L.RTS_AW_FP16)).astype(np.float64)
    assert np.max(np.abs(got16 - f64)) < 0.05


def test_aw_bands_full_width():
    """1920 wide, 200 rows, r=16: the weight workspace holds 64 rows of the
    FP32 form, so the image runs in four bands; pixels on and beside the
    band edges and the image borders match the FP64 result."""
    W, H, D, r = 1920, 200, 64, 16
    p = gen.make_pair(W, H, D, 20, 3.0, 12)
    print(f"[case] aw bands {W}x{H} D={D} r={r}", flush=True)
    ll, lr = p.lab()
    c = oracle.ad_color(p.left, p.right, D, 1, 40.0, False)
    rng = np.random.default_rng(3)
    rows = np.array([0, 1, 62, 63, 64, 65, 127, 128, 191, 192, 198, 199])
    cols = np.concatenate([[0, 1, 63, 64, W - 2, W - 1], rng.integers(0, W, 14)])
    ys, xs = (a.ravel() for a in np.meshgrid(rows, cols, indexing="ij"))
    f64 = oracle.aw_l2r_f64(ll, lr, c, ys, xs, r, 7.0, 36.0).reshape(len(ys), D)
    for prec in (L.RTS_AW_FP32, L.RTS_AW_FP16):
        got = S.aw(ll, lr, dm(c), r, 7.0, 36.0, prec)[:, ys, xs].T.astype(np.float64)
        if prec == L.RTS_AW_FP32:
            assert np.max(np.abs(got - f64) / np.maximum(np.abs(f64), 1e-30)) <= 1e-5
        else:
            assert np.max(np.abs(got - f64)) < 0.05


@pytest.mark.parametrize("W,H,D,r", [(128, 40, 100, 8), (132, 36, 70, 5), (130, 40, 64, 16), (256, 40, 192, 16)])
def test_aw_fp16_kernel_paths(W, H, D, r):
    """The FP16 aggregation's forms: cost rows staged as 16-byte words (r 8
    and 16 at widths divisible by 4) or as single words (other radii or
    widths), disparity blocks cut short (D 70, 100), several column tiles and
    disparity blocks with the wrap at both borders (256 x 40, D 192); every
    pixel and disparity against the FP64 oracle."""
    print(f"[case] aw fp16 paths {W}x{H} D={D} r={r}", flush=True)
    p = gen.make_pair(W, H, D, 8, 3.0, 31)
    ll, lr = p.lab()
    c = oracle.ad_color(p.left, p.right, D, 1, 40.0, False)
    ys, xs = np.mgrid[0:H, 0:W]
    f64 = oracle.aw_l2r_f64(ll, lr, c, ys.ravel(), xs.ravel(), r, 7.0, 36.0).reshape(H, W, D)
    got16 = pm(S.aw(ll, lr, dm(c), r, 7.0, 36.0, L.RTS_AW_FP16)).astype(np.float64)
    assert np.isfinite(got16).all()
    assert np.max(np.abs(got16 - f64)) < 0.05


@pytest.mark.parametrize("W,H,D,r", [(128, 24, 64, 5), (136, 30, 70, 3), (128, 40, 100, 8), (256, 40, 192, 16),
                                     (200, 20, 150, 8), (384, 20, 256, 16), (192, 18, 128, 16), (256, 28, 100, 12),
                                     (136, 20, 64, 6), (16, 3, 8, 1)])
def test_aw_half_volume(W, H, D, r):
    """FP16 adaptive weights from an FP16 cost volume (one to three
    disparities per lane: D 64 -> 1, 70/100/128/256 -> 2, 150/192 -> 3;
    short disparity blocks and column tiles; radii with every staging
    offset, and radii below the kernel's (12 on the r <= 16 kernel, 6 on
    r <= 8): the FP16 volume is the float costs rounded to nearest (both
    producers), and the aggregation equals rts_aw's FP16 result bit for bit."""
    print(f"[case] aw half volume {W}x{H} D={D} r={r}", flush=True)
    p = gen.make_pair(W, H, D, 8, 3.0, 41)
    ll, lr = p.lab()
    c = S.ad_color(p.left, p.right, D, L.RTS_L2R, 40.0)
    ch = S.ad_color_half(p.left, p.right, D, L.RTS_L2R, 40.0)
    assert np.array_equal(ch.view(np.uint16), c.astype(np.float16).view(np.uint16))
    assert np.array_equal(S.volume_to_half(c).view(np.uint16), ch.view(np.uint16))
    ref16 = S.aw(ll, lr, c, r, 7.0, 36.0, L.RTS_AW_FP16)
    got = S.aw_half(ll, lr, ch, r, 7.0, 36.0)
    assert np.isfinite(got).all()
    assert np.array_equal(got.view(np.uint32), ref16.view(np.uint32))
    if W * H <= 128 * 24:
        ys, xs = np.mgrid[0:H, 0:W]
        f64 = oracle.aw_l2r_f64(ll, lr, oracle.ad_color(p.left, p.right, D, 1, 40.0, False), ys.ravel(), xs.ravel(),
                                r, 7.0, 36.0).reshape(H, W, D)
        assert np.max(np.abs(pm(got).astype(np.float64) - f64)) < 0.05


@pytest.mark.parametrize("W,H,D", [(128, 12, 64), (1920, 4, 192), (1032, 5, 100), (130, 9, 40)])
def test_ad_color_half(W, H, D):
    """The FP16 colour AD (eight pixels per thread where W is a multiple of 8
    and at least 1024, the per-pixel kernel otherwise) equals the float AD
    rounded to half, for both directions, truncation and the 1/3
    normalisation; the float AD of the wide rows equals the oracle's."""
    p = _pair(W, H, D, 43)
    for trunc, norm in ((None, False), (40.0, False), (20.0, True), (None, True)):
        for direction in (L.RTS_L2R, L.RTS_R2L):
            c = S.ad_color(p.left, p.right, D, direction, trunc, norm)
            ch = S.ad_color_half(p.left, p.right, D, direction, trunc, norm)
            assert np.array_equal(ch.view(np.uint16), c.astype(np.float16).view(np.uint16)), (trunc, norm, direction)
            if W >= 1024:
                want = oracle.ad_color(p.left, p.right, D, 1 if direction == L.RTS_L2R else -1, trunc, norm)
                assert np.array_equal(pm(c), want), (trunc, norm, direction)


def test_aw_half_volume_checks():
    """Widths that are not a multiple of 8 are refused with an error."""
    assert L.lib.rts_aw_half_supported(1920, 1080, 192) == 1
    assert L.lib.rts_aw_half_supported(1922, 1080, 192) == 0
    ll = np.zeros((8, 20, 3), np.uint8)
    with pytest.raises(L.RtsError):
        S.aw_half(ll, ll, np.zeros((4, 8, 20), np.float16), 3, 7.0, 36.0)


@pytest.mark.parametrize("Lm,tao", [(40, 30), (60, 255)])
def test_cross_long_arms(Lm, tao):
    """Arms longer than 17 (beyond the sweep's ring window) take the sweep's
    fallback paths; the short-arm flag must not be used for them, and with
    arms of at most 17 it gives the same values."""
    W, H, D = 130, 70, 24
    p = _pair(W, H, D, 21)
    gl, gr = p.gray()
    c = oracle.combine(oracle.hamming(oracle.census(gl, 4, 3), oracle.census(gr, 4, 3), 4, 3, D),
                       oracle.ad_color(p.left, p.right, D, 1), 30.0, 10.0)
    P = oracle.hprefix(c)
    al, hl, vl = S.find_cross(p.left, Lm, tao)
    ar, hr, vr = S.find_cross(p.right, Lm, tao)
    assert max(hl.max(), vl.max(), hr.max(), vr.max()) > 17
    A = oracle.cross_v(P, hl, hr, vl, vr)
    assert np.array_equal(pm(S.cross_v(dm(P), al, ar, L.RTS_CROSS_EXACT)), A)
    al, hl, vl = S.find_cross(p.left, 17, tao)
    ar, hr, vr = S.find_cross(p.right, 17, tao)
    A = oracle.cross_v(P, hl, hr, vl, vr)
    assert np.array_equal(pm(S.cross_v(dm(P), al, ar, L.RTS_CROSS_EXACT | L.RTS_CROSS_ARMS_LE17)), A)


@pytest.mark.parametrize("W,H,D,seed", SIZES)
def test_cross(W, H, D, seed):
    p = _pair(W, H, D, seed)
    for Lm, tao in ((17, 20), (5, 12), (1, 255)):
        arms, horz, vert = S.find_cross(p.left, Lm, tao)
        ho, vo = oracle.find_cross(p.left, Lm, tao)
        assert np.array_equal(horz, ho) and np.array_equal(vert, vo), (Lm, tao)
        assert np.array_equal(S.pack_arms(ho, vo), arms)
    gl, gr = p.gray()
    c = oracle.combine(oracle.hamming(oracle.census(gl, 4, 3), oracle.census(gr, 4, 3), 4, 3, D),
                       oracle.ad_color(p.left, p.right, D, 1), 30.0, 10.0)
    P = oracle.hprefix(c)
    assert np.array_equal(pm(S.hprefix(dm(c))), P)
    al, hl, vl = S.find_cross(p.left)
    ar, hr, vr = S.find_cross(p.right)
    A = oracle.cross_v(P, hl, hr, vl, vr)
    assert np.array_equal(pm(S.cross_v(dm(P), al, ar, L.RTS_CROSS_EXACT)), A)
    assert np.array_equal(pm(S.cross_v(dm(P), al, ar, L.RTS_CROSS_EXACT | L.RTS_CROSS_ARMS_LE17)), A)
    f64 = oracle.cross_f64(c, hl, hr, vl, vr)
    fastv = pm(S.cross_v(dm(P), al, ar, L.RTS_CROSS_FAST)).astype(np.float64)
    assert np.max(np.abs(fastv - A) / np.maximum(A, 1e-30)) <= 1e-5
    fused = pm(S.cross_fused(dm(c), al, ar)).astype(np.float64)
    assert np.max(np.abs(fused - f64) / np.maximum(f64, 1e-30)) <= 1e-5
    # integer costs: every path is exact
    ci = oracle.ad_color(p.left, p.right, D, 1, 24.0, False)
    Ai = oracle.cross_v(oracle.hprefix(ci), hl, hr, vl, vr)
    assert np.array_equal(pm(S.cross_v(dm(oracle.hprefix(ci)), al, ar, L.RTS_CROSS_FAST)), Ai)
    assert np.array_equal(pm(S.cross_fused(dm(ci), al, ar)), Ai)


@pytest.mark.parametrize("W,H,D,seed", [(150, 300, 16, 31), (100, 1000, 4, 32)])
def test_cross_short_arm_segments(W, H, D, seed):
    """Tall images split each column strip of the short-arm sweep into
    several row segments (4 of 75 rows, 15 of 67 with a shorter last one);
    every segment boundary must give the loop form's values bit for bit."""
    p = _pair(W, H, D, seed)
    gl, gr = p.gray()
    c = oracle.combine(oracle.hamming(oracle.census(gl, 4, 3), oracle.census(gr, 4, 3), 4, 3, D),
                       oracle.ad_color(p.left, p.right, D, 1), 30.0, 10.0)
    P = oracle.hprefix(c)
    al, hl, vl = S.find_cross(p.left)
    ar, hr, vr = S.find_cross(p.right)
    A = oracle.cross_v(P, hl, hr, vl, vr)
    assert np.array_equal(pm(S.cross_v(dm(P), al, ar, L.RTS_CROSS_EXACT | L.RTS_CROSS_ARMS_LE17)), A)


@pytest.mark.parametrize("W,H,D,seed", [(64, 1, 8, 41), (66, 17, 6, 42), (72, 52, 8, 43), (33, 52, 8, 44),
                                         (130, 87, 16, 45), (65, 140, 12, 46)])
def test_cross_short_arm_edges(W, H, D, seed):
    """Shapes at the short-arm sweep's edges: one row, fewer rows than one
    step, a last step that completes more than 32 output rows (52, 87, 140:
    the rows past the step's 32 are computed after the walk), and widths
    below, at and just above one 64-column tile."""
    p = _pair(W, H, D, seed)
    gl, gr = p.gray()
    c = oracle.combine(oracle.hamming(oracle.census(gl, 4, 3), oracle.census(gr, 4, 3), 4, 3, D),
                       oracle.ad_color(p.left, p.right, D, 1), 30.0, 10.0)
    P = oracle.hprefix(c)
    al, hl, vl = S.find_cross(p.left)
    ar, hr, vr = S.find_cross(p.right)
    A = oracle.cross_v(P, hl, hr, vl, vr)
    assert np.array_equal(pm(S.cross_v(dm(P), al, ar, L.RTS_CROSS_EXACT | L.RTS_CROSS_ARMS_LE17)), A)


def test_cross_short_arm_special_values():
    """Denormal, infinite and NaN support sums take the quotient's IEEE
    division path; zeros and ordinary values the reciprocal path.  Both must
    equal the loop form (NaN positions included)."""
    W, H, D = 96, 80, 6
    p = _pair(W, H, D, 33)
    rng = np.random.default_rng(33)
    c = rng.random((H, W, D)).astype(np.float32)
    c[:, :, 1] *= np.float32(1e-39)                 # denormal costs and sums
    c[:, :, 2] = 0.0                                # exact zeros
    c[:, :, 3] *= np.float32(3e-38)                 # quotients near the normal limit
    c[rng.random((H, W)) < 0.002, 4] = np.inf       # inf and inf - inf = NaN segments
    c[rng.random((H, W)) < 0.002, 5] = np.nan
    P = oracle.hprefix(c)
    al, hl, vl = S.find_cross(p.left)
    ar, hr, vr = S.find_cross(p.right)
    A = oracle.cross_v(P, hl, hr, vl, vr)
    got = pm(S.cross_v(dm(P), al, ar, L.RTS_CROSS_EXACT | L.RTS_CROSS_ARMS_LE17))
    assert np.isnan(A).any() and (np.abs(A[np.isfinite(A)]) < np.float32(1.2e-38)).any()
    assert np.array_equal(np.isnan(got), np.isnan(A))
    fin = ~np.isnan(A)
    assert np.array_equal(got[fin].view(np.uint32), A[fin].view(np.uint32))


@pytest.mark.parametrize("W,H,D,seed", SIZES)
def test_wta_lr(W, H, D, seed):
    rng = np.random.default_rng(seed)
    v = rng.integers(0, 6, size=(H, W, D)).astype(np.float32)   # ties everywhere
    for maximize in (False, True):
        flags = L.RTS_WTA_MAX if maximize else 0
        dl, dr, m = S.wta(dm(v), flags, True, 1)
        assert np.array_equal(dl, oracle.wta(v, maximize))
        assert np.array_equal(dr, oracle.wta_r2l_diag(v, maximize))
        assert np.array_equal(m, oracle.cross_check(dl, dr, 1))
        dl2, dr2, m2 = S.wta(dm(v), flags | L.RTS_WTA_LITERAL, True, 0)
        assert np.array_equal(dl2, oracle.wta(v, maximize, literal=True))
        assert np.array_equal(m2, oracle.cross_check(dl2, dr2, 0))
    dl, _, _ = S.wta(dm(v), 0, False)
    assert np.array_equal(dl, oracle.wta(v))
    assert np.array_equal(S.cross_check(dl, dr, 2), oracle.cross_check(dl, dr, 2))


@pytest.mark.parametrize("W,H,D,seed", SIZES[:3])
def test_refine(W, H, D, seed):
    p = _pair(W, H, D, seed)
    arms, horz, vert = S.find_cross(p.left)
    rng = np.random.default_rng(seed)
    disp = rng.integers(0, D, size=(H, W)).astype(np.uint8)
    assert np.array_equal(S.disp_refine(disp, arms, D), oracle.disp_refine(disp, horz, vert, D))
    assert np.array_equal(S.disp_refine(disp, arms, D, literal=True), oracle.disp_refine(disp, horz, vert, D, literal=True))


@pytest.mark.parametrize("W,H,D,seed", [(384, 288, 16, 12), (640, 480, 64, 13)])
def test_refine_literal_full_size(W, H, D, seed):
    """The in-place raster-order refinement (anti-diagonal wavefront) at the
    384x288 workload's size and at 640x480, on a realistic map (the WTA of
    the colour AD volume) and on a noisy one, bit-exact vs the oracle."""
    p = _pair(W, H, D, seed)
    arms, horz, vert = S.find_cross(p.left)
    disp = oracle.wta(oracle.ad_color(p.left, p.right, D, 1, 20.0))
    rng = np.random.default_rng(seed)
    noisy = np.where(rng.random((H, W)) < 0.3, rng.integers(0, D, size=(H, W)), disp).astype(np.uint8)
    for dmap in (disp, noisy):
        assert np.array_equal(S.disp_refine(dmap, arms, D, literal=True),
                              oracle.disp_refine(dmap, horz, vert, D, literal=True))
        assert np.array_equal(S.disp_refine(dmap, arms, D), oracle.disp_refine(dmap, horz, vert, D))


def test_refine_long_arms():
    """Arms beyond 17 (up to 150): the wavefront's diagonal step grows with the
    largest horizontal arm, rows wider than 64 columns take extra passes, and
    one pixel's support of 300 x 301 pixels (more than a 16-bit bin holds)
    takes the out-of-place kernel's 32-bit recount."""
    W, H, D = 320, 310, 32
    print(f"[case] refine long arms {W}x{H} D={D}", flush=True)
    rng = np.random.default_rng(14)
    horz = rng.integers(0, 4, size=(H, W, 2)).astype(np.int32)
    vert = rng.integers(0, 4, size=(H, W, 2)).astype(np.int32)
    horz[:, 100, :] = 40            # rows 81 columns wide around column 100
    vert[50:60, 100, :] = 45
    horz[:, 160, :] = 150           # rows 301 columns wide around column 160
    vert[155, 160, :] = 150
    xs, ys = np.arange(W)[None, :], np.arange(H)[:, None]
    horz[..., 0] = np.minimum(horz[..., 0], xs)            # arms stay inside the image
    horz[..., 1] = np.minimum(horz[..., 1], W - 1 - xs)
    vert[..., 0] = np.minimum(vert[..., 0], ys)
    vert[..., 1] = np.minimum(vert[..., 1], H - ys)
    arms = (horz[..., 0] | horz[..., 1] << 8 | vert[..., 0] << 16 | vert[..., 1] << 24).astype(np.uint32)
    disp = rng.integers(0, D, size=(H, W)).astype(np.uint8)
    disp[:, 150:170] = rng.integers(0, 3, size=(H, 20))    # a clear mode inside the wide support
    assert np.array_equal(S.disp_refine(disp, arms, D), oracle.disp_refine(disp, horz, vert, D))
    assert np.array_equal(S.disp_refine(disp, arms, D, literal=True),
                          oracle.disp_refine(disp, horz, vert, D, literal=True))


def test_layout_roundtrip_and_gray():
    rng = np.random.default_rng(1)
    v = rng.random((7, 13, 29)).astype(np.float32)
    from rtstereo_amd.device import DeviceArray
    from rtstereo_amd._lib import call
    a = DeviceArray.from_host(pm(v))
    b = DeviceArray((7, 13, 29), np.float32)
    call("rts_vol_to_dm", a.ptr, b.ptr, 29, 13, 7, None)
    assert np.array_equal(b.download(), v)
    c = DeviceArray((13, 29, 7), np.float32)
    call("rts_vol_to_pm", b.ptr, c.ptr, 29, 13, 7, None)
    assert np.array_equal(c.download(), pm(v))
    p = _pair(33, 17, 4, 3)
    assert np.array_equal(S.rgb_to_gray(p.left), gen.to_gray(p.left))


def test_errors_are_reported():
    from rtstereo_amd._lib import RtsError
    with pytest.raises(RtsError):
        S.ad_gray(np.zeros((4, 4), np.uint8), np.zeros((4, 4), np.uint8), 300)   # D > 256
    with pytest.raises(RtsError):
        S.box(np.zeros((2, 4, 4), np.float32), 9)                                # radius > image


@pytest.mark.parametrize("W,H,D,seed", SIZES[:3])
def test_ingest_median_and_lab(W, H, D, seed):
    p = _pair(W, H, D, seed)
    assert np.array_equal(S.median3(p.left), oracle.median3(p.left))
    lab = S.rgb_to_lab(p.left)
    ref = gen.to_lab(p.left)
    diff = np.abs(lab.astype(int) - ref.astype(int))
    assert diff.max() <= 1 and (diff == 0).mean() >= 0.999


def test_pipeline_computes_lab_on_device():
    p = _pair(96, 64, 16, 9)
    desc = PL.preset("aw", 96, 64, 16, aw_r=5)
    la, lb = gen.to_lab(p.left), gen.to_lab(p.right)
    dl1 = PL.run_host(desc, p.left, p.right, la, lb)[0]
    dl2 = PL.run_host(desc, p.left, p.right)[0]
    assert (dl1 == dl2).mean() >= 0.995

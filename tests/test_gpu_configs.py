"""The measured workloads at their own parameters: D = 256 (the uint8
ceiling), the 1280x720 census+cross workload with four H/V passes, one
2560x1440 pair at D = 256 through the batch engine, and the drop-in called
from several host threads at once.

Full-size frames are checked against the oracle on a band of rows: the
oracle runs the same pipeline on the band plus a halo wide enough that the
band's borders cannot reach the compared rows (window radii, arm lengths and
passes are all bounded), so the compared rows must agree exactly (or within
the float tolerance of the path)."""
import threading

import numpy as np
import pytest

import oracle
from rtstereo_amd import gen

pytestmark = pytest.mark.gpu

from rtstereo_amd import pipeline as PL  # noqa: E402


def _band(p, y0, y1, halo):
    lo, hi = max(0, y0 - halo), min(p.left.shape[0], y1 + halo)
    sub = slice(lo, hi)
    return (np.ascontiguousarray(p.left[sub]), np.ascontiguousarray(p.right[sub]), y0 - lo, y1 - lo)


def test_census_cross_d256_bit_exact():
    """D = 256 with winners at d = 255: a stripe of rows whose right view is
    the left view shifted by 255 columns (with the wrap the reference's cost
    reads use, seq/PixelBasedCost.cpp:72-92 and CensusTransform.cpp:39-62)."""
    W, H, D = 384, 48, 256
    print(f"[case] census_cross {W}x{H} D={D}", flush=True)
    p = gen.make_pair(W, H, D, 8, 2.5, seed=31)
    rng = np.random.default_rng(31)
    right = p.right.copy()
    shifted = np.roll(p.left[12:36], -255, axis=1).astype(np.float64) + rng.normal(0, 1.0, (24, W, 3))
    right[12:36] = np.clip(np.rint(shifted), 0, 255).astype(np.uint8)
    desc = PL.preset("census_cross", W, H, D)
    dl, dr, m, vol = PL.run_host(desc, p.left, right)
    odl, odr, om, ovol = oracle.pipeline(desc, p.left, right)
    assert np.array_equal(np.moveaxis(vol, 0, -1), ovol)
    assert np.array_equal(dl, odl) and np.array_equal(dr, odr) and np.array_equal(m, om)
    assert (dl == 255).sum() > 1000, "the stripe must produce d = 255 winners"
    # the same with the reference's WTA carry
    desc.wta_literal = 1
    dl2, _, _, _ = PL.run_host(desc, p.left, right)
    odl2, _, _, _ = oracle.pipeline(desc, p.left, right)
    assert np.array_equal(dl2, odl2)


def test_census_cross_720p_four_passes_band():
    """1280x720, D = 128, four H/V passes (the camera workload with the
    survey's iteration count): bit-exact against the oracle on rows 340..360.
    Halo 80 rows: census radius 4 plus 17 rows of arm and support reach per
    pass (4 + 4 * 17 = 72)."""
    W, H, D = 1280, 720, 128
    print(f"[case] census_cross {W}x{H} D={D} cross_iters=4", flush=True)
    p = gen.census_cross_720p()
    desc = PL.preset("census_cross", W, H, D, cross_iters=4)
    dl, dr, m, vol = PL.run_host(desc, p.left, p.right)
    y0, y1 = 340, 360
    bl, br, a, b = _band(p, y0, y1, 80)
    bdesc = PL.preset("census_cross", W, bl.shape[0], D, cross_iters=4)
    odl, odr, om, ovol = oracle.pipeline(bdesc, bl, br)
    assert np.array_equal(np.moveaxis(vol[:, y0:y1], 0, -1), ovol[a:b])
    assert np.array_equal(dl[y0:y1], odl[a:b]) and np.array_equal(dr[y0:y1], odr[a:b])
    assert np.array_equal(m[y0:y1], om[a:b])
    assert (dl == p.gt).mean() > 0.8


def test_batch_2560x1440_d256_band():
    """One 2560x1440 pair at D = 256 through rts_batch (rank 7x7 AD + LoG AD +
    NCC 7x7, 11x11 box, WTA, check): at most 0.1% of the disparities on rows
    700..712 differ from the oracle (the batch's separable box adds in a
    different order than the row-major window).  Halo 10 rows: rank and NCC
    radius 3 plus box radius 5."""
    from rtstereo_amd import batch as B
    from rtstereo_amd.device import Stream
    W, H, D = 2560, 1440, 256
    print(f"[case] batch {W}x{H} D={D}", flush=True)
    p = gen.make_pair(W, H, D, 60, 3.0, seed=41)
    s = Stream()
    pr = B.Pair(p.left, p.right, D)
    b = B.Batch(PL.preset("rank_log_ncc", 64, 64, 16), 1 << 34, True)
    b.run([pr], s.handle)
    s.sync()
    dl, dr, m = pr.results()
    b.close()
    pr.free()
    y0, y1 = 700, 712
    bl, br, a, bb = _band(p, y0, y1, 10)
    odl, odr, om, _ = oracle.pipeline(PL.preset("rank_log_ncc", W, bl.shape[0], D), bl, br)
    assert (dl[y0:y1] != odl[a:bb]).mean() <= 1e-3
    assert (dr[y0:y1] != odr[a:bb]).mean() <= 1e-3
    assert dl.max() < D and (dl == p.gt).mean() > 0.5


def _dropin_work(seq, seed):
    p = gen.make_pair(160, 96, 32, 8, 2.5, seed)
    c = oracle.ad_color(p.left, p.right, 32, 1, 20.0)
    hl, vl = oracle.find_cross(p.left)
    hr, vr = oracle.find_cross(p.right)
    box = seq.box(c, 4)
    P = oracle.hprefix(c)
    cross = seq.cross_v(P, hl, hr, vl, vr)
    return box, cross, seq.wta(box), seq.wta(cross)


def test_dropin_concurrent_threads():
    """Four host threads call boxAggregate, cross_stereo_aggregation and WTA
    at once (each thread has its own stream and device buffers); every
    output equals the same call made from one thread."""
    import rtstereo_amd
    seq = rtstereo_amd.seq_lib()
    seeds = [61, 62, 63, 64]
    print("[case] drop-in 4 threads x 3 rounds", flush=True)
    want = {s: _dropin_work(seq, s) for s in seeds}
    got, errors = {}, []

    def worker(s):
        try:
            for r in range(3):
                got[(s, r)] = _dropin_work(seq, s)
        except Exception as e:  # surfaced below
            errors.append(e)

    ts = [threading.Thread(target=worker, args=(s,)) for s in seeds]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    for (s, r), outs in got.items():
        for g, w in zip(outs, want[s]):
            assert np.array_equal(g, w), f"seed {s} round {r}"

"""End-to-end pipelines on the GPU against the oracle, and size-independent
properties at the measured full sizes."""
import numpy as np
import pytest

import oracle
from rtstereo_amd import gen
from conftest import load_golden, golden_files

pytestmark = pytest.mark.gpu

from rtstereo_amd import pipeline as PL  # noqa: E402
from rtstereo_amd import _lib as L  # noqa: E402


def _case(name, W, H, D, seed, **over):
    print(f"[case] pipeline {name} {W}x{H} D={D} seed={seed} {over}", flush=True)
    p = gen.make_pair(W, H, D, 8, 2.5, seed)
    desc = PL.preset(name, W, H, D, **over)
    return p, desc


@pytest.mark.parametrize("name,W,H,D,over", [
    ("plain", 96, 64, 16, {}),
    ("plain", 77, 45, 12, {"wta_literal": 1}),
    ("census_cross", 96, 64, 24, {}),
    ("census_cross", 150, 70, 40, {"wta_literal": 1}),
    ("census_cross", 96, 64, 24, {"cost2": 0}),
    ("census_cross", 80, 48, 16, {"cost1": L.RTS_COST_AD, "cost2": 0, "ad_trunc_enable": 1, "ad_trunc": 24.0}),
    ("census_cross", 96, 64, 24, {"cross_iters": 3}),
    ("rank_box", 80, 48, 16, {}),
    ("log_box", 80, 48, 16, {}),
    ("census_cross", 80, 48, 16, {"cost1": L.RTS_COST_LOG, "cost2": L.RTS_COST_CENSUS, "lambda1": 10.0,
                                  "lambda2": 30.0, "aggr": L.RTS_AGGR_BOX, "box_r": 2}),
    ("rank_log_ncc", 96, 64, 24, {}),
    ("rank_log_ncc", 67, 41, 13, {}),
    ("rank_log_ncc", 160, 72, 40, {"ncc_hr": 2, "ncc_wr": 4}),
])
def test_pipeline_bit_exact(name, W, H, D, over):
    p, desc = _case(name, W, H, D, 5, **over)
    dl, dr, m, vol = PL.run_host(desc, p.left, p.right)
    odl, odr, om, ovol = oracle.pipeline(desc, p.left, p.right)
    if desc.cost3:
        # The NCC term enters as 1 - ncc: where ncc is near 1 a last-bit
        # difference of the float ncc is a large relative change of that
        # term, so the bound is relative to max(|cost|, 1) (costs lie in [0, 3)).
        assert np.max(np.abs(np.moveaxis(vol, 0, -1) - ovol) / np.maximum(np.abs(ovol), 1.0)) <= 1e-5
        assert (dl != odl).mean() <= 1e-3
        return
    if desc.cost2 and not (desc.aggr == L.RTS_AGGR_CROSS):
        # generic combine uses the device expf: values within 1e-5, maps equal
        assert np.max(np.abs(np.moveaxis(vol, 0, -1) - ovol) / np.maximum(np.abs(ovol), 1e-6)) <= 1e-5
        assert (dl != odl).mean() <= 1e-3
        return
    assert np.array_equal(np.moveaxis(vol, 0, -1), ovol)
    assert np.array_equal(dl, odl)
    assert np.array_equal(dr, odr)
    assert np.array_equal(m, om)


def test_pipeline_ncc_box():
    p, desc = _case("ncc_box", 80, 48, 16, 6)
    dl, dr, m, vol = PL.run_host(desc, p.left, p.right)
    odl, odr, om, ovol = oracle.pipeline(desc, p.left, p.right)
    assert np.max(np.abs(np.moveaxis(vol, 0, -1) - ovol)) <= 1e-5
    assert (dl != odl).mean() <= 1e-3


def test_pipeline_fast_cross_tolerance():
    """FAST mode against the reference's own float values (the oracle's
    literal dataflow): <= 1e-5 relative, <= 0.1% of disparities differ."""
    p, desc = _case("census_cross", 160, 96, 32, 7, cross_exact=0)
    dl, dr, m, vol = PL.run_host(desc, p.left, p.right)
    odl, odr, om, ovol = oracle.pipeline(desc, p.left, p.right)
    got = np.moveaxis(vol, 0, -1).astype(np.float64)
    assert np.max(np.abs(got - ovol) / np.maximum(ovol, 1e-30)) <= 1e-5
    assert (dl != odl).mean() <= 1e-3 and (dr != odr).mean() <= 1e-3


def test_plain_full_size_vs_reference_fixture():
    """384x288 D=16: intended semantics bit-exact vs the oracle; against the
    reference's literal output (in-place box) the disparities agree on
    almost every pixel."""
    g = load_golden(golden_files("seq_plain_384x288.npz")[0])
    p = gen.plain_384x288()
    desc = PL.preset("plain", 384, 288, 16, wta_literal=1)
    dl, dr, m, vol = PL.run_host(desc, p.left, p.right)
    odl, _, om, _ = oracle.pipeline(desc, p.left, p.right)
    assert np.array_equal(dl, odl) and np.array_equal(m, om)
    assert (dl == g["disp_literal"]).mean() > 0.95
    assert (dl == p.gt).mean() > 0.9


def test_census_cross_fixture_bit_exact():
    """The reference's own census+AD+combine+cross+WTA output (literal carry)."""
    g = load_golden(golden_files("seq_census_cross_160x96.npz")[0])
    p = gen.make_pair(160, 96, 24, 10, 3.0, seed=11)
    desc = PL.preset("census_cross", 160, 96, 24, wta_literal=1)
    dl, dr, m, vol = PL.run_host(desc, p.left, p.right)
    import hashlib
    assert hashlib.sha256(np.ascontiguousarray(np.moveaxis(vol, 0, -1)).tobytes()).hexdigest() == str(g["vol_sha"])
    assert np.array_equal(dl, g["disp_literal"])


@pytest.mark.parametrize("exact", [1, 0])
def test_census_cross_1080p_properties(exact):
    """Full north-star size: 1920x1080, D=192.  Size-independent checks:
    left-right consistency, agreement with ground truth, determinism, and a
    sampled bit-exact comparison of the exact volume with the oracle."""
    W, H, D = 1920, 1080, 192
    p = gen.census_cross_1080p()
    print(f"[case] census_cross 1080p exact={exact}", flush=True)
    desc = PL.preset("census_cross", W, H, D, cross_exact=exact)
    dl, dr, m, vol = PL.run_host(desc, p.left, p.right)
    dl2, dr2, m2, _ = PL.run_host(desc, p.left, p.right)
    assert np.array_equal(dl, dl2) and np.array_equal(m, m2)          # deterministic
    assert dl.max() < D and dr.max() < D
    assert (dl == p.gt).mean() > 0.85
    assert m.mean() > 0.8
    # oracle on a band of rows (the full oracle at this size takes minutes)
    y0, y1 = 500, 520
    band = slice(y0 - 40, y1 + 40)
    sub = gen.StereoPair(np.ascontiguousarray(p.left[band]), np.ascontiguousarray(p.right[band]), p.gt[band], 0, D)
    gl, gr = gen.to_gray(p.left), gen.to_gray(p.right)
    cl = np.ascontiguousarray(oracle.census(gl, 4, 3)[band])
    cr = np.ascontiguousarray(oracle.census(gr, 4, 3)[band])
    c = oracle.combine(oracle.hamming(cl, cr, 4, 3, D),
                       oracle.ad_color(sub.left, sub.right, D, 1), 30.0, 10.0)
    hl, vl = oracle.find_cross(sub.left)
    hr, vr = oracle.find_cross(sub.right)
    got = np.moveaxis(vol[:, y0:y1], 0, -1)
    A = oracle.cross_v(oracle.hprefix(c), hl, hr, vl, vr)[40:60]
    if exact:
        assert np.array_equal(got, A)
    else:
        assert np.max(np.abs(got - A) / np.maximum(A, 1e-30)) <= 1e-5


@pytest.mark.parametrize("name,W,H,D,over", [
    ("census_cross", 96, 64, 24, {}),
    ("census_cross", 150, 70, 40, {"cross_iters": 2}),
    ("census_cross", 96, 64, 24, {"wta_literal": 1}),
    ("census_cross", 80, 48, 16, {"cost1": L.RTS_COST_AD, "cost2": 0, "ad_trunc_enable": 1, "ad_trunc": 24.0}),
    ("plain", 96, 64, 16, {}),
    ("aw", 80, 48, 16, {"aw_r": 5}),
    ("rank_log_ncc", 96, 64, 24, {}),
    # D >= 128: the captured frame runs in disparity chunks on three streams
    ("census_cross", 320, 96, 192, {}),
    ("census_cross", 260, 80, 130, {"wta_literal": 1}),
    ("census_cross", 256, 72, 256, {"lr_tol": 0}),
])
def test_pipeline_graph_replay(name, W, H, D, over):
    """Graph replay: runs with the same buffers and stream are captured once
    and replayed; every replay, with new images uploaded into the same
    buffers and host syncs in between, must equal the eager run of those
    images.  Other buffers fall back to eager launches and are captured on
    their second run."""
    from rtstereo_amd.device import DeviceArray, Stream
    print(f"[case] graph replay {name} {W}x{H} D={D} {over}", flush=True)
    pairs = [gen.make_pair(W, H, D, 8, 2.5, seed) for seed in (3, 4)]
    desc = PL.preset(name, W, H, D, **over)
    want = []
    for pr in pairs:
        dl, dr, m, _ = PL.run_host(desc, pr.left, pr.right)
        want.append((dl, dr, m))
    assert not np.array_equal(want[0][0], want[1][0])
    pipe = PL.Pipeline(desc)
    pipe.enable_graphs(True)
    stream = Stream()
    a = DeviceArray((H, W, 3), np.uint8)
    b = DeviceArray((H, W, 3), np.uint8)
    for i in range(6):
        pr = pairs[i % 2]
        a.upload(pr.left, stream.handle)
        b.upload(pr.right, stream.handle)
        pipe.run(a, b, stream=stream.handle)
        stream.sync()
        got = (pipe.dl.download(stream.handle), pipe.dr.download(stream.handle), pipe.mask.download(stream.handle))
        for g, w in zip(got, want[i % 2]):
            assert np.array_equal(g, w), f"run {i}"
    # other input buffers: eager first, then captured again
    a2, b2 = DeviceArray.from_host(pairs[1].left), DeviceArray.from_host(pairs[1].right)
    for i in range(3):
        pipe.run(a2, b2, stream=stream.handle)
        stream.sync()
        assert np.array_equal(pipe.dl.download(stream.handle), want[1][0]), f"new buffers, run {i}"
    pipe.close()

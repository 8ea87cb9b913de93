"""Batch engine on the GPU: mixed sizes in one run agree with the oracle and
with the per-pair pipeline, graph replay and chunking change nothing."""
import numpy as np
import pytest

import oracle
from rtstereo_amd import gen

pytestmark = pytest.mark.gpu

from rtstereo_amd import batch as B  # noqa: E402
from rtstereo_amd import pipeline as PL  # noqa: E402
from rtstereo_amd.device import Stream  # noqa: E402

SHAPES = [(96, 64, 16), (160, 96, 32), (80, 48, 24), (128, 80, 40), (96, 64, 16), (200, 60, 48)]


def _pairs(seed=50):
    return [gen.make_pair(W, H, D, 6, 2.5, seed + i) for i, (W, H, D) in enumerate(SHAPES)]


def _run(pairs, budget, graphs, reps=1, tiles=None):
    s = Stream()
    dev = [B.Pair(p.left, p.right, D) for p, (_, _, D) in zip(pairs, SHAPES)]
    b = B.Batch(PL.preset("rank_log_ncc", 64, 64, 16), budget, graphs)
    if tiles:
        b.set_tiles(*tiles)
    outs = []
    for _ in range(reps):
        b.run(dev, s.handle)
        s.sync()
        outs.append([d.results() for d in dev])
    st = b.stats()
    b.close()
    for d in dev:
        d.free()
    return outs, st


def test_batch_matches_oracle_and_pipeline():
    print("[case] batch 6 mixed pairs, graphs, one chunk", flush=True)
    pairs = _pairs()
    outs, st = _run(pairs, 1 << 30, True, reps=2)
    assert st.chunks == 1 and st.graphs_built == 1 and st.graph_replays == 1
    assert st.box_rows in (32, 52, 64) and st.box_cols in (64, 128)   # 52: balanced rows for r = 5
    assert st.cost_disparities in (8, 16, 32) and min(st.tune_ms) > 0
    for (p, (W, H, D)), (dl, dr, m), (dl2, dr2, m2) in zip(zip(pairs, SHAPES), outs[0], outs[1]):
        assert np.array_equal(dl, dl2) and np.array_equal(m, m2)            # replay reproduces the capture
        desc = PL.preset("rank_log_ncc", W, H, D)
        odl, odr, om, _ = oracle.pipeline(desc, p.left, p.right)
        pdl, pdr, pm_, _ = PL.run_host(desc, p.left, p.right)
        # separable box sums round differently from the row-major window sum
        assert (dl != odl).mean() <= 1e-3 and (dr != odr).mean() <= 1e-3
        assert (dl != pdl).mean() <= 1e-3


@pytest.mark.parametrize("tiles", [(64, 32, 8), (64, 52, 16), (64, 64, 32), (128, 32, 16), (128, 52, 8),
                                   (128, 64, 16)])
def test_batch_tiles_do_not_change_results(tiles):
    """Every box tile (64 or 128 columns by 32, 52 or 64 rows, which sets the
    LDS footprint) and cost block size the tuner may pick gives the same maps
    as the tuned run: each output is summed in the same order whatever the
    tile."""
    print(f"[case] batch 6 mixed pairs, fixed tiles {tiles}", flush=True)
    pairs = _pairs(70)
    ref, _ = _run(pairs, 1 << 30, True)
    got, st = _run(pairs, 1 << 30, True, tiles=tiles)
    assert (st.box_cols, st.box_rows, st.cost_disparities) == tiles
    for (dl, dr, m), (dl2, dr2, m2) in zip(ref[0], got[0]):
        assert np.array_equal(dl, dl2) and np.array_equal(dr, dr2) and np.array_equal(m, m2)


def test_batch_chunks_and_eager_equal_graphs():
    print("[case] batch 6 mixed pairs, 3+ chunks, eager vs graphs", flush=True)
    pairs = _pairs(70)
    need = max(B.pair_bytes(W, H, D) for W, H, D in SHAPES)
    g, st = _run(pairs, int(need * 1.6), True)
    e, st2 = _run(pairs, 1 << 30, False)
    assert st.chunks >= 3 and st2.chunks == 1 and st2.graphs_built == 0
    for (a, b_, c), (x, y, z) in zip(g[0], e[0]):
        assert np.array_equal(a, x) and np.array_equal(b_, y) and np.array_equal(c, z)


MIXED = [(96, 64, 16), (160, 96, 40), (132, 70, 72), (96, 64, 16), (200, 60, 130), (160, 96, 40)]


@pytest.mark.parametrize("name,over", [("aw", {"aw_precision": 1}), ("census_cross", {}),
                                       ("census_cross", {"wta_literal": 1, "cross_iters": 2})])
def test_pipeline_batch_matches_pipelines(name, over):
    """Adaptive-weight FP16 (33x33) and census + cross pairs of mixed sizes
    through the batch engine's pipeline path: every map equals the pair run
    alone through its own pipeline, in one chunk (captured, then replayed)
    and under a budget that forces several chunks and pipeline eviction."""
    print(f"[case] pipeline batch {name} {over}", flush=True)
    pairs = [gen.make_pair(W, H, D, 6, 2.5, 80 + i) for i, (W, H, D) in enumerate(MIXED)]
    want = [PL.run_host(PL.preset(name, W, H, D, **over), p.left, p.right)[:3] for p, (W, H, D) in zip(pairs, MIXED)]
    s = Stream()
    dev = [B.Pair(p.left, p.right, D) for p, (_, _, D) in zip(pairs, MIXED)]
    need = max(B.pipeline_bytes(PL.preset(name, W, H, D, **over)) for W, H, D in MIXED)
    for budget, reps in ((1 << 34, 2), (int(need * 1.5), 2)):
        b = B.Batch(PL.preset(name, 64, 64, 16, **over), budget, True)
        for _ in range(reps):
            for d in dev:
                for a in (d.dl, d.dr, d.mask):
                    a.fill(77)
            b.run(dev, s.handle)
            s.sync()
            for d, w in zip(dev, want):
                for g, x in zip(d.results(), w):
                    assert np.array_equal(g, x)
        st = b.stats()
        assert st.graphs_built >= 1 and (budget < (1 << 34) or (st.chunks == 1 and st.graph_replays == 1))
        if budget < (1 << 34):
            assert st.chunks >= 2
        b.close()
    for d in dev:
        d.free()

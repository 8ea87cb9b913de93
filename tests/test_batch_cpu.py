"""Batch engine argument checks that need no GPU."""
import ctypes as C

import pytest

from rtstereo_amd import _lib as L
from rtstereo_amd._lib import RtsError
from rtstereo_amd import batch
from rtstereo_amd.pipeline import preset


def test_batch_configurations():
    # other pipeline configurations run through resident per-shape pipelines
    for name, over in (("census_cross", {}), ("aw", {"aw_precision": L.RTS_AW_FP16})):
        batch.Batch(preset(name, 64, 48, 16, **over), 1 << 30).close()
    # ... but only valid ones
    bad = preset("census_cross", 64, 48, 16)
    bad.cost1, bad.cost2 = L.RTS_COST_NCC, 0
    with pytest.raises(RtsError) as e:
        batch.Batch(bad, 1 << 30)
    assert e.value.status == L.RTS_ERR_INVALID
    d = preset("rank_log_ncc", 64, 48, 16)
    d.box_r = 9
    with pytest.raises(RtsError):
        batch.Batch(d, 1 << 30)


def test_set_tiles_checks():
    b = batch.Batch(preset("rank_log_ncc", 64, 48, 16), 1 << 30)
    b.set_tiles(128, 52, 8)          # 52: 64 - 2r rounded down to a multiple of 4, r = 5
    b.set_tiles(0, 0, 0)
    for bad in ((96, 32, 16), (64, 48, 16), (64, 32, 12)):
        with pytest.raises(RtsError):
            b.set_tiles(*bad)
    b.close()


def test_pair_bytes_and_budget():
    W, H, D = 640, 480, 64
    n = batch.pair_bytes(W, H, D)
    assert n >= 2 * W * H * D * 4 + 12 * W * H + 32 * W * H   # volumes, images, cost records
    b = batch.Batch(preset("rank_log_ncc", W, H, D), n - 1)
    fake = 4096   # never dereferenced: the run is refused before any device work
    it = (batch.BatchItem * 1)(batch.BatchItem(W, H, D, fake, fake, fake, fake, fake))
    with pytest.raises(RtsError):
        L.call("rts_batch_run", b.handle, it, 1, None)
    b.close()

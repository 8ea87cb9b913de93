"""Serving loop checks that need no GPU: admission refuses requests larger
than the budget, invalid sizes are rejected at submit, and a queued request
can be cancelled (admission paused so it never starts)."""
import numpy as np
import pytest

from rtstereo_amd import _lib as L
from rtstereo_amd._lib import RtsError
from rtstereo_amd import serve
from rtstereo_amd.pipeline import preset


def _imgs(W, H):
    return np.zeros((H, W, 3), np.uint8), np.zeros((H, W, 3), np.uint8)


def test_budget_rejects_oversized_request():
    s = serve.Server(budget_bytes=1 << 19, slots=2)
    a, b = _imgs(64, 48)
    with pytest.raises(RtsError) as e:
        s.submit(preset("census_cross", 64, 48, 32), a, b)     # about 0.9 MB of volumes and buffers > 512 KiB
    assert e.value.status == L.RTS_ERR_BUDGET
    s.close()


def test_invalid_request_rejected():
    s = serve.Server(budget_bytes=1 << 30, slots=1)
    a, b = _imgs(16, 16)
    with pytest.raises(RtsError) as e:
        s.submit(preset("census_cross", 16, 16, 300), a, b)    # D > 256
    assert e.value.status == L.RTS_ERR_INVALID
    a, b = _imgs(64, 48)
    with pytest.raises(RtsError):
        s.submit(preset("aw", 64, 48, 16), a, b, left_lab=a, right_lab=None)   # one Lab image only
    s.close()


def test_cancel_queued_request():
    s = serve.Server(budget_bytes=1 << 30, slots=1)
    s.pause(True)
    a, b = _imgs(64, 48)
    t = s.submit(preset("census_cross", 64, 48, 16), a, b)
    assert s.counters().queued == 1
    s.cancel(t)
    with pytest.raises(RtsError) as e:
        s.wait(t)
    assert e.value.status == L.RTS_ERR_CANCELLED
    assert s.counters().queued == 0
    s.close()


def test_destroy_cancels_waiting_requests():
    s = serve.Server(budget_bytes=1 << 30, slots=1)
    s.pause(True)
    a, b = _imgs(64, 48)
    s.submit(preset("census_cross", 64, 48, 16), a, b)
    s.close()        # must not hang with a queued request


def test_oneoff_limit_range():
    s = serve.Server(budget_bytes=1 << 30, slots=2)
    with pytest.raises(RtsError):
        s.set_oneoff_limit(0)
    with pytest.raises(RtsError):
        s.set_oneoff_limit(3)
    s.set_oneoff_limit(2)
    s.close()

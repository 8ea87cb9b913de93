"""rts_wta_noalloc (the winner-takes-all entry point that allocates nothing;
its literal-carry scratch is a fixed per-device array) against rts_wta on
the same volumes: two calls in a row, the second smaller, in every mode."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from rtstereo_amd import _lib as L  # noqa: E402
from rtstereo_amd._lib import call  # noqa: E402
from rtstereo_amd.device import DeviceArray  # noqa: E402


def _run(name, vol, flags, tol):
    D, H, W = vol.shape
    v = DeviceArray.from_host(vol)
    outs = [DeviceArray((H, W), np.uint8) for _ in range(3)]
    for o in outs:      # stale bytes must not leak into the results
        call("rts_fill_u8", o.ptr, 77, H * W, None)
    call(name, v.ptr, outs[0].ptr, outs[1].ptr, outs[2].ptr, W, H, D, flags, tol, None)
    call("rts_device_sync")
    return [o.download() for o in outs]


@pytest.mark.parametrize("flags", [0, L.RTS_WTA_LITERAL, L.RTS_WTA_MAX | L.RTS_WTA_LITERAL, L.RTS_WTA_MAX])
def test_noalloc_equals_wta(flags):
    rng = np.random.default_rng(90 + flags)
    for (W, H, D) in [(300, 120, 48), (97, 33, 12)]:      # second call smaller
        print(f"[case] wta_noalloc flags={flags} {W}x{H} D={D}", flush=True)
        # integer-valued costs: many ties, many pixels whose d = 0 is not beaten
        vol = rng.integers(0, 6, size=(D, H, W)).astype(np.float32)
        vol[0] -= (rng.random((H, W)) < 0.4) * 10.0
        want = _run("rts_wta", vol, flags, 1)
        got = _run("rts_wta_noalloc", vol, flags, 1)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)


def test_noalloc_rejects_oversize():
    from rtstereo_amd._lib import RtsError
    with pytest.raises(RtsError):
        call("rts_wta_noalloc", None, None, None, None, 4, 70000, 2, L.RTS_WTA_LITERAL, 0, None)

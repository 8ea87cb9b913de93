"""Serving loop over the C ABI (rts_server_*), and the camera + burst
workload it is measured on.

A Server admits requests continuously: each waits until one of its slots is
free and its device footprint fits the byte budget, then runs on that slot's
HIP stream while other requests run on theirs.  Camera streams keep their
pipeline resident between frames.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import sys
import time

import numpy as np


def hw_queues_for(slots: int) -> int:
    """Hardware queues a serving process wants: room for every slot stream
    and the process's other streams (twice the slots), within the 32 the
    runtime allows."""
    return min(32, 2 * slots)


if __name__ == "__main__":
    # The serving program gives every slot's HIP stream a hardware queue of
    # its own.  HIP spreads streams round-robin over GPU_MAX_HW_QUEUES queues
    # (4 by default); a camera frame whose stream shares a queue with a
    # burst's stream waits, in queue order, behind the burst's queued kernels
    # (0.1 s and more at 40 bursts/s).  The runtime reads the variable when
    # it starts, so it is set before the library loads.
    _slots = 8
    for _i, _a in enumerate(sys.argv):
        if _a == "--slots" and _i + 1 < len(sys.argv):
            _slots = int(sys.argv[_i + 1])
        elif _a.startswith("--slots="):
            _slots = int(_a.split("=", 1)[1])
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < hw_queues_for(_slots):
        os.environ["GPU_MAX_HW_QUEUES"] = str(hw_queues_for(_slots))

from . import _lib as L
from ._lib import PipelineDesc, RtsError, call, lib


class RequestDesc(C.Structure):
    _fields_ = [("pipe", PipelineDesc), ("stream_id", C.c_int),
                ("left_rgb", C.c_void_p), ("right_rgb", C.c_void_p),
                ("left_lab", C.c_void_p), ("right_lab", C.c_void_p),
                ("dl", C.c_void_p), ("dr", C.c_void_p), ("mask", C.c_void_p)]


class RequestStats(C.Structure):
    _fields_ = [("submit_ms", C.c_double), ("start_ms", C.c_double), ("done_ms", C.c_double),
                ("footprint", C.c_size_t), ("status", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("budget_bytes", C.c_size_t), ("in_use_bytes", C.c_size_t), ("queued", C.c_int),
                ("running", C.c_int), ("resident", C.c_int), ("completed", C.c_uint64)]


_SIGS = {
    "rts_server_create": ([C.c_size_t, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "rts_server_destroy": ([C.c_void_p], C.c_int),
    "rts_server_submit": ([C.c_void_p, C.POINTER(RequestDesc), C.POINTER(C.c_uint64)], C.c_int),
    "rts_server_poll": ([C.c_void_p, C.c_uint64, C.POINTER(RequestStats)], C.c_int),
    "rts_server_wait": ([C.c_void_p, C.c_uint64, C.POINTER(RequestStats)], C.c_int),
    "rts_server_cancel": ([C.c_void_p, C.c_uint64], C.c_int),
    "rts_server_pause": ([C.c_void_p, C.c_int], C.c_int),
    "rts_server_set_oneoff_limit": ([C.c_void_p, C.c_int], C.c_int),
    "rts_server_close_stream": ([C.c_void_p, C.c_int], C.c_int),
    "rts_server_counters": ([C.c_void_p, C.POINTER(ServerStats)], C.c_int),
}
for _n, (_a, _r) in _SIGS.items():
    _f = getattr(lib, _n)
    _f.argtypes = _a
    _f.restype = _r


@dataclasses.dataclass
class Ticket:
    id: int
    kind: str
    shape: tuple
    dl: np.ndarray
    dr: np.ndarray
    mask: np.ndarray
    keep: tuple          # host inputs that must outlive the request
    stats: RequestStats | None = None
    status: int | None = None

    @property
    def latency_ms(self) -> float:
        return self.stats.done_ms - self.stats.submit_ms


class Server:
    def __init__(self, budget_bytes: int, slots: int = 8):
        p = C.c_void_p()
        call("rts_server_create", budget_bytes, slots, C.byref(p))
        self.handle = p.value

    def submit(self, desc: PipelineDesc, left: np.ndarray, right: np.ndarray, left_lab=None, right_lab=None,
               stream_id: int = -1, kind: str = "") -> Ticket:
        H, W = left.shape[:2]
        left, right = np.ascontiguousarray(left), np.ascontiguousarray(right)
        dl = np.zeros((H, W), np.uint8)
        dr = np.zeros((H, W), np.uint8)
        mask = np.zeros((H, W), np.uint8)
        req = RequestDesc()
        req.pipe = desc
        req.stream_id = stream_id
        req.left_rgb, req.right_rgb = left.ctypes.data, right.ctypes.data
        if left_lab is not None:
            left_lab = np.ascontiguousarray(left_lab)
            req.left_lab = left_lab.ctypes.data
        if right_lab is not None:
            right_lab = np.ascontiguousarray(right_lab)
            req.right_lab = right_lab.ctypes.data
        req.dl, req.dr, req.mask = dl.ctypes.data, dr.ctypes.data, mask.ctypes.data
        t = C.c_uint64()
        call("rts_server_submit", self.handle, C.byref(req), C.byref(t))
        return Ticket(t.value, kind, (W, H, desc.disp), dl, dr, mask, (left, right, left_lab, right_lab))

    def wait(self, ticket: Ticket) -> Ticket:
        st = RequestStats()
        ticket.status = lib.rts_server_wait(self.handle, ticket.id, C.byref(st))
        ticket.stats = st
        if ticket.status != L.RTS_OK:
            raise RtsError(ticket.status, f"request {ticket.id}")
        return ticket

    def cancel(self, ticket: Ticket) -> None:
        call("rts_server_cancel", self.handle, ticket.id)

    def set_oneoff_limit(self, limit: int) -> None:
        call("rts_server_set_oneoff_limit", self.handle, limit)

    def pause(self, paused: bool = True) -> None:
        call("rts_server_pause", self.handle, int(paused))

    def close_stream(self, stream_id: int) -> None:
        call("rts_server_close_stream", self.handle, stream_id)

    def counters(self) -> ServerStats:
        s = ServerStats()
        call("rts_server_counters", self.handle, C.byref(s))
        return s

    def close(self) -> None:
        if self.handle:
            lib.rts_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- workload

def camera_desc(W=1280, H=720, D=128) -> PipelineDesc:
    from .pipeline import preset
    return preset("census_cross", W, H, D)


def burst_desc(W, H, D, precision=L.RTS_AW_FP16) -> PipelineDesc:
    from .pipeline import preset
    return preset("aw", W, H, D, aw_precision=precision)


BURST_SIZES = ((640, 480), (1280, 720), (1920, 1080), (2560, 1440))
BURST_DMAX = (64, 128, 192, 256)


def make_inputs(cameras: int, frames_per_camera: int, burst_pool: int, seed: int = 100, burst_scale: float = 1.0):
    """Seeded frames: `frames_per_camera` distinct 720p/128 pairs per camera
    (cycled), and a pool of burst pairs with sizes and d_max drawn uniformly
    (optionally scaled down for tests) plus their Lab images."""
    from . import gen
    cams = [[gen.make_pair(1280, 720, 128, 40, 3.0, seed + 1000 * c + f) for f in range(frames_per_camera)]
            for c in range(cameras)]
    rng = np.random.default_rng(seed)
    bursts = []
    for i in range(burst_pool):
        w, h = BURST_SIZES[int(rng.integers(0, 4))]
        d = BURST_DMAX[int(rng.integers(0, 4))]
        w, h = max(64, int(w * burst_scale)) // 4 * 4, max(48, int(h * burst_scale))
        d = min(d, w)
        p = gen.make_pair(w, h, d, 30, 3.0, seed + 7 + i)
        bursts.append((p, p.lab()))
    return cams, bursts


def simulate(server: Server, cams, bursts, duration_s: float = 3.0, fps: float = 30.0, burst_rate: float = 40.0,
             seed: int = 0, aw_precision=L.RTS_AW_FP16):
    """Replays the arrival process in real time: every camera submits a frame
    each 1/fps (phases staggered), bursts arrive as a Poisson process.
    Returns the completed tickets in arrival order."""
    rng = np.random.default_rng(seed)
    events = []
    for c in range(len(cams)):
        t = c / (fps * len(cams))
        k = 0
        while t < duration_s:
            events.append((t, "camera", c, k))
            t += 1.0 / fps
            k += 1
    t = rng.exponential(1.0 / burst_rate)
    k = 0
    while t < duration_s:
        events.append((t, "burst", -1, k))
        t += rng.exponential(1.0 / burst_rate)
        k += 1
    events.sort(key=lambda e: e[0])
    tickets = []
    t0 = time.perf_counter()
    for (at, kind, cam, k) in events:
        dt = at - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        if kind == "camera":
            p = cams[cam][k % len(cams[cam])]
            tickets.append(server.submit(camera_desc(), p.left, p.right, stream_id=cam, kind="camera"))
        else:
            p, (la, lb) = bursts[int(rng.integers(0, len(bursts)))]
            W, H = p.left.shape[1], p.left.shape[0]
            tickets.append(server.submit(burst_desc(W, H, p.dmax, aw_precision), p.left, p.right, la, lb,
                                         kind="burst"))
    submit_wall = time.perf_counter() - t0
    for t in tickets:
        server.wait(t)
    total_wall = time.perf_counter() - t0
    return tickets, submit_wall, total_wall


def summarize(tickets, total_wall):
    out = {"requests": len(tickets), "wall_s": total_wall, "frames_per_s": len(tickets) / total_wall}
    for kind in ("camera", "burst"):
        lat = np.array([t.latency_ms for t in tickets if t.kind == kind])
        pd = sum(t.shape[0] * t.shape[1] * t.shape[2] for t in tickets if t.kind == kind)
        if len(lat):
            out[kind] = {"count": int(len(lat)), "p50_ms": float(np.percentile(lat, 50)),
                         "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
                         "pd_per_s": pd / total_wall}
    return out


def main(argv=None) -> int:
    """Serving measurement: cameras at a fixed frame rate plus Poisson
    adaptive-weight bursts, replayed in real time against one server.
    Prints one JSON object with throughput, latency percentiles per request
    kind and the agreement of served results with the same pairs run alone
    (cameras: identical maps; bursts: the FP16 form against the plain FP32
    per-frame path)."""
    import argparse
    import json
    from . import pipeline as PL
    ap = argparse.ArgumentParser()
    ap.add_argument("--cameras", type=int, default=8)
    ap.add_argument("--fps", type=float, default=30.0)
    ap.add_argument("--burst-rate", type=float, default=40.0)
    ap.add_argument("--duration", type=float, default=3.0)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--oneoff-limit", type=int, default=1, help="slots one-off bursts may hold at once")
    ap.add_argument("--frames", type=int, default=2, help="distinct frames per camera (cycled)")
    ap.add_argument("--burst-pool", type=int, default=6)
    ap.add_argument("--seed", type=int, default=100)
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    cams, bursts = make_inputs(args.cameras, args.frames, args.burst_pool, args.seed)
    gen_s = time.perf_counter() - t0
    srv = Server(int(args.budget_gb * 2**30), args.slots)
    srv.set_oneoff_limit(min(args.oneoff_limit, args.slots))
    # warm-up: one frame per camera and one of each burst, so that resident
    # pipelines exist and first-use costs stay out of the measurement
    warm = [srv.submit(camera_desc(), c[0].left, c[0].right, stream_id=i, kind="camera") for i, c in enumerate(cams)]
    for p, (la, lb) in bursts:
        W, H = p.left.shape[1], p.left.shape[0]
        warm.append(srv.submit(burst_desc(W, H, p.dmax), p.left, p.right, la, lb, kind="burst"))
    for t in warm:
        srv.wait(t)
    tickets, submit_wall, total_wall = simulate(srv, cams, bursts, args.duration, args.fps, args.burst_rate, args.seed)
    out = summarize(tickets, total_wall)
    out["offered"] = {"cameras": args.cameras, "fps": args.fps, "burst_rate": args.burst_rate,
                      "duration_s": args.duration, "submit_wall_s": submit_wall}
    # agreement with each pair run alone
    alone_cam, alone_burst = {}, {}
    cam_same, burst_agree = [], []
    for t in tickets:
        key = t.keep[0].ctypes.data
        if t.kind == "camera":
            if key not in alone_cam:
                alone_cam[key] = PL.run_host(camera_desc(), t.keep[0], t.keep[1])[0]
            cam_same.append(bool(np.array_equal(t.dl, alone_cam[key])))
        else:
            if key not in alone_burst:
                W, H, D = t.shape
                alone_burst[key] = PL.run_host(burst_desc(W, H, D, L.RTS_AW_FP32), t.keep[0], t.keep[1],
                                               t.keep[2], t.keep[3])[0]
            burst_agree.append(float((t.dl == alone_burst[key]).mean()))
    out["camera_identical_to_alone"] = float(np.mean(cam_same)) if cam_same else None
    out["burst_fp16_agreement_with_fp32_alone"] = {"mean": float(np.mean(burst_agree)),
                                                   "min": float(np.min(burst_agree))} if burst_agree else None
    c = srv.counters()
    out["server"] = {"budget_bytes": c.budget_bytes, "in_use_bytes": c.in_use_bytes, "resident": c.resident,
                     "slots": args.slots, "oneoff_limit": min(args.oneoff_limit, args.slots),
                     "gpu_max_hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))}
    out["input_generation_s"] = gen_s
    srv.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())

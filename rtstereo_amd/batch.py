"""Batched pairs of mixed size over the C ABI (rts_batch_*): the rank + LoG +
NCC cost with box aggregation, chunked under a device budget, the box stage
as one persistent work-queue launch per chunk; any other configuration
(adaptive weights, census + cross) through resident per-shape pipelines;
each chunk captured into a HIP graph and replayed."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import PipelineDesc, RtsError, call, lib
from .device import DeviceArray


class BatchItem(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("disp", C.c_int),
                ("left_rgb", C.c_void_p), ("right_rgb", C.c_void_p),
                ("dl", C.c_void_p), ("dr", C.c_void_p), ("mask", C.c_void_p)]


class BatchStats(C.Structure):
    _fields_ = [("pairs", C.c_int), ("chunks", C.c_int), ("graphs_built", C.c_int), ("graph_replays", C.c_int),
                ("box_items", C.c_ulonglong), ("arena_bytes", C.c_size_t), ("box_rows", C.c_int),
                ("box_cols", C.c_int), ("cost_disparities", C.c_int), ("tune_ms", C.c_float * 9)]


_SIGS = {
    "rts_batch_create": ([C.POINTER(PipelineDesc), C.c_size_t, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "rts_batch_destroy": ([C.c_void_p], C.c_int),
    "rts_batch_run": ([C.c_void_p, C.POINTER(BatchItem), C.c_int, C.c_void_p], C.c_int),
    "rts_batch_set_tiles": ([C.c_void_p, C.c_int, C.c_int, C.c_int], C.c_int),
    "rts_batch_get_stats": ([C.c_void_p, C.POINTER(BatchStats)], C.c_int),
    "rts_batch_pair_bytes": ([C.c_int, C.c_int, C.c_int], C.c_size_t),
}
for _n, (_a, _r) in _SIGS.items():
    _f = getattr(lib, _n)
    _f.argtypes = _a
    _f.restype = _r


def pair_bytes(W: int, H: int, D: int) -> int:
    """Device bytes of one pair in the rank + LoG + NCC batch path."""
    return int(lib.rts_batch_pair_bytes(W, H, D))


def pipeline_bytes(desc: PipelineDesc) -> int:
    """Device bytes of one resident pipeline of the batch's pipeline path
    (the other configurations): one per pair shape."""
    return int(lib.rts_pipeline_bytes(C.byref(desc)))


class Pair:
    """One pair's device buffers: colour images in, D_L / D_R / mask out."""

    def __init__(self, left: np.ndarray, right: np.ndarray, disp: int):
        H, W = left.shape[:2]
        self.shape = (W, H, disp)
        self.left, self.right = DeviceArray.from_host(left), DeviceArray.from_host(right)
        self.dl, self.dr, self.mask = (DeviceArray((H, W), np.uint8) for _ in range(3))

    def item(self) -> BatchItem:
        W, H, D = self.shape
        return BatchItem(W, H, D, self.left.ptr, self.right.ptr, self.dl.ptr, self.dr.ptr, self.mask.ptr)

    def results(self, stream=None):
        return self.dl.download(stream), self.dr.download(stream), self.mask.download(stream)

    def free(self):
        for a in (self.left, self.right, self.dl, self.dr, self.mask):
            a.free()


class Batch:
    def __init__(self, desc: PipelineDesc, budget_bytes: int, use_graphs: bool = True):
        p = C.c_void_p()
        call("rts_batch_create", C.byref(desc), budget_bytes, int(use_graphs), C.byref(p))
        self.handle = p.value

    def run(self, pairs, stream) -> None:
        arr = (BatchItem * len(pairs))(*[p.item() for p in pairs])
        call("rts_batch_run", self.handle, arr, len(pairs), stream)

    def set_tiles(self, box_cols: int, box_rows: int, cost_disparities: int) -> None:
        """Fix the tuned tile shapes (all three 0: tune again on the next run)."""
        call("rts_batch_set_tiles", self.handle, box_cols, box_rows, cost_disparities)

    def stats(self) -> BatchStats:
        st = BatchStats()
        call("rts_batch_get_stats", self.handle, C.byref(st))
        return st

    def close(self) -> None:
        if self.handle:
            lib.rts_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


CONFIGS = {
    "rank_log_ncc": ("rank_log_ncc", {}),
    "aw_fp16": ("aw", {"aw_precision": L.RTS_AW_FP16}),
    "aw_fp32": ("aw", {"aw_precision": L.RTS_AW_FP32}),
    "census_cross": ("census_cross", {}),
}


def main(argv=None) -> int:
    """Batch measurement: n seeded pairs of the uneven workload (sizes from
    640x480 to 2560x1440, d_max 64..256), inputs resident on the device.
    One run builds the graphs (and, for rank_log_ncc, tunes); the timed runs
    replay them.  The same pairs then run one at a time through resident
    per-shape pipelines of the baseline configuration (by default the same
    one; e.g. --config aw_fp16 --baseline aw_fp32 compares the batched FP16
    form with per-frame FP32).  Prints one JSON object."""
    import argparse
    import json
    import time
    from . import gen
    from . import pipeline as PL
    from .device import Event, Stream
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", choices=sorted(CONFIGS), default="rank_log_ncc")
    ap.add_argument("--baseline", choices=sorted(CONFIGS), default=None)
    ap.add_argument("--pairs", type=int, default=None, help="default 256 (32 for adaptive weights)")
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--seed", type=int, default=5)
    args = ap.parse_args(argv)
    npairs = args.pairs or (32 if args.config.startswith("aw") else 256)
    base_name = args.baseline or args.config
    shapes = gen.batch_shapes(npairs, args.seed)
    t0 = time.perf_counter()
    made = {}
    pairs = []
    for W, H, D in shapes:   # one generated pair per distinct shape, reused (generation is CPU-bound)
        if (W, H, D) not in made:
            made[(W, H, D)] = gen.make_pair(W, H, D, 20, 2.5, 1000 + len(made))
        p = made[(W, H, D)]
        pairs.append(Pair(p.left, p.right, D))
    gen_s = time.perf_counter() - t0
    pd = sum(W * H * D for W, H, D in shapes)
    s = Stream()
    name, over = CONFIGS[args.config]
    b = Batch(PL.preset(name, 64, 64, 16, **over), int(args.budget_gb * 2**30), not args.no_graphs)
    t0 = time.perf_counter()
    b.run(pairs, s.handle)
    s.sync()
    first_s = time.perf_counter() - t0
    e0, e1 = Event(), Event()
    ms = []
    for _ in range(args.runs):
        e0.record(s.handle)
        b.run(pairs, s.handle)
        e1.record(s.handle)
        s.sync()
        ms.append(e0.elapsed_ms(e1))
    st = b.stats()
    batched = [p.results() for p in pairs]
    b.close()
    bname, bover = CONFIGS[base_name]
    pipes = {}
    for W, H, D in shapes:
        if (W, H, D) not in pipes:
            pipes[(W, H, D)] = PL.Pipeline(PL.preset(bname, W, H, D, **bover))
    for p, key in zip(pairs[:1], shapes[:1]):   # warm-up
        pipes[key].run(p.left, p.right, stream=s.handle)
    s.sync()
    e0.record(s.handle)
    for p, key in zip(pairs, shapes):
        pipes[key].run(p.left, p.right, stream=s.handle)
    e1.record(s.handle)
    s.sync()
    single_ms = e0.elapsed_ms(e1)
    agree = []
    for p, key, (dl, _, _) in zip(pairs, shapes, batched):
        agree.append(float((pipes[key].results()[0] == dl).mean()))
    for q in pipes.values():
        q.close()
    out = {"config": args.config, "baseline": base_name, "pairs": npairs, "distinct_shapes": len(made), "pd": pd,
           "batch_ms": ms, "batch_pd_per_s": pd / (min(ms) / 1e3), "pairs_per_s": npairs / (min(ms) / 1e3),
           "mean_ms_per_pair": min(ms) / npairs,
           "baseline_single_pipelines_ms": single_ms, "baseline_pd_per_s": pd / (single_ms / 1e3),
           "speedup": single_ms / min(ms), "first_run_s": first_s,
           "disparity_agreement_batched_vs_baseline": {"mean": float(np.mean(agree)), "min": float(np.min(agree))},
           "chunks": st.chunks, "device_bytes": st.arena_bytes, "graphs_built": st.graphs_built,
           "graph_replays": st.graph_replays, "input_generation_s": gen_s, "graphs": not args.no_graphs}
    if args.config == "rank_log_ncc":
        out["tuned"] = {"box_rows": st.box_rows, "box_cols": st.box_cols, "cost_disparities": st.cost_disparities,
                        "tune_ms": [round(x, 3) for x in st.tune_ms]}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())

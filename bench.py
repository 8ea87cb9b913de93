"""Benchmark: pixel-disparity evaluations per second of the census + cross-based
stereo pipeline at 1920x1080, D=192, on one MI355X.

One step = one frame through every stage, from the two colour images resident
in HBM to the left and right disparity maps and the left-right mask:
gray -> census 9x7 (x2) -> cross arms L=17 tao=20 (x2) -> census Hamming + colour
AD combined with lambda 30/10 and the row prefix (one kernel) -> vertical cross
sweep -> WTA with the right view from the volume's diagonal and the
left-right check (tol 1).  One pixel-disparity evaluation (pd) is one (x, y, d)
of the left-to-right volume: W*H*D = 398,131,200 per frame.  Timed frames
replay the frame's launch sequence as one HIP graph (captured in the
warm-up); the per-stage split comes from a separate eager pass with HIP
events between the stages.

Prints one JSON line (progress lines go to stderr).  Options:
  --gpus N      only 1 is supported (one GPU per box)
  --steps K     timed frames (default 100)
  --warmup W    untimed frames (default 20)
  --cpu-seconds target length of the CPU baseline sample (default 12; 0 = skip)
  --no-aw, --no-configs, --no-batch-serve, --no-copy-probe  skip those parts

Beside the headline frame the line carries the adaptive-weight FP32/FP16
frames (`aw`), the 384x288 plain and 720p four-pass configurations
(`other_configs`), and, each measured by a child process on the same GPU,
the batched rank + LoG + NCC and adaptive-weight FP16 forms against
per-pair pipelines and a 2 s serving run (`batch_rank_log_ncc`,
`batch_aw_fp16_vs_fp32`, `serving`).

cpu_baseline is the reference's seq/ stage functions on one core;
cpu_baseline_all_cores runs the same pipeline on independent row strips, one
per host thread (up to 16).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pixel·disparity evals/s at 1080p d=192; p50/p99 frame latency vs plain"
W, H, D = 1920, 1080, 192
PEAK_HBM_GBS = 8000.0

# Algorithmic HBM bytes per pd of each stage's kernel (float volumes, s = 4):
# the fused cost + row-prefix kernel writes the prefix volume once; the
# vertical sweep reads it and writes the aggregated volume; WTA reads that.
STAGE_BYTES_PER_PD = {"cost": 4.0, "aggr": 8.0, "wta": 4.0}
# The reference dataflow per pd at one H/V iteration: cost write 4, H pass
# read + write 8, V pass read + write 8, WTA read 4.
FRAME_CONTRACT_BYTES_PER_PD = 24.0
STAGE_KERNEL = {"cost": "k_front_hscan", "aggr": "k_cross_sweep", "wta": "k_wta"}
# The frame as built moves 16 B/pd: prefix volume written (4) and read (4),
# aggregated volume written (4) and read (4).
FRAME_BYTES_PER_PD = 16.0
# HBM bytes per dispatch from rocprofv3 FETCH_SIZE / WRITE_SIZE (KiB), scaled
# by the copy kernel's reading of the same counters (FETCH x2.000, WRITE
# x1.000).  Not measured in this process (counters need the profiler): the
# figures of the profile run recorded in profiles/census_cross_1080p_d192.md,
# re-taken whenever one of these kernels changes.
STAGE_TRAFFIC = {"cost": (24430 * 2 + 1555213) * 1024, "aggr": (1212000 * 2 + 1555200) * 1024,
                 "wta": (804500 * 2 + 4050) * 1024}
TRAFFIC_SOURCE = "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, profiles/census_cross_1080p_d192.md"
# Adaptive weights: 1089 taps per pd; 4 flops per tap (weight product, the
# denominator's add, the numerator's multiply-add); the VALU issue count per
# tap without packed FP32 is 3 (FP32: product, add, fused multiply-add) or
# 1 (FP16: one packed product and one dot2 per tap pair for the numerator;
# the denominators run on the matrix cores); one VALU instruction is 64
# lanes, 256 CUs x 4 SIMDs issue one every 4 cycles at 2.4 GHz.
AW_TAPS = 33 * 33
AW_FLOP_PER_TAP = 4.0
AW_VALU_PER_TAP = {"fp32": 3.0, "fp16": 1.0}
PEAK_VECTOR_TFLOPS = 157.3
VALU_LANE_OPS_PER_S = 256 * 4 * 16 * 2.4e9


def cpu_model() -> str:
    """The host CPU's model name (from /proc/cpuinfo), for the baseline record."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(target_s: float) -> dict:
    """The reference's own seq/ stage functions (compiled from its sources
    into oracle/_ref) on a full-width strip of the same frame, one core;
    falls back to the C restatement of the oracle when the reference build
    is absent."""
    import oracle
    from rtstereo_amd import gen

    p = gen.census_cross_1080p()
    ref = oracle.ref_lib()
    kind = "reference" if ref is not None else "port"

    def run(y0: int, rows: int) -> float:
        L = np.ascontiguousarray(p.left[y0:y0 + rows])
        R = np.ascontiguousarray(p.right[y0:y0 + rows])
        gl, gr = gen.to_gray(L), gen.to_gray(R)
        t = time.perf_counter()
        if ref is not None:
            c1 = ref.hamming(ref.census(gl, 4, 3), ref.census(gr, 4, 3), 4, 3, D)
            c2 = ref.ad("pixelBasedCostL2R_Color", L, R, D, 0)
            c = ref.combine(c1, c2, 30.0, 10.0)
            hl, vl = ref.find_cross(L)
            hr, vr = ref.find_cross(R)
            A = ref.cross_v(ref.hprefix(c), hl, hr, vl, vr)
            ref.wta(A)
        else:
            c1 = oracle.hamming(oracle.census(gl, 4, 3), oracle.census(gr, 4, 3), 4, 3, D)
            c = oracle.combine(c1, oracle.ad_color(L, R, D, 1), 30.0, 10.0)
            hl, vl = oracle.find_cross(L)
            hr, vr = oracle.find_cross(R)
            oracle.wta(oracle.cross_v(oracle.hprefix(c), hl, hr, vl, vr))
        return time.perf_counter() - t

    t_cal = run(500, 48)
    rows = int(max(48, min(H - 500, 48 * target_s / max(t_cal, 1e-3))))
    t = run(500, rows)
    pd = W * rows * D
    single = {"value": pd / t, "unit": "pd/s", "cores": 1, "kind": kind,
              "sample": f"census+AD combine, cross L17/tao20, WTA on a {W}x{rows} strip (rows 500..{500 + rows - 1}) "
                        f"of the 1080p frame, D={D}: {pd} pd in {t:.2f} s, "
                        f"{'seq/ sources compiled with g++ -O2' if kind == 'reference' else 'C oracle'}, single thread",
              "frame_seconds_extrapolated": W * H * D / (pd / t), "cpu": cpu_model(),
              "host_cores": os.cpu_count()}
    # All host cores (at most 16, the share of one GPU): independent
    # 48-row strips of the same frame, one per thread (the stage functions
    # are reentrant and the foreign calls release the interpreter lock).
    from concurrent.futures import ThreadPoolExecutor
    n = max(1, min(16, os.cpu_count() or 1))
    per = max(48, int(rows / n))
    starts = [(500 + 48 * i) % (H - per) for i in range(n)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(n) as ex:
        list(ex.map(lambda y0: run(y0, per), starts))
    tw = time.perf_counter() - t0
    pdn = W * per * D * n
    parallel = {"value": pdn / tw, "unit": "pd/s", "cores": n, "kind": kind,
                "sample": f"{n} threads, each a {W}x{per} strip of the same pipeline, {pdn} pd in {tw:.2f} s wall",
                "cpu": cpu_model()}
    return single, parallel


def aw_timing(stream, rows=None) -> dict:
    """Adaptive-weight pipeline (truncated AD 40, 33x33 window, gc 7, gp 36,
    WTA + check) at 1920x1080, D=192, FP32 and FP16: mean of 3 frames after
    one warm-up, Lab derived on the device.  With `rows`, each precision's
    aggregated volume on those rows is kept (key "rows") for the accuracy
    check."""
    from rtstereo_amd import gen, pipeline as PL, _lib as L
    from rtstereo_amd.device import DeviceArray, Event
    p = gen.aw_1080p()
    a, b = DeviceArray.from_host(p.left), DeviceArray.from_host(p.right)
    res = {"workload": f"AD trunc 40 + adaptive weights 33x33 (gc 7, gp 36) + WTA/check, {W}x{H}, D={D}"}
    for name, prec in (("fp32", L.RTS_AW_FP32), ("fp16", L.RTS_AW_FP16)):
        pipe = PL.Pipeline(PL.preset("aw", W, H, D, aw_precision=prec))
        pipe.enable_timing(True)
        pipe.run(a, b, stream=stream.handle)
        e0, e1 = Event(), Event()
        e0.record(stream.handle)
        agg = []
        for _ in range(3):
            pipe.run(a, b, stream=stream.handle)
            agg.append(pipe.stage_ms()["aggr"])
        e1.record(stream.handle)
        stream.sync()
        frame = e0.elapsed_ms(e1) / 3
        aggr = float(np.mean(agg))
        taps = W * H * D * AW_TAPS
        flops = taps * AW_FLOP_PER_TAP
        valu_ms = taps * AW_VALU_PER_TAP[name] / VALU_LANE_OPS_PER_S * 1e3
        res[name] = {"frame_ms": frame, "aggregation_ms": aggr, "pd_per_s": W * H * D / (frame / 1e3),
                     "roofline": {"bound": "valu", "flop_per_tap": AW_FLOP_PER_TAP, "taps_per_pd": AW_TAPS,
                                  "algorithmic_flops": flops, "achieved": flops / (aggr / 1e3) / 1e12,
                                  "peak": PEAK_VECTOR_TFLOPS, "unit": "TFLOP/s",
                                  "frac": flops / (aggr / 1e3) / 1e12 / PEAK_VECTOR_TFLOPS,
                                  "valu_instr_per_tap": AW_VALU_PER_TAP[name], "valu_bound_ms": valu_ms,
                                  "frac_of_valu_bound": valu_ms / aggr}}
        if rows is not None:
            # the accuracy run takes the host Lab images the FP64 oracle
            # uses, so that the comparison measures the aggregation alone
            # (the device's FP32 Lab conversion may differ by one level)
            hl = p.lab()
            la, lb = DeviceArray.from_host(hl[0]), DeviceArray.from_host(hl[1])
            pipe.run(a, b, la, lb, stream=stream.handle)
            stream.sync()
            la.free()
            lb.free()
            vol = pipe.volume()
            res[name]["rows"] = np.moveaxis(vol[:, rows, :], 0, -1).astype(np.float64)
            del vol
        pipe.close()
    a.free()
    b.free()
    return res


AW_ROWS = sorted(set([0, 1, 57, 58, 103, 104, 115, 116, 207, 208, 1078, 1079] +
                     list(np.linspace(2, 1077, 40).astype(int))))


def aw_accuracy(aw: dict) -> dict:
    """The AW volumes of AW_ROWS (52 whole rows, 99,840 pixels x 192
    disparities) against the FP64 oracle, as the checker (CPU; part of the
    CPU-baseline leg): FP32 max relative error, FP16 max / mean absolute
    error, and the fraction of pixels whose winning disparity differs."""
    import oracle
    from rtstereo_amd import gen
    p = gen.aw_1080p()
    ll, lr = p.lab()
    cost = oracle.ad_color(p.left, p.right, D, 1, 40.0, False)
    f64 = oracle.aw_rows_f64(ll, lr, cost, np.array(AW_ROWS), 16, 7.0, 36.0)
    d64 = np.argmin(f64, -1)
    out = {"sample": f"{len(AW_ROWS)} whole rows ({len(AW_ROWS) * W} pixels x {D} disparities) against the FP64 oracle"}
    for name in ("fp32", "fp16"):
        got = aw[name].pop("rows")
        err = np.abs(got - f64)
        out[name] = {"max_abs_err": float(err.max()), "mean_abs_err": float(err.mean()),
                     "max_rel_err": float(np.max(err / np.abs(f64))),
                     "disparity_disagreement": float((np.argmin(got, -1) != d64).mean())}
    return out


def config_timing(stream, frames: int = 50) -> dict:
    """The two other single-pair workloads: 384x288 plain (AD trunc 20, 9x9
    box, D=16) and 1280x720 census + cross-based aggregation with four H/V
    passes (D=128).  Mean frame time over `frames` graph-replayed frames
    after two warm-up frames, and agreement with the generator's ground
    truth."""
    from rtstereo_amd import gen, pipeline as PL
    from rtstereo_amd.device import DeviceArray, Event
    res = {}
    for name, pair, preset, (Wc, Hc, Dc), over in (
            ("plain_384x288_d16", gen.plain_384x288(), "plain", (384, 288, 16), {}),
            ("census_cross_720p_d128_4pass", gen.census_cross_720p(), "census_cross", (1280, 720, 128),
             {"cross_iters": 4})):
        a, b = DeviceArray.from_host(pair.left), DeviceArray.from_host(pair.right)
        pipe = PL.Pipeline(PL.preset(preset, Wc, Hc, Dc, **over))
        pipe.enable_graphs(True)
        for _ in range(2):
            pipe.run(a, b, stream=stream.handle)
        stream.sync()
        agree = float((pipe.dl.download(stream.handle) == pair.gt).mean())
        e0, e1 = Event(), Event()
        e0.record(stream.handle)
        for _ in range(frames):
            pipe.run(a, b, stream=stream.handle)
        e1.record(stream.handle)
        stream.sync()
        ms = e0.elapsed_ms(e1) / frames
        res[name] = {"frame_ms": ms, "pd_per_s": Wc * Hc * Dc / (ms / 1e3), "gt_agreement": agree}
        pipe.close()
        a.free()
        b.free()
    return res


def child_json(module_args, timeout_s):
    """Runs `python3 -m <module_args>` as a child process (its own HIP
    context and environment) and returns the JSON object it prints, or an
    error record; the bench line never fails for it."""
    import subprocess
    t0 = time.perf_counter()
    try:
        p = subprocess.run([sys.executable, "-m", *module_args], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, timeout=timeout_s, cwd=os.path.dirname(os.path.abspath(__file__)))
    except subprocess.TimeoutExpired:
        return {"error": f"timed out after {timeout_s} s", "args": module_args}
    if p.returncode != 0:
        return {"error": f"exit status {p.returncode}", "stderr_tail": p.stderr[-500:], "args": module_args}
    try:
        out = json.loads(p.stdout.strip().splitlines()[-1])
    except (ValueError, IndexError):
        return {"error": "no JSON line", "args": module_args}
    out["command"] = "python3 -m " + " ".join(module_args)
    out["wall_s"] = time.perf_counter() - t0
    return out


def batch_and_serving():
    """The batched and served forms, each measured in a child process on the
    same GPU: 64 mixed rank + LoG + NCC pairs (640x480..2560x1440, d_max
    64..256) against the same pairs through per-shape pipelines; 16 mixed
    adaptive-weight FP16 pairs against the per-frame FP32 pipelines; and 2 s
    of serving (8 cameras x 30 fps of 720p/128 census + cross, 40 AW FP16
    bursts/s)."""
    res = {}
    progress("batch: 64 rank + LoG + NCC pairs")
    b = child_json(["rtstereo_amd.batch", "--pairs", "64", "--runs", "3"], 300)
    res["batch_rank_log_ncc"] = {k: b[k] for k in ("pairs", "pd", "batch_ms", "batch_pd_per_s", "pairs_per_s",
                                                   "baseline_single_pipelines_ms", "speedup",
                                                   "disparity_agreement_batched_vs_baseline", "tuned", "command",
                                                   "error") if k in b}
    progress("batch: 16 adaptive-weight FP16 pairs against per-frame FP32")
    b = child_json(["rtstereo_amd.batch", "--config", "aw_fp16", "--baseline", "aw_fp32", "--pairs", "16", "--runs",
                    "2"], 300)
    res["batch_aw_fp16_vs_fp32"] = {k: b[k] for k in ("pairs", "pd", "batch_ms", "batch_pd_per_s", "mean_ms_per_pair",
                                                      "baseline_single_pipelines_ms", "baseline_pd_per_s", "speedup",
                                                      "disparity_agreement_batched_vs_baseline", "command", "error")
                                    if k in b}
    progress("serving: 2 s at 40 bursts/s")
    v = child_json(["rtstereo_amd.serve", "--duration", "2", "--burst-rate", "40"], 300)
    res["serving"] = {k: v[k] for k in ("camera", "burst", "offered", "camera_identical_to_alone",
                                        "burst_fp16_agreement_with_fp32_alone", "command", "error") if k in v}
    return res


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-copy-probe", action="store_true")
    ap.add_argument("--no-aw", action="store_true", help="skip the adaptive-weight FP32/FP16 timing")
    ap.add_argument("--no-configs", action="store_true", help="skip the 384x288 plain and 720p 4-pass timings")
    ap.add_argument("--aw-only", action="store_true", help="print only the adaptive-weight FP32/FP16 timing")
    ap.add_argument("--no-batch-serve", action="store_true",
                    help="skip the batched (rank + LoG + NCC, AW FP16) and serving measurements")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (each box has one GPU, "
              "and frames shard across GPUs with no exchange)", file=sys.stderr)
        return 2

    from rtstereo_amd import gen, pipeline as PL
    from rtstereo_amd.device import DeviceArray, Event, Stream, gpu_count
    from rtstereo_amd._lib import call

    if gpu_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    if args.aw_only:
        print(json.dumps({"aw": aw_timing(Stream())}))
        return 0
    pair = gen.census_cross_1080p()
    stream = Stream()
    left, right = DeviceArray.from_host(pair.left), DeviceArray.from_host(pair.right)
    desc = PL.preset("census_cross", W, H, D)
    pipe = PL.Pipeline(desc)
    # a frame's launches are replayed as one HIP graph (captured during the
    # warm-up: the first run is eager, the second captures)
    pipe.enable_graphs(True)

    for _ in range(max(args.warmup, 2)):
        pipe.run(left, right, stream=stream.handle)
    stream.sync()
    # correctness guard on the measured configuration
    dl = pipe.dl.download(stream.handle)
    gt_agree = float((dl == pair.gt).mean())

    evs = [Event() for _ in range(args.steps + 1)]
    stage_acc: dict = {}
    call("rts_device_sync")
    t0 = time.perf_counter()
    evs[0].record(stream.handle)
    for i in range(args.steps):
        pipe.run(left, right, stream=stream.handle)
        evs[i + 1].record(stream.handle)
    stream.sync()
    wall = time.perf_counter() - t0
    call("rts_device_sync")
    frame_ms = np.array([evs[i].elapsed_ms(evs[i + 1]) for i in range(args.steps)])
    total_ms = evs[0].elapsed_ms(evs[-1])

    # per-stage durations on the same stream (separate pass so the timed
    # region above holds only pipeline work; timed runs launch eagerly)
    pipe.enable_timing(True)
    for _ in range(5):
        pipe.run(left, right, stream=stream.handle)
        for k, v in pipe.stage_ms().items():
            stage_acc.setdefault(k, []).append(v)
    stage = {k: float(np.mean(v)) for k, v in stage_acc.items()}

    pd = W * H * D
    value = pd * args.steps / (total_ms / 1e3)
    dom = max(STAGE_BYTES_PER_PD, key=lambda k: stage.get(k, 0.0))
    bytes_launch = STAGE_BYTES_PER_PD[dom] * pd
    achieved = bytes_launch / (stage[dom] / 1e3) / 1e9

    copy = None
    probes = None
    if not args.no_copy_probe:
        # 2 GiB copy, read and write probes (16-byte accesses, 8 in flight
        # per lane): the achievable rates the frame's bytes are held to
        n = 1 << 31
        a, b = DeviceArray((n,), np.uint8), DeviceArray((n,), np.uint8)
        sink = DeviceArray((4096,), np.uint32)

        def rate(fn, bytes_moved):
            fn()
            e0, e1 = Event(), Event()
            e0.record(stream.handle)
            for _ in range(10):
                fn()
            e1.record(stream.handle)
            stream.sync()
            return bytes_moved * 10 / (e0.elapsed_ms(e1) / 1e3) / 1e9

        copy = rate(lambda: call("rts_copy_kernel", a.ptr, b.ptr, n, stream.handle), 2 * n)
        probes = {"copy": copy,
                  "read": rate(lambda: call("rts_read_kernel", a.ptr, n, sink.ptr, stream.handle), n),
                  "write": rate(lambda: call("rts_write_kernel", b.ptr, n, stream.handle), n)}
        a.free()
        b.free()
        sink.free()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pd/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator: blurred noise + ramps, 60 rectangles, noise sigma 3)",
        "config": {"workload": "census9x7+colour AD combined (lambda 30/10) -> cross-based aggregation "
                               "(L 17, tao 20, one H/V pass) -> WTA + right view + L/R check (tol 1), "
                               f"{W}x{H}, D={D}",
                   "arithmetic": "bit-identical to the reference's float order", "width": W, "height": H, "disparities": D, "pd_per_frame": pd},
        "frames_per_s": 1e3 / (total_ms / args.steps),
        "latency_ms": {"p50": float(np.percentile(frame_ms, 50)), "p99": float(np.percentile(frame_ms, 99)),
                       "max": float(frame_ms.max())},
        "wall_s": wall,
        "stage_ms": stage,
        "gt_agreement": gt_agree,
        "roofline": {"bound": "hbm", "kernel": STAGE_KERNEL[dom], "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": STAGE_TRAFFIC[dom],
                     "bytes_per_pd": STAGE_BYTES_PER_PD[dom], "algorithmic_bytes_per_launch": bytes_launch,
                     "kernel_ms": stage[dom], "traffic_source": TRAFFIC_SOURCE},
        # The whole frame on the bytes it moves (16 B/pd: prefix volume
        # written and read, aggregated volume written and read).
        "roofline_frame": {"bytes_per_pd": FRAME_BYTES_PER_PD,
                           "achieved": FRAME_BYTES_PER_PD * pd / (total_ms / args.steps / 1e3) / 1e9,
                           "peak": PEAK_HBM_GBS, "unit": "GB/s",
                           "frac": FRAME_BYTES_PER_PD * pd / (total_ms / args.steps / 1e3) / 1e9 / PEAK_HBM_GBS},
        # The whole frame against the reference dataflow's compulsory HBM
        # bytes (cost volume write, H and V pass read + write, WTA read:
        # 24 B/pd at one iteration).  The fused front end never moves the
        # cost volume, so this is an equivalent rate, not measured traffic.
        "roofline_frame_equivalent": {"bytes_per_pd": FRAME_CONTRACT_BYTES_PER_PD,
                                      "achieved": FRAME_CONTRACT_BYTES_PER_PD * pd / (total_ms / args.steps / 1e3) / 1e9,
                                      "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                      "frac": FRAME_CONTRACT_BYTES_PER_PD * pd / (total_ms / args.steps / 1e3) / 1e9
                                      / PEAK_HBM_GBS},
        "copy_kernel_gbs": copy,
        "bandwidth_probes_gbs": probes,
    }
    if probes:
        # The frame's floor on this GPU's measured rates: it reads 8 B/pd
        # (prefix and aggregated volumes) and writes 8 B/pd (the same two).
        floor_ms = (8.0 * pd / (probes["read"] * 1e9) + 8.0 * pd / (probes["write"] * 1e9)) * 1e3
        out["roofline_frame_measured"] = {"read_bytes_per_pd": 8.0, "write_bytes_per_pd": 8.0,
                                          "read_gbs": probes["read"], "write_gbs": probes["write"],
                                          "floor_ms": floor_ms, "frac": floor_ms / (total_ms / args.steps)}
    if not args.no_aw:
        progress("adaptive weights FP32 / FP16")
        out["aw"] = aw_timing(stream, AW_ROWS if args.cpu_seconds > 0 else None)
    if not args.no_configs:
        progress("other configurations")
        out["other_configs"] = config_timing(stream)
    if not args.no_batch_serve:
        out.update(batch_and_serving())
    if args.cpu_seconds > 0:
        progress("CPU baseline")
        cb, cbn = cpu_baseline(args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["cpu_baseline_all_cores"] = cbn
        out["speedup_vs_cpu"] = value / cb["value"]
        if "aw" in out:
            out["aw"]["accuracy"] = aw_accuracy(out["aw"])
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

#!/usr/bin/env python3
"""SIRT slice-iterations/s at 1024 x 1024 x 720 angles x 1448 bins on one MI355X.

Workload (synthetic, deterministic): 20 random ellipses per slice (seeded
centres in the unit disk, semi-axes U(0.02, 0.3), angles U(0, pi), additive
attenuations U(-0.2, 0.5), clamped >= 0) on a 1024 x 1024 grid, forward
projected with the Joseph projector at 720 angles over [0, pi) onto 1448 bins,
plus Gaussian noise of sigma = 1% of the sinogram maximum.

A step is one SIRT iteration (relaxation 0.25) over a batch of --batch such
slices resident in HBM: sinograms (measured data and residual) in FP16, images
in FP32 (--storage f16-sino; f16 also stores the images in FP16), FP32
interpolation and accumulation: one forward- and one back-projection launch
over every slice, replayed from a captured hipGraph.  The plain FP32 path (the SIRT task, one
slice per process() call, reference-compatible arithmetic) is measured beside
it, and accuracy is reported against the FP64 evaluation of the same
operators.

Serving leg (default run): the reconstruction server (continuous batching)
takes SIRT requests of the same shape (20 iterations each) arriving as a
Poisson process at --serve-qps; p50 / p99 request latency (submission to
result on the host) and the served slice-iterations/s are reported beside the
batched throughput.  --serve-c5 runs the split-Bregman TV serving workload
instead (image edge N in {512, 1024, 2048}, angles U{90..900}, bins
ceil(N sqrt 2), lambda ~ logU(1e-3, 1e-1), mu = 1, 30 outer x 5 inner sweeps,
FP16 sinograms with FP32 images and math) and prints its own JSON line.

Ray-weight floor: --ray-weight-min (default 0, the reference's inv: every
ray with FP(1) > 0 is weighted 1 / FP(1)).  With a floor, rays that cross
less than that fraction of a pixel (image-corner clips) get weight 0, and the
FP64 reference uses the same floor.  The accuracy leg reports the batched
path against the FP64 SIRT both without the floor and with 0.1, and the line
carries an explicit gate: rel-L2 <= 1e-3 against the FP64 SIRT of the
headline's own setting.

--line measures SIRT with the line projector (the projector's model "line":
exact ray-pixel chord lengths) on one slice of the same shape and prints its
own JSON line.

Prints one JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
import argparse
import json
import math
import os
import queue
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ufo-ir_amd"))

METRIC = "slice-iters/s SIRT 1024²×720 ang, batched vs plain @relL2≤1e-3; p99 latency"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# LDS read rate of the projectors' 8-byte reads, one instruction each:
# 139.5 bytes per clock per CU measured (uir_probe.hip's read loop) x 256 CUs x
# 2.4 GHz.  (16-byte reads reach 216 only when the 16 lanes of a pass read 16
# consecutive entries; at the projectors' spaced addresses they take twice
# the clocks of an 8-byte read.)
LDS_PEAK_TBS = 139.5 * 256 * 2.4e9 / 1e12


def ellipses(seed, count=20):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        r = np.sqrt(rng.random())
        ang = 2 * np.pi * rng.random()
        out.append((r * np.cos(ang), r * np.sin(ang), rng.uniform(0.02, 0.3), rng.uniform(0.02, 0.3),
                    rng.uniform(0, np.pi), rng.uniform(-0.2, 0.5)))
    return out


def phantom(n, seed):
    c = (np.arange(n) + 0.5 - n / 2.0) / (n / 2.0)
    X, Y = np.meshgrid(c, -c)
    img = np.zeros((n, n), np.float32)
    for (x0, y0, a, b, t, v) in ellipses(seed):
        xr = (X - x0) * np.cos(t) + (Y - y0) * np.sin(t)
        yr = -(X - x0) * np.sin(t) + (Y - y0) * np.cos(t)
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += np.float32(v)
    np.maximum(img, 0, out=img)
    return img


def percentile(v, q):
    return float(np.percentile(np.asarray(v), q))


RAY_WEIGHT_MIN = 0.1


def run_serving(uir, storage, make, qps, n_requests, seed):
    """Poisson arrivals at `qps` of requests make(k) -> (request, sinogram);
    returns latency percentiles and served slice-iterations/s."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, n_requests)
    work = [make(k) for k in range(n_requests)]
    tickets = queue.Queue()
    with uir.Server(0, storage, hbm_reserve=4 << 30) as srv:
        def producer():
            t_next = time.perf_counter()
            for k in range(n_requests):
                t_next += gaps[k]
                dt = t_next - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                tickets.put(srv.submit(*work[k]))
            tickets.put(None)

        th = threading.Thread(target=producer)
        th.start()
        stats = []
        while True:
            t = tickets.get()
            if t is None:
                break
            stats.append(srv.wait(t, timeout=600)[1])
        th.join()
        info = srv.info()
    lat = [(s["t_done"] - s["t_submit"]) * 1e3 for s in stats]
    t0 = min(s["t_submit"] for s in stats)
    t1 = max(s["t_done"] for s in stats)
    return {"qps": qps, "requests": n_requests, "duration_s": round(t1 - t0, 2),
            "latency_ms": {"p50": round(percentile(lat, 50), 2), "p99": round(percentile(lat, 99), 2),
                           "mean": round(float(np.mean(lat)), 2), "max": round(max(lat), 2)},
            "slice_iters_per_s": round(info["slice_iterations"] / (t1 - t0), 1),
            "mean_batch": round(info["slice_iterations"] / max(info["steps"], 1), 2),
            "max_batch": max(s["max_batch"] for s in stats),
            "steps": info["steps"]}


def serve_c5(args, uir):
    """Split-Bregman TV serving at a fixed Poisson rate (one JSON line)."""
    rng = np.random.default_rng(20261205)
    sizes = [512, 1024, 2048]
    pool = {}
    for n in sizes:
        d = math.ceil(n * math.sqrt(2))
        # content does not change the work: smooth synthetic sinograms, one per size
        t = np.linspace(-1, 1, d, dtype=np.float32)
        th = np.linspace(0, np.pi, 900, dtype=np.float32)[:, None]
        pool[n] = (np.maximum(0.0, 1 - (t[None, :] - 0.3 * np.cos(th)) ** 2) * n / 4).astype(np.float32)

    def make(k):
        n = sizes[rng.integers(0, 3)]
        a = int(rng.integers(90, 901))
        d = math.ceil(n * math.sqrt(2))
        lam = float(10 ** rng.uniform(-3, -1))
        req = uir.make_request(n, d, a, float(np.float32(np.pi / a)), "sbtv-sirt", iterations=30,
                               inner_iterations=5, relaxation=1.0, lam=lam, mu=1.0,
                               ray_weight_min=RAY_WEIGHT_MIN)
        return req, np.ascontiguousarray(pool[n][:a])

    storage = args.c5_storage
    out = run_serving(uir, storage, make, args.serve_qps_c5, args.serve_requests_c5, 20261206)
    print("c5: served %d requests" % args.serve_requests_c5, file=sys.stderr, flush=True)

    # accuracy: requests of the workload's kind (full 30 x 5 sweeps) served
    # together, against the FP64 oracle of the method
    import oracle as O
    checks = [(512, 200, 0.01), (512, 400, 0.05), (1024, 120, 0.002)]
    reqs = []
    for (n, a, lam) in checks:
        d = math.ceil(n * math.sqrt(2))
        reqs.append((uir.make_request(n, d, a, float(np.float32(np.pi / a)), "sbtv-sirt", iterations=30,
                                      inner_iterations=5, relaxation=1.0, lam=lam, mu=1.0,
                                      ray_weight_min=RAY_WEIGHT_MIN), np.ascontiguousarray(pool[n][:a])))
    with uir.Server(0, storage) as srv:
        tickets = [srv.submit(r, sn) for r, sn in reqs]
        imgs = [srv.wait(t, timeout=600)[0] for t in tickets]
    acc = []
    for (n, a, lam), (r, sn), img in zip(checks, reqs, imgs):
        d = math.ceil(n * math.sqrt(2))
        g = O.Geometry(n, d, a, np.float32(np.pi / a))
        ref = O.sbsirt64(g, sn, 30, 5, relax=1.0, lam=lam, mu=1.0, ray_weight_min=RAY_WEIGHT_MIN)
        acc.append({"n": n, "angles": a, "lambda": lam, "rel_l2_vs_fp64": O.rel_l2(img, ref),
                    "psnr_db": round(O.psnr(img, ref), 2)})
        print("c5: accuracy N %d, %d angles: %.3e" % (n, a, acc[-1]["rel_l2_vs_fp64"]), file=sys.stderr, flush=True)
    line = {"metric": "split-Bregman TV serving: p50/p99 latency and slice-iters/s", "value": out["slice_iters_per_s"],
            "unit": "slice-iters/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": {"f16": "f16 storage, f32 math", "f16-sino": "f16 sinograms, f32 images, f32 math"}.get(
                storage, storage),
            "data": "synthetic smooth sinograms (content does not change the work)",
            "config": {"workload": "SBTV_SIRT requests, N in {512, 1024, 2048}, angles U{90..900}, bins ceil(N sqrt 2), "
                                   "lambda logU(1e-3, 1e-1), mu 1, 30 outer x 5 inner sweeps, Poisson arrivals",
                       "qps": args.serve_qps_c5, "requests": args.serve_requests_c5, "storage": storage},
            "serving": out,
            "accuracy": {"requests": acc, "reference": "FP64 SBTV_SIRT oracle (orc64_sbsirt), same sweeps, "
                                                       "ray_weight_min %g" % RAY_WEIGHT_MIN}}
    print(json.dumps(line), flush=True)


def sart_c2(args, uir):
    """Plain SART on one 1024 x 1024 slice, 720 angles, 1448 bins, --c2-iters
    iterations (30), relaxation 0.5 (the reference-compatible SART task, FP32).
    Beside the time: accuracy against the oracle's FP32 SART (the reference's
    arithmetic) and its FP64 evaluation, the roofline of the per-angle
    projection kernels (single-angle launches through the kernel-level ABI,
    the kernels the sweep runs), and the oracle's SART on the host cores."""
    import oracle as O
    res = uir.Resources(0)
    n, a, d = 1024, 720, 1448
    step = float(np.float32(np.pi / a))
    proj = uir.Task("parallel-projector", step=step, is_forward=True, angles_num=a, image_size=n,
                    detectors_num=d).setup(res)
    img = uir.DeviceArray.from_numpy(phantom(n, 20261202))
    s = uir.DeviceArray((a, d))
    proj.process(img, s)
    sino = s.numpy().astype(np.float64)
    sino += np.random.default_rng(20261202).normal(0.0, 0.01 * sino.max(), sino.shape)
    sino32 = sino.astype(np.float32)
    ds = uir.DeviceArray.from_numpy(sino32)
    out = uir.DeviceArray((n, n))
    pj = uir.Task("parallel-projector", step=step, is_forward=False, image_size=n)
    iters = args.c2_iters
    sart = uir.Task("sart", projector=pj, relaxation_factor=0.5, num_iterations=iters).setup(res)
    sart.process(ds, out)                      # warm-up (captures the sweep graph)
    res.synchronize()
    timer = uir.Timer()
    timer.start(res.stream)
    sart.process(ds, out)
    timer.stop(res.stream)
    ms = timer.elapsed_ms()
    got = out.numpy()
    print("c2: %d SART iterations in %.1f ms" % (iters, ms), file=sys.stderr, flush=True)

    # per-launch durations of the sweep's single-angle FP and BP (HIP events on
    # the task's stream around each launch of one sweep)
    g = O.Geometry(n, d, a, np.float32(step))
    fp_us, bp_us = sart.time_kernels(ds, out, stride=1)    # after the result was read
    # algorithmic bytes per launch: FP reads the image and the angle's row and
    # writes the row; BP reads the row and the image and writes the image
    fp_bytes = n * n * 4 + 2 * d * 4
    bp_bytes = 2 * n * n * 4 + d * 4
    k = ("k_fp_ref_seg", fp_us, fp_bytes) if fp_us >= bp_us else ("k_bp_ref_angle", bp_us, bp_bytes)
    gbs = k[2] / (k[1] * 1e-6) / 1e9
    roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": k[0],
            "launch_us": {"fp": round(fp_us, 2), "bp": round(bp_us, 2)},
            "bytes_per_launch": {"fp": fp_bytes, "bp": bp_bytes},
            "what": "one angle per launch (the sweep's structure, ufo-ir-sart-task.c:208-220): the FP streams "
                    "the image once, the BP reads and writes it; every launch of one sweep bracketed by "
                    "events (uir_task_time_kernels), so the figures include each launch's start-up"}
    print("c2: FP %.2f us, BP %.2f us per angle" % (fp_us, bp_us), file=sys.stderr, flush=True)

    # accuracy (full iteration count) and the CPU baseline
    t0 = time.time()
    ref32 = O.sart(g, sino32, iters, 0.5)
    t_ref32 = time.time() - t0
    print("c2: oracle FP32 SART in %.1f s" % t_ref32, file=sys.stderr, flush=True)
    t0 = time.time()
    ref64 = O.sart64(g, sino32.astype(np.float64), iters, 0.5)
    print("c2: oracle FP64 SART in %.1f s" % (time.time() - t0), file=sys.stderr, flush=True)
    cpu_it_s = iters / t_ref32
    acc = {"iterations": iters,
           "rel_l2_vs_oracle_fp32": O.rel_l2(got, ref32), "bit_exact_vs_oracle_fp32": bool(np.array_equal(got, ref32)),
           "rel_l2_vs_fp64": O.rel_l2(got, ref64), "psnr_vs_fp64_db": round(O.psnr(got, ref64), 2),
           "oracle_fp32_rel_l2_vs_fp64": O.rel_l2(ref32, ref64),
           "note": "1% noise and the reference's ray weights (1 / FP(1), up to ~8e3 on corner-grazing rays) make "
                   "FP32 and FP64 differ; the GPU path is the reference's FP32 arithmetic"}
    line = {"metric": "plain SART, %d iterations" % iters, "value": round(ms, 2), "unit": "ms", "n_gpus": 1,
            "higher_is_better": False, "dtype": "f32", "data": "synthetic: 20 random ellipses, 1% Gaussian noise",
            "config": {"workload": "SART task, 1024x1024, 720 angles over [0, pi), 1448 bins, relaxation 0.5, "
                                   "ordered single-angle subsets, %d iterations" % iters},
            "ms_per_iteration": round(ms / iters, 3), "us_per_angle": round(ms / iters / a * 1000, 2),
            "roofline": roof, "accuracy": acc,
            "cpu_baseline": {"value": round(cpu_it_s, 4), "unit": "iterations/s", "cores": O.num_threads(),
                             "kind": "port",
                             "sample": "the oracle's FP32 SART (the reference's arithmetic), same slice, "
                                       "%d iterations, OpenMP over the host's cores" % iters}}
    print(json.dumps(line), flush=True)


def line_samples(n, d, a, step, axis):
    """(ray, image line) pairs of the line model whose two taps touch the
    image: the FP's algorithmic samples.  The model's geometry: the float
    LUT pair scaled to unit length (components below 2^-20 taken as 0); a ray
    advances one image line per sample along its major axis, at position
    v = v0 + i step across it, and touches the image while -1 <= v < n."""
    ang = (np.arange(a, dtype=np.float32) * np.float32(step)).astype(np.float64)
    sn = np.sin(ang).astype(np.float32).astype(np.float64)
    cs = np.cos(ang).astype(np.float32).astype(np.float64)
    r = np.hypot(cs, sn)
    cs, sn = cs / r, sn / r
    tiny = 2.0 ** -20
    cs, sn = (np.where(np.abs(cs) < tiny, 0.0, np.where(np.abs(sn) < tiny, np.sign(cs), cs)),
              np.where(np.abs(sn) < tiny, 0.0, np.where(np.abs(cs) < tiny, np.sign(sn), sn)))
    h = n / 2.0
    t = np.arange(d) + 0.5 - axis
    total = 0
    for k in range(a):
        c, s = cs[k], sn[k]
        if abs(s) > abs(c):
            st, v0 = c / s, (0.5 - h) * c / s + h - 0.5 - t / s
        else:
            st, v0 = s / c, (0.5 - h) * s / c + h - 0.5 + t / c
        if st == 0.0:
            total += int(np.sum((v0 >= -1) & (v0 < n))) * n
            continue
        lo, hi = (-1.0 - v0) / st, (n - v0) / st
        if st > 0:
            first, end = np.ceil(lo), np.ceil(hi)
        else:
            first, end = np.floor(hi) + 1, np.floor(lo) + 1
        total += int(np.sum(np.maximum(0, np.minimum(n, end) - np.maximum(0, first))))
    return total


def line_sirt(args, uir):
    """SIRT with the line projector (exact ray-pixel chord lengths, the
    projector's model "line") at --n x --n, --angles angles, --dets bins:
    slice-iterations/s of a --batch-slice batch in --storage (`value`) and of
    the SIRT task, one FP32 slice per process() (`plain_fp32`); the durations
    of the FP and BP launches (the table-driven, LDS-staged kernels in their
    line mode) from HIP events on the stream they run on; accuracy of both
    paths after --accuracy-iters iterations against the oracle's FP64 SIRT
    over the same model, and that oracle on the host cores."""
    import oracle as O
    res = uir.Resources(0)
    n, a, d = args.n, args.angles, args.dets
    step = float(np.float32(np.pi / a))
    axis = d / 2.0
    g = O.Geometry(n, d, a, np.float32(step), model="line")
    fwd = uir.Task("parallel-projector", model="line", step=step, is_forward=True, angles_num=a, image_size=n,
                   detectors_num=d).setup(res)
    img = uir.DeviceArray.from_numpy(phantom(n, 20261203))
    assert fwd.requisition(img) == (a, d)
    s = uir.DeviceArray((a, d))
    fwd.process(img, s)
    sino = s.numpy().astype(np.float64)
    sino += np.random.default_rng(20261203).normal(0.0, 0.01 * sino.max(), sino.shape)
    sino32 = sino.astype(np.float32)
    ds = uir.DeviceArray.from_numpy(sino32)
    out = uir.DeviceArray((n, n))
    pj = uir.Task("parallel-projector", model="line", step=step, is_forward=False, image_size=n)
    relax = 0.25
    sirt = uir.Task("sirt", projector=pj, relaxation_factor=relax, num_iterations=1).setup(res)
    sirt.process(ds, out)
    timer = uir.Timer()
    k = max(4, args.steps // 3)
    times = {}
    for iters in (1, k + 1):
        sirt["num_iterations"] = iters
        res.synchronize()
        timer.start(res.stream)
        sirt.process(ds, out)
        timer.stop(res.stream)
        times[iters] = timer.elapsed_ms()
    ms_it = (times[k + 1] - times[1]) / k
    print("line: SIRT %.3f ms per iteration" % ms_it, file=sys.stderr, flush=True)

    # the projections alone, each repeated `reps` times between two events
    bwd = uir.Task("parallel-projector", model="line", step=step, is_forward=False, angles_num=a, image_size=n,
                   detectors_num=d).setup(res)
    assert bwd.requisition(s) == (n, n)
    reps = 10
    fwd.forward(img, s)
    timer.start(res.stream)
    for _ in range(reps):
        fwd.forward(img, s)
    timer.stop(res.stream)
    fp_ms = timer.elapsed_ms() / reps
    bwd.backward(s, out)
    timer.start(res.stream)
    for _ in range(reps):
        bwd.backward(s, out)
    timer.stop(res.stream)
    bp_ms = timer.elapsed_ms() / reps
    fp_samples = line_samples(n, d, a, step, axis)
    bp_samples = n * n * a
    print("line: task FP %.3f ms (%d samples), BP %.3f ms (%d samples)" % (fp_ms, fp_samples, bp_ms, bp_samples),
          file=sys.stderr, flush=True)
    sirt["num_iterations"] = args.accuracy_iters
    sirt.process(ds, out)
    plain_img = out.numpy()

    # ---- batched: --batch slices of the same shape in --storage, one FP and
    # one BP launch per iteration over every slice, replayed from a hipGraph
    spec = dict(n=n, n_dets=d, n_angles=a, step=step, relaxation=relax, model="line")
    bt = uir.Batch(res, [spec] * args.batch, args.storage)
    for i in range(args.batch):
        bt.set_sinogram(i, sino32)
    bfp_ms, bbp_ms = bt.time_kernels(5)
    b1 = uir.Batch(res, [spec], args.storage)
    b1.set_sinogram(0, sino32)
    b1.reset()
    b1.iterate(args.accuracy_iters, use_graph=False)
    batch_img = b1.image(0)
    b1.close()
    bt.reset()
    bt.iterate(1)
    for _ in range(args.warmup):
        bt.iterate(1)
    res.synchronize()
    timer.start(res.stream)
    for _ in range(args.steps):
        bt.iterate(1)
    timer.stop(res.stream)
    total_ms = timer.elapsed_ms()
    ms_per_step = total_ms / args.steps
    value = args.batch * args.steps / (total_ms / 1000.0)
    print("line: batched %s, %d slices: %.3f ms per iteration (FP %.3f, BP %.3f ms per launch)"
          % (args.storage, args.batch, ms_per_step, bfp_ms, bbp_ms), file=sys.stderr, flush=True)

    # 8 flop per two-tap sample: two weights (one FMA each) and two FMAs into the sum
    es = 2 if args.storage in ("f16", "bf16", "f16-sino", "bf16-sino") else 4
    ei = 2 if args.storage in ("f16", "bf16") else 4
    nv, ns = n * n, a * d
    kname, kms, ks = ("fp", bfp_ms, fp_samples) if bfp_ms >= bbp_ms else ("bp", bbp_ms, bp_samples)
    tfl = 8.0 * ks * args.batch / (kms * 1e-3) / 1e12
    kbytes = args.batch * ((nv * ei + 2 * ns * es + ns * 4) if kname == "fp" else (ns * es + 2 * nv * ei + nv * 4))
    gbs = kbytes / (kms * 1e-3) / 1e9
    roof = {"bound": "valu", "achieved": round(tfl, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tfl / FP32_PEAK_TFLOPS, 4), "traffic": None,
            "kernel": "k_fp_pairs (line mode)" if kname == "fp" else "k_bp_tab (line mode)",
            "ms": {"fp": round(bfp_ms, 4), "bp": round(bbp_ms, 4)},
            "samples_per_slice": {"fp": fp_samples, "bp": bp_samples},
            "note": "one launch over the %d-slice batch; FP32 vector work, 8 flop per two-tap sample (two chord "
                    "weights, two FMAs into the sum); the taps are read from LDS-staged image lines (FP) or "
                    "sinogram rows (BP); hbm frac %.4f at %.1f GB/s algorithmic (%d bytes: images, sinograms "
                    "and weights streamed once)" % (args.batch, gbs / HBM_PEAK_GBS, gbs, kbytes)}
    acc = cpu = None
    if not args.no_cpu:
        t0 = time.time()
        ref = O.sirt64(g, sino32.astype(np.float64), args.accuracy_iters, relax)
        t_ref = time.time() - t0
        print("line: oracle FP64 SIRT, %d iterations in %.1f s" % (args.accuracy_iters, t_ref), file=sys.stderr,
              flush=True)
        acc = {"iterations": args.accuracy_iters,
               "batched_%s_rel_l2_vs_fp64" % args.storage: O.rel_l2(batch_img, ref),
               "batched_%s_psnr_db" % args.storage: round(O.psnr(batch_img, ref), 2),
               "plain_fp32_rel_l2_vs_fp64": O.rel_l2(plain_img, ref),
               "gate": {"rel_l2_max": 1e-3, "pass": bool(O.rel_l2(batch_img, ref) <= 1e-3)},
               "reference": "the oracle's FP64 SIRT over the line model (orc64_sirt), same data"}
        cpu = {"value": round(args.accuracy_iters / t_ref, 4), "unit": "slice-iterations/s",
               "cores": O.num_threads(), "kind": "port",
               "sample": "the oracle's FP64 SIRT over the line model, the same slice, %d iterations (the ray and "
                         "pixel weights, two more projections, included), OpenMP over the host's cores"
                         % args.accuracy_iters}
    line = {"metric": "SIRT slice-iterations/s with the line projector, batched vs plain", "value": round(value, 2),
            "unit": "slice-iterations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 math, %s storage" % args.storage,
            "data": "synthetic: 20 random ellipses, line FP, 1% Gaussian noise",
            "config": {"workload": "batched SIRT with the line projector, %d slices of %dx%d, %d angles over [0, pi), "
                                   "%d bins, relaxation %g" % (args.batch, n, n, a, d, relax),
                       "batch": args.batch, "storage": args.storage},
            "plain_fp32": {"value": round(1000.0 / ms_it, 2), "unit": "slice-iterations/s",
                           "ms_per_iteration": round(ms_it, 4),
                           "kernels_ms": {"fp": round(fp_ms, 4), "bp": round(bp_ms, 4)},
                           "path": "SIRT task with a line projector, one slice per process(), FP32"},
            "roofline": roof, "accuracy": acc, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def batch_c3(args, uir):
    """Batched SIRT over 512 slices of uneven size (N in {256, 384, 512, 768,
    1024, 2048}, angles U{180..1800}, bins ceil(N sqrt 2)), 100 iterations from
    a captured hipGraph, FP32 and FP16 storage; accuracy against the FP64
    oracle on the slices with N <= 384."""
    import oracle as O
    res = uir.Resources(0)
    rng = np.random.default_rng(20261207)
    sizes = [256, 384, 512, 768, 1024, 2048]
    specs = []
    for _ in range(args.c3_slices):
        n = sizes[rng.integers(0, len(sizes))]
        a = int(rng.integers(180, 1801))
        specs.append(dict(n=n, n_dets=math.ceil(n * math.sqrt(2)), n_angles=a, step=float(np.float32(np.pi / a)),
                          relaxation=0.25))
    # sinograms: one random-ellipse phantom per image size, projected at each
    # slice's own angles by the reference-compatible GPU projector
    t0 = time.time()
    imgs = {n: uir.DeviceArray.from_numpy(phantom(n, 20261300 + n)) for n in sizes}
    sinos = []
    for sp in specs:
        n, a, d = sp["n"], sp["n_angles"], sp["n_dets"]
        pj = uir.Task("parallel-projector", step=sp["step"], is_forward=True, angles_num=a, image_size=n,
                      detectors_num=d).setup(res)
        s = uir.DeviceArray((a, d))
        pj.process(imgs[n], s)
        sinos.append(s.numpy())
        s.free()
    for im in imgs.values():
        im.free()
    t_data = time.time() - t0
    print("c3: %d sinograms in %.1f s" % (len(specs), t_data), file=sys.stderr, flush=True)
    # FP64 references of the first slices with N <= 384
    t0 = time.time()
    check = [i for i, sp in enumerate(specs) if sp["n"] <= 384][:3]
    refs = {}
    for i in check:
        sp = specs[i]
        g = O.Geometry(sp["n"], sp["n_dets"], sp["n_angles"], np.float32(sp["step"]))
        refs[i] = O.sirt64(g, sinos[i].astype(np.float64), args.c3_iters, 0.25)
    t_ref = time.time() - t0
    print("c3: FP64 references of %d slices in %.1f s" % (len(check), t_ref), file=sys.stderr, flush=True)
    out = {}
    for storage in ("f32", "f16", "f16-sino"):
        bt = uir.Batch(res, specs, storage)
        for i, sn in enumerate(sinos):
            bt.set_sinogram(i, sn)
        info = bt.info()
        bt.reset()
        bt.iterate(1)                          # capture
        bt.reset()
        res.synchronize()
        timer = uir.Timer()
        timer.start(res.stream)
        for _ in range(args.c3_iters):
            bt.iterate(1)
        timer.stop(res.stream)
        ms = timer.elapsed_ms()
        errs = [O.rel_l2(bt.image(i), refs[i]) for i in check]
        print("c3: %s %d iterations in %.1f ms" % (storage, args.c3_iters, ms), file=sys.stderr, flush=True)
        out[storage] = {"ms_total": round(ms, 1), "ms_per_iteration": round(ms / args.c3_iters, 3),
                        "slice_iters_per_s": round(len(specs) * args.c3_iters / (ms / 1000.0), 1),
                        "rel_l2_vs_fp64_max": max(errs) if errs else None, "checked_slices": len(errs),
                        "device_bytes": info["device_bytes"], "fp_items": info["fp_items"],
                        "bp_items": info["bp_items"], "geometries": info["n_geometries"]}
        bt.close()
    line = {"metric": "batched SIRT over uneven slices", "value": out["f16-sino"]["slice_iters_per_s"],
            "unit": "slice-iters/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f16 sinograms, f32 images, f32 math",
            "data": "synthetic: 20 random ellipses per slice, Joseph FP (reference-compatible GPU projector)",
            "config": {"workload": "%d slices, N in {256..2048}, angles U{180..1800}, bins ceil(N sqrt 2), "
                                   "%d SIRT iterations from a captured hipGraph" % (len(specs), args.c3_iters),
                       "slices": len(specs), "iterations": args.c3_iters},
            "f32": out["f32"], "f16": out["f16"], "f16-sino": out["f16-sino"], "data_gen_s": round(t_data, 1),
            "fp64_oracle_s": round(t_ref, 1)}
    print(json.dumps(line), flush=True)


def ellipsoid_slices(n, n_slices, seed, k=32):
    """A piecewise-constant 3-D phantom of k random ellipsoids in [-1, 1]^3
    (centres U(-0.6, 0.6), semi-axes U(0.08, 0.5), rotation about z U(0, pi),
    additive attenuations U(-0.2, 0.5), clamped >= 0), sampled on n x n pixel
    centres at n_slices z planes; yields one (n, n) FP32 slice at a time."""
    rng = np.random.default_rng(seed)
    cen = rng.uniform(-0.6, 0.6, (k, 3))
    ax = rng.uniform(0.08, 0.5, (k, 3))
    rot = rng.uniform(0.0, np.pi, k)
    val = rng.uniform(-0.2, 0.5, k).astype(np.float32)
    c = (np.arange(n) + 0.5 - n / 2.0) / (n / 2.0)
    for zi in range(n_slices):
        z = (zi + 0.5 - n_slices / 2.0) / (n_slices / 2.0)
        img = np.zeros((n, n), np.float32)
        for e in range(k):
            h = 1.0 - ((z - cen[e, 2]) / ax[e, 2]) ** 2
            if h <= 0:
                continue
            a, b = ax[e, 0] * np.sqrt(h), ax[e, 1] * np.sqrt(h)
            r = max(a, b)
            lo = lambda v: max(0, int((v - r + 1) * n / 2) - 1)
            hi = lambda v: min(n, int((v + r + 1) * n / 2) + 2)
            x0, x1, y0, y1 = lo(cen[e, 0]), hi(cen[e, 0]), lo(-cen[e, 1]), hi(-cen[e, 1])
            if x0 >= x1 or y0 >= y1:
                continue
            X, Y = np.meshgrid(c[x0:x1], -c[y0:y1])
            xr = (X - cen[e, 0]) * np.cos(rot[e]) + (Y - cen[e, 1]) * np.sin(rot[e])
            yr = -(X - cen[e, 0]) * np.sin(rot[e]) + (Y - cen[e, 1]) * np.cos(rot[e])
            img[y0:y1, x0:x1][(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += val[e]
        np.maximum(img, 0, out=img)
        yield img


def asdpocs_c4(args, uir):
    """Batched ASD-POCS on a stack of independent 1024 x 1024 slices of a 3-D
    random-ellipsoid phantom, sparse view (120 angles over [0, pi), 1448
    bins), Poisson noise at I0 = 5e3 with 0.01 attenuation per pixel unit,
    40 iterations of one SART pass and 20 TV steps, FP64 norm reductions on
    the device.  Beside it: the plain per-slice task path on a few slices,
    accuracy against the FP64 oracle on one slice, and the FP32 oracle on
    the host cores as the CPU baseline."""
    import oracle as O
    res = uir.Resources(0)
    n, a, d = 1024, 120, 1448
    S, iters = args.c4_slices, args.c4_iters
    step = float(np.float32(np.pi / a))
    spec = dict(n=n, n_dets=d, n_angles=a, step=step, axis_position=-1.0, relaxation=0.25,
                ray_weight_min=RAY_WEIGHT_MIN)
    params = dict(beta=1.0, beta_red=0.995, alpha=0.2, alpha_red=0.95, r_max=0.95, ng=20, df_iterations=1,
                  positive_constraint=True, compat=bool(args.c4_compat))
    i0, mu = 5e3, 0.01
    rng = np.random.default_rng(20261204)
    t0 = time.time()
    proj = uir.Task("parallel-projector", step=step, is_forward=True, angles_num=a, image_size=n,
                    detectors_num=d).setup(res)
    dimg, dsino = uir.DeviceArray((n, n)), uir.DeviceArray((a, d))
    sinos = []
    for k, img in enumerate(ellipsoid_slices(n, S, 20261204)):
        dimg.upload(img, res.stream)
        proj.process(dimg, dsino)
        p = dsino.numpy(res.stream).astype(np.float64)
        counts = np.maximum(rng.poisson(i0 * np.exp(-mu * p)), 1)
        sinos.append((-np.log(counts / i0) / mu).astype(np.float32))
        if k % 128 == 127:
            print("c4: %d / %d sinograms" % (k + 1, S), file=sys.stderr, flush=True)
    dimg.free()
    dsino.free()
    t_data = time.time() - t0
    out = {}
    for storage in args.c4_storage.split(","):
        tile = uir.tune_geometry(res, spec, storage, reps=3)["best"]   # the residual FP's tile shape
        bt = uir.Batch(res, [spec] * S, storage)
        for k, sn in enumerate(sinos):
            bt.set_sinogram(k, sn)
        bt.asdpocs(1, **params)                # allocates the method's buffers, captures the iteration graph
        res.synchronize()
        timer = uir.Timer()
        timer.start(res.stream)
        bt.asdpocs(iters, **params)
        timer.stop(res.stream)
        ms = timer.elapsed_ms()
        print("c4: %s %d slices x %d iterations in %.1f ms" % (storage, S, iters, ms), file=sys.stderr, flush=True)
        out[storage] = {"ms_total": round(ms, 1), "ms_per_iteration": round(ms / iters, 2),
                        "slice_iters_per_s": round(S * iters / (ms / 1000.0), 1),
                        "device_bytes": bt.info()["device_bytes"], "tile": tile}
        if storage == args.c4_storage.split(",")[0]:
            diag0 = bt.asdpocs_diagnostics(0)
        out[storage]["image0"] = bt.image(0)
        bt.close()
    # the plain path: the asdpocs task, one slice per process() call
    proj_b = uir.Task("parallel-projector", step=step, is_forward=False, image_size=n)
    sart = uir.Task("sart", num_iterations=1)
    task = uir.Task("asdpocs", projector=proj_b, df_minimizer=sart, num_iterations=iters,
                    compat=bool(args.c4_compat)).setup(res)
    n_plain = min(S, args.c4_plain_slices)
    dout = uir.DeviceArray((n, n))
    ds = [uir.DeviceArray.from_numpy(sinos[k]) for k in range(n_plain)]
    task.process(ds[0], dout)
    res.synchronize()
    t1 = time.perf_counter()
    for k in range(n_plain):
        task.process(ds[k], dout)
    res.synchronize()
    plain_s = (time.perf_counter() - t1) / n_plain
    # accuracy of slice 0 against the FP64 oracle, every storage mode
    g = O.Geometry(n, d, a, np.float32(step))
    t2 = time.time()
    kw = {k: v for k, v in params.items() if k not in ("df_iterations", "compat")}
    ref = O.asdpocs64(g, sinos[0].astype(np.float64), iters, df_iters=1, textbook=not args.c4_compat,
                      ray_weight_min=RAY_WEIGHT_MIN, **kw)
    t_ref = time.time() - t2
    for storage in out:
        img0 = out[storage].pop("image0")
        out[storage]["rel_l2_vs_fp64_slice0"] = O.rel_l2(img0, ref)
        out[storage]["psnr_vs_fp64_slice0_db"] = round(O.psnr(img0, ref), 2)
    # CPU baseline: the FP32 oracle (C, OpenMP) on slice 0, two iterations
    t3 = time.perf_counter()
    O.asdpocs(g, sinos[0], 2, df_iters=1, textbook=not args.c4_compat, **kw)
    cpu_s_per_it = (time.perf_counter() - t3) / 2
    main_storage = args.c4_storage.split(",")[0]
    line = {"metric": "batched ASD-POCS, sparse-view stack", "value": out[main_storage]["slice_iters_per_s"],
            "unit": "slice-iters/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": {"f32": "f32", "f16-sino": "f16 sinograms, f32 images, f32 math"}.get(main_storage, main_storage),
            "data": "synthetic: 32 random ellipsoids, Joseph FP, Poisson counts at I0 = 5e3, 0.01 per pixel unit",
            "config": {"workload": "%d slices of 1024x1024, 120 angles over [0, pi), 1448 bins; %d ASD-POCS "
                                   "iterations of 1 SART pass + 20 TV steps (%s), FP64 device reductions"
                                   % (S, iters, "reference arithmetic" if args.c4_compat else "as published"),
                       "slices": S, "iterations": iters, "compat": bool(args.c4_compat)},
            "storage": out,
            "plain": {"s_per_slice": round(plain_s, 4), "slices_timed": n_plain,
                      "slice_iters_per_s": round(iters / plain_s, 1)},
            "speedup_vs_plain": round(out[main_storage]["slice_iters_per_s"] / (iters / plain_s), 2),
            "diagnostics_slice0_last": [float(v) for v in diag0[-1]] if len(diag0) else None,
            "cpu_baseline": {"value": round(1.0 / cpu_s_per_it, 3), "unit": "slice-iters/s",
                             "cores": O.num_threads(), "kind": "port",
                             "sample": "FP32 oracle ASD-POCS, slice 0, 2 iterations"},
            "data_gen_s": round(t_data, 1), "fp64_oracle_s": round(t_ref, 1)}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--storage", default="f16-sino", choices=["f16", "bf16", "f32", "f16-sino", "bf16-sino"],
                    help="image and sinogram storage (f16-sino: FP16 sinograms, FP32 images)")
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--angles", type=int, default=720)
    ap.add_argument("--dets", type=int, default=1448)
    ap.add_argument("--accuracy-iters", type=int, default=10)
    ap.add_argument("--ray-weight-min", type=float, default=0.0,
                    help="ray-weight floor of the batched and served SIRT (0: the reference's inv)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the FP64 accuracy check")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-qps", type=float, default=100.0)
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-iters", type=int, default=20)
    ap.add_argument("--serve-c5", action="store_true", help="run the split-Bregman TV serving workload only")
    ap.add_argument("--serve-qps-c5", type=float, default=10.0)
    ap.add_argument("--serve-requests-c5", type=int, default=200)
    ap.add_argument("--c5-storage", default="f16-sino", choices=["f16", "f16-sino", "f32"],
                    help="storage of the split-Bregman serving workload (f16-sino: FP16 sinograms, FP32 images; "
                         "f16 images random-walk to ~6e-3 from FP64 over 30 x 5 sweeps)")
    ap.add_argument("--c2", action="store_true", help="time plain SART on 1024x1024, 720 angles, 30 iterations")
    ap.add_argument("--c2-iters", type=int, default=30)
    ap.add_argument("--line", action="store_true", help="SIRT with the line projector (exact chord lengths) on "
                    "one --n x --n slice")
    ap.add_argument("--c3", action="store_true", help="batched SIRT over 512 slices of uneven size")
    ap.add_argument("--c3-slices", type=int, default=512)
    ap.add_argument("--c3-iters", type=int, default=100)
    ap.add_argument("--c4", action="store_true", help="batched ASD-POCS on a sparse-view stack of 1024^2 slices")
    ap.add_argument("--c4-slices", type=int, default=1024)
    ap.add_argument("--c4-iters", type=int, default=40)
    ap.add_argument("--c4-storage", default="f32,f16-sino", help="comma-separated storage modes (first is `value`)")
    ap.add_argument("--c4-compat", action="store_true", help="the reference's ASD-POCS arithmetic")
    ap.add_argument("--c4-plain-slices", type=int, default=4)
    ap.add_argument("--tile", default=None,
                    help="fp_angles,bp_angles,bp_rows: use this tile shape instead of tuning (no probe launches)")
    ap.add_argument("--ceilings", action="store_true",
                    help="measure the LDS read rate and the FP32 FMA issue rate of the device, print them as one "
                         "JSON line and exit")
    ap.add_argument("--profile-pass", action="store_true",
                    help="for counter collection: build the batch, run 3 iterations with direct launches and one "
                         "512 MiB device-to-device copy (the calibration of the byte counters), then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (slices shard across GPUs with no "
              "exchange; see the design notes).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import uir
    if uir.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.serve_c5:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        serve_c5(args, uir)
        return
    if args.c2:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sart_c2(args, uir)
        return
    if args.c4:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        asdpocs_c4(args, uir)
        return
    if args.line:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        line_sirt(args, uir)
        return
    if args.c3:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        batch_c3(args, uir)
        return
    res = uir.Resources(0)
    if args.ceilings:
        print(json.dumps({"ceilings": uir.probe_ceilings(res)}))
        return
    n, a, d = args.n, args.angles, args.dets
    step = float(np.float32(np.pi / a))
    relax = 0.25

    # ---- synthetic data: phantoms projected by the reference-compatible projector
    t0 = time.time()
    n_distinct = min(4, args.batch)
    proj = uir.Task("parallel-projector", step=step, is_forward=True, angles_num=a, image_size=n,
                    detectors_num=d).setup(res)
    sinos = []
    rng = np.random.default_rng(20261203)
    for k in range(n_distinct):
        img = uir.DeviceArray.from_numpy(phantom(n, 20261203 + k))
        s = uir.DeviceArray((a, d))
        proj.process(img, s)
        sino = s.numpy().astype(np.float64)
        sino += rng.normal(0.0, 0.01 * sino.max(), sino.shape)
        sinos.append(sino.astype(np.float32))
        img.free()
        s.free()
    t_data = time.time() - t0

    if args.profile_pass:
        if args.tile:
            uir.tune_set(dict(n=n, n_dets=d, n_angles=a, step=step), args.storage,
                         *(int(v) for v in args.tile.split(",")))
        else:
            uir.tune_geometry(res, dict(n=n, n_dets=d, n_angles=a, step=step), args.storage, reps=3)
        specs = [dict(n=n, n_dets=d, n_angles=a, step=step, relaxation=relax,
                      ray_weight_min=args.ray_weight_min)] * args.batch
        bt = uir.Batch(res, specs, args.storage)
        for i in range(args.batch):
            bt.set_sinogram(i, sinos[i % n_distinct])
        bt.reset()
        bt.iterate(3, use_graph=False)
        res.synchronize()
        # calibration kernel: in-place scale of 512 MiB (reads and writes 512 MiB)
        nbytes = 512 << 20
        buf = uir.DeviceArray((nbytes // 4,))
        uir.check(uir.load().uir_op_mul_scalar(buf.ptr, nbytes // 4, 1.0, None))
        res.synchronize()
        print("profile pass: batch %d x %dx%d, %d angles, %d bins, %s; 3 iterations; calibration %d bytes"
              % (args.batch, n, n, a, d, args.storage, nbytes), flush=True)
        return

    # ---- plain FP32 path: the SIRT task, one slice per process()
    ds = uir.DeviceArray.from_numpy(sinos[0])
    out = uir.DeviceArray((n, n))
    pj = uir.Task("parallel-projector", step=step, is_forward=False, image_size=n)
    sirt = uir.Task("sirt", projector=pj, relaxation_factor=relax, num_iterations=1).setup(res)
    sirt.process(ds, out)
    timer = uir.Timer()
    k_plain = max(4, args.steps // 3)
    times = {}
    for iters in (1, k_plain + 1):
        sirt["num_iterations"] = iters
        timer.start(res.stream)
        sirt.process(ds, out)
        timer.stop(res.stream)
        times[iters] = timer.elapsed_ms()
    plain_ms = (times[k_plain + 1] - times[1]) / k_plain
    plain_rate = 1000.0 / plain_ms
    sirt["num_iterations"] = args.accuracy_iters
    sirt.process(ds, out)
    plain_img = out.numpy()

    # ---- batched path: tile shapes tuned for this geometry first (or given)
    geo = dict(n=n, n_dets=d, n_angles=a, step=step)
    if args.tile:
        fa, ba, br = (int(v) for v in args.tile.split(","))
        uir.tune_set(geo, args.storage, fa, ba, br)
        tune = {"best": {"fp_angles": fa, "bp_angles": ba, "bp_rows": br}}
    else:
        tune = uir.tune_geometry(res, geo, args.storage, reps=3)
    specs = [dict(n=n, n_dets=d, n_angles=a, step=step, relaxation=relax,
                  ray_weight_min=args.ray_weight_min)] * args.batch
    bt = uir.Batch(res, specs, args.storage)
    for i in range(args.batch):
        bt.set_sinogram(i, sinos[i % n_distinct])
    info = bt.info()
    fp_ms, bp_ms = bt.time_kernels(5)
    # accuracy images: slice 0 without the floor (the reference's inv) and with
    # 0.1, each from a one-slice batch of the same storage and tile shape
    acc_imgs = {}
    for rmin in sorted({0.0, 0.1, args.ray_weight_min}):
        b1 = uir.Batch(res, [dict(specs[0], ray_weight_min=rmin)], args.storage)
        b1.set_sinogram(0, sinos[0])
        b1.reset()
        b1.iterate(args.accuracy_iters, use_graph=False)
        acc_imgs[rmin] = b1.image(0)
        b1.close()
    bt.reset()
    bt.iterate(1)                         # capture the one-iteration graph
    for _ in range(args.warmup):
        bt.iterate(1)
    res.synchronize()
    step_ms = []
    t_all = uir.Timer()
    t_all.start(res.stream)
    for _ in range(args.steps):
        timer.start(res.stream)
        bt.iterate(1)
        timer.stop(res.stream)
        step_ms.append(timer.elapsed_ms())
    t_all.stop(res.stream)
    total_ms = t_all.elapsed_ms()
    res.synchronize()
    ms_per_step = total_ms / args.steps
    value = args.batch * args.steps / (total_ms / 1000.0)

    # ---- roofline of the dominant kernel (per launch over the whole batch)
    es = 2 if args.storage in ("f16", "bf16", "f16-sino", "bf16-sino") else 4   # sinograms
    ei = 2 if args.storage in ("f16", "bf16") else 4                             # images
    nv, ns = n * n, a * d
    fp_bytes = args.batch * (nv * ei + ns * es + ns * 4 + ns * es)     # x, b, R read; r written
    bp_bytes = args.batch * (ns * es + nv * ei + nv * ei)              # r, x read; x written
    fp_flops = 4.0 * info["fp_samples"]
    bp_flops = 4.0 * info["bp_samples"]
    dom = ("fp", fp_ms, fp_bytes, fp_flops) if fp_ms >= bp_ms else ("bp", bp_ms, bp_bytes, bp_flops)
    kname, kms, kbytes, kflops = dom
    gbs = kbytes / (kms * 1e-3) / 1e9
    tfl = kflops / (kms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (kname, args.storage))
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch_b%d" % args.batch)
    if tfl / FP32_PEAK_TFLOPS >= gbs / HBM_PEAK_GBS:
        roof = {"bound": "valu", "binding": "lds", "achieved": round(tfl, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(tfl / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                "kernel": "k_fp_pairs" if kname == "fp" else "k_bp_tab",
                "note": "FP32 vector ALU work on the FP32 vector peak (no MFMA: no contraction), 4 flop per "
                        "interpolation sample; hbm frac %.4f at %.1f GB/s algorithmic; the binding resource is "
                        "LDS reads, one 8-byte (e, d) pair per sample (see lds)" % (gbs / HBM_PEAK_GBS, gbs)}
    else:
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_fp_pairs" if kname == "fp" else "k_bp_tab"}

    ksamples = info["fp_samples"] if kname == "fp" else info["bp_samples"]
    lds_tbs = 8.0 * ksamples / (kms * 1e-3) / 1e12
    roof["lds"] = {"achieved": round(lds_tbs, 2), "peak": round(LDS_PEAK_TBS, 2), "unit": "TB/s",
                   "frac": round(lds_tbs / LDS_PEAK_TBS, 4),
                   "what": "LDS bytes read by the sampling loop, 8 per sample (staging writes not counted)"}
    try:
        ceil = uir.probe_ceilings(res)
        roof["lds"]["probe"] = {"tb_s": ceil["lds_tb_s"], "bytes_per_clock_cu": ceil["lds_bytes_per_clock_cu"],
                                "fp32_fma_tflop_s": ceil["fp32_tflop_s"],
                                "fma_wave64_per_clock_cu": ceil["fma_wave64_per_clock_cu"],
                                "cvt_wave64_per_clock_cu": ceil["cvt_wave64_per_clock_cu"],
                                "what": "measured on this device: a read-only LDS loop, an FMA loop and a float-to-int loop"}
    except uir.UirError as e:
        roof["lds"]["probe"] = {"error": str(e)}

    # ---- CPU baseline and FP64 accuracy (the oracle, on this host's cores)
    cpu = None
    acc = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        g = O.Geometry(n, d, a, np.float32(step))
        tt = []
        for iters in (1, 3):
            t1 = time.time()
            O.sirt(g, sinos[0], iters, relax)
            tt.append(time.time() - t1)
        cpu_rate = 2.0 / max(tt[1] - tt[0], 1e-9)
        cpu = {"value": round(cpu_rate, 4), "unit": "slice-iters/s", "cores": O.num_threads(), "kind": "port",
               "sample": "1 slice at 1024x1024x720x1448, ref-compat FP32 SIRT, (t(3 it) - t(1 it)) / 2"}
        t1 = time.time()
        ref64 = {}
        for rmin in acc_imgs:
            ref64[rmin] = O.sirt64(g, sinos[0].astype(np.float64), args.accuracy_iters, relax, ray_weight_min=rmin)
            print("accuracy: FP64 SIRT, %d iterations, ray_weight_min %g: %.1f s"
                  % (args.accuracy_iters, rmin, time.time() - t1), file=sys.stderr, flush=True)
        t_ref = time.time() - t1
        st = args.storage
        acc = {"iterations": args.accuracy_iters,
               "batched_%s_rel_l2_vs_fp64" % st: O.rel_l2(acc_imgs[args.ray_weight_min], ref64[args.ray_weight_min]),
               "batched_%s_psnr_db" % st: round(O.psnr(acc_imgs[args.ray_weight_min], ref64[args.ray_weight_min]), 2),
               "batched_ray_weight_min": args.ray_weight_min,
               "batched_%s_rel_l2_vs_fp64_unfloored" % st: O.rel_l2(acc_imgs[0.0], ref64[0.0]),
               "batched_%s_psnr_db_unfloored" % st: round(O.psnr(acc_imgs[0.0], ref64[0.0]), 2),
               "batched_%s_rel_l2_vs_fp64_floor_0.1" % st: O.rel_l2(acc_imgs[0.1], ref64[0.1]),
               "unfloored_reference": "FP64 SIRT with the reference's inv (ufo-ir-sirt-task.c:180-185)",
               "plain_fp32_rel_l2_vs_fp64": O.rel_l2(plain_img, ref64[0.0]),
               "plain_fp32_psnr_db": round(O.psnr(plain_img, ref64[0.0]), 2),
               "plain_reference": "FP64 SIRT without a ray weight floor (the reference's inv)",
               "fp64_oracle_s": round(t_ref, 1)}
        gate_err = acc["batched_%s_rel_l2_vs_fp64" % st]
        acc["gate"] = {"bar": 1e-3, "rel_l2": gate_err, "passed": bool(gate_err <= 1e-3),
                       "what": "batched result at the headline's ray-weight floor against the FP64 SIRT with "
                               "the same floor, %d iterations" % args.accuracy_iters}

    # ---- serving: Poisson arrivals of SIRT requests of the same shape
    serving = None
    if not args.no_serve:
        del bt
        def make(k):
            return (uir.make_request(n, d, a, step, "sirt", iterations=args.serve_iters, relaxation=relax,
                                     ray_weight_min=args.ray_weight_min), sinos[k % n_distinct])
        serving = run_serving(uir, args.storage, make, args.serve_qps, args.serve_requests, 20261204)
        serving["workload"] = "SIRT requests of %dx%d, %d angles, %d bins, %d iterations each, Poisson arrivals" % (
            n, n, a, d, args.serve_iters)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "slice-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f16-sino": "f16 sinograms, f32 images, f32 math", "bf16-sino": "bf16 sinograms, f32 images, f32 math"}
                 .get(args.storage, "%s storage, f32 math" % args.storage),
        "data": "synthetic: 20 random ellipses per slice, Joseph FP, 1%% Gaussian noise (%d distinct)" % n_distinct,
        "config": {"workload": "batched SIRT, %d slices of %dx%d, %d angles over [0, pi), %d bins, relaxation %.2f"
                               % (args.batch, n, n, a, d, relax),
                   "batch": args.batch, "image": n, "angles": a, "bins": d, "storage": args.storage,
                   "tile": tune["best"],
                   # consecutive slices of the geometry share BP items (slice pairs) when the
                   # paired table still fills the device
                   "bp_slice_pairs": bool(info["bp_items"] * 2 <= args.batch * ((n + 63) // 64)
                                          * ((n + tune["best"]["bp_rows"] - 1) // tune["best"]["bp_rows"]))},
        "plain_fp32": {"value": round(plain_rate, 2), "unit": "slice-iters/s", "ms_per_iteration": round(plain_ms, 4),
                       "path": "SIRT task, one slice per process(), reference-compatible FP32"},
        "speedup_vs_plain": round(value / plain_rate, 3),
        "latency_ms": {"p50": round(percentile(step_ms, 50), 4), "p99": round(percentile(step_ms, 99), 4),
                       "what": "one batched SIRT iteration (all slices)"},
        "serving": serving,
        "kernels_ms": {"fp": round(fp_ms, 4), "bp": round(bp_ms, 4)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "accuracy": acc,
        "data_gen_s": round(t_data, 1),
        "device_bytes": info["device_bytes"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

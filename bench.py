#!/usr/bin/env python3
"""SIRT slice-iterations/s at 1024 x 1024 x 720 angles x 1448 bins on one MI355X.

Workload (synthetic, deterministic): 20 random ellipses per slice (seeded
centres in the unit disk, semi-axes U(0.02, 0.3), angles U(0, pi), additive
attenuations U(-0.2, 0.5), clamped >= 0) on a 1024 x 1024 grid, forward
projected with the Joseph projector at 720 angles over [0, pi) onto 1448 bins,
plus Gaussian noise of sigma = 1% of the sinogram maximum.

A step is one SIRT iteration (relaxation 0.25) over a batch of --batch such
slices resident in HBM in FP16 storage with FP32 interpolation and
accumulation: one forward- and one back-projection launch over every slice,
replayed from a captured hipGraph.  The plain FP32 path (the SIRT task, one
slice per process() call, reference-compatible arithmetic) is measured beside
it, and accuracy is reported against the FP64 evaluation of the same
operators.

Prints one JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ufo-ir_amd"))

METRIC = "slice-iters/s SIRT 1024²×720 ang, batched vs plain @relL2≤1e-3; p99 latency"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


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


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--storage", default="f16", choices=["f16", "bf16", "f32"])
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--angles", type=int, default=720)
    ap.add_argument("--dets", type=int, default=1448)
    ap.add_argument("--accuracy-iters", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the FP64 accuracy check")
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
    res = uir.Resources(0)
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
        specs = [dict(n=n, n_dets=d, n_angles=a, step=step, relaxation=relax)] * args.batch
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

    # ---- batched path
    specs = [dict(n=n, n_dets=d, n_angles=a, step=step, relaxation=relax)] * args.batch
    bt = uir.Batch(res, specs, args.storage)
    for i in range(args.batch):
        bt.set_sinogram(i, sinos[i % n_distinct])
    info = bt.info()
    fp_ms, bp_ms = bt.time_kernels(5)
    bt.reset()
    bt.iterate(args.accuracy_iters, use_graph=False)
    batched_img = bt.image(0)
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
    es = 2 if args.storage in ("f16", "bf16") else 4
    nv, ns = n * n, a * d
    fp_bytes = args.batch * (nv * es + ns * es + ns * 4 + ns * es)     # x, b, R read; r written
    bp_bytes = args.batch * (ns * es + nv * es + nv * es)              # r, x read; x written
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
        roof = {"bound": "mfma", "achieved": round(tfl, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tfl / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                "kernel": "k_fp_tab" if kname == "fp" else "k_bp_tab",
                "note": "compute-bound on the FP32 vector ALU (no MFMA); 4 flop per interpolation sample; "
                        "hbm frac %.4f at %.1f GB/s algorithmic" % (gbs / HBM_PEAK_GBS, gbs)}
    else:
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_fp_tab" if kname == "fp" else "k_bp_tab"}

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
        ref64 = O.sirt64(g, sinos[0].astype(np.float64), args.accuracy_iters, relax)
        t_ref = time.time() - t1
        acc = {"iterations": args.accuracy_iters,
               "batched_%s_rel_l2_vs_fp64" % args.storage: O.rel_l2(batched_img, ref64),
               "batched_%s_psnr_db" % args.storage: round(O.psnr(batched_img, ref64), 2),
               "plain_fp32_rel_l2_vs_fp64": O.rel_l2(plain_img, ref64),
               "plain_fp32_psnr_db": round(O.psnr(plain_img, ref64), 2),
               "fp64_oracle_s": round(t_ref, 1)}

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
        "dtype": "%s-storage/f32-math" % args.storage,
        "data": "synthetic: 20 random ellipses per slice, Joseph FP, 1%% Gaussian noise (%d distinct)" % n_distinct,
        "config": {"workload": "batched SIRT, %d slices of %dx%d, %d angles over [0, pi), %d bins, relaxation %.2f"
                               % (args.batch, n, n, a, d, relax),
                   "batch": args.batch, "image": n, "angles": a, "bins": d, "storage": args.storage},
        "plain_fp32": {"value": round(plain_rate, 2), "unit": "slice-iters/s", "ms_per_iteration": round(plain_ms, 4),
                       "path": "SIRT task, one slice per process(), reference-compatible FP32"},
        "speedup_vs_plain": round(value / plain_rate, 3),
        "latency_ms": {"p50": round(percentile(step_ms, 50), 4), "p99": round(percentile(step_ms, 99), 4),
                       "what": "one batched SIRT iteration (all slices)"},
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

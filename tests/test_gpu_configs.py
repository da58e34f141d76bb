"""GPU parity at the workloads' full sizes, against the oracle.

  - the headline batch: 1024 x 1024 slices, 720 angles, 1448 bins, 1 % noise,
    FP16 sinograms (f16-sino), 10 SIRT iterations, with and without the ray
    weight floor, as a BP slice pair and as single slices, against the FP64
    SIRT (ufo-ir-sirt-task.c:176-215); the same shape with the line model;
  - plain SART on 1024 x 1024 / 720 / 1448 with 1 % noise, 2 sweeps,
    relaxation 0.5, against the reference-compatible FP32 oracle
    (ufo-ir-sart-task.c:193-220);
  - one batch of uneven slices holding the largest sizes of the uneven-slice
    workload (2048 x 2048 with 1800 angles, 1024 x 1024, 384 x 384), in FP32
    and f16-sino, against the FP64 SIRT;
  - split-Bregman TV requests (SBTV_SIRT) at N = 512, 1024 and 2048 served
    together, against the FP64 oracle of that method.

Every case prints its name and shape before touching the device."""
import numpy as np
import pytest

import oracle as O
import uir
from phantoms import ellipse_phantom

pytestmark = pytest.mark.gpu


def _geom(n, d, a, axis=-1.0):
    return O.Geometry(n, d, a, np.float32(np.pi / a), axis)


def _noisy(g, seed, sigma=0.01):
    b = O.fp64(g, ellipse_phantom(g.n, seed=seed).astype(np.float64))
    return b + np.random.default_rng(seed).normal(0.0, sigma * b.max(), b.shape)


def _spec(n, d, a, relax, rmin=0.0):
    return dict(n=n, n_dets=d, n_angles=a, step=float(np.float32(np.pi / a)), relaxation=relax,
                ray_weight_min=rmin)


@pytest.mark.parametrize("rmin", [0.0, 0.1])
def test_headline_batch_f16_sino_matches_fp64(res, rmin, monkeypatch):
    """Two slices of the metric's shape, FP16 sinograms, 10 iterations; the
    FP64 reference applies the same ray-weight floor (0 is the reference's
    inv, ufo-ir-basic-ops.cl:18-36).  The two slices run once as a BP slice
    pair on 64-row tiles (the headline's path at 32 slices) and once as
    single slices on 32-row tiles (what the default thresholds give two
    slices), with bit-identical images."""
    n, d, a, iters, relax = 1024, 1448, 720, 10, 0.25
    print("case headline f16-sino", n, d, a, "ray_weight_min", rmin, flush=True)
    g = _geom(n, d, a)
    sinos = [_noisy(g, 20261203 + k) for k in range(2)]
    images, items = {}, {}
    for pairs in ("0", "1000000000"):
        monkeypatch.setenv("UIR_PAIR_MIN_ITEMS", pairs)
        bt = uir.Batch(res, [_spec(n, d, a, relax, rmin)] * 2, "f16-sino")
        items[pairs] = bt.info()["bp_items"]
        for k, s in enumerate(sinos):
            bt.set_sinogram(k, s.astype(np.float32))
        bt.reset()
        bt.iterate(iters)
        images[pairs] = [bt.image(k) for k in range(2)]
        bt.close()
    assert 4 * items["0"] == items["1000000000"]  # paired 64-row tiles, then single 32-row ones
    for k in range(2):
        assert np.array_equal(images["0"][k], images["1000000000"][k]), k
    bt_images = images["0"]
    for k, s in enumerate(sinos):
        # the reference starts from the FP32 data handed to the batch (the
        # device's FP16 rounding of it counts against the batched path)
        ref = O.sirt64(g, s.astype(np.float32).astype(np.float64), iters, relax, ray_weight_min=rmin)
        got = bt_images[k]
        err = O.rel_l2(got, ref)
        print("slice", k, "rel-L2 vs FP64 %.3e, PSNR %.1f dB" % (err, O.psnr(got, ref)), flush=True)
        assert err <= 1e-3


def test_line_headline_shape_matches_fp64(res, monkeypatch):
    """The line model at the metric's shape: two slices with the same data run
    as a BP slice pair (mode 4), FP16 sinograms, 10 iterations, against the
    oracle's FP64 SIRT over the model; both slices of the pair give the same
    image."""
    n, d, a, iters, relax = 1024, 1448, 720, 10, 0.25
    print("case line headline f16-sino", n, d, a, flush=True)
    monkeypatch.setenv("UIR_PAIR_MIN_ITEMS", "0")
    g = O.Geometry(n, d, a, np.float32(np.pi / a), model="line")
    sino = _noisy(g, 20261205).astype(np.float32)
    bt = uir.Batch(res, [dict(_spec(n, d, a, relax), model="line")] * 2, "f16-sino")
    assert bt.info()["bp_items"] == 256
    for k in range(2):
        bt.set_sinogram(k, sino)
    bt.reset()
    bt.iterate(iters)
    got = [bt.image(k) for k in range(2)]
    assert np.array_equal(got[0], got[1])
    ref = O.sirt64(g, sino.astype(np.float64), iters, relax)
    err = O.rel_l2(got[0], ref)
    print("line rel-L2 vs FP64 %.3e" % err, flush=True)
    assert err <= 1e-3


def test_sart_task_full_shape_matches_oracle(res):
    """The reference-compatible SART task at 1024 x 1024 / 720 / 1448."""
    n, d, a, iters, relax = 1024, 1448, 720, 2, 0.5
    print("case sart task", n, d, a, iters, flush=True)
    g = _geom(n, d, a)
    b = _noisy(g, 20261202).astype(np.float32)
    proj = uir.Task("parallel-projector", step=float(np.float32(np.pi / a)), is_forward=False, image_size=n)
    sart = uir.Task("sart", projector=proj, num_iterations=iters, relaxation_factor=relax).setup(res)
    ds = uir.DeviceArray.from_numpy(b)
    out = uir.DeviceArray(sart.requisition(ds))
    sart.process(ds, out)
    got = out.numpy()
    ref = O.sart(g, b, iters, relax)
    err = O.rel_l2(got, ref)
    print("SART rel-L2 vs oracle %.3e" % err, flush=True)
    assert err <= 1e-4


MIXED = [(2048, 2897, 1800), (1024, 1449, 720), (384, 544, 300)]


@pytest.fixture(scope="module")
def mixed_refs():
    iters = 2
    data, refs = [], []
    for k, (n, d, a) in enumerate(MIXED):
        g = _geom(n, d, a)
        b = O.fp64(g, ellipse_phantom(n, seed=300 + k).astype(np.float64)).astype(np.float32)
        data.append(b)
        refs.append(O.sirt64(g, b.astype(np.float64), iters, 0.25))
    return iters, data, refs


@pytest.mark.parametrize("storage,tol", [("f32", 2e-5), ("f16-sino", 1e-3)])
def test_mixed_batch_largest_slices_match_fp64(res, mixed_refs, storage, tol):
    iters, data, refs = mixed_refs
    print("case mixed batch", storage, MIXED, flush=True)
    bt = uir.Batch(res, [_spec(n, d, a, 0.25) for (n, d, a) in MIXED], storage)
    for k, b in enumerate(data):
        bt.set_sinogram(k, b)
    bt.reset()
    bt.iterate(iters)
    for k, ref in enumerate(refs):
        err = O.rel_l2(bt.image(k), ref)
        print(storage, MIXED[k], "rel-L2 vs FP64 %.3e" % err, flush=True)
        assert err <= tol


def test_sbtv_sirt_serving_large_requests_match_fp64():
    """SBTV_SIRT requests of the split-Bregman serving workload's sizes, served
    together with FP16 sinograms and FP32 images, the workload's storage
    (6 outer x 5 inner sweeps to keep the FP64 oracle short; the workload runs
    30 x 5).  With FP16 images the rounding of the image after every sweep
    random-walks: 9.6e-4 after these 30 sweeps, about 6e-3 after 150."""
    outer, inner, mu = 6, 5, 1.0
    cases = [(512, 725, 400, 0.002), (1024, 1449, 300, 0.02), (2048, 2897, 90, 0.08)]
    reqs, sinos = [], []
    for k, (n, d, a, lam) in enumerate(cases):
        g = _geom(n, d, a)
        b = _noisy(g, 500 + k).astype(np.float32)
        req = uir.make_request(n, d, a, float(np.float32(np.pi / a)), "sbtv-sirt", iterations=outer,
                               inner_iterations=inner, relaxation=1.0, lam=lam, mu=mu, ray_weight_min=0.1)
        reqs.append(req)
        sinos.append(b)
    print("case sbtv-sirt serving", [c[:3] for c in cases], flush=True)
    with uir.Server(0, "f16-sino") as srv:
        tickets = [srv.submit(r, s) for r, s in zip(reqs, sinos)]
        out = [srv.wait(t, timeout=120) for t in tickets]
    for k, (n, d, a, lam) in enumerate(cases):
        g = _geom(n, d, a)
        ref = O.sbsirt64(g, sinos[k], outer, inner, relax=1.0, lam=lam, mu=mu, ray_weight_min=0.1)
        err = O.rel_l2(out[k][0], ref)
        print("SBTV_SIRT", cases[k][:3], "rel-L2 vs FP64 %.3e" % err, flush=True)
        assert out[k][1]["steps"] == outer * inner
        assert err <= 1e-3

"""Reconstruction server (uir_server_*): continuous batching of requests of
uneven geometry and method.  Each request's result must not depend on what
shares its steps (bit-identical alone and in a mix, and equal to the batch
engine's), SBTV_SIRT must match its FP64 oracle, admission must respect the
device budget and the slot limit, and requests arriving while others run must
join the running steps."""
import threading
import time

import numpy as np
import pytest

import oracle as O
import uir
from phantoms import ellipse_phantom

pytestmark = pytest.mark.gpu

GEOMS = [  # (n, n_dets, n_angles)
    (128, 182, 90),
    (96, 136, 120),
    (160, 227, 75),
    (64, 91, 45),
]


def _step(a):
    return float(np.float32(np.pi / a))


def _data(n, d, a, seed, noise=0.0):
    g = O.Geometry(n, d, a, np.float32(np.pi / a))
    b = O.fp64(g, ellipse_phantom(n, seed=seed).astype(np.float64))
    if noise:
        b = b + np.random.default_rng(seed).normal(0, noise * b.max(), b.shape)
    return g, b


def _sirt(n, d, a, iters, relax=0.5, posc=False):
    return uir.make_request(n, d, a, _step(a), "sirt", iterations=iters, relaxation=relax,
                            positive_constraint=posc)


def _sb(n, d, a, outer, inner, lam=0.05, mu=1.0, relax=1.0):
    return uir.make_request(n, d, a, _step(a), "sbtv-sirt", iterations=outer, inner_iterations=inner,
                            relaxation=relax, lam=lam, mu=mu)


def _run(server, reqs, sinos):
    tickets = [server.submit(r, s) for r, s in zip(reqs, sinos)]
    return [server.wait(t, timeout=120) for t in tickets]


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_serve_sirt_equals_batch(res, storage, monkeypatch):
    """Served SIRT equals the batch engine's SIRT bit for bit.  The last
    request repeats the first one's geometry: with the pairing threshold at 0
    the server runs the two as a BP slice pair whenever both are admitted,
    the batch (where they are not adjacent) as single slices."""
    monkeypatch.setenv("UIR_PAIR_MIN_ITEMS", "0")
    iters = 6
    geoms = GEOMS + [GEOMS[0]]
    sinos = [_data(n, d, a, 10 + i)[1].astype(np.float32) for i, (n, d, a) in enumerate(geoms)]
    with uir.Server(0, storage) as srv:
        got = _run(srv, [_sirt(n, d, a, iters) for (n, d, a) in geoms], sinos)
    specs = [dict(n=n, n_dets=d, n_angles=a, step=_step(a), relaxation=0.5) for (n, d, a) in geoms]
    bt = uir.Batch(res, specs, storage)
    for i, s in enumerate(sinos):
        bt.set_sinogram(i, s)
    bt.reset()
    bt.iterate(iters, use_graph=False)
    for i in range(len(geoms)):
        assert np.array_equal(got[i][0], bt.image(i)), geoms[i]
        assert got[i][1]["steps"] == iters


@pytest.mark.parametrize("storage,tol", [("f32", 2e-5), ("f16-sino", 1e-3)])
def test_serve_line_sirt_equals_batch_and_fp64(res, storage, tol):
    """SIRT requests of the line model served beside joseph requests (one of
    them of the same geometry, so the geometry cache must tell the models
    apart): each equals the batch engine's slice bit for bit and the oracle's
    FP64 SIRT of its own model."""
    iters = 8
    cases = [(128, 182, 90, "line", 0.1), (128, 182, 90, "joseph", 0.1), (97, 137, 60, "line", 0.1)]
    print("case serve line + joseph", storage, cases, flush=True)
    sinos, refs = [], []
    for i, (n, d, a, m, rmin) in enumerate(cases):
        g = O.Geometry(n, d, a, np.float32(np.pi / a), model=m)
        b = O.fp64(g, ellipse_phantom(n, seed=730 + i).astype(np.float64))
        b = (b + np.random.default_rng(730 + i).normal(0.0, 0.01 * b.max(), b.shape)).astype(np.float32)
        sinos.append(b)
        refs.append(O.sirt64(g, b.astype(np.float64), iters, 0.5, ray_weight_min=rmin))
    reqs = [uir.make_request(n, d, a, _step(a), "sirt", iterations=iters, relaxation=0.5, ray_weight_min=rmin,
                             model=m) for (n, d, a, m, rmin) in cases]
    with uir.Server(0, storage) as srv:
        got = _run(srv, reqs, sinos)
    specs = [dict(n=n, n_dets=d, n_angles=a, step=_step(a), relaxation=0.5, ray_weight_min=rmin, model=m)
             for (n, d, a, m, rmin) in cases]
    bt = uir.Batch(res, specs, storage)
    for i, b in enumerate(sinos):
        bt.set_sinogram(i, b)
    bt.reset()
    bt.iterate(iters, use_graph=False)
    for i, c in enumerate(cases):
        err = O.rel_l2(got[i][0], refs[i])
        print(storage, c, "rel-L2 vs FP64 %.3e" % err, flush=True)
        assert np.array_equal(got[i][0], bt.image(i)), c
        assert got[i][1]["steps"] == iters
        assert err <= tol


def test_serve_result_same_alone_and_in_mix():
    n, d, a = GEOMS[0]
    _, b = _data(n, d, a, 21, noise=0.01)
    sb_alone = _sb(n, d, a, 4, 3)
    with uir.Server(0, "f16") as srv:
        (alone_sirt, st1), = _run(srv, [_sirt(n, d, a, 9)], [b.astype(np.float32)])
        (alone_sb, _), = _run(srv, [sb_alone], [b.astype(np.float32)])
        assert st1["max_batch"] == 1
        mix_reqs, mix_sinos = [], []
        for i, (nn, dd, aa) in enumerate(GEOMS[1:]):
            _, bb = _data(nn, dd, aa, 30 + i)
            mix_reqs += [_sirt(nn, dd, aa, 5 + 2 * i), _sb(nn, dd, aa, 2, 2, lam=0.02 * (i + 1))]
            mix_sinos += [bb.astype(np.float32)] * 2
        mix_reqs += [_sirt(n, d, a, 9), _sb(n, d, a, 4, 3)]
        mix_sinos += [b.astype(np.float32)] * 2
        # A long request runs first, so the mix arrives while steps are being
        # taken and joins them (without it, the mix's sharing of steps would
        # depend on how fast the first request is admitted and run).
        nb, db, ab = 256, 363, 180
        _, bb = _data(nb, db, ab, 40)
        blocker = srv.submit(_sirt(nb, db, ab, 200), bb.astype(np.float32))
        out = _run(srv, mix_reqs, mix_sinos)
        srv.wait(blocker, timeout=120)
    assert min(o[1]["max_batch"] for o in out) > 1
    assert np.array_equal(out[-2][0], alone_sirt)
    assert np.array_equal(out[-1][0], alone_sb)


# Noisy data: rays that clip an image corner by a tiny fraction of a pixel
# have FP(1) near 0, so R = 1/FP(1) magnifies the noise they carry and the
# FP32 rounding of their weights into the corner pixels (in the FP64
# evaluation too, up to 4x the phantom's maximum).  Noisy cases therefore run
# with ray_weight_min = 0.1 (rays crossing less than a tenth of a pixel get
# weight 0) in both the GPU request and the oracle.
@pytest.mark.parametrize("storage,noise,rmin,tol", [("f32", 0.0, 0.0, 2e-5), ("f16", 0.0, 0.0, 1e-3),
                                                    ("f32", 0.01, 0.1, 2e-5), ("f16", 0.01, 0.1, 1e-3)])
def test_serve_sbtv_sirt_matches_fp64(storage, noise, rmin, tol):
    outer, inner, lam, mu, relax = 5, 4, 0.05, 1.0, 1.0
    reqs, sinos, refs = [], [], []
    for i, (n, d, a) in enumerate(GEOMS[:3]):
        g, b = _data(n, d, a, 40 + i, noise=noise)
        req = _sb(n, d, a, outer, inner, lam=lam * (i + 1), mu=mu, relax=relax)
        req.slice.ray_weight_min = rmin
        reqs.append(req)
        sinos.append(b.astype(np.float32))
        refs.append(O.sbsirt64(g, b.astype(np.float32), outer, inner, relax=relax, lam=lam * (i + 1), mu=mu,
                               ray_weight_min=rmin))
    with uir.Server(0, storage) as srv:
        out = _run(srv, reqs, sinos)
    for i in range(len(reqs)):
        err = O.rel_l2(out[i][0], refs[i])
        print(storage, "noise", noise, GEOMS[i], "SBTV_SIRT rel-L2 vs FP64", err)
        assert err <= tol
        assert out[i][1]["steps"] == outer * inner


@pytest.mark.parametrize("storage,tol", [("f32", 2e-5), ("f16-sino", 1e-3)])
def test_serve_sbtv_sirt_line_model_matches_fp64(storage, tol):
    """Split-Bregman requests of the line model (its preconditioner is an
    image, P = 1 / (mu BP(1) + 8 lambda) per pixel, formed in the BP's
    epilogue from the pixel weights) beside a joseph one of the same shape:
    each matches the oracle's FP64 method of its own model."""
    outer, inner, mu, relax = 5, 4, 1.0, 1.0
    cases = [(128, 182, 90, "line", 0.05), (97, 137, 60, "line", 0.2), (128, 182, 90, "joseph", 0.05)]
    print("case serve sbtv line", storage, cases, flush=True)
    reqs, sinos, refs = [], [], []
    for i, (n, d, a, m, lam) in enumerate(cases):
        g = O.Geometry(n, d, a, np.float32(np.pi / a), model=m)
        b = O.fp64(g, ellipse_phantom(n, seed=770 + i).astype(np.float64))
        b = (b + np.random.default_rng(770 + i).normal(0.0, 0.01 * b.max(), b.shape)).astype(np.float32)
        req = uir.make_request(n, d, a, _step(a), "sbtv-sirt", iterations=outer, inner_iterations=inner,
                               relaxation=relax, lam=lam, mu=mu, ray_weight_min=0.1, model=m)
        reqs.append(req)
        sinos.append(b)
        refs.append(O.sbsirt64(g, b.astype(np.float64), outer, inner, relax=relax, lam=lam, mu=mu,
                               ray_weight_min=0.1))
    with uir.Server(0, storage) as srv:
        out = _run(srv, reqs, sinos)
    for i, c in enumerate(cases):
        err = O.rel_l2(out[i][0], refs[i])
        print(storage, c, "SBTV_SIRT rel-L2 vs FP64 %.3e" % err, flush=True)
        assert err <= tol
        assert out[i][1]["steps"] == outer * inner


def test_serve_sbtv_sirt_is_not_plain_sirt():
    """With a TV weight the result differs from SIRT with the same sweeps and
    has less total variation (the TV term acts)."""
    n, d, a = GEOMS[0]
    _, b = _data(n, d, a, 50, noise=0.03)
    with uir.Server(0, "f32") as srv:
        (tv, _), (plain, _) = _run(srv, [_sb(n, d, a, 6, 4, lam=0.5, mu=1.0), _sirt(n, d, a, 24, relax=1.0)],
                                   [b.astype(np.float32)] * 2)

    def total_variation(u):
        return np.sum(np.hypot(u - np.roll(u, 1, 1), u - np.roll(u, 1, 0)))

    assert O.rel_l2(tv, plain) > 1e-3
    assert total_variation(tv) < total_variation(plain)


def test_serve_admission_respects_budget():
    """A budget that holds one request at a time serialises them (each runs
    alone) and gives the same results as an unconstrained server; a request
    larger than the budget is refused at submission."""
    reqs = [_sirt(n, d, a, 4) for (n, d, a) in GEOMS[:3]]
    sinos = [_data(n, d, a, 60 + i)[1].astype(np.float32) for i, (n, d, a) in enumerate(GEOMS[:3])]
    fps = [uir.request_footprint(r, "f16") for r in reqs]
    budget = int(max(fps) * 1.2)
    with uir.Server(0, "f16", hbm_budget=budget) as srv:
        tight = _run(srv, reqs, sinos)
        big = _sirt(1024, 1448, 720, 1)
        with pytest.raises(uir.UirError) as e:
            srv.submit(big, np.zeros((720, 1448), np.float32))
        assert e.value.status == 2
        info = srv.info()
        assert info["used"] <= budget and info["active"] == 0 and info["pending"] == 0
    assert all(o[1]["max_batch"] == 1 for o in tight)
    with uir.Server(0, "f16") as srv:
        free = _run(srv, reqs, sinos)
    for a_, b_ in zip(tight, free):
        assert np.array_equal(a_[0], b_[0])


def test_serve_slot_limit():
    reqs = [_sirt(n, d, a, 3) for (n, d, a) in GEOMS] * 2
    sinos = [_data(n, d, a, 70 + i)[1].astype(np.float32) for i, (n, d, a) in enumerate(GEOMS)] * 2
    with uir.Server(0, "f16", max_active=2) as srv:
        out = _run(srv, reqs, sinos)
    assert max(o[1]["max_batch"] for o in out) == 2


def test_serve_continuous_batching_overlaps_arrivals():
    """Short requests submitted while a long one runs are admitted into its
    steps and finish first; timestamps are ordered."""
    n, d, a = 1024, 1448, 720        # about 0.7 ms per step alone: 300 steps outlast the arrivals
    b = np.random.default_rng(80).random((a, d)).astype(np.float32)
    short = [(GEOMS[3], _data(*GEOMS[3], 81)[1].astype(np.float32))]
    with uir.Server(0, "f16") as srv:
        t_long = srv.submit(_sirt(n, d, a, 300), b)
        time.sleep(0.03)
        ts = [srv.submit(_sirt(*gm, 5), s) for gm, s in short * 3]
        outs = [srv.wait(t, timeout=120) for t in ts]
        long_img, long_st = srv.wait(t_long, timeout=120)
        info = srv.info()
    for img, st in outs:
        assert st["t_submit"] <= st["t_admit"] <= st["t_done"]
        assert st["t_done"] < long_st["t_done"]
        assert st["max_batch"] >= 2
    assert long_st["steps"] == 300
    assert info["completed"] == 4 and info["slice_iterations"] == 300 + 3 * 5


def test_serve_concurrent_submitters():
    """Several client threads submit and wait at once."""
    results = {}
    errors = []

    def client(k):
        try:
            n, d, a = GEOMS[k % len(GEOMS)]
            _, b = _data(n, d, a, 90 + k)
            t = srv.submit(_sirt(n, d, a, 3 + k % 3), b.astype(np.float32))
            results[k] = srv.wait(t, timeout=120)
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    with uir.Server(0, "f16") as srv:
        th = [threading.Thread(target=client, args=(k,)) for k in range(8)]
        for t in th:
            t.start()
        for t in th:
            t.join(180)
    assert not errors, errors
    assert len(results) == 8


def test_serve_rejects_bad_requests():
    with uir.Server(0, "f16") as srv:
        for bad in (_sirt(0, 10, 10, 1), _sirt(64, 91, 45, 0), _sirt(64, 91, 45, 2, relax=1.5),
                    _sb(64, 91, 45, 2, 0), _sb(64, 91, 45, 2, 2, lam=0.0)):
            with pytest.raises(uir.UirError) as e:
                srv.submit(bad, np.zeros((max(bad.slice.n_angles, 1), max(bad.slice.n_dets, 1)), np.float32))
            assert e.value.status == 1
        with pytest.raises(uir.UirError):
            srv.wait(12345, timeout=0.1)

"""The autotuner and the tile-shape variants on the GPU: every variant agrees
with the FP64 oracle and with the default shape to FP32 rounding (the BP's
shape changes no bit at all; the FP's moves its ray anchors), batches
mixing geometries of different shapes launch one table per shape with each
slice's result unchanged, and the server's autotune uses the tuned shape."""
import itertools

import numpy as np
import pytest

import oracle as O
import uir
from phantoms import ellipse_phantom

pytestmark = pytest.mark.gpu

VARIANTS = list(itertools.product((16, 8, 32), (32, 16), (64, 32)))


def _spec(n, d, a, relax=0.5):
    return dict(n=n, n_dets=d, n_angles=a, step=float(np.float32(np.pi / a)), relaxation=relax)


def _sino(n, d, a, seed):
    g = O.Geometry(n, d, a, np.float32(np.pi / a))
    return g, O.fp64(g, ellipse_phantom(n, seed=seed).astype(np.float64))


def _run(res, specs, sinos, storage, iters):
    bt = uir.Batch(res, specs, storage)
    for i, s in enumerate(sinos):
        bt.set_sinogram(i, s.astype(np.float32))
    bt.reset()
    bt.iterate(iters, use_graph=False)
    out = [bt.image(i) for i in range(len(specs))]
    bt.close()
    return out


def test_tune_geometry_records_the_fastest(res):
    uir.tune_clear()
    spec = _spec(256, 363, 180)
    r = uir.tune_geometry(res, spec, "f16", reps=2)
    print("tune 256/363/180 f16:", r)
    assert all(t > 0 for t in r["fp_ms"]) and all(t > 0 for t in r["bp_ms"])
    fp_best = (16, 8, 32)[int(np.argmin(r["fp_ms"]))]
    bp_best = [(32, 64), (32, 32), (16, 64), (16, 32)][int(np.argmin(r["bp_ms"]))]
    assert r["best"] == {"fp_angles": fp_best, "bp_angles": bp_best[0], "bp_rows": bp_best[1]}
    assert uir.tune_get(spec, "f16") == (r["best"], True)
    uir.tune_clear()


@pytest.mark.parametrize("fa,ba,br", VARIANTS)
def test_every_variant_matches_fp64_and_default(res, fa, ba, br):
    n, d, a = 160, 227, 75
    spec = _spec(n, d, a)
    g, b = _sino(n, d, a, 11)
    ref64 = O.sirt64(g, b, 12, 0.5)
    uir.tune_clear()
    default, = _run(res, [spec], [b], "f32", 12)
    uir.tune_set(spec, "f32", fa, ba, br)
    got, = _run(res, [spec], [b], "f32", 12)
    uir.tune_clear()
    assert O.rel_l2(got, ref64) <= 2e-5
    assert O.rel_l2(got, default) <= 2e-6


def test_mixed_shapes_in_one_batch(res):
    geoms = [(128, 182, 90), (96, 136, 120), (160, 227, 75)]
    shapes = [(8, 16, 32), (16, 32, 64), (32, 32, 32)]
    specs = [_spec(*gm) for gm in geoms]
    sinos = [_sino(*gm, 20 + i)[1] for i, gm in enumerate(geoms)]
    uir.tune_clear()
    for sp, sh in zip(specs, shapes):
        uir.tune_set(sp, "f16", *sh)
    mixed = _run(res, specs, sinos, "f16", 6)
    alone = [_run(res, [sp], [sn], "f16", 6)[0] for sp, sn in zip(specs, sinos)]
    bt = uir.Batch(res, specs, "f16")
    info = bt.info()
    bt.close()
    uir.tune_clear()
    for m, a_ in zip(mixed, alone):
        assert np.array_equal(m, a_)
    assert info["fp_items"] > 0 and info["bp_items"] > 0


def test_server_autotune_uses_the_tuned_shape(res):
    n, d, a = 128, 182, 90
    spec = _spec(n, d, a)
    _, b = _sino(n, d, a, 30)
    uir.tune_clear()
    with uir.Server(0, "f16", autotune=True) as srv:
        t = srv.submit(uir.make_request(n, d, a, spec["step"], "sirt", iterations=5, relaxation=0.5),
                       b.astype(np.float32))
        served, _ = srv.wait(t, timeout=120)
    tile, tuned = uir.tune_get(spec, "f16")
    assert tuned
    batched, = _run(res, [spec], [b], "f16", 5)      # the batch plans with the same tuned shape
    uir.tune_clear()
    assert np.array_equal(served, batched)


@pytest.mark.parametrize("storage", ["f32", "f16-sino", "bf16"])
def test_bp_shape_does_not_change_results(res, monkeypatch, storage):
    """The BP's shape changes the launch, not the arithmetic: a pixel's BP
    window and sample coordinates are those of its 64 x 64 super-tile and its
    angle sum runs in angle order.  With the FP's shape fixed, every BP shape
    (32 or 16 staged angles, tiles of 64 or 32 rows), with slice pairs
    (UIR_PAIR_MIN_ITEMS=0: the tuned rows) and without (a huge threshold,
    which also selects 32-row tiles for a table that small), gives the same
    bits: joseph with a centred and an off-centre axis (the clamped
    coordinate), and the line model."""
    cases = [  # (n, d, a, axis_position, model), two slices each
        (160, 227, 75, -1.0, "joseph"),
        (97, 137, 60, 70.3, "joseph"),
        (128, 182, 90, -1.0, "line"),
    ]
    specs, sinos = [], []
    for k, (n, d, a, ax, m) in enumerate(cases):
        g = O.Geometry(n, d, a, np.float32(np.pi / a), ax)
        for j in range(2):
            specs.append(dict(_spec(n, d, a), axis_position=ax, model=m, positive_constraint=j == 1,
                              ray_weight_min=0.1 if m == "line" else 0.0))
            sinos.append(O.fp64(g, ellipse_phantom(n, seed=40 + 2 * k + j).astype(np.float64)))
    print("case bp shapes", storage, cases, flush=True)
    got = {}
    for pairs in ("0", "1000000000"):
        monkeypatch.setenv("UIR_PAIR_MIN_ITEMS", pairs)
        for ba, br in [(32, 64), (32, 32), (16, 64), (16, 32)]:
            uir.tune_clear()
            for sp in specs[:4]:
                uir.tune_set(sp, storage, 16, ba, br)
            got[(pairs, ba, br)] = _run(res, specs, sinos, storage, 4)
    uir.tune_clear()
    ref = got[("0", 32, 64)]
    for key, imgs in got.items():
        for i in range(len(specs)):
            assert np.array_equal(imgs[i], ref[i]), (key, i)

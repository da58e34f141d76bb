"""Batched SIRT engine (uir_batch_*): accuracy against the FP64 evaluation of
the same operators, independence of each slice's result from the batch mix,
and graph replay equal to direct launches."""
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
]


def _spec(n, d, a, relax=0.5, posc=False):
    return dict(n=n, n_dets=d, n_angles=a, step=float(np.float32(np.pi / a)), axis_position=-1.0,
                relaxation=relax, positive_constraint=posc)


def _data(n, d, a, seed):
    g = O.Geometry(n, d, a, np.float32(np.pi / a))
    x = ellipse_phantom(n, seed=seed).astype(np.float64)
    b = O.fp64(g, x)
    return g, b


@pytest.mark.parametrize("storage,tol", [("f32", 2e-5), ("f16", 1e-3), ("f16-sino", 1e-3), ("bf16-sino", 3e-3)])
def test_batch_matches_fp64(res, storage, tol):
    iters = 20
    specs = [_spec(*gm) for gm in GEOMS]
    bt = uir.Batch(res, specs, storage)
    refs = []
    for i, (n, d, a) in enumerate(GEOMS):
        g, b = _data(n, d, a, 100 + i)
        bt.set_sinogram(i, b.astype(np.float32))
        refs.append(O.sirt64(g, b, iters, 0.5))
    bt.reset()
    bt.iterate(iters)
    for i in range(len(GEOMS)):
        err = O.rel_l2(bt.image(i), refs[i])
        print(storage, GEOMS[i], "rel-L2 vs FP64", err)
        assert err <= tol


@pytest.mark.parametrize("storage", ["f32", "f16-sino", "f16", "bf16-sino"])
def test_batch_slice_pairs_equal_single_slices(res, storage, monkeypatch):
    """Consecutive slices of one geometry share their BP work items (slice
    pairs: one coordinate, address and 16-byte LDS read per pixel-angle serve
    both slices).  Each slice's image must equal, bit for bit, that slice
    reconstructed alone (a batch of one: no pair), including a slice left
    unpaired between pairs, an off-centre axis (the BP's clamped coordinate),
    sparse angles (the single-value FP), distinct relaxations and POSC within
    a pair, and pairs of the line model (shared chord weights).  These
    batches are too small for the default pairing threshold (4 workgroups
    per CU after pairing), so the test lowers it to 0."""
    monkeypatch.setenv("UIR_PAIR_MIN_ITEMS", "0")
    iters = 7
    cases = [  # (n, d, a, axis_position, relax, posc, model)
        (128, 182, 90, -1.0, 0.5, False, "joseph"),
        (128, 182, 90, -1.0, 0.3, True, "joseph"),
        (128, 182, 90, -1.0, 0.5, False, "joseph"),
        (97, 137, 60, 70.3, 0.5, False, "joseph"),
        (97, 137, 60, 70.3, 0.7, False, "joseph"),
        (256, 363, 24, -1.0, 0.5, False, "joseph"),
        (256, 363, 24, -1.0, 0.5, True, "joseph"),
        (128, 182, 90, -1.0, 0.5, False, "line"),
        (128, 182, 90, -1.0, 0.4, True, "line"),
        (97, 137, 60, 40.0, 0.5, False, "line"),
        (97, 137, 60, 40.0, 0.5, False, "line"),
    ]
    print("case slice pairs", storage, cases, flush=True)
    specs = [dict(n=n, n_dets=d, n_angles=a, step=float(np.float32(np.pi / a)), axis_position=ax, relaxation=rl,
                  positive_constraint=pc, model=m, ray_weight_min=0.1 if m == "line" else 0.0)
             for (n, d, a, ax, rl, pc, m) in cases]
    sinos = []
    for i, (n, d, a, ax, rl, pc, m) in enumerate(cases):
        g = O.Geometry(n, d, a, np.float32(np.pi / a), ax)
        b = O.fp64(g, ellipse_phantom(n, seed=900 + i).astype(np.float64))
        sinos.append((b + np.random.default_rng(900 + i).normal(0.0, 0.01 * b.max(), b.shape)).astype(np.float32))
    bt = uir.Batch(res, specs, storage)
    for i, b in enumerate(sinos):
        bt.set_sinogram(i, b)
    bt.reset()
    bt.iterate(iters)
    for i, c in enumerate(cases):
        solo = uir.Batch(res, [specs[i]], storage)
        solo.set_sinogram(0, sinos[i])
        solo.reset()
        solo.iterate(iters)
        assert np.array_equal(bt.image(i), solo.image(0)), (i, c)


def test_batch_pairs_slices_by_default_only_at_scale(res, monkeypatch):
    """With the default threshold (4 workgroups per CU), a BP table that
    would hold no more items than that takes tiles of 32 rows instead of 64,
    and BP slice pairs are used only while the paired table keeps at least
    that many: ten 1024 x 1024 slices of one geometry pair up on 64-row tiles
    (half the BP items), one or two take 32-row tiles unpaired.  The results
    equal the single-slice ones bit for bit."""
    monkeypatch.delenv("UIR_PAIR_MIN_ITEMS", raising=False)
    n, d, a, iters = 1024, 1448, 90, 2
    spec = dict(n=n, n_dets=d, n_angles=a, step=float(np.float32(np.pi / a)), relaxation=0.5)
    print("case default pairing", n, d, a, flush=True)
    rng = np.random.default_rng(5)
    sinos = [rng.uniform(0.0, 50.0, (a, d)).astype(np.float32) for _ in range(10)]
    fill = 4 * uir.probe_ceilings(res)["n_cu"]
    tiles = (n // 64) ** 2  # 64 x 64 tiles of one slice (the untuned BP tile)

    def expected(k):  # BP items of k slices: the two rules
        per = 2 * tiles if k * tiles <= fill else tiles
        paired = (k // 2 + k % 2) * per
        return paired if paired >= fill else k * per

    solo = uir.Batch(res, [spec], "f16-sino")
    big = uir.Batch(res, [spec] * 10, "f16-sino")
    small = uir.Batch(res, [spec] * 2, "f16-sino")
    assert 10 * tiles > fill and 5 * tiles >= fill and 2 * tiles <= fill  # the cases meant
    assert solo.info()["bp_items"] == expected(1) == 2 * tiles
    assert big.info()["bp_items"] == expected(10) == 5 * tiles
    assert small.info()["bp_items"] == expected(2) == 4 * tiles
    small.close()
    for i, b in enumerate(sinos):
        big.set_sinogram(i, b)
    big.reset()
    big.iterate(iters)
    for i, b in enumerate(sinos):
        solo.set_sinogram(0, b)
        solo.reset()
        solo.iterate(iters)
        assert np.array_equal(big.image(i), solo.image(0)), i


@pytest.mark.parametrize("storage,tol", [("f32", 2e-5), ("f16", 1e-3)])
def test_batch_noisy_with_ray_weight_floor_matches_fp64(res, storage, tol):
    """1% noise: with ray_weight_min = 0.1 in both the batch and the oracle the
    corner-clipping rays drop out and the FP32 / FP16 results stay at the
    clean-data tolerance."""
    iters = 20
    specs = [dict(_spec(*gm), ray_weight_min=0.1) for gm in GEOMS]
    bt = uir.Batch(res, specs, storage)
    refs = []
    for i, (n, d, a) in enumerate(GEOMS):
        g, b = _data(n, d, a, 300 + i)
        b = (b + np.random.default_rng(i).normal(0, 0.01 * b.max(), b.shape)).astype(np.float32)
        bt.set_sinogram(i, b)
        refs.append(O.sirt64(g, b, iters, 0.5, ray_weight_min=0.1))
    bt.reset()
    bt.iterate(iters)
    for i in range(len(GEOMS)):
        err = O.rel_l2(bt.image(i), refs[i])
        print(storage, GEOMS[i], "noisy rel-L2 vs FP64", err)
        assert err <= tol


def test_long_runs_need_fp32_images(res):
    """100 iterations on 1%-noise data: rounding the iterate to FP16 every
    iteration random-walks past the 1e-3 bar (about 3e-3 here), while FP16
    sinograms with an FP32 image stay well inside it (about 3e-4)."""
    iters = 100
    (n, d, a) = GEOMS[0]
    g, b = _data(n, d, a, 400)
    b = (b + np.random.default_rng(4).normal(0, 0.01 * b.max(), b.shape)).astype(np.float32)
    ref = O.sirt64(g, b, iters, 0.5, ray_weight_min=0.1)
    errs = {}
    for storage in ("f16", "f16-sino"):
        bt = uir.Batch(res, [dict(_spec(n, d, a), ray_weight_min=0.1)], storage)
        bt.set_sinogram(0, b)
        bt.reset()
        bt.iterate(iters)
        errs[storage] = O.rel_l2(bt.image(0), ref)
        bt.close()
    print("100 iterations rel-L2 vs FP64:", errs)
    assert errs["f16-sino"] <= 5e-4
    assert errs["f16-sino"] < errs["f16"] / 4


def test_batch_result_independent_of_mix(res):
    iters = 5
    (n, d, a) = GEOMS[0]
    g, b = _data(n, d, a, 7)
    alone = uir.Batch(res, [_spec(n, d, a)], "f16")
    alone.set_sinogram(0, b.astype(np.float32))
    alone.reset()
    alone.iterate(iters)
    mix_specs = [_spec(*GEOMS[1]), _spec(n, d, a), _spec(*GEOMS[2]), _spec(n, d, a, relax=0.25)]
    mix = uir.Batch(res, mix_specs, "f16")
    for i, s in enumerate(mix_specs):
        _, bb = _data(s["n"], s["n_dets"], s["n_angles"], 50 + i)
        mix.set_sinogram(i, bb.astype(np.float32))
    mix.set_sinogram(1, b.astype(np.float32))
    mix.reset()
    mix.iterate(iters)
    assert np.array_equal(alone.image(0), mix.image(1))


def test_graph_replay_equals_direct_launches(res):
    specs = [_spec(*gm) for gm in GEOMS]
    out = []
    for use_graph in (False, True):
        bt = uir.Batch(res, specs, "f32")
        for i, (n, d, a) in enumerate(GEOMS):
            _, b = _data(n, d, a, 200 + i)
            bt.set_sinogram(i, b.astype(np.float32))
        bt.reset()
        for _ in range(3):          # several replays of the same graph
            bt.iterate(4, use_graph=use_graph)
        out.append([bt.image(i) for i in range(len(GEOMS))])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_positive_constraint_and_info(res):
    n, d, a = GEOMS[0]
    bt = uir.Batch(res, [_spec(n, d, a, relax=1.0, posc=True)], "f32")
    g, b = _data(n, d, a, 3)
    rng = np.random.default_rng(0)
    bt.set_sinogram(0, (b + rng.normal(0, 0.05 * b.max(), b.shape)).astype(np.float32))
    bt.reset()
    bt.iterate(10)
    assert bt.image(0).min() >= 0.0
    info = bt.info()
    # 128 x 128: 2 x 2 tiles of 64 x 64, too few to fill the device, so 2 x 4
    # tiles of 32 rows
    assert info["n_slices"] == 1 and info["fp_items"] > 0 and info["bp_items"] == 8
    assert info["bp_samples"] == a * n * n


def test_batch_rejects_bad_specs(res):
    with pytest.raises(uir.UirError):
        uir.Batch(res, [_spec(0, 10, 10)], "f32")
    with pytest.raises(uir.UirError):
        uir.Batch(res, [_spec(64, 91, 30, relax=1.5)], "f32")


@pytest.mark.parametrize("shift", [4.5, -6.0, 4.3, -6.7])
def test_batch_shifted_axis_matches_fp64(res, shift):
    """axis_position off the detector centre: the detectors the reference
    skips (projector-parallel-joseph.cl:62-66) carry no weight, and the BP
    window shrinks (:180-182), on both sides.  Shifts of a multiple of 1/2
    put the clamp edges on bin centres (the BP clamps bins); other shifts put
    them inside a bin (the BP clamps the coordinate, :203-210)."""
    iters = 10
    n, d, a = GEOMS[0]
    axis = d / 2.0 + shift
    spec = dict(_spec(n, d, a), axis_position=axis)
    g = O.Geometry(n, d, a, np.float32(np.pi / a), axis)
    b = O.fp64(g, ellipse_phantom(n, seed=800).astype(np.float64))
    bt = uir.Batch(res, [spec], "f32")
    bt.set_sinogram(0, b.astype(np.float32))
    bt.reset()
    bt.iterate(iters)
    ref = O.sirt64(g, b, iters, 0.5)
    err = O.rel_l2(bt.image(0), ref)
    print("axis shift", shift, "rel-L2 vs FP64", err)
    assert err <= 2e-5


@pytest.mark.parametrize("storage,tol", [("f32", 2e-5), ("f16-sino", 1e-3)])
def test_batch_sparse_and_dense_angles_in_one_batch(res, storage, tol):
    """A sparse-angle slice (one angle step moves edge rays by ~34 pixels: the
    FP stages single values with wider windows) next to dense-angle slices
    (staged (e, d) pairs): both FP kinds run in one iteration and each slice
    matches FP64."""
    geoms = [(256, 363, 24), (128, 182, 90), (192, 272, 16)]
    specs = [_spec(*gm) for gm in geoms]
    assert specs[0]["step"] * 256 > 15 and specs[1]["step"] * 128 < 15
    bt = uir.Batch(res, specs, storage)
    refs = []
    for i, (n, d, a) in enumerate(geoms):
        g, b = _data(n, d, a, 700 + i)
        bt.set_sinogram(i, b.astype(np.float32))
        refs.append(O.sirt64(g, b, 10, 0.5))
    bt.reset()
    bt.iterate(10)
    for i in range(len(geoms)):
        err = O.rel_l2(bt.image(i), refs[i])
        print(storage, geoms[i], "rel-L2 vs FP64", err)
        assert err <= tol


@pytest.mark.parametrize("storage,tol", [("f32", 2e-5), ("f16-sino", 1e-3), ("f16", 1e-3), ("bf16-sino", 3e-3)])
def test_batch_line_and_joseph_slices_match_fp64(res, storage, tol):
    """Slices of both projection models in one batch (the line model's FP and
    BP run as separate launches of the table kernels' line mode, with their own
    ray weights and per-pixel weights): each matches the FP64 SIRT of its own
    model; a line slice with the ray-weight floor and an odd image edge (rays
    along pixel edges at 90 degrees) included.  The joseph slice takes the
    floor: on noisy data its corner-grazing rays (weights up to ~1e4) leave
    FP32 and FP64 apart without it (test_batch_noisy_with_ray_weight_floor...)."""
    iters = 20
    cases = [(128, 182, 90, "line", 0.0), (96, 136, 120, "joseph", 0.1), (97, 137, 60, "line", 0.1),
             (160, 227, 75, "line", 0.0)]
    specs = [dict(_spec(n, d, a), model=m, ray_weight_min=rmin) for (n, d, a, m, rmin) in cases]
    print("case batch line + joseph", storage, cases, flush=True)
    bt = uir.Batch(res, specs, storage)
    refs = []
    for i, (n, d, a, m, rmin) in enumerate(cases):
        g = O.Geometry(n, d, a, np.float32(np.pi / a), model=m)
        b = O.fp64(g, ellipse_phantom(n, seed=700 + i).astype(np.float64))
        b = b + np.random.default_rng(700 + i).normal(0.0, 0.01 * b.max(), b.shape)
        b32 = b.astype(np.float32)
        bt.set_sinogram(i, b32)
        refs.append(O.sirt64(g, b32.astype(np.float64), iters, 0.5, ray_weight_min=rmin))
    bt.reset()
    bt.iterate(iters)
    for i, c in enumerate(cases):
        err = O.rel_l2(bt.image(i), refs[i])
        print(storage, c, "rel-L2 vs FP64", err, flush=True)
        assert err <= tol
    # a line slice alone in a batch gives the same image as in the mix
    solo = uir.Batch(res, [specs[0]], storage)
    g0 = O.Geometry(128, 182, 90, np.float32(np.pi / 90), model="line")
    b0 = O.fp64(g0, ellipse_phantom(128, seed=700).astype(np.float64))
    b0 = (b0 + np.random.default_rng(700).normal(0.0, 0.01 * b0.max(), b0.shape)).astype(np.float32)
    solo.set_sinogram(0, b0)
    solo.reset()
    solo.iterate(iters)
    assert np.array_equal(solo.image(0), bt.image(0))


def test_batch_asdpocs_refuses_mixed_models(res):
    """Batched ASD-POCS runs one data step for the whole batch: slices of the
    two projection models together are refused."""
    bt = uir.Batch(res, [dict(_spec(64, 91, 30), model="line"), _spec(64, 91, 30)], "f32")
    with pytest.raises(uir.UirError) as e:
        bt.asdpocs(1)
    assert e.value.status == 4

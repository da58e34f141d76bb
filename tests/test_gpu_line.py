"""GPU parity of the line projection model (the projector's model "line":
exact ray-pixel intersection lengths; ufo-ir-parallel-projector-task.c:
123-128,370 select a projector by model, and the reference ships only
joseph) against the oracle's FP64 line model, whose weights are pinned in
test_oracle_line_pins.py.

  - FP and BP at odd and even image edges (rays along pixel edges at 0 and 90
    degrees), an off-centre axis, a full circle and the metric's shape, to
    1e-5 relative L2 (FP32 arithmetic against FP64);
  - the BP is the transpose of the FP, <A x, y> = <x, A^T y>, to FP32 rounding;
  - projector subsets over angle ranges that span both ray families;
  - SIRT (with and without POSC), SART (its sweep captured and replayed) and
    ASD-POCS as published, each with a line projector, against the oracle's
    FP64 methods over the same model, to 1e-4.

Every case prints its name and shape before touching the device."""
import numpy as np
import pytest

import oracle as O
import uir
from phantoms import ellipse_phantom

pytestmark = pytest.mark.gpu

FP_TOL = 1e-5
METHOD_TOL = 1e-4


def _geom(n, d, a, axis=-1.0, turn=1.0):
    return O.Geometry(n, d, a, np.float32(turn * np.pi / a), axis, model="line")


def _projector(g, forward, **extra):
    return uir.Task("parallel-projector", model="line", step=float(g.step), is_forward=forward,
                    angles_num=g.n_angles, image_size=g.n, detectors_num=g.n_dets, axis_position=g.axis,
                    **extra)


def _fp(res, g, x):
    p = _projector(g, True).setup(res)
    dx = uir.DeviceArray.from_numpy(x.astype(np.float32))
    assert p.requisition(dx) == (g.n_angles, g.n_dets)
    out = uir.DeviceArray((g.n_angles, g.n_dets))
    p.process(dx, out)
    return out.numpy()


def _bp(res, g, y):
    p = _projector(g, False).setup(res)
    ds = uir.DeviceArray.from_numpy(y.astype(np.float32))
    assert p.requisition(ds) == (g.n, g.n)
    out = uir.DeviceArray((g.n, g.n))
    p.process(ds, out)
    return out.numpy()


SHAPES = [  # (image edge, detectors, angles, axis, fraction of a half turn)
    (33, 50, 20, -1.0, 1.0),       # odd edge: the 0 and 90 degree rays run along pixel edges
    (64, 91, 45, -1.0, 1.0),
    (97, 137, 60, 70.3, 1.0),      # off-centre axis (every ray is active in this model)
    (64, 64, 120, -1.0, 2.0),      # full circle: five family runs
    (256, 363, 180, -1.0, 1.0),
    (1024, 1448, 720, -1.0, 1.0),  # the metric's shape
]


@pytest.mark.parametrize("n,d,a,axis,turn", SHAPES)
def test_line_fp_matches_fp64(res, n, d, a, axis, turn):
    print("case line fp", n, d, a, axis, turn, flush=True)
    g = _geom(n, d, a, axis, turn)
    x = ellipse_phantom(n, seed=n + a)
    got = _fp(res, g, x)
    ref = O.fp64(g, x.astype(np.float64))
    err = O.rel_l2(got, ref)
    print("rel-L2 %.3e" % err, flush=True)
    assert err <= FP_TOL, (err, np.abs(got - ref).max())


@pytest.mark.parametrize("n,d,a,axis,turn", SHAPES)
def test_line_bp_matches_fp64(res, n, d, a, axis, turn):
    print("case line bp", n, d, a, axis, turn, flush=True)
    g = _geom(n, d, a, axis, turn)
    y = np.random.default_rng(n).random((a, d))
    got = _bp(res, g, y)
    ref = O.bp64(g, y.astype(np.float32).astype(np.float64))
    err = O.rel_l2(got, ref)
    print("rel-L2 %.3e" % err, flush=True)
    assert err <= FP_TOL, (err, np.abs(got - ref).max())


@pytest.mark.parametrize("n,d,a,axis", [(64, 91, 45, -1.0), (128, 128, 90, 70.3), (512, 725, 360, -1.0)])
def test_line_bp_is_the_transpose_of_fp(res, n, d, a, axis):
    print("case line transpose", n, d, a, axis, flush=True)
    g = _geom(n, d, a, axis)
    rng = np.random.default_rng(a)
    x = rng.standard_normal((n, n)).astype(np.float32)
    y = rng.standard_normal((a, d)).astype(np.float32)
    lhs = float(np.sum(_fp(res, g, x).astype(np.float64) * y))
    rhs = float(np.sum(x.astype(np.float64) * _bp(res, g, y)))
    # sums of n * a * d products of unit-variance terms: FP32 rounding of the
    # projections is ~1e-7 relative per entry
    scale = float(np.sqrt(np.sum(_fp(res, g, np.abs(x)).astype(np.float64) ** 2) * np.sum(y.astype(np.float64) ** 2)))
    assert abs(lhs - rhs) <= 1e-6 * scale, (lhs, rhs, scale)


def test_line_subsets_match_fp64(res):
    """subset_fp over angles [30, 75) (both families) writes those rows of
    r + FP; subset_bp adds relax * BP of those angles."""
    n, d, a = 96, 136, 90
    print("case line subsets", n, d, a, flush=True)
    g = _geom(n, d, a)
    x = ellipse_phantom(n, seed=8)
    y0 = np.random.default_rng(8).random((a, d)).astype(np.float32)
    p = _projector(g, True).setup(res)
    p.requisition(uir.DeviceArray((n, n)))
    dx = uir.DeviceArray.from_numpy(x)
    ds = uir.DeviceArray.from_numpy(y0)
    p.subset_fp(dx, ds, (30, 45, 0))
    got = ds.numpy()
    full = O.fp64(g, x.astype(np.float64))
    want = y0.astype(np.float64).copy()
    want[30:75] += full[30:75]
    assert np.array_equal(got[:30], y0[:30]) and np.array_equal(got[75:], y0[75:])
    assert O.rel_l2(got[30:75], want[30:75]) <= FP_TOL
    # BP of the same angles: the full BP of a sinogram that is zero elsewhere
    p["relaxation"] = 0.5
    v0 = np.random.default_rng(9).random((n, n)).astype(np.float32)
    dv = uir.DeviceArray.from_numpy(v0)
    p.subset_bp(dv, uir.DeviceArray.from_numpy(y0), (30, 45, 0))
    ymask = np.zeros((a, d))
    ymask[30:75] = y0[30:75]
    ref = v0 + 0.5 * O.bp64(g, ymask)
    assert O.rel_l2(dv.numpy(), ref) <= FP_TOL


def _data(g, seed, sigma=0.01):
    b = O.fp64(g, ellipse_phantom(g.n, seed=seed).astype(np.float64))
    return (b + np.random.default_rng(seed).normal(0.0, sigma * b.max(), b.shape)).astype(np.float32)


def _method(res, g, method, **props):
    proj = uir.Task("parallel-projector", model="line", step=float(g.step), is_forward=False, image_size=g.n,
                    axis_position=g.axis)
    return uir.Task(method, projector=proj, **props).setup(res)


def _run(m, b, n):
    ds = uir.DeviceArray.from_numpy(b)
    assert m.requisition(ds) == (n, n)
    out = uir.DeviceArray((n, n))
    m.process(ds, out)
    return out.numpy()


@pytest.mark.parametrize("n,d,a,axis,iters,posc", [(256, 363, 180, -1.0, 10, False), (97, 137, 60, 70.3, 20, True)])
def test_line_sirt_matches_fp64(res, n, d, a, axis, iters, posc):
    print("case line sirt", n, d, a, axis, iters, posc, flush=True)
    g = _geom(n, d, a, axis)
    b = _data(g, 40 + n)
    m = _method(res, g, "sirt", num_iterations=iters, relaxation_factor=0.5, positive_constraint=posc)
    got = _run(m, b, n)
    ref = O.sirt64(g, b.astype(np.float64), iters, 0.5, posc)
    err = O.rel_l2(got, ref)
    print("rel-L2 %.3e" % err, flush=True)
    assert err <= METHOD_TOL


@pytest.mark.parametrize("n,d,a,axis,iters", [(128, 181, 90, -1.0, 3), (65, 65, 40, 30.5, 2)])
def test_line_sart_matches_fp64(res, n, d, a, axis, iters):
    """Sweeps after the first replay the captured sweep; a second process()
    on the same buffers starts again from x = 0 and replays the cached graph."""
    print("case line sart", n, d, a, axis, iters, flush=True)
    g = _geom(n, d, a, axis)
    b = _data(g, 50 + n)
    m = _method(res, g, "sart", num_iterations=iters, relaxation_factor=0.5)
    ds = uir.DeviceArray.from_numpy(b)
    assert m.requisition(ds) == (n, n)
    out = uir.DeviceArray((n, n))
    m.process(ds, out)
    got = out.numpy()
    ref = O.sart64(g, b.astype(np.float64), iters, 0.5)
    err = O.rel_l2(got, ref)
    print("rel-L2 %.3e" % err, flush=True)
    assert err <= METHOD_TOL
    m.process(ds, out)
    assert np.array_equal(out.numpy(), got)


def test_line_asdpocs_textbook_matches_fp64(res):
    n = d = 96
    a = 45
    print("case line asdpocs", n, d, a, flush=True)
    g = _geom(n, d, a)
    b = _data(g, 61, sigma=0.02)
    proj = uir.Task("parallel-projector", model="line", step=float(g.step), is_forward=False)
    dfm = uir.Task("sart", num_iterations=1)
    m = uir.Task("asdpocs", projector=proj, df_minimizer=dfm, num_iterations=4, compat=False, ng=10, beta=0.9,
                 beta_red=0.9).setup(res)
    got = _run(m, b, n)
    ref = O.asdpocs64(g, b.astype(np.float64), 4, df="sart", df_iters=1, ng=10, textbook=True, beta=0.9,
                      beta_red=0.9)
    err = O.rel_l2(got, ref)
    print("rel-L2 %.3e" % err, flush=True)
    assert err <= METHOD_TOL


EDGES = [  # (image edge, detectors, angles, axis): degenerate sizes and axes
    (1, 1, 1, -1.0),
    (2, 3, 1, -1.0),
    (5, 2, 7, -1.0),
    (8, 1, 4, -1.0),
    (3, 9, 180, -1.0),
    (16, 23, 12, 0.0),      # axis at the first detector's edge
    (16, 23, 12, 40.0),     # axis beyond the detector: most rays miss the image
    (65, 93, 33, 46.25),
]


@pytest.mark.parametrize("n,d,a,axis", EDGES)
def test_line_edge_shapes_match_fp64(res, n, d, a, axis):
    """Tiny images, one detector, one angle, axes at or beyond the detector's
    ends (the task path accepts any axis, as the reference does): FP, BP and
    10 SIRT iterations against the FP64 line model."""
    print("case line edge", n, d, a, axis, flush=True)
    g = _geom(n, d, a, axis)
    rng = np.random.default_rng(n * 1000 + d)
    x = rng.random((n, n))
    y = rng.random((a, d))
    fp = _fp(res, g, x)
    bp = _bp(res, g, y)
    fref = O.fp64(g, x.astype(np.float32).astype(np.float64))
    bref = O.bp64(g, y.astype(np.float32).astype(np.float64))
    assert np.allclose(fp, fref, rtol=1e-5, atol=1e-5 * max(1.0, np.abs(fref).max())), np.abs(fp - fref).max()
    assert np.allclose(bp, bref, rtol=1e-5, atol=1e-5 * max(1.0, np.abs(bref).max())), np.abs(bp - bref).max()
    b = fref.astype(np.float32)
    m = _method(res, g, "sirt", num_iterations=10, relaxation_factor=0.5)
    got = _run(m, b, n)
    ref = O.sirt64(g, b.astype(np.float64), 10, 0.5)
    assert np.allclose(got, ref, rtol=1e-4, atol=1e-5 * max(1.0, np.abs(ref).max())), np.abs(got - ref).max()
    m = _method(res, g, "sart", num_iterations=2, relaxation_factor=0.5)
    got = _run(m, b, n)
    ref = O.sart64(g, b.astype(np.float64), 2, 0.5)
    assert np.allclose(got, ref, rtol=1e-4, atol=1e-5 * max(1.0, np.abs(ref).max())), np.abs(got - ref).max()

"""GPU parity of the Joseph projector pair against the ref-compat oracle
(bar: <= 1e-5 relative L2 in FP32, projector-parallel-joseph.cl semantics)."""
import numpy as np
import pytest

import oracle as O
import uir
from phantoms import ellipse_phantom, shepp_logan

pytestmark = pytest.mark.gpu

TOL = 1e-5

SHAPES = [  # (image edge, detectors, angles, step)
    (367, 367, 180, np.float32(np.pi / 180)),     # config-1 shape on the reference D x D grid
    (256, 367, 180, np.float32(np.pi / 180)),     # config-1 native grid
    (1024, 1448, 720, np.float32(np.pi / 720)),   # config-2 / metric shape, native grid
    (1448, 1448, 720, np.float32(np.pi / 720)),   # config-2 reference grid
    (96, 96, 90, np.float32(0.034906585)),        # json-sample step (2 degrees)
    (64, 64, 120, np.float32(2 * np.pi / 120)),   # full circle: five direction runs
]


def _projector(n, d, a, step, forward, **extra):
    p = uir.Task("parallel-projector", step=float(step), is_forward=forward, angles_num=a,
                 image_size=n, detectors_num=d, **extra)
    return p


@pytest.mark.parametrize("n,d,a,step", SHAPES)
def test_fp_matches_oracle(res, n, d, a, step):
    print("case fp", n, d, a, flush=True)
    g = O.Geometry(n, d, a, step)
    x = ellipse_phantom(n, seed=n + a)
    ref = O.fp(g, x)
    p = _projector(n, d, a, step, True).setup(res)
    dx = uir.DeviceArray.from_numpy(x)
    assert p.requisition(dx) == (a, d)
    out = uir.DeviceArray((a, d))
    p.process(dx, out)
    got = out.numpy()
    err = O.rel_l2(got, ref)
    assert err <= TOL, (err, np.abs(got - ref).max())


@pytest.mark.parametrize("n,d,a,step", SHAPES)
def test_bp_matches_oracle(res, n, d, a, step):
    print("case bp", n, d, a, flush=True)
    g = O.Geometry(n, d, a, step)
    sino = O.fp(g, ellipse_phantom(n, seed=3 * n))
    ref = O.bp(g, sino)
    p = _projector(n, d, a, step, False).setup(res)
    ds = uir.DeviceArray.from_numpy(sino)
    assert p.requisition(ds) == (n, n)
    out = uir.DeviceArray((n, n))
    p.process(ds, out)
    got = out.numpy()
    err = O.rel_l2(got, ref)
    assert err <= TOL, (err, np.abs(got - ref).max())


def test_luts_and_runs_match_oracle(res):
    a, step = 720, np.float32(np.pi / 720)
    g = O.Geometry(1448, 1448, a, step)
    p = _projector(1448, 1448, a, step, False).setup(res)
    p.requisition(uir.DeviceArray((a, 1448)))
    s, c = p.luts()
    assert np.array_equal(s, g.sin) and np.array_equal(c, g.cos)
    assert p.subsets() == g.runs == [(0, 181, 1), (181, 359, 0), (540, 180, 1)]


@pytest.mark.parametrize("axis", [180.0, 190.25])
def test_off_centre_axis(res, axis):
    """axis_position != n_dets / 2 switches detectors off in FP and shrinks the
    BP clamp window (projector-parallel-joseph.cl:62-66,180-182)."""
    n = d = 367
    a, step = 180, np.float32(np.pi / 180)
    g = O.Geometry(n, d, a, step, axis=axis)
    x = shepp_logan(n, extent=256)
    p = _projector(n, d, a, step, True, axis_position=axis).setup(res)
    out = uir.DeviceArray((a, d))
    p.process(uir.DeviceArray.from_numpy(x), out)
    ref = O.fp(g, x)
    assert O.rel_l2(out.numpy(), ref) <= TOL
    q = _projector(n, d, a, step, False, axis_position=axis).setup(res)
    vol = uir.DeviceArray((n, n))
    q.process(uir.DeviceArray.from_numpy(ref), vol)
    assert O.rel_l2(vol.numpy(), O.bp(g, ref)) <= TOL


def test_subset_fp_bp_accumulate(res):
    """subset_fp / subset_bp accumulate with correction_scale / relaxation
    (ufo-ir-parallel-projector-task.c:167-195)."""
    n = d = 128
    a, step = 60, np.float32(np.pi / 60)
    g = O.Geometry(n, d, a, step)
    x = ellipse_phantom(n, seed=11)
    base = np.random.default_rng(1).random((a, d)).astype(np.float32)
    p = _projector(n, d, a, step, False).setup(res)
    p.requisition(uir.DeviceArray((a, d)))
    p["correction_scale"] = 2.5
    p["relaxation"] = 0.75
    dx = uir.DeviceArray.from_numpy(x)
    ds = uir.DeviceArray.from_numpy(base)
    for sub in g.angle_subsets[::7]:
        p.subset_fp(dx, ds, sub)
    ref = base.copy()
    for sub in g.angle_subsets[::7]:
        ref = O.fp(g, x, sino=ref, corr=2.5, subsets=[sub])
    assert O.rel_l2(ds.numpy(), ref) <= TOL
    vol = uir.DeviceArray.from_numpy(x)
    for sub in g.runs:
        p.subset_bp(vol, ds, sub)
    vref = O.bp(g, ref, vol=x, relax=0.75)
    assert O.rel_l2(vol.numpy(), vref) <= TOL

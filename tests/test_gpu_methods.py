"""GPU parity of the method tasks against the ref-compat oracle
(bar: <= 1e-4 relative L2 after a fixed iteration count)."""
import numpy as np
import pytest

import oracle as O
import uir
from phantoms import ellipse_phantom, shepp_logan

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _setup(res, method, n, d, a, step, **props):
    proj = uir.Task("parallel-projector", step=float(step), is_forward=False, image_size=n)
    m = uir.Task(method, projector=proj, **props).setup(res)
    return m, proj


def _run(m, sino, n):
    ds = uir.DeviceArray.from_numpy(sino)
    assert m.requisition(ds) == (n, n)
    out = uir.DeviceArray((n, n))
    m.process(ds, out)
    return out.numpy(), ds


@pytest.mark.parametrize("n,d,a,iters,relax,posc", [
    (367, 367, 180, 50, 1.0, True),      # config 1 on the reference grid
    (256, 367, 180, 20, 1.0, True),      # config 1 native grid
    (1024, 1448, 720, 3, 0.25, False),   # metric shape
])
def test_sirt(res, n, d, a, iters, relax, posc):
    print("case sirt", n, d, a, iters, flush=True)
    step = np.float32(np.pi / a)
    g = O.Geometry(n, d, a, step)
    b = O.fp(g, shepp_logan(n, extent=min(n, 256)) if n <= 367 else ellipse_phantom(n, seed=5))
    ref = O.sirt(g, b, iters, relax, posc)
    m, _ = _setup(res, "sirt", n, d, a, step, num_iterations=iters, relaxation_factor=relax,
                  positive_constraint=posc)
    got, _ = _run(m, b, n)
    assert O.rel_l2(got, ref) <= TOL


@pytest.mark.parametrize("n,d,a,iters,relax", [(128, 128, 90, 3, 0.5), (96, 136, 60, 2, 0.25)])
def test_sart(res, n, d, a, iters, relax):
    step = np.float32(np.pi / a)
    g = O.Geometry(n, d, a, step)
    b = O.fp(g, ellipse_phantom(n, seed=9))
    ref = O.sart(g, b, iters, relax)
    m, _ = _setup(res, "sart", n, d, a, step, num_iterations=iters, relaxation_factor=relax)
    got, _ = _run(m, b, n)
    assert O.rel_l2(got, ref) <= TOL


@pytest.mark.parametrize("n,d,a,axis,iters", [
    (1100, 1556, 36, -1.0, 2),   # rays of 18 blocks: two passes of a wave's 16
    (97, 137, 30, -1.0, 3),      # odd edge: partial last block, per-pixel BP stores
    (128, 128, 90, 70.3, 2),     # off-centre axis: inactive detectors, shrunken BP window
    (64, 91, 24, -1.0, 4),
])
def test_sart_bit_exact(res, n, d, a, axis, iters):
    """The SART sweep's single-angle kernels (wave-per-ray FP from the block
    start table, four-pixel BP) run the reference's operations in its order:
    the task equals the oracle bit for bit."""
    print("case sart bit-exact", n, d, a, axis, flush=True)
    step = np.float32(np.pi / a)
    g = O.Geometry(n, d, a, step, axis)
    b = O.fp(g, ellipse_phantom(n, seed=19))
    b = (b + np.random.default_rng(19).normal(0, 0.01 * b.max(), b.shape)).astype(np.float32)
    ref = O.sart(g, b, iters, 0.5)
    proj = uir.Task("parallel-projector", step=float(step), is_forward=False, image_size=n, axis_position=axis)
    m = uir.Task("sart", projector=proj, num_iterations=iters, relaxation_factor=0.5).setup(res)
    got, _ = _run(m, b, n)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("df", ["sart", "sirt"])
def test_asdpocs(res, df):
    n = d = 96
    a, step = 45, np.float32(np.pi / 45)
    g = O.Geometry(n, d, a, step)
    b = O.fp(g, ellipse_phantom(n, seed=4))
    ref, diag = O.asdpocs(g, b, 4, df=df, df_iters=2, diag=True)
    proj = uir.Task("parallel-projector", step=float(step), is_forward=False)
    dfm = uir.Task(df, num_iterations=2)
    m = uir.Task("asdpocs", projector=proj, df_minimizer=dfm, num_iterations=4).setup(res)
    got, _ = _run(m, b, n)
    assert O.rel_l2(got, ref) <= TOL
    gd = m.asdpocs_diagnostics()
    assert gd.shape == (4, 4)
    # dd, dp, dg are L1 norms (FP64 on the device vs sequential float on the host)
    assert np.allclose(gd, diag, rtol=1e-3, atol=1e-4), (gd, diag)


def test_sbtv(res):
    n = d = 64
    a, step = 45, np.float32(np.pi / 45)
    g = O.Geometry(n, d, a, step)
    b = O.fp(g, ellipse_phantom(n, seed=21))
    ref, bnorm = O.sbtv(g, b, 2, lam=0.1, mu=0.5)
    m, _ = _setup(res, "sbtv", n, d, a, step, num_iterations=2, **{"lambda": 0.1, "mu": 0.5})
    got, ds = _run(m, b, n)
    assert O.rel_l2(ds.numpy(), bnorm) <= 1e-6   # input normalised in place
    assert O.rel_l2(got, ref) <= TOL


@pytest.mark.parametrize("df", ["sart", "sirt"])
def test_asdpocs_textbook(res, df):
    """compat = false: ASD-POCS as published (warm-started data step, L2
    distances, ng normalised TV steps, output after TV) against the oracle's
    textbook variant."""
    n = d = 96
    a, step = 45, np.float32(np.pi / 45)
    g = O.Geometry(n, d, a, step)
    rng = np.random.default_rng(31)
    b = O.fp(g, ellipse_phantom(n, seed=4)) + rng.normal(0, 0.5, (a, d)).astype(np.float32)
    ref, diag = O.asdpocs(g, b, 5, df=df, df_iters=1, diag=True, ng=10, textbook=True, beta=0.9, beta_red=0.9)
    proj = uir.Task("parallel-projector", step=float(step), is_forward=False)
    dfm = uir.Task(df, num_iterations=1)
    m = uir.Task("asdpocs", projector=proj, df_minimizer=dfm, num_iterations=5, compat=False, ng=10, beta=0.9,
                 beta_red=0.9).setup(res)
    got, _ = _run(m, b, n)
    assert O.rel_l2(got, ref) <= TOL
    assert np.allclose(m.asdpocs_diagnostics(), diag, rtol=1e-4), (m.asdpocs_diagnostics(), diag)
    # and it is not the reference's result
    m2 = uir.Task("asdpocs", projector=proj, df_minimizer=uir.Task(df, num_iterations=1), num_iterations=5,
                  ng=10, beta=0.9, beta_red=0.9).setup(res)
    got2, _ = _run(m2, b, n)
    assert O.rel_l2(got2, ref) > 1e-2


def test_sbtv_true_dot(res):
    """compat = false: CGS with an inner product over every element."""
    n = d = 64
    a, step = 45, np.float32(np.pi / 45)
    g = O.Geometry(n, d, a, step)
    b = O.fp(g, ellipse_phantom(n, seed=21))
    ref, _ = O.sbtv(g, b, 2, lam=0.1, mu=0.5, dot_mode=1)
    m, _ = _setup(res, "sbtv", n, d, a, step, num_iterations=2, compat=False, **{"lambda": 0.1, "mu": 0.5})
    got, _ = _run(m, b, n)
    assert O.rel_l2(got, ref) <= TOL
    compat_ref, _ = O.sbtv(g, b, 2, lam=0.1, mu=0.5)
    assert O.rel_l2(got, compat_ref) > 10 * TOL

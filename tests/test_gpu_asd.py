"""Batched ASD-POCS (uir_batch_asdpocs): both variants against the FP64
oracle (ufo-ir-asdpocs-task.c semantics for compat, the published method
otherwise), per-slice diagnostics, independence of a slice's result from the
batch mix, graph replay equal to direct launches, and the single-angle
projector launches against the FP64 operators."""
import numpy as np
import pytest

import oracle as O
import uir
from phantoms import ellipse_phantom

pytestmark = pytest.mark.gpu

GEOMS = [  # (n, n_dets, n_angles): uneven sizes and angle counts in one batch
    (96, 136, 40),
    (64, 91, 24),
    (128, 182, 30),
]
RMIN = 0.1


def _spec(n, d, a, relax=0.25):
    return dict(n=n, n_dets=d, n_angles=a, step=float(np.float32(np.pi / a)), axis_position=-1.0,
                relaxation=relax, positive_constraint=False, ray_weight_min=RMIN)


def _data(n, d, a, seed, noise=0.02):
    g = O.Geometry(n, d, a, np.float32(np.pi / a))
    b = O.fp64(g, ellipse_phantom(n, seed=seed).astype(np.float64))
    b = b + np.random.default_rng(seed).normal(0, noise * b.max(), b.shape)
    return g, b.astype(np.float32)


def _batch(res, storage, geoms=GEOMS, seed0=500):
    bt = uir.Batch(res, [_spec(*gm) for gm in geoms], storage)
    data = []
    for i, gm in enumerate(geoms):
        g, b = _data(*gm, seed0 + i)
        bt.set_sinogram(i, b)
        data.append((g, b))
    return bt, data


KW = dict(beta=0.9, beta_red=0.9, alpha=0.2, alpha_red=0.95, r_max=0.95)


@pytest.mark.parametrize("compat,storage,tol", [
    (True, "f32", 1e-4), (False, "f32", 1e-4), (True, "f16-sino", 1e-3),
    # The published variant's normalised TV steps (smoothed TV, eps 1e-3) turn
    # in flat regions on differences far below the data: the FP16 rounding of
    # the SART residual rows (5e-4 relative) grows to a few 1e-3 in 4 iterations.
    (False, "f16-sino", 1e-2)])
def test_batch_asdpocs_matches_fp64(res, compat, storage, tol):
    """The oracle runs on the sinogram the batch holds (rounded to FP16 for
    f16-sino): what is measured is the arithmetic, not the data's rounding."""
    iters, ng = 4, 10
    bt, data = _batch(res, storage)
    bt.asdpocs(iters, ng=ng, compat=compat, **KW)
    for i, (g, b) in enumerate(data):
        held = b.astype(np.float16) if storage == "f16-sino" else b
        ref, diag = O.asdpocs64(g, held.astype(np.float64), iters, ng=ng, df_iters=1, textbook=not compat,
                                ray_weight_min=RMIN, df_relax_initial=0.25, diag=True, **KW)
        got = bt.image(i)
        err = O.rel_l2(got, ref)
        print("compat" if compat else "textbook", storage, GEOMS[i], "rel-L2 vs FP64", err, flush=True)
        assert err <= tol
        gd = bt.asdpocs_diagnostics(i)
        assert gd.shape == (iters, 4)
        assert np.allclose(gd, diag, rtol=max(20 * tol, 2e-3), atol=1e-6), (gd, diag)


# The published variant's normalised TV steps amplify the data step's FP32
# differences (the line model's chord weights scale coordinate rounding by
# 1 / (M m)): 1e-5 to 1e-4 after 4 iterations, against 2e-6 to 4e-6 for the
# reference variant on the same slices.
@pytest.mark.parametrize("compat,storage,tol", [(True, "f32", 1e-4), (False, "f32", 3e-4), (True, "f16-sino", 1e-3)])
def test_batch_asdpocs_line_model_matches_fp64(res, compat, storage, tol):
    """A batch of the line model: its SART data step runs the model's FP rows
    and BP update per angle over every slice (uirk_sart_line), its dd the
    model's FP table; each slice against the oracle's FP64 ASD-POCS of the
    line model, diagnostics included."""
    iters, ng = 4, 10
    geoms = GEOMS + [(97, 137, 33)]
    print("case batch asdpocs line", compat, storage, geoms, flush=True)
    bt = uir.Batch(res, [dict(_spec(*gm), model="line") for gm in geoms], storage)
    data = []
    for i, (n, d, a) in enumerate(geoms):
        g = O.Geometry(n, d, a, np.float32(np.pi / a), model="line")
        b = O.fp64(g, ellipse_phantom(n, seed=540 + i).astype(np.float64))
        b = (b + np.random.default_rng(540 + i).normal(0, 0.02 * b.max(), b.shape)).astype(np.float32)
        bt.set_sinogram(i, b)
        data.append((g, b))
    bt.asdpocs(iters, ng=ng, compat=compat, **KW)
    for i, (g, b) in enumerate(data):
        held = b.astype(np.float16) if storage == "f16-sino" else b
        ref, diag = O.asdpocs64(g, held.astype(np.float64), iters, ng=ng, df_iters=1, textbook=not compat,
                                ray_weight_min=RMIN, df_relax_initial=0.25, diag=True, **KW)
        err = O.rel_l2(bt.image(i), ref)
        print("line", "compat" if compat else "textbook", storage, geoms[i], "rel-L2 vs FP64 %.3e" % err,
              flush=True)
        assert err <= tol
        assert np.allclose(bt.asdpocs_diagnostics(i), diag, rtol=max(20 * tol, 2e-3), atol=1e-6)


def test_batch_asdpocs_slice_independent_of_mix(res):
    """A slice's image is the same alone and in a batch of other geometries."""
    bt, _ = _batch(res, "f32")
    bt.asdpocs(3, compat=False, ng=8, **KW)
    together = [bt.image(i) for i in range(len(GEOMS))]
    for i, gm in enumerate(GEOMS):
        alone, _ = _batch(res, "f32", geoms=[gm], seed0=500 + i)
        alone.asdpocs(3, compat=False, ng=8, **KW)
        assert np.array_equal(alone.image(0), together[i]), GEOMS[i]


def test_batch_asdpocs_graph_equals_direct(res):
    bt, _ = _batch(res, "f32")
    bt.asdpocs(3, use_graph=False, compat=False, ng=6, **KW)
    direct = [bt.image(i) for i in range(len(GEOMS))]
    bt.asdpocs(3, use_graph=True, compat=False, ng=6, **KW)
    for i in range(len(GEOMS)):
        assert np.array_equal(bt.image(i), direct[i])
    # a second replay of the same graph starts from zero again
    bt.asdpocs(3, use_graph=True, compat=False, ng=6, **KW)
    for i in range(len(GEOMS)):
        assert np.array_equal(bt.image(i), direct[i])


def test_batch_asdpocs_without_tv_is_sart(res):
    """alpha = 0 and no positivity: the textbook loop is warm-started SART
    with the beta schedule, i.e. plain SART when beta_red = 1."""
    bt, data = _batch(res, "f32")
    bt.asdpocs(3, compat=False, alpha=0.0, beta=0.5, beta_red=1.0, positive_constraint=False)
    for i, (g, b) in enumerate(data):
        ref = O.asdpocs64(g, b.astype(np.float64), 3, alpha=0.0, beta=0.5, beta_red=1.0, positive_constraint=False,
                          df_iters=1, textbook=True, ray_weight_min=RMIN)
        assert O.rel_l2(bt.image(i), ref) <= 2e-5


@pytest.mark.parametrize("n,d,a", [(1024, 1448, 120)])
def test_batch_asdpocs_sparse_view_stack_shape(res, n, d, a):
    """The sparse-view stack geometry (1024^2, 120 angles, 1448 bins), two
    iterations of the textbook variant against FP64.  alpha = 0.05: at the
    default 0.2 the normalised TV steps on this image are chaotic (the FP64
    method itself turns a 1e-7 data perturbation into 3e-4 after two
    iterations), so FP32 rounding alone moves the result by ~2e-3."""
    g, b = _data(n, d, a, 77, noise=0.01)
    kw = dict(KW, alpha=0.05)
    bt = uir.Batch(res, [_spec(n, d, a)], "f32")
    bt.set_sinogram(0, b)
    bt.asdpocs(2, compat=False, **kw)
    ref = O.asdpocs64(g, b.astype(np.float64), 2, df_iters=1, textbook=True, ray_weight_min=RMIN, **kw)
    err = O.rel_l2(bt.image(0), ref)
    print("1024^2 x 120 rel-L2 vs FP64", err, flush=True)
    assert err <= 1e-4


def _poisson_data(n, d, a, seed, i0=5e3, mu=0.01):
    """Sparse-view stack data: Joseph projection of a random-ellipse slice,
    Poisson counts at I0 with mu attenuation per pixel unit, back to line
    integrals."""
    g = O.Geometry(n, d, a, np.float32(np.pi / a))
    p = O.fp64(g, ellipse_phantom(n, seed=seed).astype(np.float64))
    counts = np.maximum(np.random.default_rng(seed).poisson(i0 * np.exp(-mu * p)), 1)
    return g, (-np.log(counts / i0) / mu).astype(np.float32)


@pytest.mark.parametrize("compat,storage,alpha,iters,tol", [
    # the reference's variant at the default alpha: its output is POSC of the
    # last SART restart (ufo-ir-asdpocs-task.c:457-468), not chaotic
    (True, "f32", 0.2, 3, 1e-4),
    (True, "f16-sino", 0.2, 3, 1e-3),
    # the published variant (alpha 0.05: see the previous test) on FP16
    # sinograms: its L2-normalised TV steps turn, in flat regions, on
    # differences far below the FP16 rounding of the SART residual rows
    # (measured 3.7e-3 here; the same bar as test_batch_asdpocs_matches_fp64)
    (False, "f16-sino", 0.05, 2, 1e-2),
])
def test_batch_asdpocs_sparse_view_stack_variants(res, compat, storage, alpha, iters, tol):
    """Both variants at 1024^2, 120 angles, 1448 bins on Poisson data
    (I0 = 5e3), in FP32 and with FP16 sinograms, against FP64 from the FP32
    data handed to the batch."""
    n, d, a = 1024, 1448, 120
    print("case asdpocs stack shape", "compat" if compat else "textbook", storage, "alpha", alpha, flush=True)
    g, b = _poisson_data(n, d, a, 91)
    kw = dict(KW, alpha=alpha)
    bt = uir.Batch(res, [_spec(n, d, a)] * 2, storage)
    bt.set_sinogram(0, b)
    bt.set_sinogram(1, b[::-1].copy())
    bt.asdpocs(iters, compat=compat, **kw)
    ref = O.asdpocs64(g, b.astype(np.float64), iters, df_iters=1, textbook=not compat, ray_weight_min=RMIN,
                      df_relax_initial=0.25, **kw)
    err = O.rel_l2(bt.image(0), ref)
    print("rel-L2 vs FP64 %.3e, PSNR %.1f dB" % (err, O.psnr(bt.image(0), ref)), flush=True)
    assert err <= tol


@pytest.mark.parametrize("storage,tol", [("f16", 2e-2), ("bf16", 1e-1)])
def test_batch_asdpocs_half_images(res, storage, tol):
    """FP16 / BF16 image storage runs the same kernels on half-precision
    images (the second SART buffer then holds half data too); the iterate's
    rounding at every SART angle random-walks, hence the wide bar."""
    bt, data = _batch(res, storage)
    bt.asdpocs(3, compat=False, ng=6, **KW)
    for i, (g, b) in enumerate(data):
        held = b.astype(np.float16) if storage == "f16" else b
        ref = O.asdpocs64(g, held.astype(np.float64), 3, ng=6, df_iters=1, textbook=True, ray_weight_min=RMIN,
                          df_relax_initial=0.25, **KW)
        err = O.rel_l2(bt.image(i), ref)
        print(storage, GEOMS[i], "rel-L2 vs FP64", err, flush=True)
        assert err <= tol


@pytest.mark.parametrize("shift", [4.5, -6.0, 3.3])
def test_batch_asdpocs_shifted_axis(res, shift):
    """The fused SART step and the residual rows with axis_position off the
    detector centre (inactive detectors, shrunken BP window)."""
    n, d, a = GEOMS[0]
    axis = d / 2.0 + shift
    g = O.Geometry(n, d, a, np.float32(np.pi / a), axis)
    b = O.fp64(g, ellipse_phantom(n, seed=801).astype(np.float64))
    b = (b + np.random.default_rng(5).normal(0, 0.02 * b.max(), b.shape)).astype(np.float32)
    bt = uir.Batch(res, [dict(_spec(n, d, a), axis_position=axis)], "f32")
    bt.set_sinogram(0, b)
    bt.asdpocs(3, compat=False, ng=8, **KW)
    ref = O.asdpocs64(g, b.astype(np.float64), 3, ng=8, df_iters=1, textbook=True, ray_weight_min=RMIN,
                      df_relax_initial=0.25, **KW)
    err = O.rel_l2(bt.image(0), ref)
    print("axis shift", shift, "rel-L2 vs FP64", err, flush=True)
    assert err <= 1e-4

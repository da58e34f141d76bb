"""Pins of the CPU oracle (oracle/uir_oracle.c) by facts that do not come from
retyping its formulas: closed forms of the geometry, analytic line integrals,
mathematical identities of the operators, dense-matrix restatements of the
method loops and an independent solver.  The reference has no tests, fixtures
or golden vectors of its own and its OpenCL kernels cannot run here, so these
are what pin the restatement.  Each test names the plausible mistake it
catches."""
import numpy as np
import pytest

import oracle as O
from phantoms import gaussian_blobs, gaussian_blobs_radon

PI = np.pi


def geom(n, d, a, step=None, axis=-1.0):
    return O.Geometry(n, d, a, np.float32(PI / a) if step is None else step, axis)


# ---------------------------------------------------------------- geometry --

@pytest.mark.parametrize("a,runs", [
    (180, [(0, 46, 1), (46, 89, 0), (135, 45, 1)]),
    (720, [(0, 181, 1), (181, 359, 0), (540, 180, 1)]),
    (120, [(0, 31, 1), (31, 59, 0), (90, 30, 1)]),
])
def test_direction_runs_closed_form(a, runs):
    """|sin| <= |cos| (vertical) exactly on [0, 45] and [135, 180) degrees,
    ties at 45 / 135 going to vertical (catches a swapped comparison or a
    strict inequality)."""
    assert geom(16, 16, a).runs == runs


def test_json_sample_step_runs():
    """json-sample.json:29 uses step 0.034906585 (2 degrees) and 90 angles."""
    g = O.Geometry(16, 16, 90, np.float32(0.034906585))
    assert g.runs == [(0, 23, 1), (23, 45, 0), (68, 22, 1)]


def test_lut_is_float_product_then_double_trig():
    """create_lut_buffer: rad = (float)i * step in float, then sin/cos in double
    rounded to float (catches a double product)."""
    a = 720
    step = np.float32(PI / a)
    g = geom(16, 16, a, step)
    i = np.arange(a, dtype=np.float32)
    want = np.sin((i * step).astype(np.float64)).astype(np.float32)
    assert np.array_equal(g.sin, want)
    dbl = np.sin(np.arange(a, dtype=np.float64) * np.float64(step)).astype(np.float32)
    assert not np.array_equal(want, dbl)   # the two conventions do differ here


def test_axis_default_is_half_n_dets():
    assert geom(8, 367, 10).axis == np.float32(183.5)
    assert geom(8, 367, 10, axis=100.25).axis == np.float32(100.25)


# --------------------------------------------------------- forward project --

def test_fp_theta0_is_column_sums():
    """At theta = 0 the vertical-family ray of detector s walks column s at
    pixel centres: FP = column sums (catches a transposed image or a shifted
    detector)."""
    n = 96
    g = geom(n, n, 4, step=np.float32(PI / 4))
    x = np.random.default_rng(0).random((n, n)).astype(np.float32)
    s = O.fp(g, x)
    assert np.allclose(s[0], x.astype(np.float64).sum(0), rtol=2e-6)


def test_fp_theta90_is_reversed_row_sums():
    """At theta = 90 degrees the horizontal-family ray of detector s walks row
    n-1-s (catches a sign error in the detector step)."""
    n = 96
    g = geom(n, n, 2)
    x = np.random.default_rng(1).random((n, n)).astype(np.float32)
    s = O.fp(g, x)
    assert np.allclose(s[1], x.astype(np.float64).sum(1)[::-1], rtol=1e-5)


def test_fp_inactive_detectors_with_shifted_axis():
    """axis_position a != n_dets / 2: detectors s > 2a produce nothing
    (projector-parallel-joseph.cl:62-66)."""
    n = 64
    g = geom(n, n, 30, axis=25.0)
    s = O.fp(g, np.ones((n, n), np.float32))
    assert np.all(s[:, 51:] == 0) and np.any(s[:, :50] != 0)


@pytest.mark.parametrize("n,d,a", [(128, 182, 60), (96, 96, 45)])
def test_fp_matches_analytic_radon_of_gaussians(n, d, a):
    """Joseph sums of a smooth image equal max(|sin|, |cos|) times its line
    integral along X cos - Y sin = t', with t' = (s + 0.5 - axis) + 0.5 cos
    (horizontal family) or - 0.5 sin (vertical family): the half-pixel offset
    of the reference's start coordinate.  Catches a wrong sign of theta, a
    swapped X / Y, a wrong axis or the missing half-pixel term."""
    blobs = [(10.0, -7.0, 6.0, 1.0), (-20.0, 15.0, 8.0, 0.5), (25.0, 22.0, 5.0, 0.7)]
    g = geom(n, d, a)
    img = gaussian_blobs(n, blobs).astype(np.float32)
    s = O.fp(g, img).astype(np.float64)
    th = np.arange(a) * np.float64(np.float32(PI / a))
    sn, cs = np.sin(th), np.cos(th)
    vert = np.abs(sn) <= np.abs(cs)
    t = np.arange(d) + 0.5 - g.axis
    off = np.where(vert, -0.5 * sn, 0.5 * cs)
    m = np.maximum(np.abs(sn), np.abs(cs))
    want = m[:, None] * gaussian_blobs_radon(blobs, th, t[None, :] + off[:, None])
    err = O.rel_l2(s, want)
    assert err < 1e-2       # Joseph's linear-interpolation error at these widths
    # without the half-pixel term, with its sign flipped or with theta negated
    # the analytic match is clearly worse
    for wrong_off, sign in ((np.zeros(a), 1), (-off, 1), (off, -1)):
        bad = m[:, None] * gaussian_blobs_radon(blobs, sign * th, t[None, :] + wrong_off[:, None])
        assert O.rel_l2(s, bad) > 2.5 * err


def test_fp_linear_and_fp64_agree():
    """FP is linear; the FP32 restatement and the FP64 evaluation agree to
    float rounding (catches a dropped term in either)."""
    n, d, a = 64, 91, 37
    g = geom(n, d, a)
    rng = np.random.default_rng(2)
    x, y = rng.random((n, n)).astype(np.float32), rng.random((n, n)).astype(np.float32)
    lhs = O.fp(g, 2 * x + 3 * y)
    rhs = 2 * O.fp(g, x).astype(np.float64) + 3 * O.fp(g, y)
    assert O.rel_l2(lhs, rhs) < 1e-5
    assert O.rel_l2(O.fp(g, x), O.fp64(g, x)) < 2e-5


# ------------------------------------------------------------ backproject --

def test_bp_of_ones_is_angle_count():
    """BP(1) = A exactly: the clamp keeps both taps on the detector
    (projector-parallel-joseph.cl:203-212)."""
    g = geom(80, 113, 57)
    v = O.bp(g, np.ones((57, 113), np.float32))
    assert np.all(v == 57.0)


def test_bp_theta0_broadcasts_bin_x():
    """One angle at theta = 0: pixel (x, y) receives bin x (catches x / y swap)."""
    n = 48
    g = geom(n, n, 4, step=np.float32(PI / 4))
    row = np.random.default_rng(3).random(n).astype(np.float32)
    sino = np.zeros((4, n), np.float32)
    sino[0] = row
    v = O.bp(g, sino, subsets=[(0, 1, 1)])
    assert np.array_equal(v, np.tile(row, (n, 1)))


def test_bp_matches_analytic_profile():
    """BP of a sinogram sampled from smooth profiles G_theta(t) equals
    sum_theta G_theta(X cos - Y sin) (catches a wrong sign or axis in the
    detector coordinate of a pixel)."""
    n, d, a = 96, 136, 48
    g = geom(n, d, a)
    th = np.arange(a) * np.float64(np.float32(PI / a))
    t = np.arange(d) + 0.5 - g.axis
    c = 8.0 * np.sin(3 * th)[:, None]
    sino = np.exp(-((t[None, :] - c) ** 2) / (2 * 9.0 ** 2))
    v = O.bp(g, sino.astype(np.float32)).astype(np.float64)
    X = np.arange(n) + 0.5 - n / 2
    XX, YY = np.meshgrid(X, X)
    want = np.zeros((n, n))
    for k in range(a):
        u = XX * np.cos(th[k]) - YY * np.sin(th[k])
        want += np.exp(-((u - c[k, 0]) ** 2) / (2 * 9.0 ** 2))
    assert O.rel_l2(v, want) < 2e-3


def test_bp_clamps_to_edge_bins():
    """Pixels projecting beyond the detector read the edge bins (the reference
    clamps the coordinate instead of using the border colour)."""
    n, d = 64, 40
    g = geom(n, d, 3, step=np.float32(PI / 3))
    sino = np.zeros((3, d), np.float32)
    sino[0] = np.linspace(1.0, 2.0, d, dtype=np.float32)
    v = O.bp(g, sino, subsets=[(0, 1, 1)])
    assert np.all(v[:, :10] == sino[0, 0]) and np.all(v[:, -10:] == sino[0, -1])


def test_fp_bp_nearly_adjoint_for_smooth_data():
    """The reference FP sums samples without ray-length weighting, so per angle
    it is max(|sin|, |cos|) times the line integral while BP is the plain
    transpose of the line integral: <FP x, y> ~ <x, BP(m y)>.  The pair is not
    exactly adjoint (half-pixel offset), so equality is approximate, but a
    sign or orientation error, or a ray-length weighting, breaks it."""
    n, d, a = 96, 136, 60
    g = geom(n, d, a)
    x = gaussian_blobs(n, [(8.0, -5.0, 9.0, 1.0), (-15.0, 10.0, 7.0, 0.6)])
    th = np.arange(a) * np.float64(np.float32(PI / a))
    m = np.maximum(np.abs(np.sin(th)), np.abs(np.cos(th)))
    t = np.arange(d) + 0.5 - g.axis
    y = np.exp(-((t[None, :] - 10 * np.cos(th)[:, None]) ** 2) / (2 * 12.0 ** 2))
    lhs = float(np.sum(O.fp(g, x.astype(np.float32)).astype(np.float64) * y))
    rhs = float(np.sum(x * O.bp(g, (m[:, None] * y).astype(np.float32))))
    assert abs(lhs - rhs) < 5e-3 * abs(lhs)
    plain = float(np.sum(x * O.bp(g, y.astype(np.float32))))
    assert abs(lhs - plain) > 5 * abs(lhs - rhs)


# ------------------------------------------------- reductions and stencils --

def test_dot_compat_closed_form():
    """ufo-ir-basic-ops-processor.c:186-193 sums a[p*i] b[p*i] over p < H,
    i < W: for a[k] = b[k] = k that is (sum p^2)(sum i^2), not sum k^2."""
    w, h = 12, 9
    a = np.arange(w * h, dtype=np.float32).reshape(h, w)
    want = float(np.sum(np.arange(h) ** 2) * np.sum(np.arange(w) ** 2))
    assert O.dot_compat(a, a) == pytest.approx(want, rel=1e-6)
    assert want != float(np.sum(a.astype(np.float64) ** 2))
    assert O.l2_compat(a) == pytest.approx(np.sqrt(want), rel=1e-6)


def test_l1_norm_and_normalize():
    rng = np.random.default_rng(4)
    v = rng.standard_normal((33, 47)).astype(np.float32)
    assert O.l1_norm(v) == pytest.approx(float(np.abs(v.astype(np.float64)).sum()), rel=1e-5)
    nv = O.normalize(v)
    assert nv.min() == pytest.approx(0.0, abs=1e-7) and nv.max() == pytest.approx(1.0, abs=1e-6)
    ref = (v.astype(np.float64) - v.min()) / (v.max() - v.min())
    assert np.allclose(nv, ref, atol=1e-6)


def test_difference_operators_are_mutual_adjoints():
    """Dxt / Dyt are the exact transposes of the periodic Dx / Dy
    (ufo-ir-gradient-processor.cl); catches a wrong wrap index."""
    rng = np.random.default_rng(5)
    u = rng.standard_normal((13, 17)).astype(np.float32)
    v = rng.standard_normal((13, 17)).astype(np.float32)
    for fwd, adj in ((O.dx, O.dxt), (O.dy, O.dyt)):
        lhs = float(np.sum(fwd(u).astype(np.float64) * v))
        rhs = float(np.sum(u.astype(np.float64) * adj(v)))
        assert lhs == pytest.approx(rhs, rel=1e-5)
    assert np.allclose(O.dx(u), u - np.roll(u, 1, axis=1), atol=1e-6)
    assert np.allclose(O.dy(u), u - np.roll(u, 1, axis=0), atol=1e-6)


def smoothed_tv(img, eps=1e-3):
    """TV(f) = sum sqrt(eps + (f(x,y) - f(x-1,y))^2 + (f(x,y) - f(x,y+1))^2)
    over the image grown by one zero row above and one zero column to the
    right (FP64)."""
    h, w = img.shape
    p = np.zeros((h + 2, w + 2))
    p[1:-1, 1:-1] = img
    c = p[0:h + 1, 1:w + 2]                       # rows y = -1 .. h-1, columns x = 0 .. w
    return float(np.sum(np.sqrt(eps + (c - p[0:h + 1, 0:w + 1]) ** 2 + (c - p[1:h + 2, 1:w + 2]) ** 2)))


def tv_gradient_fd(f, hstep=1e-6):
    """Central finite differences of smoothed_tv."""
    num = np.zeros_like(f)
    for y in range(f.shape[0]):
        for x in range(f.shape[1]):
            e = np.zeros_like(f)
            e[y, x] = hstep
            num[y, x] = (smoothed_tv(f + e) - smoothed_tv(f - e)) / (2 * hstep)
    return num


@pytest.mark.parametrize("which", ["fp32", "fp64"])
def test_l1_grad_is_gradient_of_smoothed_tv(which):
    """l1_grad (ufo-math-tvstd-method.cl) and its FP64 restatement are the
    gradient of the smoothed TV (eps = 1e-3) with a zero border; checked
    against central finite differences in FP64 (catches a swapped neighbour
    or a wrong sign in any of the three terms)."""
    rng = np.random.default_rng(6)
    f = rng.standard_normal((7, 9)) if which == "fp32" else rng.standard_normal((9, 9))
    if which == "fp32":
        g = O.l1_grad(f.astype(np.float32)).astype(np.float64)
        f = f.astype(np.float32).astype(np.float64)
    else:
        g = O.l1_grad64(f)
    num = tv_gradient_fd(f)
    assert np.allclose(g, num, atol=2e-4 if which == "fp32" else 1e-6)


def test_sbtv_shrink_is_isotropic_prox():
    """The update_db shrinkage is the proximal map of ||.||_2 / lambda:
    |d| = max(|t| - 1/lambda, 0) and d parallel to t."""
    rng = np.random.default_rng(7)
    tx, ty = rng.standard_normal(1000).astype(np.float32), rng.standard_normal(1000).astype(np.float32)
    lam = 2.5
    dx, dy = O.sbtv_shrink(tx, ty, lam)
    t = np.hypot(tx.astype(np.float64), ty)
    d = np.hypot(dx.astype(np.float64), dy)
    assert np.allclose(d, np.maximum(t - 1 / lam, 0), atol=1e-6)
    assert np.allclose(dx * ty - dy * tx, 0, atol=1e-6)


# ---------------------------------------------------- methods as matrices --

def dense(g):
    """System matrices of the oracle's FP (rows: rays) and BP (rows: pixels)."""
    n, d, a = g.n, g.n_dets, g.n_angles
    Afp = np.zeros((a * d, n * n))
    for k in range(n * n):
        e = np.zeros(n * n, np.float32)
        e[k] = 1
        Afp[:, k] = O.fp(g, e.reshape(n, n)).ravel()
    Abp = np.zeros((n * n, a * d))
    for k in range(a * d):
        e = np.zeros(a * d, np.float32)
        e[k] = 1
        Abp[:, k] = O.bp(g, e.reshape(a, d)).ravel()
    return Afp, Abp


@pytest.fixture(scope="module")
def tiny():
    g = geom(10, 14, 12)
    Afp, Abp = dense(g)
    x = np.random.default_rng(8).random((10, 10))
    b = Afp @ x.ravel()
    return g, Afp, Abp, b


def test_sirt_loop_as_matrix_iteration(tiny):
    """x <- x + C (relax A^T (R (b - A x))), R = 1/A1, C = 1/A^T1, x0 = 0
    (ufo-ir-sirt-task.c:176-215); catches a misplaced weight or relaxation."""
    g, Afp, Abp, b = tiny
    R = 1 / np.where(Afp.sum(1) != 0, Afp.sum(1), np.inf)
    C = 1 / Abp.sum(1)
    x = np.zeros(Afp.shape[1])
    for _ in range(7):
        x = x + C * (0.7 * (Abp @ (R * (b - Afp @ x))))
    got = O.sirt(g, b.reshape(g.n_angles, g.n_dets).astype(np.float32), 7, 0.7)
    assert O.rel_l2(got.ravel(), x) < 2e-5


def test_sart_loop_as_matrix_iteration(tiny):
    """Per angle i in order: x <- x + relax A_i^T (R_i (b_i - A_i x)), no pixel
    weights (ufo-ir-sart-task.c:193-220)."""
    g, Afp, Abp, b = tiny
    d = g.n_dets
    R = 1 / np.where(Afp.sum(1) != 0, Afp.sum(1), np.inf)
    x = np.zeros(Afp.shape[1])
    for _ in range(3):
        for i in range(g.n_angles):
            rows = slice(i * d, (i + 1) * d)
            x = x + 0.5 * (Abp[:, rows] @ (R[rows] * (b[rows] - Afp[rows] @ x)))
    got = O.sart(g, b.reshape(g.n_angles, d).astype(np.float32), 3, 0.5)
    assert O.rel_l2(got.ravel(), x) < 2e-5


@pytest.mark.parametrize("df", ["sart", "sirt"])
def test_asdpocs_output_is_posc_of_last_df_minimizer(df):
    """The output is copied before the TV steps and the df-minimizer restarts
    from 0 (ufo-ir-asdpocs-task.c:460-492): output = POSC(df(b; beta_{K-1}))."""
    g = geom(32, 46, 24)
    rng = np.random.default_rng(9)
    b = O.fp(g, rng.random((32, 32)).astype(np.float32)) + rng.normal(0, 0.3, (24, 46)).astype(np.float32)
    K, beta, red = 4, 0.9, 0.8
    out = O.asdpocs(g, b, K, beta=beta, beta_red=red, df=df, df_iters=3)
    relax = np.float32(beta)
    for _ in range(K - 1):
        relax = np.float32(relax * np.float32(red))
    last = O.sart(g, b, 3, relax) if df == "sart" else O.sirt(g, b, 3, relax)
    assert np.array_equal(out, np.maximum(last, 0))


def asd_problem(n=24, d=34, a=20, seed=11, noise=0.3):
    g = geom(n, d, a)
    rng = np.random.default_rng(seed)
    x = np.zeros((n, n))
    x[n // 4:3 * n // 4, n // 4:3 * n // 4] = 1.0
    x[n // 3:n // 2, n // 3:2 * n // 3] += 0.5
    b = O.fp(g, x.astype(np.float32)) + rng.normal(0, noise, (a, d)).astype(np.float32)
    return g, b


def test_asdpocs_textbook_without_tv_is_warm_sart_chain(tiny):
    """Textbook ASD-POCS with alpha = 0 takes TV steps of length 0, so it is
    the warm-started chain x <- POSC(SART_{beta_k}(x)) with beta_k = beta
    red^k, one matrix-form SART pass per iteration (catches a data step that
    restarts from zero, or a relaxation schedule off by one)."""
    g, Afp, Abp, b = tiny
    d, nn = g.n_dets, g.n * g.n
    rng = np.random.default_rng(12)
    bn = b + rng.normal(0, 0.5, b.shape)
    bs = bn.reshape(g.n_angles, d)

    def chain(Afp, Abp):
        R = 1 / np.where(Afp.sum(1) != 0, Afp.sum(1), np.inf)
        x = np.zeros(Afp.shape[1])
        relax = np.float32(0.9)
        for _ in range(4):
            for i in range(g.n_angles):
                rows = slice(i * d, (i + 1) * d)
                x = x + float(relax) * (Abp[:, rows] @ (R[rows] * (bn[rows] - Afp[rows] @ x)))
            x = np.maximum(x, 0)
            relax = np.float32(relax * np.float32(0.7))
        return x

    x = chain(Afp, Abp)
    got = O.asdpocs(g, bs.astype(np.float32), 4, beta=0.9, beta_red=0.7, alpha=0.0, df_iters=1, textbook=True)
    assert O.rel_l2(got.ravel(), x) < 2e-5
    # the FP64 restatement, against the matrices of its own (exact-coordinate) operators
    A64 = np.stack([O.fp64(g, np.eye(nn)[k].reshape(g.n, g.n)).ravel() for k in range(nn)], 1)
    B64 = np.stack([O.bp64(g, np.eye(g.n_angles * d)[k].reshape(g.n_angles, d)).ravel()
                    for k in range(g.n_angles * d)], 1)
    got64 = O.asdpocs64(g, bs, 4, beta=0.9, beta_red=0.7, alpha=0.0, df_iters=1, textbook=True)
    assert O.rel_l2(got64.ravel(), chain(A64, B64)) < 1e-7   # beta_k: float vs double products
    # the reference variant restarts from zero every iteration: a different image
    ref = O.asdpocs(g, bs.astype(np.float32), 4, beta=0.9, beta_red=0.7, alpha=0.0, df_iters=1)
    assert O.rel_l2(ref.ravel(), x) > 1e-3


def test_asdpocs_textbook_one_step_closed_form():
    """K = 1, ng = 1: x = POSC(SART(0)), dtgv = alpha |x|_2 and the output is
    x - dtgv g / |g|_2 with g the finite-difference gradient of the smoothed
    TV (catches an L1 normalisation, output before the TV step, or dp taken
    from the wrong image)."""
    g, b = asd_problem()
    alpha = 0.3

    def one_step(x):
        grad = tv_gradient_fd(x)
        return x - alpha * np.linalg.norm(x) * grad / np.linalg.norm(grad)

    x = np.maximum(O.sart64(g, b.astype(np.float64), 1, 1.0), 0)
    got64, diag = O.asdpocs64(g, b.astype(np.float64), 1, beta=1.0, alpha=alpha, ng=1, df_iters=1, textbook=True,
                              diag=True)
    assert O.rel_l2(got64.ravel(), one_step(x).ravel()) < 1e-6
    assert abs(diag[0, 1] - np.linalg.norm(x)) < 1e-9 * np.linalg.norm(x)
    x32 = np.maximum(O.sart(g, b, 1, 1.0), 0).astype(np.float64)
    got = O.asdpocs(g, b, 1, beta=1.0, alpha=alpha, ng=1, df_iters=1, textbook=True)
    assert O.rel_l2(got.ravel(), one_step(x32).ravel()) < 1e-4


def test_asdpocs_textbook_tv_steps_have_length_dtgv():
    """Each normalised TV step moves the image by exactly dtgv in L2, so with
    ng = 1 the TV distance dg equals dtgv in every iteration, and with ng = 3
    it is at most 3 dtgv (triangle inequality); dtgv shrinks by alpha_red exactly when dg >
    r_max dp and dd > 1e-3 (catches an unnormalised or L1-normalised step)."""
    g, b = asd_problem(seed=13)
    for ng in (1, 3):
        out, d = O.asdpocs64(g, b.astype(np.float64), 6, beta=0.8, beta_red=0.9, alpha=0.5, ng=ng, df_iters=1,
                             textbook=True, diag=True)
        dd, dp, dg, dtgv = d.T
        if ng == 1:
            assert np.allclose(dg, dtgv, rtol=1e-9)
        else:
            assert np.all(dg <= 3 * dtgv * (1 + 1e-9))
        red = float(np.float32(0.95))          # r_max and alpha_red are float parameters
        shrink = (dg[:-1] > red * dp[:-1]) & (dd[:-1] > 1e-3)
        assert shrink.any() and np.allclose(dtgv[1:], np.where(shrink, dtgv[:-1] * red, dtgv[:-1]), rtol=1e-12)


def test_sbtv_operator_matrix(tiny):
    """processA = mu A^T A + lambda (Dy^T Dy + Dx^T Dx) with periodic
    differences (ufo-ir-sbtv-task.c:587-635)."""
    g, Afp, Abp, _ = tiny
    n = g.n
    lam, mu = 0.3, 0.7
    M = np.zeros((n * n, n * n))
    for k in range(n * n):
        e = np.zeros(n * n, np.float32)
        e[k] = 1
        M[:, k] = O.sbtv_apply(g, e.reshape(n, n), lam, mu).ravel()
    I = np.eye(n)
    Dd = I - np.roll(I, 1, axis=0)            # (D u)[i] = u[i] - u[i-1], periodic
    Dx = np.kron(I, Dd)
    Dy = np.kron(Dd, I)
    want = mu * (Abp @ Afp) + lam * (Dy.T @ Dy + Dx.T @ Dx)
    assert np.allclose(M, want, atol=2e-5 * np.abs(want).max())


def test_cgs_matches_scipy_with_a_true_dot():
    """With the plain dot product the reference's MATLAB-style CGS iterates as
    scipy's CGS does (same recurrence); the reference's own strided "dot" is
    pinned separately by its closed form."""
    from scipy.sparse.linalg import cgs
    g = geom(8, 12, 10)
    n = g.n
    lam, mu = 0.2, 1.0
    M = np.zeros((n * n, n * n))
    for k in range(n * n):
        e = np.zeros(n * n, np.float32)
        e[k] = 1
        M[:, k] = O.sbtv_apply(g, e.reshape(n, n), lam, mu).ravel()
    rhs = np.random.default_rng(10).standard_normal(n * n)
    x0 = np.zeros(n * n)
    # Not converged after `it` iterations, the reference returns its iterate of
    # smallest recurrence residual (xmin, ufo-ir-sbtv-task.c:554-572), like
    # MATLAB's cgs: the output must be one of scipy's CGS iterates x_0..x_it
    # and close to the smallest true residual among them.
    iterates = [x0] + [cgs(M, rhs, x0=x0, rtol=0.0, atol=0.0, maxiter=j)[0] for j in range(1, 7)]
    resid = [np.linalg.norm(rhs - M @ x) for x in iterates]
    assert resid[3] == min(resid[:4])                  # an unambiguous best iterate
    for it in (1, 2, 3, 4, 6):
        got, flag = O.sbtv_cgs(g, rhs.reshape(n, n), x0.reshape(n, n), it, lam, mu, dot_mode=1)
        assert flag != 0
        errs = [O.rel_l2(got.ravel(), x) if np.any(x) else float(np.abs(got).max()) for x in iterates[:it + 1]]
        assert min(errs) < 1e-3, (it, errs)
        assert np.linalg.norm(rhs - M @ got.ravel()) <= 2 * min(resid[:it + 1])
        if it == 3:
            assert int(np.argmin(errs)) == 3


# ------------------- split-Bregman TV with SIRT sweeps (this build's method) --

@pytest.mark.parametrize("model", ["joseph", "line"])
def test_sbsirt_without_tv_is_sirt(model):
    """lambda = 0 removes the TV term: the sweeps are SIRT with relaxation w and
    pixel weight 1/A (joseph) or 1/BP(1) with the line SIRT's floor (catches
    a wrong preconditioner or a missing R)."""
    g = O.Geometry(24, 34, 30, np.float32(PI / 30), -1.0, model=model)
    f = O.fp64(g, gaussian_blobs(24, [(-3.0, 2.0, 4.0, 1.0), (5.0, -4.0, 2.5, 0.6)]))
    got = O.sbsirt64(g, f, 4, 3, relax=0.8, lam=0.0, mu=2.0)
    want = O.sirt64(g, f, 12, relax=0.8)
    assert O.rel_l2(got, want) < 1e-12


def test_sb_tv_is_negative_gradient_of_coupling_term():
    """tv = D^T (e - D u) is minus the gradient of 1/2 |e - D u|^2 with the
    periodic backward differences of ufo-ir-gradient-processor.cl (catches a
    transposed or sign-flipped difference)."""
    rng = np.random.default_rng(11)
    n = 6
    u, ex, ey = (rng.standard_normal((n, n)) for _ in range(3))

    def J(v):
        dx = v - np.roll(v, 1, axis=1)
        dy = v - np.roll(v, 1, axis=0)
        return 0.5 * (np.sum((ex - dx) ** 2) + np.sum((ey - dy) ** 2))

    num = np.zeros((n, n))
    h = 1e-6
    for k in range(n * n):
        e = np.zeros(n * n)
        e[k] = h
        num.ravel()[k] = (J(u + e.reshape(n, n)) - J(u - e.reshape(n, n))) / (2 * h)
    assert np.allclose(O.sb_tv64(u, ex, ey), -num, atol=1e-6)


def test_sb_shrink_matches_reference_update_db():
    """The shrink + Bregman update equals update_db's shrinkage
    (ufo-ir-sbtv-task.c:360-402, the FP32 restatement pinned above) applied to
    g = D u + b, with b <- g - d and e = d - b."""
    rng = np.random.default_rng(12)
    n = 16
    u = rng.standard_normal((n, n))
    bx, by = rng.standard_normal((n, n)) * 0.3, rng.standard_normal((n, n)) * 0.3
    lam = 1.7
    nbx, nby, ex, ey = O.sb_shrink64(u, bx, by, lam)
    gx = u - np.roll(u, 1, axis=1) + bx
    gy = u - np.roll(u, 1, axis=0) + by
    dx, dy = O.sbtv_shrink(gx.astype(np.float32).ravel(), gy.astype(np.float32).ravel(), lam)
    assert np.allclose(gx - nbx, dx.reshape(n, n), atol=1e-5)
    assert np.allclose(gy - nby, dy.reshape(n, n), atol=1e-5)
    assert np.allclose(ex, (gx - nbx) - nbx, atol=1e-12)


@pytest.mark.parametrize("model", ["joseph", "line"])
def test_sbsirt_as_matrix_iteration(model):
    """The whole loop from dense FP64 system matrices and periodic difference
    matrices: inner sweeps u += w P (mu B R (f - A u) + lambda D^T (e - D u)),
    then the shrink (catches a shrink after the last outer iteration, a
    swapped mu / lambda, or a sweep count off by one).  P = 1 / (mu A + 8
    lambda) for joseph; for the line model, per pixel, 1 / (mu BP(1) + 8
    lambda) with BP(1) = the row sums of B, taken as 0 below 2^-10 (catches a
    scalar P or the joseph angle count in its place)."""
    n, d, a = 8, 12, 10
    g = O.Geometry(n, d, a, np.float32(PI / a), -1.0, model=model)
    A = np.stack([O.fp64(g, np.eye(n * n)[k].reshape(n, n)).ravel() for k in range(n * n)], axis=1)
    B = np.stack([O.bp64(g, np.eye(a * d)[k].reshape(a, d)).ravel() for k in range(a * d)], axis=1)
    I = np.eye(n)
    d1 = I - np.roll(I, 1, axis=0)   # (d1 v)[x] = v[x] - v[x-1], periodic
    Dx = np.kron(I, d1)              # along x (fastest index)
    Dy = np.kron(d1, I)              # along y
    x_true = np.random.default_rng(13).random(n * n)
    f = A @ x_true
    R = 1 / np.where(A.sum(1) != 0, A.sum(1), np.inf)
    w, lam, mu, outer, inner = 0.9, 0.3, 1.5, 3, 2
    if model == "line":
        b1 = B.sum(1)
        P = 1 / (mu * np.where(b1 >= 2.0 ** -10, b1, 0.0) + 8 * lam)
        assert np.ptp(P) > 0.01 * P.max()  # not a constant in disguise
    else:
        P = 1 / (mu * a + 8 * lam)
    u = np.zeros(n * n)
    bx = np.zeros(n * n)
    by = np.zeros(n * n)
    ex = np.zeros(n * n)
    ey = np.zeros(n * n)
    for k in range(outer):
        for _ in range(inner):
            tv = Dx.T @ (ex - Dx @ u) + Dy.T @ (ey - Dy @ u)
            u = u + w * P * (mu * (B @ (R * (f - A @ u))) + lam * tv)
        if k + 1 < outer:
            gx, gy = Dx @ u + bx, Dy @ u + by
            s = np.hypot(gx, gy)
            t = np.maximum(s - 1 / lam, 0) / np.maximum(s, 1e-12)
            bx, by = gx - t * gx, gy - t * gy
            ex, ey = t * gx - bx, t * gy - by
    got = O.sbsirt64(g, f.reshape(a, d), outer, inner, relax=w, lam=lam, mu=mu)
    assert O.rel_l2(got.ravel(), u) < 1e-12


def test_sirt64_ray_weight_floor_as_matrix_iteration():
    """ray_weight_min drops exactly the rays whose FP(1) is below it (R = 0, as
    for rays that miss the image); catches a floor applied to the pixel
    weights or compared against R instead of FP(1)."""
    n, d, a = 8, 13, 10
    g = geom(n, d, a)
    A = np.stack([O.fp64(g, np.eye(n * n)[k].reshape(n, n)).ravel() for k in range(n * n)], axis=1)
    B = np.stack([O.bp64(g, np.eye(a * d)[k].reshape(a, d)).ravel() for k in range(a * d)], axis=1)
    rs = A.sum(1)
    tau = float(np.sort(rs[rs > 0])[5]) + 1e-9      # drops the six weakest rays
    R = np.where((rs != 0) & (rs >= tau), 1 / np.where(rs != 0, rs, 1), 0)
    C = 1 / B.sum(1)
    f = np.random.default_rng(14).random(a * d)
    x = np.zeros(n * n)
    for _ in range(5):
        x = x + C * (0.8 * (B @ (R * (f - A @ x))))
    got = O.sirt64(g, f.reshape(a, d), 5, 0.8, ray_weight_min=tau)
    assert O.rel_l2(got.ravel(), x) < 1e-12
    assert O.rel_l2(got, O.sirt64(g, f.reshape(a, d), 5, 0.8)) > 1e-3

"""Pins of the oracle's second projection model, "line" (the projector's
`model` property, ufo-ir-parallel-projector-task.c:123-128,370; the reference
ships only joseph).  Ray s at angle theta is the line X cos - Y sin = t_s,
t_s = s + 1/2 - axis, and its weight on a pixel is the length of that line
inside the pixel's unit square.

  - the trapezoid formula of the weight equals the length found by clipping
    the line against the square (Liang-Barsky), at random offsets and angles;
  - the FP equals a brute-force sum over every pixel of clipped lengths;
  - FP(1) is the length of each ray inside the whole image square;
  - at theta = 0 with the bins on the columns, the FP is the column sums;
  - the BP is the exact transpose of the FP (<A x, y> = <x, A^T y>);
  - SIRT over the model equals the dense iteration x += relax C A^T R (b - A x)
    with C = 1 / BP(1) only where BP(1) >= 2^-10: pixels that rays merely
    graze keep their value."""
import numpy as np
import pytest

import oracle as O


def clip_length(t, c, s, x0, x1, y0, y1):
    """Length of the line X c - Y s = t inside [x0, x1] x [y0, y1] (Liang-Barsky
    on the point (t c, -t s) + l (s, c), a unit-speed parametrisation).  A line
    parallel to a side lies in the box when its offset t is in [t0, t1), the
    box's extent along the normal (c, -s): a ray along the edge of two
    squares belongs to one of them."""
    px, py, vx, vy = t * c, -t * s, s, c
    lo, hi = -np.inf, np.inf
    for p, v, a, b, sig in ((px, vx, x0, x1, c), (py, vy, y0, y1, -s)):
        if v == 0.0:
            if not min(sig * a, sig * b) <= t < max(sig * a, sig * b):
                return 0.0
            continue
        l1, l2 = (a - p) / v, (b - p) / v
        lo, hi = max(lo, min(l1, l2)), min(hi, max(l1, l2))
    return max(0.0, hi - lo)


def chord(d, c, s):
    return O.lib().orc64_line_chord(float(d), float(c), float(s))


def test_chord_is_the_clipped_length():
    rng = np.random.default_rng(0)
    for _ in range(3000):
        th = rng.uniform(0, np.pi)
        c, s = np.cos(th), np.sin(th)
        d = rng.uniform(-1.0, 1.0)
        # the unit square centred at the origin seen from the line at t = d
        want = clip_length(d, c, s, -0.5, 0.5, -0.5, 0.5)
        assert abs(chord(d, c, s) - want) <= 1e-12, (th, d)
    # theta = 0: the square spans d in [-1/2, 1/2), so a ray along an edge counts once
    assert chord(-0.5, 1.0, 0.0) == 1.0 and chord(0.5, 1.0, 0.0) == 0.0
    # 45 degrees: a triangle of height sqrt(2)
    assert abs(chord(0.0, np.sqrt(0.5), np.sqrt(0.5)) - np.sqrt(2)) <= 1e-12


def _geom(n, d, a, axis=-1.0, scale=1.0):
    return O.Geometry(n, d, a, np.float32(scale * np.pi / a), axis, model="line")


def _dir(g, k):
    """The model's direction: the float LUT pair scaled to unit length, a
    component below 2^-20 taken as 0 (float cos(pi/2) is -4.4e-8)."""
    c, s = float(g.cos[k]), float(g.sin[k])
    r = np.hypot(c, s)
    c, s = c / r, s / r
    if abs(c) < 2.0 ** -20:
        return 0.0, float(np.sign(s))
    if abs(s) < 2.0 ** -20:
        return float(np.sign(c)), 0.0
    return c, s


@pytest.mark.parametrize("n,d,a,axis", [(12, 17, 13, -1.0), (10, 10, 9, -1.0), (9, 14, 11, 7.3), (9, 12, 8, -1.0)])
def test_fp_is_brute_force_clipped_lengths(n, d, a, axis):
    g = _geom(n, d, a, axis)
    x = np.random.default_rng(n).random((n, n))
    got = O.fp64(g, x)
    h = n / 2.0
    want = np.zeros((a, d))
    for k in range(a):
        c, s = _dir(g, k)
        for b in range(d):
            t = b + 0.5 - g.axis
            acc = 0.0
            for yy in range(n):
                for xx in range(n):
                    X, Y = xx + 0.5 - h, yy + 0.5 - h
                    if abs(t - (X * c - Y * s)) < 1.0:  # the only pixels the line can cross
                        acc += clip_length(t, c, s, X - 0.5, X + 0.5, Y - 0.5, Y + 0.5) * x[yy, xx]
            want[k, b] = acc
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n,d,a,axis", [(64, 91, 45, -1.0), (48, 48, 30, 20.25), (33, 50, 20, -1.0)])
def test_fp_of_ones_is_the_chord_through_the_image(n, d, a, axis):
    g = _geom(n, d, a, axis)
    got = O.fp64(g, np.ones((n, n)))
    h = n / 2.0
    want = np.array([[clip_length(b + 0.5 - g.axis, *_dir(g, k), -h, h, -h, h)
                      for b in range(d)] for k in range(a)])
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_right_angle_is_row_sums():
    """The float LUT's 90 degrees (cos = -4.4e-8) is the axis-aligned model:
    ray s at t_s = s + 1/2 - axis is the line Y = -t_s, i.e. row N/2 - 1 - s
    for D = N and the centred axis, and with N odd the rays run along pixel
    edges and take the row whose square spans [Y - 1/2, Y + 1/2)."""
    for n in (20, 21):
        g = _geom(n, n, 4)
        assert float(g.cos[2]) != 0.0 and abs(float(g.cos[2])) < 1e-7
        x = np.random.default_rng(n).random((n, n))
        got = O.fp64(g, x)[2]
        rows = x.sum(axis=1)
        h = n / 2.0
        for b in range(n):
            t = b + 0.5 - g.axis
            # d = t + Y must be in [-1/2, 1/2): Y in [-t - 1/2, -t + 1/2)
            y = int(np.floor(-t - 0.5 + h))
            want = rows[y] if 0 <= y < n else 0.0
            assert abs(got[b] - want) <= 1e-13 * n, (n, b)


def test_theta0_is_column_sums():
    n = d = 20
    g = _geom(n, d, 4)                         # angle 0 is exactly theta = 0
    x = np.random.default_rng(2).random((n, n))
    assert float(g.sin[0]) == 0.0 and float(g.cos[0]) == 1.0
    assert np.allclose(O.fp64(g, x)[0], x.sum(axis=0), rtol=1e-14, atol=0)


@pytest.mark.parametrize("n,d,a,axis", [(64, 91, 45, -1.0), (40, 40, 36, 23.7), (31, 50, 17, -1.0)])
def test_bp_is_the_exact_transpose(n, d, a, axis):
    g = _geom(n, d, a, axis)
    rng = np.random.default_rng(a)
    x = rng.standard_normal((n, n))
    y = rng.standard_normal((a, d))
    lhs = float(np.sum(O.fp64(g, x) * y))
    rhs = float(np.sum(x * O.bp64(g, y)))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs))
    # negative control: the joseph pair is not a transpose
    gj = O.Geometry(n, d, a, np.float32(np.pi / a), axis)
    lj, rj = float(np.sum(O.fp64(gj, x) * y)), float(np.sum(x * O.bp64(gj, y)))
    assert abs(lj - rj) > 1e-6 * (abs(lj) + abs(rj))


def test_sirt_pixel_weight_floor():
    """One detector whose ray passes 1e-5 beside the pixel corners at 45
    degrees: the pixels it grazes have BP(1) in (0, 2^-10) and are never
    updated; the others follow the dense iteration."""
    n, d, a = 4, 1, 4
    g = O.Geometry(n, d, a, np.float32(np.pi / a), 0.5 + 1e-5, model="line")
    nv, ns = n * n, a * d
    A = np.zeros((ns, nv))
    for j in range(nv):
        e = np.zeros(nv)
        e[j] = 1.0
        A[:, j] = O.fp64(g, e.reshape(n, n)).ravel()
    colsum = A.sum(axis=0)
    grazed = (colsum > 0) & (colsum < 2.0 ** -10)
    assert grazed.any() and (colsum >= 2.0 ** -10).any()
    assert np.allclose(O.bp64(g, np.ones((a, d))).ravel(), colsum, rtol=1e-12, atol=1e-15)
    rowsum = A.sum(axis=1)
    R = np.where(rowsum != 0, 1.0 / np.where(rowsum != 0, rowsum, 1.0), 0.0)
    C = np.where(colsum >= 2.0 ** -10, 1.0 / np.where(colsum > 0, colsum, 1.0), 0.0)
    b = np.random.default_rng(5).random(ns) + 1.0
    x = np.zeros(nv)
    for _ in range(6):
        x = x + 0.5 * C * (A.T @ (R * (b - A @ x)))
    got = O.sirt64(g, b.reshape(a, d), 6, 0.5).ravel()
    assert np.allclose(got, x, rtol=1e-12, atol=1e-12)
    assert np.all(got[grazed] == 0.0) and np.any(got[colsum >= 2.0 ** -10] != 0.0)

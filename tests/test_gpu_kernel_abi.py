"""Kernel-level C ABI (uir_fp_joseph / uir_bp_joseph / uir_op_*), the
reference's inner boundary (projector-parallel-joseph.cl:37-45,159-167,
ufo-ir-basic-ops.cl), against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import uir
from phantoms import ellipse_phantom

pytestmark = pytest.mark.gpu


def _dev(a):
    return uir.DeviceArray.from_numpy(np.ascontiguousarray(a, np.float32))


@pytest.mark.parametrize("n,a", [(256, 180), (367, 180)])
def test_fp_bp_joseph_per_subset(res, n, a):
    step = np.float32(np.pi / a)
    g = O.Geometry(n, n, a, step)
    x = ellipse_phantom(n, seed=2)
    L = uir.load()
    dx, dsin, dcos = _dev(x), _dev(g.sin[None, :]), _dev(g.cos[None, :])
    sino = _dev(np.zeros((a, n)))
    dims = uir.GeometryDims(n, n, n, a)
    for (o, k, d) in g.runs:
        uir.check(L.uir_fp_joseph(dx.ptr, sino.ptr, sino.ptr, dsin.ptr, dcos.ptr, dims, g.axis,
                                  uir.Subset(o, k, d), -1.0, None))
    ref = O.fp(g, x, corr=-1.0)
    assert np.array_equal(sino.numpy(), ref)      # same IEEE operations, same order: bit-exact
    vol = _dev(x)
    for (o, k, d) in g.runs:
        uir.check(L.uir_bp_joseph(vol.ptr, vol.ptr, sino.ptr, 0.5, dsin.ptr, dcos.ptr, dims, g.axis,
                                  uir.Subset(o, k, d), None))
    assert np.array_equal(vol.numpy(), O.bp(g, ref, vol=x, relax=0.5))


@pytest.mark.parametrize("n,d,a,axis", [(256, 256, 60, -1.0), (200, 283, 45, -1.0), (128, 128, 30, 70.0),
                                        (1100, 1556, 6, -1.0)])
def test_fp_joseph_single_angle_split_is_bit_identical(res, n, d, a, axis):
    """Subsets of <= 4 angles take the split kernel (blocks of a ray over 16
    threads, replayed float stepping); it must equal the one-thread-per-ray
    kernel bit for bit, inactive detectors (shifted axis) included."""
    step = np.float32(np.pi / a)
    g = O.Geometry(n, d, a, step, axis)
    x = ellipse_phantom(n, seed=4)
    L = uir.load()
    dx, dsin, dcos = _dev(x), _dev(g.sin[None, :]), _dev(g.cos[None, :])
    r0 = np.random.default_rng(5).random((a, d)).astype(np.float32)
    per_angle, per_run = _dev(r0), _dev(r0)
    dims = uir.GeometryDims(n, n, d, a)
    for (o, k, dr) in g.runs:
        for ang in range(o, o + k):
            uir.check(L.uir_fp_joseph(dx.ptr, per_angle.ptr, per_angle.ptr, dsin.ptr, dcos.ptr, dims, g.axis,
                                      uir.Subset(ang, 1, dr), -1.0, None))
        for o2 in range(o, o + k, 5):       # blocks of 5 angles take the one-thread-per-ray kernel
            uir.check(L.uir_fp_joseph(dx.ptr, per_run.ptr, per_run.ptr, dsin.ptr, dcos.ptr, dims, g.axis,
                                      uir.Subset(o2, min(5, o + k - o2), dr), -1.0, None))
    assert np.array_equal(per_angle.numpy(), per_run.numpy())
    # same IEEE operations in the same order as the CPU restatement: bit-exact
    assert np.array_equal(per_angle.numpy(), O.fp(g, x, sino=r0, corr=-1.0))


def test_elementwise_ops(res):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((37, 53)).astype(np.float32)
    b = rng.standard_normal((37, 53)).astype(np.float32)
    a[0, :5] = 0
    n = a.size
    L = uir.load()
    da, db, do = _dev(a), _dev(b), _dev(np.zeros_like(a))
    cases = [
        (lambda: L.uir_op_mul(da.ptr, db.ptr, do.ptr, n, None), a * b),
        (lambda: L.uir_op_add(da.ptr, db.ptr, do.ptr, n, None), a + b),
        (lambda: L.uir_op_deduction(da.ptr, db.ptr, do.ptr, n, None), a - b),
        (lambda: L.uir_op_deduction2(da.ptr, db.ptr, 0.3, do.ptr, n, None), a - np.float32(0.3) * b),
        (lambda: L.uir_op_add2(da.ptr, db.ptr, 0.3, do.ptr, n, None), a + np.float32(0.3) * b),
        (lambda: L.uir_op_inv(da.ptr, do.ptr, n, None),
         np.where(a != 0, np.float32(1) / np.where(a != 0, a, 1), 0).astype(np.float32)),
        (lambda: L.uir_op_positive_constraint(da.ptr, do.ptr, n, None), np.maximum(a, 0)),
        (lambda: L.uir_op_max_ew(da.ptr, db.ptr, do.ptr, n, None), np.maximum(a, b)),
    ]
    for fn, want in cases:
        uir.check(fn())
        assert np.array_equal(do.numpy(), want)


def test_reductions_and_stencils(res):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((40, 48)).astype(np.float32)
    b = rng.standard_normal((40, 48)).astype(np.float32)
    L = uir.load()
    da, db = _dev(a), _dev(b)
    r = C.c_float()
    uir.check(L.uir_op_dot_compat(da.ptr, db.ptr, 48, 40, C.byref(r), None))
    assert r.value == O.dot_compat(a, b)                 # bit-exact restatement
    uir.check(L.uir_op_l2_compat(da.ptr, 48, 40, C.byref(r), None))
    assert r.value == O.l2_compat(a)
    uir.check(L.uir_op_l1_norm(da.ptr, 48, 40, C.byref(r), None))
    assert abs(r.value - O.l1_norm(a)) <= 1e-5 * O.l1_norm(a)
    out = _dev(np.zeros_like(a))
    for name, fn in (("uir_op_dx", O.dx), ("uir_op_dxt", O.dxt), ("uir_op_dy", O.dy), ("uir_op_dyt", O.dyt),
                     ("uir_op_l1_grad", O.l1_grad)):
        uir.check(getattr(L, name)(da.ptr, out.ptr, 48, 40, None))
        assert np.array_equal(out.numpy(), fn(a)), name
    v = _dev(a)
    uir.check(L.uir_op_normalization(v.ptr, a.size, None))
    assert np.array_equal(v.numpy(), O.normalize(a))

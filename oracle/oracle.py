"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product library never uses it.

Arrays are numpy float32 (ref-compat FP32 restatement) or float64 (FP64
evaluation of the same operators), C-contiguous; images are [y][x] (N x N),
sinograms [angle][det].
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Subset(C.Structure):
    _fields_ = [("offset", C.c_uint32), ("n", C.c_uint32), ("direction", C.c_int32)]


class Geom(C.Structure):
    _fields_ = [("n", C.c_int), ("n_dets", C.c_int), ("n_angles", C.c_int), ("axis", C.c_float),
                ("sin_lut", C.c_void_p), ("cos_lut", C.c_void_p), ("model", C.c_int)]


MODELS = {"joseph": 0, "line": 1}


class AsdPocsParams(C.Structure):
    _fields_ = [("beta", C.c_float), ("beta_red", C.c_float), ("alpha", C.c_float),
                ("alpha_red", C.c_float), ("r_max", C.c_float), ("positive_constraint", C.c_int),
                ("df_is_sirt", C.c_int), ("df_iters", C.c_uint32), ("df_relax_initial", C.c_float),
                ("ng", C.c_uint32), ("textbook", C.c_int)]


def build():
    """Compile liboracle.so with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        G = C.POINTER(Geom)
        L.orc_lut.argtypes = [C.c_uint32, C.c_float, f32p, f32p]
        L.orc_axis_resolve.argtypes = [C.c_float, C.c_uint32]
        L.orc_axis_resolve.restype = C.c_float
        L.orc_runs.argtypes = [f32p, f32p, C.c_uint32, C.POINTER(Subset)]
        L.orc_runs.restype = C.c_int
        L.orc_angle_subsets.argtypes = [f32p, f32p, C.c_uint32, C.POINTER(Subset)]
        L.orc_angle_subsets.restype = C.c_int
        L.orc_fp_subset.argtypes = [G, f32p, f32p, Subset, C.c_float]
        L.orc_bp_subset.argtypes = [G, f32p, f32p, Subset, C.c_float]
        L.orc_fp.argtypes = [G, f32p, f32p, C.POINTER(Subset), C.c_int, C.c_float]
        L.orc_bp.argtypes = [G, f32p, f32p, C.POINTER(Subset), C.c_int, C.c_float]
        for name in ("orc_op_inv",):
            getattr(L, name).argtypes = [f32p, C.c_size_t]
        L.orc_op_set.argtypes = [f32p, C.c_size_t, C.c_float]
        for name in ("orc_op_mul", "orc_op_add", "orc_op_sub", "orc_mul_ew", "orc_max_ew", "orc_div_ew"):
            getattr(L, name).argtypes = [f32p, f32p, f32p, C.c_size_t]
        for name in ("orc_op_sub_scaled", "orc_op_add_scaled"):
            getattr(L, name).argtypes = [f32p, f32p, C.c_float, f32p, C.c_size_t]
        L.orc_op_mul_rows.argtypes = [f32p, f32p, f32p, C.c_size_t, C.c_uint32, C.c_uint32]
        L.orc_op_posc.argtypes = [f32p, f32p, C.c_size_t]
        L.orc_l1_norm.argtypes = [f32p, C.c_size_t, C.c_size_t]
        L.orc_l1_norm.restype = C.c_float
        L.orc_dot_compat.argtypes = [f32p, f32p, C.c_size_t, C.c_size_t]
        L.orc_dot_compat.restype = C.c_float
        L.orc_l2_compat.argtypes = [f32p, C.c_size_t, C.c_size_t]
        L.orc_l2_compat.restype = C.c_float
        L.orc_mul_scalar.argtypes = [f32p, C.c_size_t, C.c_float]
        L.orc_normalize.argtypes = [f32p, C.c_size_t]
        L.orc_sqrt.argtypes = [f32p, C.c_size_t]
        for name in ("orc_dx", "orc_dxt", "orc_dy", "orc_dyt", "orc_l1_grad"):
            getattr(L, name).argtypes = [f32p, f32p, C.c_int, C.c_int]
        L.orc_sirt.argtypes = [G, f32p, f32p, C.c_uint32, C.c_float, C.c_int]
        L.orc_sart.argtypes = [G, f32p, f32p, C.c_uint32, C.c_float]
        L.orc_asdpocs.argtypes = [G, f32p, f32p, C.c_uint32, C.POINTER(AsdPocsParams), C.c_void_p]
        L.orc_sbtv.argtypes = [G, f32p, f32p, C.c_uint32, C.c_float, C.c_float, C.c_int]
        L.orc_sbtv_cgs.argtypes = [G, f32p, f32p, f32p, C.c_uint32, C.c_float, C.c_float, C.c_int]
        L.orc_sbtv_cgs.restype = C.c_int
        L.orc_sbtv_apply.argtypes = [G, f32p, f32p, C.c_float, C.c_float]
        L.orc_cgs_dense.argtypes = [f32p, C.c_uint32, C.c_uint32, f32p, f32p, f32p, C.c_uint32, C.c_int]
        L.orc_cgs_dense.restype = C.c_int
        L.orc64_line_chord.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc64_line_chord.restype = C.c_double
        L.orc64_line_dir.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_sbtv_shrink.argtypes = [f32p, f32p, C.c_size_t, C.c_float, f32p, f32p]
        L.orc64_fp.argtypes = [G, f64p, f64p]
        L.orc64_bp.argtypes = [G, f64p, f64p]
        L.orc64_sirt.argtypes = [G, f64p, f64p, C.c_uint32, C.c_double, C.c_int, C.c_double]
        L.orc64_sart.argtypes = [G, f64p, f64p, C.c_uint32, C.c_double]
        L.orc64_l1_grad.argtypes = [f64p, f64p, C.c_int]
        L.orc64_asdpocs.argtypes = [G, f64p, f64p, C.c_uint32, C.POINTER(AsdPocsParams), C.c_double, C.c_void_p]
        L.orc64_sb_tv.argtypes = [f64p, f64p, f64p, f64p, C.c_int]
        L.orc64_sb_shrink.argtypes = [f64p, f64p, f64p, f64p, f64p, C.c_int, C.c_double]
        L.orc64_sbsirt.argtypes = [G, f64p, f64p, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_int,
                                   C.c_double]
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def num_threads():
    return lib().orc_num_threads()


class Geometry:
    """Parallel-beam geometry as the reference projector builds it.

    n: image edge (the reference always uses n == n_dets);
    step: angular step in radians (projector "step" property);
    axis: projector "axis_position" (-1 selects n_dets / 2).
    """

    def __init__(self, n, n_dets, n_angles, step, axis=-1.0, model="joseph"):
        self.n, self.n_dets, self.n_angles = int(n), int(n_dets), int(n_angles)
        self.model = model
        self.step = float(step)
        self.sin = np.empty(self.n_angles, np.float32)
        self.cos = np.empty(self.n_angles, np.float32)
        lib().orc_lut(self.n_angles, self.step, self.sin, self.cos)
        self.axis = lib().orc_axis_resolve(float(axis), self.n_dets)
        self._g = Geom(self.n, self.n_dets, self.n_angles, self.axis,
                       self.sin.ctypes.data, self.cos.ctypes.data, MODELS[model])
        self.runs = self._subsets(lib().orc_runs)
        self.angle_subsets = self._subsets(lib().orc_angle_subsets)

    def _subsets(self, fn):
        buf = (Subset * max(self.n_angles, 1))()
        k = fn(self.sin, self.cos, self.n_angles, buf)
        return [(buf[i].offset, buf[i].n, buf[i].direction) for i in range(k)]

    @property
    def ptr(self):
        return C.byref(self._g)

    def vol_shape(self):
        return (self.n, self.n)

    def sino_shape(self):
        return (self.n_angles, self.n_dets)


def _subset_array(subsets):
    arr = (Subset * len(subsets))()
    for i, (o, n, d) in enumerate(subsets):
        arr[i] = Subset(o, n, d)
    return arr


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def fp(geom, vol, sino=None, corr=1.0, subsets=None):
    """Ref-compat FP: sino <- sino + corr * FP(vol) over the direction runs."""
    vol = _f32(vol)
    sino = np.zeros(geom.sino_shape(), np.float32) if sino is None else _f32(sino).copy()
    subs = geom.runs if subsets is None else subsets
    arr = _subset_array(subs)
    lib().orc_fp(geom.ptr, vol, sino, arr, len(subs), float(corr))
    return sino


def bp(geom, sino, vol=None, relax=1.0, subsets=None):
    """Ref-compat BP: vol <- vol + relax * BP(sino), per run as the reference."""
    sino = _f32(sino)
    vol = np.zeros(geom.vol_shape(), np.float32) if vol is None else _f32(vol).copy()
    subs = geom.runs if subsets is None else subsets
    arr = _subset_array(subs)
    lib().orc_bp(geom.ptr, vol, sino, arr, len(subs), float(relax))
    return vol


def fp64(geom, vol):
    vol = np.ascontiguousarray(vol, dtype=np.float64)
    out = np.zeros(geom.sino_shape(), np.float64)
    lib().orc64_fp(geom.ptr, vol, out)
    return out


def bp64(geom, sino):
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    out = np.zeros(geom.vol_shape(), np.float64)
    lib().orc64_bp(geom.ptr, out, sino)
    return out


def sirt(geom, b, iters, relax=0.25, posc=False):
    x = np.zeros(geom.vol_shape(), np.float32)
    lib().orc_sirt(geom.ptr, _f32(b), x, int(iters), float(relax), int(bool(posc)))
    return x


def sart(geom, b, iters, relax=0.25):
    x = np.zeros(geom.vol_shape(), np.float32)
    lib().orc_sart(geom.ptr, _f32(b), x, int(iters), float(relax))
    return x


def sirt64(geom, b, iters, relax=0.25, posc=False, ray_weight_min=0.0):
    """SIRT in FP64; rays with FP(1) < ray_weight_min get weight 0."""
    x = np.zeros(geom.vol_shape(), np.float64)
    lib().orc64_sirt(geom.ptr, np.ascontiguousarray(b, np.float64), x, int(iters), float(relax), int(bool(posc)),
                     float(ray_weight_min))
    return x


def sart64(geom, b, iters, relax=0.25):
    x = np.zeros(geom.vol_shape(), np.float64)
    lib().orc64_sart(geom.ptr, np.ascontiguousarray(b, np.float64), x, int(iters), float(relax))
    return x


def sbsirt64(geom, f, outer, inner, relax=1.0, lam=0.1, mu=1.0, posc=False, ray_weight_min=0.0):
    """Split-Bregman TV with SIRT-preconditioned inner sweeps, FP64."""
    u = np.zeros(geom.vol_shape(), np.float64)
    lib().orc64_sbsirt(geom.ptr, np.ascontiguousarray(f, np.float64), u, int(outer), int(inner), float(relax),
                       float(lam), float(mu), int(bool(posc)), float(ray_weight_min))
    return u


def sb_tv64(u, ex, ey):
    """tv = Dx^T (ex - Dx u) + Dy^T (ey - Dy u) (FP64, periodic differences)."""
    n = u.shape[0]
    tv = np.zeros((n, n), np.float64)
    c = lambda a: np.ascontiguousarray(a, np.float64)
    lib().orc64_sb_tv(c(u), c(ex), c(ey), tv, n)
    return tv


def sb_shrink64(u, bx, by, lam):
    """One shrink + Bregman update from (u, bx, by): returns (bx, by, ex, ey)."""
    n = u.shape[0]
    bx = np.array(bx, np.float64)
    by = np.array(by, np.float64)
    ex = np.zeros((n, n), np.float64)
    ey = np.zeros((n, n), np.float64)
    lib().orc64_sb_shrink(np.ascontiguousarray(u, np.float64), bx, by, ex, ey, n, float(lam))
    return bx, by, ex, ey


def _asd_params(beta, beta_red, alpha, alpha_red, r_max, positive_constraint, df, df_iters, df_relax_initial, ng,
                textbook):
    return AsdPocsParams(beta, beta_red, alpha, alpha_red, r_max, int(bool(positive_constraint)),
                         1 if df == "sirt" else 0, int(df_iters), df_relax_initial, int(ng), int(bool(textbook)))


def asdpocs(geom, b, iters, beta=1.0, beta_red=0.995, alpha=0.2, alpha_red=0.95, r_max=0.95,
            positive_constraint=True, df="sart", df_iters=10, df_relax_initial=0.25, diag=False, ng=20,
            textbook=False):
    """Ref-compat FP32 ASD-POCS; textbook=True is the published variant
    (warm-started data step, L2 distances, normalised TV steps, ng of them,
    output after TV)."""
    p = _asd_params(beta, beta_red, alpha, alpha_red, r_max, positive_constraint, df, df_iters, df_relax_initial,
                    ng, textbook)
    out = np.zeros(geom.vol_shape(), np.float32)
    d = np.zeros(4 * max(int(iters), 1), np.float32)
    lib().orc_asdpocs(geom.ptr, _f32(b), out, int(iters), C.byref(p), d.ctypes.data if diag else None)
    return (out, d.reshape(-1, 4)[:iters]) if diag else out


def asdpocs64(geom, b, iters, beta=1.0, beta_red=0.995, alpha=0.2, alpha_red=0.95, r_max=0.95,
              positive_constraint=True, df="sart", df_iters=10, df_relax_initial=0.25, ng=20, textbook=False,
              ray_weight_min=0.0, diag=False):
    """FP64 ASD-POCS with exact coordinates (both variants of asdpocs)."""
    p = _asd_params(beta, beta_red, alpha, alpha_red, r_max, positive_constraint, df, df_iters, df_relax_initial,
                    ng, textbook)
    out = np.zeros(geom.vol_shape(), np.float64)
    d = np.zeros(4 * max(int(iters), 1), np.float64)
    lib().orc64_asdpocs(geom.ptr, np.ascontiguousarray(b, np.float64), out, int(iters), C.byref(p),
                        float(ray_weight_min), d.ctypes.data if diag else None)
    return (out, d.reshape(-1, 4)[:iters]) if diag else out


def l1_grad64(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    lib().orc64_l1_grad(x, out, x.shape[0])
    return out


def sbtv(geom, b, iters, lam=0.1, mu=0.5, dot_mode=0):
    """Returns (u, normalised input); the reference normalises its input in place."""
    bb = _f32(b).copy()
    u = np.zeros(geom.vol_shape(), np.float32)
    lib().orc_sbtv(geom.ptr, bb, u, int(iters), float(lam), float(mu), int(dot_mode))
    return u, bb


def sbtv_apply(geom, x, lam, mu):
    out = np.zeros(geom.vol_shape(), np.float32)
    lib().orc_sbtv_apply(geom.ptr, _f32(x), out, float(lam), float(mu))
    return out


def sbtv_cgs(geom, rhs, x0, max_iter, lam, mu, dot_mode=0):
    x = np.zeros(geom.vol_shape(), np.float32)
    flag = lib().orc_sbtv_cgs(geom.ptr, _f32(rhs), x, _f32(x0), int(max_iter), float(lam), float(mu), int(dot_mode))
    return x, flag


def cgs_dense(M, b, x0, max_iter, w, h, dot_mode=0):
    """The oracle's CGS (ufo-ir-sbtv-task.c:404-585) on a dense operator M
    ((w h) x (w h), float rows summed in order); returns (x, flag)."""
    M, b, x0 = _f32(M), _f32(b).ravel(), _f32(x0).ravel()
    x = np.zeros_like(b)
    flag = lib().orc_cgs_dense(M, int(w), int(h), b, x, x0, int(max_iter), int(dot_mode))
    return x, flag


def sbtv_shrink(tx, ty, lam):
    tx, ty = _f32(tx), _f32(ty)
    dx, dy = np.zeros_like(tx), np.zeros_like(ty)
    lib().orc_sbtv_shrink(tx, ty, tx.size, float(lam), dx, dy)
    return dx, dy


def l1_norm(v):
    v = _f32(v)
    h, w = v.shape
    return lib().orc_l1_norm(v, w, h)


def dot_compat(a, b):
    a, b = _f32(a), _f32(b)
    h, w = a.shape
    return lib().orc_dot_compat(a, b, w, h)


def l2_compat(a):
    a = _f32(a)
    h, w = a.shape
    return lib().orc_l2_compat(a, w, h)


def normalize(v):
    v = _f32(v).copy()
    lib().orc_normalize(v, v.size)
    return v


def _stencil(name, v):
    v = _f32(v)
    out = np.zeros_like(v)
    h, w = v.shape
    getattr(lib(), name)(v, out, w, h)
    return out


def dx(v):
    return _stencil("orc_dx", v)


def dxt(v):
    return _stencil("orc_dxt", v)


def dy(v):
    return _stencil("orc_dy", v)


def dyt(v):
    return _stencil("orc_dyt", v)


def l1_grad(v):
    return _stencil("orc_l1_grad", v)


def rel_l2(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(a - ref) / (den if den > 0 else 1.0))


def psnr(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    mse = float(np.mean((a - ref) ** 2))
    peak = float(np.max(np.abs(ref)))
    if mse == 0:
        return float("inf")
    return 10.0 * np.log10(peak * peak / mse)

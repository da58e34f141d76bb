"""FP work-table planner (uirk_plan_fp, host code, no device needed).

The planner cuts each direction run into angle blocks per 64-detector block
so that every chunk's footprint fits the staging window: the planner bounds
32-line chunks (the single-value kernel's), and the pair kernel's 16-line
chunks lie inside them.  These tests
recompute, in float64 and independently of the planner's own corner
argument, what the FP kernels do with each item:
  - every (angle, detector block) is planned exactly once, inside one run;
  - the window the kernel stages for each chunk (hull of all active rays over
    the chunk's lines, floor(hi - 0.5) - floor(lo - 0.5) + 4 positions) fits
    the limit the planner was given, for chunks starting at the item's l_lo;
  - every sample of every active ray that can touch the image lies in the
    item's line range [l_lo, l_hi)."""
import ctypes as C

import numpy as np
import pytest

import uir

LINES, DETS = (32, 16), 64  # chunk lines: single-value kernel, pair kernel


class Item(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("slice", "dir", "a0", "na", "d0", "l_lo", "l_hi", "pad")]


class Vec(C.Structure):
    _fields_ = [("v", C.POINTER(Item)), ("n", C.c_size_t), ("cap", C.c_size_t)]


class Subset(C.Structure):
    _fields_ = [("offset", C.c_uint32), ("n", C.c_uint32), ("direction", C.c_int32)]


def _plan(n, d, a, step, axis_prop, max_angles, wmax):
    L = uir.load()
    s = np.zeros(a, np.float32)
    c = np.zeros(a, np.float32)
    fp = C.POINTER(C.c_float)
    L.uir_geom_lut.argtypes = [C.c_uint32, C.c_float, fp, fp]
    L.uir_geom_lut(a, step, s.ctypes.data_as(fp), c.ctypes.data_as(fp))
    L.uir_geom_axis.restype = C.c_float
    L.uir_geom_axis.argtypes = [C.c_float, C.c_uint32]
    axis = L.uir_geom_axis(axis_prop, d)
    # direction runs, computed here: 1 = vertical (|sin| <= |cos|), 0 = horizontal
    dirs = (np.abs(s) <= np.abs(c)).astype(int)
    subs = []
    i = 0
    while i < a:
        j = i
        while j < a and dirs[j] == dirs[i]:
            j += 1
        subs.append((i, j - i, int(dirs[i])))
        i = j
    arr = (Subset * len(subs))(*[Subset(o, k, dr) for o, k, dr in subs])
    vec = Vec()
    L.uirk_plan_fp.restype = C.c_int
    L.uirk_plan_fp.argtypes = [C.POINTER(Vec), C.c_int, C.c_int, C.c_int, C.c_float, fp, fp, C.POINTER(Subset),
                               C.c_uint32, C.c_int, C.c_int]
    assert L.uirk_plan_fp(C.byref(vec), 0, n, d, axis, s.ctypes.data_as(fp), c.ctypes.data_as(fp), arr,
                          len(subs), max_angles, wmax) == 0
    items = [(it.dir, it.a0, it.na, it.d0, it.l_lo, it.l_hi) for it in vec.v[:vec.n]]
    C.CDLL(None).free(C.cast(vec.v, C.c_void_p))
    return items, subs, s.astype(np.float64), c.astype(np.float64), float(axis)


def _positions(dirn, s, c, dets, axis, n, lines):
    """Ray positions p[angle, det, line] of the fast path (exact coordinates)."""
    if dirn == 0:
        det_step, step = -1.0 / s, c / s
    else:
        det_step, step = 1.0 / c, s / c
    t = dets[None, :] + 0.5 - axis
    p0 = t * det_step[:, None] - 0.5 * n * step[:, None] + 0.5 * n
    return p0[:, :, None] + lines[None, None, :] * step[:, None, None]


def _active(dets, d, axis):
    req = 2.0 * np.float32(axis)
    diff = np.float32(d) - req
    on = dets < d
    if diff < 0:
        on &= ~(dets.astype(np.float32) < abs(diff))
    if diff > 0:
        on &= ~(dets.astype(np.float32) > req)
    return on


GEOMS = [  # (n, n_dets, n_angles, step scale, axis property)
    (96, 136, 40, 1.0, -1.0),
    (128, 182, 90, 1.0, -1.0),
    (160, 227, 75, 1.0, -1.0),
    (256, 363, 24, 1.0, -1.0),
    (192, 272, 180, 1.0, 140.5),   # axis shifted right of the detector centre
    (192, 272, 180, 1.0, 128.0),   # and left
    (200, 283, 17, 0.37, -1.0),    # angles over part of [0, pi)
]


@pytest.mark.parametrize("wmax", [192, 256])
@pytest.mark.parametrize("max_angles", [8, 16, 32])
@pytest.mark.parametrize("geom", GEOMS)
def test_plan_covers_each_ray_once_and_windows_fit(geom, max_angles, wmax):
    n, d, a, scale, axis_prop = geom
    step = float(np.float32(scale * np.pi / a))
    items, subs, s, c, axis = _plan(n, d, a, step, axis_prop, max_angles, wmax)
    seen = np.zeros((a, (d + DETS - 1) // DETS), int)
    run_of = np.zeros(a, int)
    for r, (o, k, _) in enumerate(subs):
        run_of[o:o + k] = r
    lines_all = np.arange(n, dtype=np.float64)
    for dirn, a0, na, d0, l_lo, l_hi in items:
        assert 1 <= na <= max_angles and d0 % DETS == 0 and 0 <= l_lo <= l_hi <= n
        assert run_of[a0] == run_of[a0 + na - 1] and subs[run_of[a0]][2] == dirn
        seen[a0:a0 + na, d0 // DETS] += 1
        dets = np.arange(d0, min(d0 + DETS, d))
        dets = dets[_active(dets, d, axis)]
        if dets.size == 0:
            continue
        ang = np.arange(a0, a0 + na)
        p = _positions(dirn, s[ang], c[ang], dets, axis, n, lines_all)
        touch = (p >= -0.5) & (p < n + 0.5)
        ls = np.nonzero(touch.any(axis=(0, 1)))[0]
        if ls.size:
            assert l_lo <= ls.min() and ls.max() < l_hi, (dirn, a0, na, d0, l_lo, l_hi, ls.min(), ls.max())
        for cl in LINES:
            for l0 in range(l_lo, l_hi, cl):
                nl = min(cl, l_hi - l0)
                ends = p[:, :, [l0, l0 + nl - 1]]
                lo, hi = ends.min(), ends.max()
                if hi < -1.0 or lo > n + 1.0:
                    continue  # the kernel skips chunks outside the image
                w = np.floor(hi - 0.5) - np.floor(lo - 0.5) + 4
                assert w <= wmax, (dirn, a0, na, d0, cl, l0, w)
    assert (seen == 1).all()


def test_wide_detector_blocks_cut_only_where_needed():
    """At 1024^2 / 720 angles / 1448 bins with 16-angle items and the pair
    window, only the outer detector blocks get fewer than 16 angles."""
    n, d, a = 1024, 1448, 720
    items, _, _, _, _ = _plan(n, d, a, float(np.float32(np.pi / a)), -1.0, 16, 192)
    full = {}
    for _, _, na, d0, _, _ in items:
        full.setdefault(d0, []).append(na)
    mean = {d0: np.mean(v) for d0, v in full.items()}
    centre = [mean[d0] for d0 in mean if abs(d0 + 32 - d / 2) < 256]
    edge = [mean[d0] for d0 in mean if abs(d0 + 32 - d / 2) > 600]
    assert min(centre) > 14.0 and min(edge) < min(centre)


def _active_bins(d, axis):
    L = uir.load()
    L.uir_geom_active_bins.restype = C.c_int
    L.uir_geom_active_bins.argtypes = [C.c_uint32, C.c_float, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lo, hi = C.c_int(), C.c_int()
    ok = L.uir_geom_active_bins(d, axis, C.byref(lo), C.byref(hi))
    return ok, lo.value, hi.value


@pytest.mark.parametrize("d", [363, 1448])
def test_active_bins_window(d):
    # projector-parallel-joseph.cl:174-182: a centred axis activates every bin;
    # a shifted axis drops bins on the side away from it (2 |axis - D/2| of them)
    assert _active_bins(d, d / 2) == (1, 0, d - 1)
    ok, lo, hi = _active_bins(d, d / 2 + 10.0)
    assert ok and lo == 20 and hi == d - 1
    ok, lo, hi = _active_bins(d, d / 2 - 10.0)
    assert ok and lo == 0 and hi == d - 1 - 20
    # the extreme axes that still leave one detector active
    assert _active_bins(d, 0.5)[0] == 1 and _active_bins(d, d - 0.5)[0] == 1
    # axes that leave none: the fast kernels' staging would clamp to an empty
    # range, so the batch and serving paths refuse them
    for axis in (0.0, 0.25, d - 0.25, float(d), 1.5 * d):
        ok, lo, hi = _active_bins(d, axis)
        assert not ok, (axis, lo, hi)


@pytest.mark.parametrize("axis", [0.0, 0.25, 363.0, 600.0])
def test_request_with_no_active_detector_is_refused(axis):
    req = uir.make_request(256, 363, 180, float(np.pi / 180), axis_position=axis)
    with pytest.raises(uir.UirError) as e:
        uir.request_footprint(req)
    assert "no active detector" in str(e.value)
    # the default (-1, centred) and an in-range shift are accepted
    assert uir.request_footprint(uir.make_request(256, 363, 180, float(np.pi / 180))) > 0
    assert uir.request_footprint(uir.make_request(256, 363, 180, float(np.pi / 180), axis_position=190.0)) > 0


def test_server_line_model_requests():
    """The server takes SIRT and split-Bregman requests of the line model,
    whose geometry holds the model's angle table and N x N pixel weights on
    top of joseph's ray weights; an unknown model is invalid."""
    n, d, a = 256, 363, 180
    req = uir.make_request(n, d, a, float(np.pi / 180), model="line")
    joseph = uir.request_footprint(uir.make_request(n, d, a, float(np.pi / 180)))
    align = lambda b: (b + 255) // 256 * 256
    # 8 floats per angle, the pixel weights, and 72 bytes per angle
    # (uirk_line_angle: 2 doubles, 11 floats, 3 int32)
    extra = align(a * 32) + align(n * n * 4) + align(a * 72)
    assert uir.request_footprint(req) == joseph + extra
    sb = uir.make_request(n, d, a, float(np.pi / 180), method="sbtv-sirt", model="line")
    sb_joseph = uir.make_request(n, d, a, float(np.pi / 180), method="sbtv-sirt")
    assert uir.request_footprint(sb) == uir.request_footprint(sb_joseph) + extra
    req.slice.model = 7
    with pytest.raises(uir.UirError) as e:
        uir.request_footprint(req)
    assert e.value.status == 1
    assert uir.slice_spec(dict(n=8, n_dets=8, n_angles=4, step=0.1, model="line")).model == 1
    assert uir.slice_spec(dict(n=8, n_dets=8, n_angles=4, step=0.1)).model == 0


def _kernel_windows_and_samples(n, d, a, axis_prop, items, s, c, axis, max_items=None):
    """Replays, in FP32 where the kernel rounds, what k_fp_pairs computes for
    each chunk of each item: the staging window from the item's corner rays
    (its first and last active detector at each angle, at the chunk's first
    and last line), and every active ray's entry index at every line of the
    chunk (the coordinate anchored per chunk from double, stepped by float
    additions, rounded by rne_index).  Returns the worst margins: the smallest
    entry position and the largest position minus the staged width (both
    must keep the reads inside [0, W)), and the widest window."""
    f32 = np.float32
    kPH, kPL, kPW = 96, 16, 192
    req = f32(axis) * f32(2)
    diff = f32(d) - req
    rot = diff / f32(2)
    tbase = 0.5 + float(rot)
    thalf = float(f32(0.5) * f32(d))
    worst_lo, worst_hi, widest = 10 ** 9, -10 ** 9, 0
    pick = items if max_items is None else items[::max(1, len(items) // max_items)]
    for dirn, a0, na, d0, l_lo, l_hi in pick:
        dets = np.arange(d0, d0 + DETS)
        on_det = _active(np.minimum(dets, d - 1), d, axis) & (dets < d)
        if not on_det.any():
            continue
        ang = np.arange(a0, a0 + na)
        st, ct = s[ang], c[ang]
        hor = dirn == 0
        det_step = -1.0 / st if hor else 1.0 / ct
        sdv = ct / st if hor else st / ct
        # ray anchors of the active rays: p0 = ((tbase + s) - thalf) det_step - N/2 sd + N/2
        sa = dets[on_det].astype(np.float64)
        p0 = ((tbase + sa[None, :]) - thalf) * det_step[:, None] - 0.5 * n * sdv[:, None] + 0.5 * n
        sdf = sdv.astype(f32)
        s_first, s_last = sa.min(), sa.max()
        cp = np.stack([((tbase + sx) - thalf) * det_step - 0.5 * n * sdv + 0.5 * n for sx in (s_first, s_last)])
        for l0 in range(l_lo, l_hi, kPL):
            nl = min(kPL, l_hi - l0)
            nm1 = f32(nl - 1)
            av = (cp + float(l0) * sdv[None, :]).astype(f32)
            bv = av + nm1 * sdf[None, :]
            lo, hi = min(av.min(), bv.min()), max(av.max(), bv.max())
            pmin = int(np.floor(lo - f32(0.5))) - 1
            W = int(np.floor(hi - f32(0.5))) + 3 - pmin
            if W > kPW:
                raise AssertionError("window overflow", dirn, a0, na, d0, l0, W)
            if pmin + W <= 0 or pmin >= n:
                continue
            widest = max(widest, W)
            off = (f32(1.0) + f32(pmin)) + f32(kPH)
            u = (p0 + float(l0) * sdv[:, None]).astype(f32) - off
            for _ in range(nl):
                idx = (u + f32(12582912.0)).view(np.int32) - 0x4B400000
                pos = idx + kPH
                worst_lo = min(worst_lo, int(pos.min()))
                worst_hi = max(worst_hi, int(pos.max()) - W)
                u = u + sdf[:, None]
    return worst_lo, worst_hi, widest


@pytest.mark.parametrize("n,d,a,axis_prop,max_items", [
    (256, 363, 180, -1.0, None),
    (192, 272, 180, 140.5, None),
    (1024, 1448, 720, -1.0, 160),
    (2048, 2897, 900, -1.0, 60),     # items of more than 64 chunks: the 64-chunk window batches
])
def test_pair_kernel_reads_stay_in_the_staged_window(n, d, a, axis_prop, max_items):
    items, _, s, c, axis = _plan(n, d, a, float(np.float32(np.pi / a)), axis_prop, 16, 192)
    lo, hi, widest = _kernel_windows_and_samples(n, d, a, axis_prop, items, s, c, axis, max_items)
    assert lo >= 0, lo          # no entry left of the window
    assert hi <= -1, hi         # no entry at or right of position W
    assert widest <= 192


def _fma32(a, b, c):
    """float32 fused multiply-add: the float32 product is exact in float64."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(np.float32)


def _bp_sample_entries(n, d, a, axis_prop, line):
    """The fast BP's sample entries, recomputed in float32 the way the kernel
    forms them (uir_fast.hip, k_bp_tab): per angle and 64 x 64 super-tile the
    window start umin from the super-tile's corners, the window coordinate
    base at its corner rounded once from double, and each pixel's coordinate
    fma(-ys, sin, fma(lane, cos, base)), at the four corner pixels (the
    coordinate is affine in the pixel, so they bound it up to a rounding).
    Returns the smallest entry index, the largest index read (the line
    model's second tap included) and the widest window the kernel's own
    check sees."""
    f32 = np.float32
    L = uir.load()
    s = np.zeros(a, np.float32)
    c = np.zeros(a, np.float32)
    fp = C.POINTER(C.c_float)
    L.uir_geom_lut.argtypes = [C.c_uint32, C.c_float, fp, fp]
    L.uir_geom_lut(a, float(np.float32(np.pi / a)), s.ctypes.data_as(fp), c.ctypes.data_as(fp))
    L.uir_geom_axis.restype = C.c_float
    L.uir_geom_axis.argtypes = [C.c_float, C.c_uint32]
    axis = f32(L.uir_geom_axis(axis_prop, d))
    D, N = f32(d), f32(n)
    diff = D - axis * f32(2)
    half = (D - axis) if diff < 0 else axis
    edge0, edge1 = axis - half + f32(0.5), axis + half - f32(0.5)
    shift = D / f32(2) - diff / f32(2)
    clamp_u = not line and (np.floor(edge0 - f32(0.5)) != edge0 - f32(0.5) or
                            np.floor(edge1 - f32(0.5)) != edge1 - f32(0.5))
    t0 = np.arange(0, n, 64, dtype=np.float32) + f32(0.5) - N / f32(2)     # super-tile corners
    cx0, cy0 = [m.ravel()[None, :] for m in np.meshgrid(t0, t0)]
    cx1, cy1 = cx0 + f32(63), cy0 + f32(63)
    st, ct = s[:, None], c[:, None]
    us = [cx * ct - cy * st for cx in (cx0, cx1) for cy in (cy0, cy1)]       # float32, no contraction
    umn = np.minimum(np.minimum(us[0], us[1]), np.minimum(us[2], us[3])) + shift
    umx = np.maximum(np.maximum(us[0], us[1]), np.maximum(us[2], us[3])) + shift
    if clamp_u:
        umn, umx = np.clip(umn, edge0, edge1), np.clip(umx, edge0, edge1)
    umin = np.floor(umn - f32(0.5)).astype(np.int64) - 1
    widest = int((np.floor(umx - f32(0.5)).astype(np.int64) + 2 - umin).max())
    base = ((cx0.astype(np.float64) * ct.astype(np.float64) - cy0.astype(np.float64) * st.astype(np.float64)) +
            float(shift) - (0.5 if line else 1.0) - umin).astype(np.float32)
    lo, hi = 1 << 30, -(1 << 30)
    for lane in (0, 63):
        for ys in (0, 63):
            w = _fma32(-f32(ys), st, _fma32(f32(lane), ct, base))
            if clamp_u:
                w = np.clip(w, f32(edge0 - f32(1)) - umin.astype(np.float32),
                            f32(edge1 - f32(1)) - umin.astype(np.float32))
            for dw in (-2e-4, 2e-4):  # a rounding either way
                x = w.astype(np.float64) + dw
                i = np.floor(x) if line else np.rint(x)
                lo, hi = min(lo, int(i.min())), max(hi, int(i.max()) + (1 if line else 0))
    return lo, hi, widest


@pytest.mark.parametrize("n,d,a,axis_prop,line", [
    (1024, 1448, 720, -1.0, False),
    (2048, 2897, 900, -1.0, False),
    (97, 137, 60, 70.3, False),       # an edge inside a bin: the clamped coordinate
    (256, 363, 180, 140.5, False),
    (192, 272, 180, -1.0, True),
    (1024, 1448, 720, -1.0, True),
    (97, 137, 60, 40.0, True),
])
def test_bp_reads_stay_in_the_staged_window(n, d, a, axis_prop, line):
    """Every sample of the fast BP reads an entry of its angle's staged
    window (100 entries: the kernel stages bins umin .. umin + 99), for
    every pixel of every 64 x 64 super-tile, the 32-row tiles included (they
    take their super-tile's window and coordinates)."""
    lo, hi, widest = _bp_sample_entries(n, d, a, axis_prop, line)
    assert widest < 100, widest  # the kernel's own footprint check passes
    assert lo >= 0, lo
    assert hi <= 99, hi

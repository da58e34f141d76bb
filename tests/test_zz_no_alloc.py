"""The allocation-free strided "dot" (uir_op_dot_compat_noalloc) against the
allocating entry point it mirrors (uir_op_dot_compat,
ufo-ir-basic-ops-processor.c:148-196) and the oracle: bit-identical, on two
calls in a row with the second smaller, through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import uir

pytestmark = pytest.mark.gpu


def _dot_existing(L, da, db, w, h):
    r = C.c_float()
    uir.check(L.uir_op_dot_compat(da.ptr, db.ptr, w, h, C.byref(r), None))
    return np.float32(r.value)


def _dot_noalloc(L, da, db, w, h, dres):
    uir.check(L.uir_op_dot_compat_noalloc(da.ptr, db.ptr, w, h, dres.ptr, None))
    uir.check(L.uir_stream_synchronize(None))
    return dres.numpy()[0]


def test_noalloc_dot_matches_allocating_dot(res):
    L = uir.load()
    rng = np.random.default_rng(11)
    dres = uir.DeviceArray((1,))
    out = []
    # the second call is smaller: its partials must not mix with the first's
    for (h, w) in ((4096, 512), (1000, 300), (65536, 3)):
        print("case noalloc dot", h, "x", w, flush=True)
        a = rng.standard_normal((h, w)).astype(np.float32)
        b = rng.standard_normal((h, w)).astype(np.float32)
        da, db = uir.DeviceArray.from_numpy(a), uir.DeviceArray.from_numpy(b)
        got = _dot_noalloc(L, da, db, w, h, dres)
        want = _dot_existing(L, da, db, w, h)
        assert got == want, (h, w, got, want)
        assert got == np.float32(O.dot_compat(a, b))     # the bit-exact restatement
        out.append(got)
    assert out[0] != out[1]


def test_noalloc_dot_refuses_heights_beyond_its_array(res):
    L = uir.load()
    dres = uir.DeviceArray((1,))
    a = uir.DeviceArray((65537, 1))
    with pytest.raises(uir.UirError) as e:
        uir.check(L.uir_op_dot_compat_noalloc(a.ptr, a.ptr, 1, 65537, dres.ptr, None))
    assert e.value.status == 1 and "65536" in str(e.value)
    with pytest.raises(uir.UirError) as e:
        uir.check(L.uir_op_dot_compat_noalloc(a.ptr, a.ptr, 1, 0, dres.ptr, None))
    assert e.value.status == 1

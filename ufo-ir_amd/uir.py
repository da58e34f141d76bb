"""Python mirror of libuir's C ABI (include/uir.h), over ctypes.

It exposes the reference's plugin surface: tasks created by plugin name
("parallel-projector", "sirt", "sart", "asdpocs", "sbtv") with the reference's
property names, set up on a device, asked for their requisition and run with
process(), plus the kernel-level entry points.  Device memory is owned by
DeviceArray (uir_device_malloc); numpy is used on the host side only.

The library must be built (make -C ufo-ir_amd, or __graft_entry__.build());
load() raises if libuir.so is missing, and nothing here falls back to CPU code.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libuir.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "uir.h")

UIR_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "OUT_OF_MEMORY", 3: "DEVICE", 4: "UNSUPPORTED",
                5: "SETUP", 6: "NOT_FOUND", 7: "BUSY", 8: "CANCELLED", 9: "TIMEOUT", 10: "NO_DEVICE", 11: "IO"}
DTYPE_F32, DTYPE_F16, DTYPE_BF16 = 0, 1, 2


class UirError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), message))
        self.status = status


class Buffer(C.Structure):
    _fields_ = [("data", C.c_void_p), ("dims", C.c_uint32 * 2), ("dtype", C.c_int)]


class Requisition(C.Structure):
    _fields_ = [("n_dims", C.c_uint32), ("dims", C.c_uint64 * 2)]


class Subset(C.Structure):
    _fields_ = [("offset", C.c_uint32), ("n", C.c_uint32), ("direction", C.c_int32)]


class GeometryDims(C.Structure):
    _fields_ = [("height", C.c_uint64), ("width", C.c_uint64), ("n_dets", C.c_uint64), ("n_angles", C.c_uint64)]


class SliceSpec(C.Structure):
    _fields_ = [("n", C.c_uint32), ("n_dets", C.c_uint32), ("n_angles", C.c_uint32), ("step", C.c_float),
                ("axis_position", C.c_float), ("relaxation", C.c_float), ("positive_constraint", C.c_int32),
                ("ray_weight_min", C.c_float), ("model", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("slice", SliceSpec), ("method", C.c_uint32), ("iterations", C.c_uint32),
                ("inner_iterations", C.c_uint32), ("lambda_", C.c_float), ("mu", C.c_float)]


class ServerConfig(C.Structure):
    _fields_ = [("storage", C.c_int), ("max_active", C.c_uint32), ("hbm_budget", C.c_uint64),
                ("hbm_reserve", C.c_uint64), ("autotune", C.c_int32)]


class Tile(C.Structure):
    _fields_ = [("fp_angles", C.c_int32), ("bp_angles", C.c_int32), ("bp_rows", C.c_int32)]


class TuneResult(C.Structure):
    _fields_ = [("best", Tile), ("fp_ms", C.c_float * 3), ("bp_ms", C.c_float * 4), ("probe_slices", C.c_uint32)]


class RequestStats(C.Structure):
    _fields_ = [("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("device_bytes", C.c_uint64), ("steps", C.c_uint32), ("max_batch", C.c_uint32)]


class ServerInfo(C.Structure):
    _fields_ = [("budget", C.c_uint64), ("used", C.c_uint64), ("active", C.c_uint32), ("pending", C.c_uint32),
                ("steps", C.c_uint64), ("completed", C.c_uint64), ("busy_s", C.c_double),
                ("slice_iterations", C.c_double)]


class JsonReport(C.Structure):
    _fields_ = [("frames", C.c_uint32), ("tasks", C.c_uint32), ("seconds", C.c_double)]


class BatchInfo(C.Structure):
    _fields_ = [("device_bytes", C.c_uint64), ("n_slices", C.c_uint32), ("n_geometries", C.c_uint32),
                ("fp_items", C.c_uint32), ("bp_items", C.c_uint32), ("fp_grid", C.c_uint32),
                ("bp_grid", C.c_uint32), ("fp_samples", C.c_double), ("bp_samples", C.c_double)]


class Ceilings(C.Structure):
    _fields_ = [("n_cu", C.c_int), ("clock_ghz", C.c_double), ("lds_bytes_per_clock_cu", C.c_double),
                ("fma_wave64_per_clock_cu", C.c_double), ("cvt_wave64_per_clock_cu", C.c_double),
                ("lds_ms", C.c_double), ("fma_ms", C.c_double)]


class AsdPocsParams(C.Structure):
    _fields_ = [("beta", C.c_float), ("beta_red", C.c_float), ("alpha", C.c_float), ("alpha_red", C.c_float),
                ("r_max", C.c_float), ("ng", C.c_uint32), ("df_iterations", C.c_uint32),
                ("positive_constraint", C.c_int32), ("compat", C.c_int32)]


_lib = None


def load():
    """Load libuir.so (built in-tree); raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libuir.so is not built (%s); run make -C ufo-ir_amd" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, st, sz, u32, f32 = C.c_void_p, C.c_int, C.c_size_t, C.c_uint32, C.c_float
    L.uir_last_error.restype = C.c_char_p
    L.uir_status_string.restype = C.c_char_p
    L.uir_abi_version.restype = u32
    L.uir_device_count.argtypes = [C.POINTER(C.c_int)]
    L.uir_resources_new.argtypes = [C.c_int, C.POINTER(vp)]
    L.uir_resources_free.argtypes = [vp]
    L.uir_resources_free.restype = None
    L.uir_resources_stream.argtypes = [vp]
    L.uir_resources_stream.restype = vp
    L.uir_resources_synchronize.argtypes = [vp]
    L.uir_device_malloc.argtypes = [sz, C.POINTER(vp)]
    L.uir_device_free.argtypes = [vp]
    for n in ("uir_memcpy_htod", "uir_memcpy_dtoh", "uir_memcpy_dtod"):
        getattr(L, n).argtypes = [vp, vp, sz, vp]
    L.uir_stream_synchronize.argtypes = [vp]
    L.uir_mem_get_info.argtypes = [C.POINTER(sz), C.POINTER(sz)]
    L.uir_timer_new.argtypes = [C.POINTER(vp)]
    L.uir_timer_free.argtypes = [vp]
    L.uir_timer_free.restype = None
    L.uir_timer_start.argtypes = [vp, vp]
    L.uir_timer_stop.argtypes = [vp, vp]
    L.uir_timer_elapsed_ms.argtypes = [vp, C.POINTER(f32)]
    for n in ("uir_ir_parallel_projector_task_new", "uir_ir_sirt_task_new", "uir_ir_sart_task_new",
              "uir_ir_asdpocs_task_new", "uir_ir_sbtv_task_new"):
        getattr(L, n).restype = vp
        getattr(L, n).argtypes = []
    L.uir_task_new.argtypes = [C.c_char_p]
    L.uir_task_new.restype = vp
    L.uir_task_ref.argtypes = [vp]
    L.uir_task_ref.restype = vp
    L.uir_task_unref.argtypes = [vp]
    L.uir_task_unref.restype = None
    L.uir_task_plugin.argtypes = [vp]
    L.uir_task_plugin.restype = C.c_char_p
    L.uir_task_set_float.argtypes = [vp, C.c_char_p, f32]
    L.uir_task_set_uint.argtypes = [vp, C.c_char_p, u32]
    L.uir_task_set_bool.argtypes = [vp, C.c_char_p, C.c_int]
    L.uir_task_set_string.argtypes = [vp, C.c_char_p, C.c_char_p]
    L.uir_task_set_object.argtypes = [vp, C.c_char_p, vp]
    L.uir_task_get_float.argtypes = [vp, C.c_char_p, C.POINTER(f32)]
    L.uir_task_get_uint.argtypes = [vp, C.c_char_p, C.POINTER(u32)]
    L.uir_task_get_bool.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int)]
    L.uir_task_get_string.argtypes = [vp, C.c_char_p, C.c_char_p, sz]
    L.uir_task_get_object.argtypes = [vp, C.c_char_p, C.POINTER(vp)]
    L.uir_task_setup.argtypes = [vp, vp]
    L.uir_task_get_requisition.argtypes = [vp, C.POINTER(Buffer), C.POINTER(Requisition)]
    L.uir_task_process.argtypes = [vp, C.POINTER(Buffer), C.POINTER(Buffer)]
    L.uir_task_time_kernels.argtypes = [vp, C.POINTER(Buffer), C.POINTER(Buffer), u32, C.POINTER(f32),
                                        C.POINTER(f32)]
    L.uir_projector_forward.argtypes = [vp, C.POINTER(Buffer), C.POINTER(Buffer)]
    L.uir_projector_backward.argtypes = [vp, C.POINTER(Buffer), C.POINTER(Buffer)]
    L.uir_parallel_projector_subset_fp.argtypes = [vp, C.POINTER(Buffer), C.POINTER(Buffer), C.POINTER(Subset)]
    L.uir_parallel_projector_subset_bp.argtypes = [vp, C.POINTER(Buffer), C.POINTER(Buffer), C.POINTER(Subset)]
    L.uir_parallel_projector_get_luts.argtypes = [vp, C.POINTER(C.POINTER(f32)), C.POINTER(C.POINTER(f32)),
                                                  C.POINTER(u32)]
    L.uir_parallel_projector_get_subsets.argtypes = [vp, C.POINTER(C.POINTER(Subset)), C.POINTER(u32)]
    L.uir_asdpocs_get_diagnostics.argtypes = [vp, C.POINTER(C.POINTER(f32)), C.POINTER(u32)]
    L.uir_fp_joseph.argtypes = [vp, vp, vp, vp, vp, GeometryDims, f32, Subset, f32, vp]
    L.uir_bp_joseph.argtypes = [vp, vp, vp, f32, vp, vp, GeometryDims, f32, Subset, vp]
    L.uir_batch_new.argtypes = [vp, C.c_int, C.POINTER(SliceSpec), u32, C.POINTER(vp)]
    L.uir_batch_free.argtypes = [vp]
    L.uir_batch_free.restype = None
    L.uir_batch_set_sinogram.argtypes = [vp, u32, vp, C.c_int]
    L.uir_batch_reset.argtypes = [vp]
    L.uir_batch_iterate.argtypes = [vp, u32, C.c_int]
    L.uir_asdpocs_params_default.argtypes = [C.POINTER(AsdPocsParams)]
    L.uir_asdpocs_params_default.restype = None
    L.uir_batch_asdpocs.argtypes = [vp, u32, C.POINTER(AsdPocsParams), C.c_int]
    L.uir_batch_asdpocs_diagnostics.argtypes = [vp, u32, vp, u32, C.POINTER(u32)]
    L.uir_batch_get_image.argtypes = [vp, u32, vp, C.c_int]
    L.uir_batch_get_info.argtypes = [vp, C.POINTER(BatchInfo)]
    L.uir_batch_time_kernels.argtypes = [vp, u32, C.POINTER(f32), C.POINTER(f32)]
    L.uir_clock.restype = C.c_double
    L.uir_clock.argtypes = []
    L.uir_server_new.argtypes = [C.c_int, C.POINTER(ServerConfig), C.POINTER(vp)]
    L.uir_server_submit.argtypes = [vp, C.POINTER(Request), vp, C.POINTER(C.c_uint64)]
    L.uir_server_wait.argtypes = [vp, C.c_uint64, vp, C.c_double, C.POINTER(RequestStats)]
    L.uir_server_get_info.argtypes = [vp, C.POINTER(ServerInfo)]
    L.uir_request_footprint.argtypes = [C.POINTER(Request), C.c_int, C.POINTER(C.c_uint64)]
    L.uir_server_free.argtypes = [vp]
    L.uir_server_free.restype = None
    L.uir_task_property_type.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int)]
    L.uir_tune_geometry.argtypes = [vp, C.POINTER(SliceSpec), C.c_int, u32, C.POINTER(TuneResult)]
    L.uir_tune_get.argtypes = [C.POINTER(SliceSpec), C.c_int, C.POINTER(Tile)]
    L.uir_tune_set.argtypes = [C.POINTER(SliceSpec), C.c_int, C.POINTER(Tile)]
    L.uir_tune_clear.argtypes = []
    L.uir_tune_clear.restype = None
    L.uir_tiff_info.argtypes = [C.c_char_p, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]
    L.uir_tiff_read.argtypes = [C.c_char_p, u32, vp, sz, C.POINTER(u32), C.POINTER(u32)]
    L.uir_tiff_write.argtypes = [C.c_char_p, vp, u32, u32, C.c_int]
    L.uir_json_run.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(JsonReport)]
    L.uir_op_set.argtypes = [vp, sz, f32, vp]
    L.uir_op_inv.argtypes = [vp, vp, sz, vp]
    for n in ("uir_op_mul", "uir_op_add", "uir_op_deduction", "uir_op_max_ew", "uir_op_div_ew"):
        getattr(L, n).argtypes = [vp, vp, vp, sz, vp]
    for n in ("uir_op_deduction2", "uir_op_add2"):
        getattr(L, n).argtypes = [vp, vp, f32, vp, sz, vp]
    L.uir_op_mul_rows.argtypes = [vp, vp, vp, u32, u32, u32, vp]
    L.uir_op_positive_constraint.argtypes = [vp, vp, sz, vp]
    L.uir_op_mul_scalar.argtypes = [vp, sz, f32, vp]
    L.uir_op_sqrt.argtypes = [vp, sz, vp]
    L.uir_op_normalization.argtypes = [vp, sz, vp]
    L.uir_op_l1_norm.argtypes = [vp, u32, u32, C.POINTER(f32), vp]
    L.uir_op_dot_compat.argtypes = [vp, vp, u32, u32, C.POINTER(f32), vp]
    L.uir_op_dot_compat_noalloc.argtypes = [vp, vp, u32, u32, vp, vp]
    L.uir_op_l2_compat.argtypes = [vp, u32, u32, C.POINTER(f32), vp]
    L.uir_op_dot_f64.argtypes = [vp, vp, sz, C.POINTER(C.c_double), vp]
    for n in ("uir_op_dx", "uir_op_dxt", "uir_op_dy", "uir_op_dyt", "uir_op_l1_grad"):
        getattr(L, n).argtypes = [vp, vp, u32, u32, vp]
    L.uirk_probe_ceilings.argtypes = [vp, C.POINTER(Ceilings)]
    _lib = L
    return L


def check(status):
    if status != UIR_OK:
        raise UirError(status, load().uir_last_error().decode(errors="replace"))


def device_count():
    n = C.c_int(0)
    load().uir_device_count(C.byref(n))
    return n.value


def abi_version():
    return load().uir_abi_version()


def mem_get_info():
    f, t = C.c_size_t(0), C.c_size_t(0)
    check(load().uir_mem_get_info(C.byref(f), C.byref(t)))
    return f.value, t.value


class Resources:
    """A device plus its in-order stream (UfoResources + UfoGpuNode queue)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(load().uir_resources_new(int(device), C.byref(h)))
        self.handle = h
        self.device = device

    @property
    def stream(self):
        return load().uir_resources_stream(self.handle)

    def synchronize(self):
        check(load().uir_resources_synchronize(self.handle))

    def close(self):
        if self.handle:
            load().uir_resources_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_NP_OF = {DTYPE_F32: np.float32, DTYPE_F16: np.float16, DTYPE_BF16: np.uint16}


class DeviceArray:
    """A 2-D device buffer ([height][width]) with UfoRequisition dims order."""

    def __init__(self, shape, dtype=DTYPE_F32):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = dtype
        self.itemsize = 2 if dtype in (DTYPE_F16, DTYPE_BF16) else 4
        self.nbytes = int(np.prod(self.shape)) * self.itemsize
        p = C.c_void_p()
        check(load().uir_device_malloc(self.nbytes, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr, stream=None):
        arr = np.ascontiguousarray(arr, dtype=np.float32)
        d = cls(arr.shape)
        check(load().uir_memcpy_htod(d.ptr, arr.ctypes.data, d.nbytes, stream))
        return d

    @classmethod
    def zeros(cls, shape, stream=None):
        d = cls(shape)
        check(load().uir_op_set(d.ptr, int(np.prod(shape)), 0.0, stream))
        check(load().uir_stream_synchronize(stream))
        return d

    def numpy(self, stream=None):
        out = np.empty(self.shape, dtype=_NP_OF[self.dtype])
        check(load().uir_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes, stream))
        return out

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr, dtype=_NP_OF[self.dtype])
        assert arr.nbytes == self.nbytes
        check(load().uir_memcpy_htod(self.ptr, arr.ctypes.data, self.nbytes, stream))

    def buffer(self):
        b = Buffer()
        b.data = self.ptr
        b.dims[0] = self.shape[1] if len(self.shape) > 1 else self.shape[0]
        b.dims[1] = self.shape[0] if len(self.shape) > 1 else 1
        b.dtype = self.dtype
        return b

    @property
    def size(self):
        return int(np.prod(self.shape))

    def free(self):
        if self.ptr:
            load().uir_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _c(s):
    return s.encode() if isinstance(s, str) else s


class Task:
    """A reference task plugin: Task('sirt', num_iterations=50, relaxation_factor=1.0)."""

    def __init__(self, plugin, **props):
        h = load().uir_task_new(_c(plugin))
        if not h:
            raise UirError(6, load().uir_last_error().decode())
        self.handle = h
        self._children = []
        self.set(**props)

    def set(self, **props):
        for k, v in props.items():
            self[k] = v
        return self

    def __setitem__(self, name, v):
        L = load()
        n = _c(name)
        if isinstance(v, Task):
            check(L.uir_task_set_object(self.handle, n, v.handle))
            self._children.append(v)
        elif isinstance(v, bool):
            check(L.uir_task_set_bool(self.handle, n, int(v)))
        elif isinstance(v, str):
            check(L.uir_task_set_string(self.handle, n, _c(v)))
        elif isinstance(v, int) and not isinstance(v, bool):
            st = L.uir_task_set_uint(self.handle, n, v) if v >= 0 else 1
            if st != UIR_OK:
                check(L.uir_task_set_float(self.handle, n, float(v)))
        else:
            check(L.uir_task_set_float(self.handle, n, float(v)))

    def get_float(self, name):
        v = C.c_float()
        check(load().uir_task_get_float(self.handle, _c(name), C.byref(v)))
        return v.value

    def get_uint(self, name):
        v = C.c_uint32()
        check(load().uir_task_get_uint(self.handle, _c(name), C.byref(v)))
        return v.value

    def get_bool(self, name):
        v = C.c_int()
        check(load().uir_task_get_bool(self.handle, _c(name), C.byref(v)))
        return bool(v.value)

    def get_string(self, name):
        buf = C.create_string_buffer(256)
        check(load().uir_task_get_string(self.handle, _c(name), buf, 256))
        return buf.value.decode()

    @property
    def plugin(self):
        return load().uir_task_plugin(self.handle).decode()

    def setup(self, res):
        check(load().uir_task_setup(self.handle, res.handle))
        self._res = res
        return self

    def requisition(self, inp):
        r = Requisition()
        b = inp.buffer()
        check(load().uir_task_get_requisition(self.handle, C.byref(b), C.byref(r)))
        return (int(r.dims[1]), int(r.dims[0]))  # (height, width)

    def process(self, inp, out):
        bi, bo = inp.buffer(), out.buffer()
        check(load().uir_task_process(self.handle, C.byref(bi), C.byref(bo)))

    def time_kernels(self, inp, out, stride=1):
        """(FP, BP) average launch durations in us over one timed sweep (sart)."""
        f, b = C.c_float(), C.c_float()
        bi, bo = inp.buffer(), out.buffer()
        check(load().uir_task_time_kernels(self.handle, C.byref(bi), C.byref(bo), int(stride), C.byref(f),
                                           C.byref(b)))
        return f.value, b.value

    def luts(self):
        s, c, n = C.POINTER(C.c_float)(), C.POINTER(C.c_float)(), C.c_uint32()
        check(load().uir_parallel_projector_get_luts(self.handle, C.byref(s), C.byref(c), C.byref(n)))
        return (np.ctypeslib.as_array(s, (n.value,)).copy(), np.ctypeslib.as_array(c, (n.value,)).copy())

    def subsets(self):
        p, n = C.POINTER(Subset)(), C.c_uint32()
        check(load().uir_parallel_projector_get_subsets(self.handle, C.byref(p), C.byref(n)))
        return [(p[i].offset, p[i].n, p[i].direction) for i in range(n.value)]

    def forward(self, vol, sino):
        bi, bo = vol.buffer(), sino.buffer()
        check(load().uir_projector_forward(self.handle, C.byref(bi), C.byref(bo)))

    def backward(self, sino, vol):
        bi, bo = sino.buffer(), vol.buffer()
        check(load().uir_projector_backward(self.handle, C.byref(bi), C.byref(bo)))

    def subset_fp(self, vol, sino, subset):
        bv, bs = vol.buffer(), sino.buffer()
        s = Subset(*subset)
        check(load().uir_parallel_projector_subset_fp(self.handle, C.byref(bv), C.byref(bs), C.byref(s)))

    def subset_bp(self, vol, sino, subset):
        bv, bs = vol.buffer(), sino.buffer()
        s = Subset(*subset)
        check(load().uir_parallel_projector_subset_bp(self.handle, C.byref(bv), C.byref(bs), C.byref(s)))

    def asdpocs_diagnostics(self):
        p, n = C.POINTER(C.c_float)(), C.c_uint32()
        check(load().uir_asdpocs_get_diagnostics(self.handle, C.byref(p), C.byref(n)))
        if n.value == 0:
            return np.zeros((0, 4), np.float32)
        return np.ctypeslib.as_array(p, (n.value * 4,)).reshape(-1, 4).copy()

    def close(self):
        if self.handle:
            load().uir_task_unref(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Timer:
    def __init__(self):
        h = C.c_void_p()
        check(load().uir_timer_new(C.byref(h)))
        self.handle = h

    def start(self, stream):
        check(load().uir_timer_start(self.handle, stream))

    def stop(self, stream):
        check(load().uir_timer_stop(self.handle, stream))

    def elapsed_ms(self):
        v = C.c_float()
        check(load().uir_timer_elapsed_ms(self.handle, C.byref(v)))
        return v.value

    def __del__(self):
        try:
            load().uir_timer_free(self.handle)
        except Exception:
            pass


def header_functions(path=HEADER_PATH):
    """Names of the functions include/uir.h declares."""
    import re
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(uir_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


DTYPES = {"f32": DTYPE_F32, "f16": DTYPE_F16, "bf16": DTYPE_BF16,
          "f16-sino": 3, "bf16-sino": 4}   # sinograms FP16 / BF16, images FP32


MODELS = {"joseph": 0, "line": 1}


def slice_spec(s):
    """A uir_slice_spec from a dict (n, n_dets, n_angles, step, axis_position,
    relaxation, positive_constraint, ray_weight_min, model: "joseph" or
    "line")."""
    return SliceSpec(int(s["n"]), int(s["n_dets"]), int(s["n_angles"]), float(s["step"]),
                     float(s.get("axis_position", -1.0)), float(s.get("relaxation", 0.25)),
                     int(bool(s.get("positive_constraint", False))), float(s.get("ray_weight_min", 0.0)),
                     MODELS[s.get("model", "joseph")])


def _tile_dict(t):
    return {"fp_angles": t.fp_angles, "bp_angles": t.bp_angles, "bp_rows": t.bp_rows}


def tune_geometry(res, spec, storage="f16", reps=3):
    """Times every tile shape on a probe of the geometry; records and returns the best."""
    r = TuneResult()
    sp = slice_spec(spec)
    check(load().uir_tune_geometry(res.handle, C.byref(sp), DTYPES[storage], int(reps), C.byref(r)))
    return {"best": _tile_dict(r.best), "fp_ms": list(r.fp_ms), "bp_ms": list(r.bp_ms),
            "probe_slices": r.probe_slices}


def tune_get(spec, storage="f16"):
    """(tile dict, tuned?) for a geometry."""
    t = Tile()
    sp = slice_spec(spec)
    rc = load().uir_tune_get(C.byref(sp), DTYPES[storage], C.byref(t))
    if rc not in (UIR_OK, 6):
        check(rc)
    return _tile_dict(t), rc == UIR_OK


def tune_set(spec, storage, fp_angles, bp_angles, bp_rows):
    sp = slice_spec(spec)
    t = Tile(int(fp_angles), int(bp_angles), int(bp_rows))
    check(load().uir_tune_set(C.byref(sp), DTYPES[storage], C.byref(t)))


def tune_clear():
    load().uir_tune_clear()


def probe_ceilings(res):
    """Measured LDS read rate and FP32 FMA issue rate of the device (uir_probe.hip).

    Returns bytes and wave64 FMAs per clock per CU at the device's reported
    clock, and the chip-wide TB/s and TFLOP/s they imply."""
    c = Ceilings()
    rc = load().uirk_probe_ceilings(res.stream, C.byref(c))
    if rc != 0:
        raise UirError(rc, "uirk_probe_ceilings failed")
    hz = c.clock_ghz * 1e9
    return {"n_cu": c.n_cu, "clock_ghz": c.clock_ghz,
            "lds_bytes_per_clock_cu": round(c.lds_bytes_per_clock_cu, 2),
            "lds_tb_s": round(c.lds_bytes_per_clock_cu * c.n_cu * hz / 1e12, 2),
            "fma_wave64_per_clock_cu": round(c.fma_wave64_per_clock_cu, 3),
            "fp32_tflop_s": round(c.fma_wave64_per_clock_cu * 64 * 2 * c.n_cu * hz / 1e12, 1),
            "cvt_wave64_per_clock_cu": round(c.cvt_wave64_per_clock_cu, 3),
            "lds_ms": round(c.lds_ms, 3), "fma_ms": round(c.fma_ms, 3)}


class Batch:
    """Batched SIRT over slices of uneven geometry (uir_batch_*).

    specs: list of dicts with n, n_dets, n_angles, step, axis_position (-1),
    relaxation, positive_constraint.  storage: "f32", "f16" or "bf16"."""

    def __init__(self, res, specs, storage="f32"):
        arr = (SliceSpec * len(specs))()
        for i, s in enumerate(specs):
            arr[i] = slice_spec(s)
        h = C.c_void_p()
        check(load().uir_batch_new(res.handle, DTYPES[storage], arr, len(specs), C.byref(h)))
        self.handle = h
        self.specs = [dict(s) for s in specs]
        self.res = res

    def set_sinogram(self, i, sino):
        if isinstance(sino, DeviceArray):
            check(load().uir_batch_set_sinogram(self.handle, i, sino.ptr, 1))
        else:
            a = np.ascontiguousarray(sino, np.float32)
            check(load().uir_batch_set_sinogram(self.handle, i, a.ctypes.data, 0))

    def reset(self):
        check(load().uir_batch_reset(self.handle))

    def iterate(self, iterations, use_graph=True):
        check(load().uir_batch_iterate(self.handle, int(iterations), int(bool(use_graph))))

    def asdpocs(self, iterations, use_graph=True, **params):
        """Batched ASD-POCS (uir_batch_asdpocs); params: beta, beta_red,
        alpha, alpha_red, r_max, ng, df_iterations, positive_constraint,
        compat (defaults of uir_asdpocs_params_default)."""
        p = AsdPocsParams()
        load().uir_asdpocs_params_default(C.byref(p))
        for k, v in params.items():
            if k not in dict(AsdPocsParams._fields_):
                raise TypeError("unknown ASD-POCS parameter %r" % k)
            setattr(p, k, int(v) if isinstance(v, bool) else v)
        check(load().uir_batch_asdpocs(self.handle, int(iterations), C.byref(p), int(bool(use_graph))))

    def asdpocs_diagnostics(self, i, max_rows=100000):
        """(iterations, 4) rows dd, dp, dg, dtgv of slice i from the last asdpocs()."""
        n = C.c_uint32()
        check(load().uir_batch_asdpocs_diagnostics(self.handle, i, None, 0, C.byref(n)))
        rows = np.zeros((min(n.value, max_rows), 4), np.float32)
        if len(rows):
            check(load().uir_batch_asdpocs_diagnostics(self.handle, i, rows.ctypes.data, len(rows), C.byref(n)))
        return rows

    def image(self, i):
        n = int(self.specs[i]["n"])
        out = np.empty((n, n), np.float32)
        check(load().uir_batch_get_image(self.handle, i, out.ctypes.data, 0))
        return out

    def info(self):
        inf = BatchInfo()
        check(load().uir_batch_get_info(self.handle, C.byref(inf)))
        return {k: getattr(inf, k) for k, _ in BatchInfo._fields_}

    def time_kernels(self, reps=10):
        """Average FP and BP launch durations (ms).  Runs `reps` FP and BP
        launches, so it changes the images: reset() afterwards."""
        a, b = C.c_float(), C.c_float()
        check(load().uir_batch_time_kernels(self.handle, int(reps), C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if self.handle:
            load().uir_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


METHODS = {"sirt": 0, "sbtv-sirt": 1}
STATUS_TIMEOUT = 9


def make_request(n, n_dets, n_angles, step, method="sirt", iterations=10, inner_iterations=1, relaxation=0.25,
                 lam=0.1, mu=1.0, axis_position=-1.0, positive_constraint=False, ray_weight_min=0.0,
                 model="joseph"):
    """A uir_request (method "sirt" or "sbtv-sirt"; model "joseph" or
    "line")."""
    return Request(SliceSpec(int(n), int(n_dets), int(n_angles), float(step), float(axis_position),
                             float(relaxation), int(bool(positive_constraint)), float(ray_weight_min),
                             MODELS[model]),
                   METHODS[method], int(iterations), int(inner_iterations), float(lam), float(mu))


def request_footprint(req, storage="f16"):
    v = C.c_uint64()
    check(load().uir_request_footprint(C.byref(req), DTYPES[storage], C.byref(v)))
    return v.value


def clock():
    return load().uir_clock()


class Server:
    """Continuous-batching reconstruction server (uir_server_*).

    submit() copies the request and its FP32 sinogram and returns a ticket;
    wait(ticket) returns (image, stats).  ctypes releases the GIL inside the
    calls, so submitters and waiters may run in different Python threads."""

    def __init__(self, device=0, storage="f16", max_active=0, hbm_budget=0, hbm_reserve=0, autotune=False):
        cfg = ServerConfig(DTYPES[storage], int(max_active), int(hbm_budget), int(hbm_reserve), int(bool(autotune)))
        h = C.c_void_p()
        check(load().uir_server_new(int(device), C.byref(cfg), C.byref(h)))
        self.handle = h
        self.sizes = {}

    def submit(self, req, sinogram):
        a = np.ascontiguousarray(sinogram, np.float32)
        if a.shape != (req.slice.n_angles, req.slice.n_dets):
            raise ValueError("sinogram shape %s, request wants %s" % (a.shape, (req.slice.n_angles, req.slice.n_dets)))
        t = C.c_uint64()
        check(load().uir_server_submit(self.handle, C.byref(req), a.ctypes.data, C.byref(t)))
        self.sizes[t.value] = int(req.slice.n)
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """(image, stats) of a submitted request; raises UirError with the
        request's status (TIMEOUT leaves the ticket valid)."""
        n = self.sizes.get(ticket)
        img = np.empty((n, n), np.float32) if n else None
        st = RequestStats()
        rc = load().uir_server_wait(self.handle, C.c_uint64(ticket), img.ctypes.data if n else None,
                                    float(timeout), C.byref(st))
        if rc != STATUS_TIMEOUT:
            self.sizes.pop(ticket, None)
        check(rc)
        return img, {k: getattr(st, k) for k, _ in RequestStats._fields_}

    def info(self):
        inf = ServerInfo()
        check(load().uir_server_get_info(self.handle, C.byref(inf)))
        return {k: getattr(inf, k) for k, _ in ServerInfo._fields_}

    def close(self):
        if self.handle:
            load().uir_server_free(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- frame I/O and the graph runner ----

PROP_TYPES = {0: "float", 1: "uint", 2: "bool", 3: "string", 4: "object"}


def tiff_info(path):
    """(pages, width, height) of a TIFF file."""
    n, w, h = C.c_uint32(), C.c_uint32(), C.c_uint32()
    check(load().uir_tiff_info(os.fsencode(path), C.byref(n), C.byref(w), C.byref(h)))
    return n.value, w.value, h.value


def tiff_read(path, page=0):
    """Page `page` of a TIFF file as a float32 array [height][width]."""
    w, h = C.c_uint32(), C.c_uint32()
    check(load().uir_tiff_read(os.fsencode(path), int(page), None, 0, C.byref(w), C.byref(h)))
    out = np.empty((h.value, w.value), np.float32)
    check(load().uir_tiff_read(os.fsencode(path), int(page), out.ctypes.data, out.size, C.byref(w), C.byref(h)))
    return out


def tiff_write(path, image, append=False):
    a = np.ascontiguousarray(image, np.float32)
    if a.ndim != 2:
        raise ValueError("a 2-D image is expected")
    check(load().uir_tiff_write(os.fsencode(path), a.ctypes.data, a.shape[1], a.shape[0], int(bool(append))))


def json_run(graph, base_dir=None, device=0):
    """Runs a ufo-core JSON graph (text or dict); returns the run report."""
    import json as _json
    text = graph if isinstance(graph, str) else _json.dumps(graph)
    rep = JsonReport()
    check(load().uir_json_run(text.encode(), os.fsencode(base_dir) if base_dir else None, int(device),
                              C.byref(rep)))
    return {k: getattr(rep, k) for k, _ in JsonReport._fields_}

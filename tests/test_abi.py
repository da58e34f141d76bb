"""The C-ABI library loads and exports every function include/uir.h declares;
host-side behaviour that needs no device (tasks, properties, errors)."""
import ctypes as C

import pytest

import uir


def test_library_exports_every_declared_function():
    lib = uir.load()
    names = uir.header_functions()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version():
    assert uir.abi_version() == 1


@pytest.mark.parametrize("plugin", ["parallel-projector", "sirt", "sart", "asdpocs", "sbtv",
                                    "parallel_projector", "SIRT"])
def test_factories(plugin):
    t = uir.Task(plugin)
    assert t.plugin == plugin.lower().replace("_", "-")


def test_unknown_plugin():
    with pytest.raises(uir.UirError) as e:
        uir.Task("fbp")
    assert e.value.status == 6


def test_projector_defaults_and_ranges():
    # ufo-ir-projector-task.c:68-110, ufo-ir-parallel-projector-task.c:123-154
    p = uir.Task("parallel-projector")
    assert p.get_float("axis_position") == -1.0
    assert p.get_float("step") == 0.0
    assert p.get_float("relaxation") == 1.0
    assert p.get_float("correction_scale") == 1.0
    assert p.get_bool("is_forward") is True
    assert p.get_string("model") == "joseph"
    assert p.get_uint("angles_num") == 0
    p["model"] = "JOSEPH"                         # g_ascii_strdown
    assert p.get_string("model") == "joseph"
    p["axis-position"] = 100.5                    # '-' and '_' both accepted
    assert p.get_float("axis_position") == 100.5
    with pytest.raises(uir.UirError):
        p["step"] = -1.0                          # below the GParamSpec minimum
    assert p.get_float("step") == 0.0
    with pytest.raises(uir.UirError):
        p["correction_scale"] = -1.0              # G_MINFLOAT minimum
    with pytest.raises(uir.UirError) as e:
        p["no_such_property"] = 1.0
    assert e.value.status == 6


def test_method_defaults_and_ranges():
    s = uir.Task("sirt")
    assert s.get_uint("num-iterations") == 10      # ufo-ir-method-task.c:74-79
    assert abs(s.get_float("relaxation_factor") - 0.25) < 1e-7
    with pytest.raises(uir.UirError):
        s["relaxation_factor"] = 1.5               # [0, 1]
    with pytest.raises(uir.UirError):
        s["num-iterations"] = 0                    # >= 1
    a = uir.Task("asdpocs")
    assert a.get_bool("positive_constraint") is True
    assert a.get_uint("ng") == 20
    for name, v in (("beta", 1.0), ("beta_red", 0.995), ("alpha", 0.2), ("alpha_red", 0.95), ("r_max", 0.95)):
        assert abs(a.get_float(name) - v) < 1e-7
    b = uir.Task("sbtv")
    assert abs(b.get_float("lambda") - 0.1) < 1e-7 and abs(b.get_float("mu") - 0.5) < 1e-7


def test_object_properties_type_checked():
    s = uir.Task("sirt")
    with pytest.raises(uir.UirError):
        s["projector"] = uir.Task("sart")
    a = uir.Task("asdpocs")
    with pytest.raises(uir.UirError):
        a["df_minimizer"] = uir.Task("parallel-projector")
    a["df_minimizer"] = uir.Task("sart")


def test_setup_without_device_or_projector_fails_cleanly():
    lib = uir.load()
    s = uir.Task("sirt")
    # uir_task_setup with a NULL resources handle
    assert lib.uir_task_setup(s.handle, None) == 1
    assert b"NULL" in lib.uir_last_error()


def test_status_strings():
    lib = uir.load()
    lib.uir_status_string.argtypes = [C.c_int]
    assert lib.uir_status_string(0) == b"ok"
    assert lib.uir_status_string(2) == b"out of memory"

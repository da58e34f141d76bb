"""The JSON graph runner on the GPU: json-sample.json's chain (read ->
asdpocs with a projector and a SART df-minimizer -> write, one TIFF per
frame) gives, frame by frame, the oracle's ASD-POCS (ufo-ir-asdpocs-task.c:
411-509 over ufo-ir-sart-task.c:173-228) and what the task ABI gives
directly; the sample's own properties (10 iterations of a 100-sweep SART
df-minimizer) on one frame against the oracle; a SIRT chain with a
multi-page output against the oracle's SIRT; the projector's model property
through the graph ("line" against the oracle's FP64 line model, an unknown
model refused at setup); and the uir-runjson program."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import uir
from phantoms import ellipse_phantom

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STEP = 0.034906585          # json-sample.json:29 (2 degrees)
N_ANG, N_DET = 90, 128


def _sinos(k):
    g = O.Geometry(N_DET, N_DET, N_ANG, np.float32(STEP))
    return [O.fp(g, ellipse_phantom(N_DET, seed=500 + i)).astype(np.float32) for i in range(k)]


def _graph(out_name, method_node):
    return {"nodes": [{"plugin": "read", "name": "read", "properties": {"path": "sino.tif"}},
                      {"plugin": "write", "name": "write", "properties": {"filename": out_name}},
                      dict(method_node, name="m", package="ir")],
            "edges": [{"from": {"name": "read"}, "to": {"name": "m"}},
                      {"from": {"name": "m"}, "to": {"name": "write"}}]}


def _direct(res, task, sino):
    s = uir.DeviceArray.from_numpy(sino)
    out = uir.DeviceArray((N_DET, N_DET))
    task.process(s, out)
    img = out.numpy()
    s.free()
    out.free()
    return img


def test_json_sample_chain_matches_task_api(res, tmp_path):
    sinos = _sinos(2)
    for i, s in enumerate(sinos):
        uir.tiff_write(str(tmp_path / "sino.tif"), s, append=i > 0)
    node = {"plugin": "asdpocs", "properties": {
        "projector": {"plugin": "parallel-projector", "name": "parallel-projector", "package": "ir",
                      "properties": {"model": "joseph", "is_forward": False, "step": STEP}},
        "df_minimizer": {"plugin": "sart", "package": "ir", "name": "sart_minimizer",
                         "properties": {"relaxation_factor": 0.25, "num-iterations": 3}},
        "num-iterations": 3}}
    rep = uir.json_run(_graph("out/out%03i.tif", node), base_dir=str(tmp_path))
    assert rep["frames"] == 2 and rep["tasks"] == 1
    proj = uir.Task("parallel-projector", model="joseph", is_forward=False, step=STEP)
    df = uir.Task("sart", relaxation_factor=0.25, num_iterations=3)
    asd = uir.Task("asdpocs", projector=proj, df_minimizer=df, num_iterations=3).setup(res)
    g = O.Geometry(N_DET, N_DET, N_ANG, np.float32(STEP))
    for i, s in enumerate(sinos):
        got = uir.tiff_read(str(tmp_path / "out" / ("out%03d.tif" % i)))
        assert np.array_equal(got, _direct(res, asd, s))
        ref = O.asdpocs(g, s, 3, df="sart", df_iters=3, df_relax_initial=0.25)
        assert O.rel_l2(got, ref) <= 1e-4


def test_json_sample_properties_match_oracle(tmp_path):
    """json-sample.json:17-44 as written (asdpocs num-iterations 10, SART
    df-minimizer relaxation 0.25 and num-iterations 100, projector step 2
    degrees) on one 128-bin frame of 90 angles."""
    print("case json-sample properties, one frame", flush=True)
    node = {"plugin": "asdpocs", "properties": {
        "projector": {"plugin": "parallel-projector", "name": "parallel-projector", "package": "ir",
                      "properties": {"model": "joseph", "is_forward": False, "step": STEP}},
        "df_minimizer": {"plugin": "sart", "package": "ir", "name": "sart_minimizer",
                         "properties": {"relaxation_factor": 0.25, "num-iterations": 100}},
        "num-iterations": 10}}
    graph = _graph("out/out%03i.tif", node)
    graph["nodes"][0]["properties"]["path"] = "sino01.tif"
    sino = _sinos(1)[0]
    uir.tiff_write(str(tmp_path / "sino01.tif"), sino)
    rep = uir.json_run(graph, base_dir=str(tmp_path))
    assert rep["frames"] == 1 and rep["tasks"] == 1
    got = uir.tiff_read(str(tmp_path / "out" / "out000.tif"))
    g = O.Geometry(N_DET, N_DET, N_ANG, np.float32(STEP))
    ref = O.asdpocs(g, sino, 10, df="sart", df_iters=100, df_relax_initial=0.25)
    err = O.rel_l2(got, ref)
    print("json-sample ASD-POCS rel-L2 vs oracle %.3e" % err, flush=True)
    assert err <= 1e-4


def test_sirt_chain_multipage_and_cli(res, tmp_path):
    sinos = _sinos(3)
    for i, s in enumerate(sinos):
        uir.tiff_write(str(tmp_path / "sino.tif"), s, append=i > 0)
    node = {"plugin": "sirt", "properties": {
        "projector": {"plugin": "parallel-projector", "properties": {"is-forward": False, "step": STEP}},
        "relaxation-factor": 0.5, "num_iterations": 7}}
    g = _graph("recon.tif", node)
    g["nodes"][0]["properties"]["start"] = 1            # frames 1 and 2
    rep = uir.json_run(g, base_dir=str(tmp_path))
    assert rep["frames"] == 2
    assert uir.tiff_info(str(tmp_path / "recon.tif")) == (2, N_DET, N_DET)
    proj = uir.Task("parallel-projector", is_forward=False, step=STEP)
    sirt = uir.Task("sirt", projector=proj, relaxation_factor=0.5, num_iterations=7).setup(res)
    gm = O.Geometry(N_DET, N_DET, N_ANG, np.float32(STEP))
    for page, i in enumerate((1, 2)):
        got = uir.tiff_read(str(tmp_path / "recon.tif"), page)
        assert np.array_equal(got, _direct(res, sirt, sinos[i]))
        assert O.rel_l2(got, O.sirt(gm, sinos[i], 7, 0.5)) <= 1e-4
    # the command-line runner on the same graph, from the graph's directory
    g["nodes"][1]["properties"]["filename"] = "cli-%02d.tif"
    (tmp_path / "g.json").write_text(json.dumps(g))
    exe = os.path.join(ROOT, "ufo-ir_amd", "uir-runjson")
    r = subprocess.run([exe, "g.json"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "2 frame(s)" in r.stdout
    for page in range(2):
        assert np.array_equal(uir.tiff_read(str(tmp_path / ("cli-%02d.tif" % page))),
                              uir.tiff_read(str(tmp_path / "recon.tif"), page))


def test_graph_projector_model_line(tmp_path):
    """model "line" in the graph's projector node (the property of
    ufo-ir-parallel-projector-task.c:123-128) selects the line model: the SIRT
    chain equals the oracle's FP64 SIRT over that model; an unknown model
    fails the run at setup, as the reference fails to load its kernels."""
    print("case graph line model", flush=True)
    gl = O.Geometry(N_DET, N_DET, N_ANG, np.float32(STEP), model="line")
    sino = O.fp64(gl, ellipse_phantom(N_DET, seed=510).astype(np.float64)).astype(np.float32)
    uir.tiff_write(str(tmp_path / "sino.tif"), sino)
    node = {"plugin": "sirt", "properties": {
        "projector": {"plugin": "parallel-projector", "package": "ir",
                      "properties": {"model": "line", "is_forward": False, "step": STEP}},
        "relaxation_factor": 0.5, "num-iterations": 20}}
    rep = uir.json_run(_graph("out/line%03i.tif", node), base_dir=str(tmp_path))
    assert rep["frames"] == 1
    got = uir.tiff_read(str(tmp_path / "out" / "line000.tif"))
    ref = O.sirt64(gl, sino.astype(np.float64), 20, 0.5)
    err = O.rel_l2(got, ref)
    print("graph line SIRT rel-L2 vs FP64 %.3e" % err, flush=True)
    assert err <= 1e-4
    node["properties"]["projector"]["properties"]["model"] = "cone"
    with pytest.raises(uir.UirError) as e:
        uir.json_run(_graph("out/cone%03i.tif", node), base_dir=str(tmp_path))
    assert "cone" in str(e.value)

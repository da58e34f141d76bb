"""Tile-shape cache of the autotuner (host side, no device needed)."""
import pytest

import uir

SPEC = dict(n=128, n_dets=182, n_angles=90, step=0.0349066, relaxation=0.5)


def test_untuned_geometry_gets_the_default_shape():
    uir.tune_clear()
    tile, tuned = uir.tune_get(SPEC, "f16")
    assert not tuned and tile == {"fp_angles": 16, "bp_angles": 32, "bp_rows": 64}


def test_set_get_per_geometry_and_dtype():
    uir.tune_clear()
    uir.tune_set(SPEC, "f16", 8, 16, 32)
    assert uir.tune_get(SPEC, "f16") == ({"fp_angles": 8, "bp_angles": 16, "bp_rows": 32}, True)
    assert uir.tune_get(SPEC, "f32")[1] is False                     # dtype is part of the key
    assert uir.tune_get(dict(SPEC, n_angles=91), "f16")[1] is False  # so is the geometry
    uir.tune_set(SPEC, "f16", 16, 32, 32)                            # overwrite
    assert uir.tune_get(SPEC, "f16")[0]["bp_rows"] == 32
    uir.tune_clear()
    assert uir.tune_get(SPEC, "f16")[1] is False


@pytest.mark.parametrize("tile", [(12, 32, 64), (64, 32, 64), (16, 8, 64), (16, 32, 48), (0, 0, 0)])
def test_only_built_variants_are_accepted(tile):
    with pytest.raises(uir.UirError) as e:
        uir.tune_set(SPEC, "f16", *tile)
    assert e.value.status == 1

"""Ceiling probes (uir_probe.hip): the LDS read rate and the FP32 FMA issue
rate that the bench's roofline quotes.  The bounds are physical sanity
limits of a CDNA compute unit (LDS: 32 banks x 4 bytes, so at most twice
128 bytes per clock even with a doubled data path; FMA: four SIMDs, at most
one wave64 instruction each per clock), not performance targets."""
import pytest

import uir

pytestmark = pytest.mark.gpu


def test_probe_ceilings_are_physical(res):
    c = uir.probe_ceilings(res)
    assert c["n_cu"] > 0 and c["clock_ghz"] > 0
    assert 0 < c["lds_bytes_per_clock_cu"] <= 256, c
    assert 0 < c["fma_wave64_per_clock_cu"] <= 4, c
    assert 0 < c["cvt_wave64_per_clock_cu"] <= 4, c
    assert c["lds_ms"] > 0 and c["fma_ms"] > 0

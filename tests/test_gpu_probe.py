"""The device random-access rates the reservation rounds are priced at:
measured, finite and ordered the way the memory system orders them (an
atomic read-modify-write is no cheaper than a plain read of the same word)."""
import math

import pytest

import mcpm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("record_bytes", [16, 32])
def test_probe_rates_are_measured(record_bytes):
    print(f"probe record_bytes={record_bytes}", flush=True)
    r = mcpm.probe_random_access(1 << 22, record_bytes)
    for k, v in r.items():
        assert math.isfinite(v) and 1e8 < v < 1e13, (k, v)
    assert r["atomic_min_word"] <= 1.5 * r["gather_word"]

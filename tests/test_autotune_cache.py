"""The autotuner's on-disk cache (CPU): keys, record / lookup round trip."""

from __future__ import annotations

import json

from warpmpc_amd import autotune
from warpmpc_amd.autotune import LaunchConfig


def test_record_and_lookup_round_trip(tmp_path):
    path = tmp_path / "tune.json"
    k1 = autotune.cache_key(0x1234, 65536, False, "AMD Instinct MI355X")
    k2 = autotune.cache_key(0x1234, 65536, True, "AMD Instinct MI355X")
    assert k1 != k2 and "0000000000001234" in k1
    assert autotune.lookup(k1, path) is None
    c1 = LaunchConfig(ipw=64, lds_budget=40960, specialized=True)
    c2 = LaunchConfig(ipw=16, lds_budget=98304, specialized=False)
    autotune.record(k1, c1, 1.5, path)
    autotune.record(k2, c2, 2.5, path)
    assert autotune.lookup(k1, path) == c1
    assert autotune.lookup(k2, path) == c2
    data = json.loads(path.read_text())
    assert data[k1]["ms"] == 1.5


def test_damaged_cache_reads_as_empty(tmp_path):
    path = tmp_path / "tune.json"
    path.write_text("{not json")
    assert autotune.lookup("x", path) is None
    autotune.record("x", LaunchConfig(ipw=8, lds_budget=40960, specialized=False), 1.0, path)
    assert autotune.lookup("x", path).ipw == 8

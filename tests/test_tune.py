"""Autotuner host logic: buckets, defaults outside the table, the JSON cache
(no device work)."""
import json

from fastembedr_amd import tune


def test_buckets_and_rules():
    assert tune.bucket_of(1999) == -1
    assert tune.bucket_of(2000) == 0
    assert tune.bucket_of(4999) == 0 and tune.bucket_of(5000) == 1
    assert tune.bucket_of(99999) == 4 and tune.bucket_of(100000) == 5
    assert tune.bucket_of(250000) == -1
    assert tune.cpu_rule_grid(49999) == 128 and tune.cpu_rule_grid(50000) == 256


def test_shape_defaults_and_cache_roundtrip(tmp_path):
    path = str(tmp_path / "tune.json")
    t = tune.Autotuner(cache_path=path)
    assert t.shape_for(3000) == {"grid_size": 0, "attraction": 0, "pair_workgroups": 0}
    t.table[0] = {"grid_size": 128, "attraction": 3, "pair_workgroups": 256, "n": 3162}
    t.save()
    data = json.load(open(path))
    assert data["key"] == t._key() and "0" in data["buckets"]
    t2 = tune.Autotuner(cache_path=path)
    assert t2.shape_for(4000) == {"grid_size": 128, "attraction": 3, "pair_workgroups": 256}
    assert t2.shape_for(6000)["grid_size"] == 0
    # another k is another table
    t3 = tune.Autotuner(k=30, cache_path=path)
    assert t3.table == {}

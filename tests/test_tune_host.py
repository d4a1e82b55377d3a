"""Autotuner shape table at the C ABI without a GPU: set / get / entries /
clear round trip, T binned by ceil(log2 T), shapes the tile kernels cannot
run rejected, JSON save / load."""
import pytest

from ParallelMCMC import ArgumentError, Tune


def test_table_round_trip(tmp_path):
    Tune.clear()
    assert Tune.lookup("scan_diag", "f32", 64, 1 << 20) is None
    Tune.set_entry("scan_diag", "f32", 64, 1 << 20, tc=64, parts=2, single_pass=1, us=12.5)
    Tune.set_entry("deer", "f64", 100, 4096, tc=16, parts=2)
    got = Tune.lookup("scan_diag", "f32", 64, (1 << 19) + 1)  # same bin: ceil(log2 T) = 20
    assert got == {"tc": 64, "parts": 2, "single_pass": 1, "us": 12.5}
    assert Tune.lookup("scan_diag", "f64", 64, 1 << 20) is None  # other precision
    assert Tune.lookup("deer", "f64", 100, 3000)["tc"] == 16
    assert Tune.lookup("deer", "f64", 100, 5000) is None  # other bin
    ents = Tune.entries()
    assert len(ents) == 2
    path = tmp_path / "tune.json"
    Tune.save(str(path))
    Tune.clear()
    assert Tune.entries() == []
    Tune.load(str(path))
    assert sorted((e["kind"], e["D"], e["tc"]) for e in Tune.entries()) == [("deer", 100, 16), ("scan_diag", 64, 64)]
    Tune.clear()


def test_shapes_that_cannot_run_are_rejected():
    Tune.clear()
    with pytest.raises(ArgumentError):  # 8 parts of 64 rows exceed the part arrays (P * D <= 256)
        Tune.set_entry("scan_diag", "f32", 64, 4096, tc=64, parts=8)
    with pytest.raises(ArgumentError):  # more parts than columns
        Tune.set_entry("deer", "f32", 4, 4096, tc=2, parts=4)
    with pytest.raises(ArgumentError):  # beyond the LDS budget
        Tune.set_entry("deer", "f32", 100, 1 << 20, tc=4096, parts=1)
    with pytest.raises(ArgumentError):  # register tiles: 16 <= D <= 256
        Tune.set_entry("scan_diag", "f32", 8, 4096, tc=128, parts=1, single_pass=2)
    with pytest.raises(ArgumentError):
        Tune.set_entry("scan_diag", "f32", 300, 4096, tc=64, parts=1, single_pass=2)
    with pytest.raises(ArgumentError):  # 256-column register tiles are FP32 only
        Tune.set_entry("scan_diag", "f64", 64, 4096, tc=256, parts=1, single_pass=2)
    Tune.set_entry("scan_diag", "f64", 64, 4096, tc=128, parts=1, single_pass=2)
    with pytest.raises(ArgumentError):  # persistent register tiles: 64 or 128 columns
        Tune.set_entry("scan_diag", "f32", 64, 4096, tc=256, parts=1, single_pass=3)
    with pytest.raises(ArgumentError):  # register tiles: 64, 128 or 256 columns
        Tune.set_entry("scan_diag", "f32", 128, 4096, tc=32, parts=1, single_pass=2)
    Tune.set_entry("scan_diag", "f32", 128, 4096, tc=64, parts=1, single_pass=3)
    Tune.set_entry("scan_diag", "f32", 100, 4096, tc=128, parts=1, single_pass=2)
    Tune.set_entry("scan_diag", "f64", 256, 4096, tc=128, parts=1, single_pass=3)
    Tune.set_entry("scan_diag", "f32", 64, 4096, tc=128, parts=1, single_pass=3)
    assert Tune.lookup("scan_diag", "f32", 64, 4096)["single_pass"] == 3
    Tune.clear()
    assert Tune.entries() == []

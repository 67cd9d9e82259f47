"""Autotuner pieces that need no GPU: the cell-grid rule with the 3x3 stencil's preconditions
(a grid of at least 3 x 3 cells, cells no narrower than 2 r_max + cutoff; reference
base_point_particle_kernels.cu:72-94 relies on both without checking), and the JSON cache."""
import json
import math

import numpy as np
import pytest

from dpmd_amd import _lib as L
from dpmd_amd import autotune as T
from dpmd_amd import workloads as W


def test_cell_grid_rule_and_preconditions():
    rng = np.random.default_rng(3)
    for _ in range(200):
        Lx, Ly = rng.uniform(5.0, 400.0, 2)
        rmax, cut = rng.uniform(0.3, 1.0), rng.uniform(0.0, 0.8)
        f = rng.choice([1.0, 1.25, 1.5, 2.0])
        side = f * (2 * rmax + cut)
        ex, ey = math.floor(Lx / side), math.floor(Ly / side)
        if ex < 3 or ey < 3:
            with pytest.raises(L.DpmdError) as e:
                T.cell_grid(Lx, Ly, rmax, cut, f)
            assert "3 x 3" in str(e.value)
            continue
        nx, ny = T.cell_grid(Lx, Ly, rmax, cut, f)
        assert (nx, ny) == (ex, ey)
        # the stencil's two preconditions
        assert nx >= 3 and ny >= 3
        assert Lx / nx >= 2 * rmax + cut and Ly / ny >= 2 * rmax + cut


def test_cell_grid_rejects_bad_arguments():
    for args in [(10, 10, 0.5, 0.3, 0.9), (-1, 10, 0.5, 0.3, 1.0), (10, 10, 0.0, 0.3, 1.0), (10, 10, 0.5, -0.1, 1.0)]:
        with pytest.raises(L.DpmdError):
            T.cell_grid(*args)
    # exactly three cells per axis is the smallest legal grid
    assert T.cell_grid(3 * 1.7, 3 * 1.7 + 1e-9, 0.7, 0.3, 1.0) == (3, 3)
    with pytest.raises(L.DpmdError):
        T.cell_grid(3 * 1.7 - 1e-9, 10.0, 0.7, 0.3, 1.0)


def test_cell_grid_factor_one_matches_oracle_grid():
    """At cell_factor 1 the tuner's grid is the default grid the oracle derives."""
    import oracle as orc
    b = W.uneven_batch(S=12, nmin=64, nmax=600, seed=5)
    nx, ny = orc.OracleDisk(b).cells()
    for s in range(b.S):
        lo, hi = b.off[s], b.off[s + 1]
        assert T.cell_grid(b.Lx[s], b.Ly[s], b.rad[lo:hi].max(), b.cutoff[s], 1.0) == (nx[s], ny[s])


def test_cache_round_trip(tmp_path):
    p = str(tmp_path / "tune.json")
    assert T.cache_load(p, "sig-a") is None          # a missing file is an empty cache
    a = dict(skin=0.2, cell_factor=1.0, block=128, max_batch=64, fused_vv=1)
    b = dict(skin=0.45, cell_factor=1.25, block=512, max_batch=128, fused_vv=0)
    T.cache_store(p, "sig-a", a, 1.5e10)
    T.cache_store(p, 'sig "b"|S=4', b, 2.0e9)
    got = T.cache_load(p, "sig-a")
    assert got is not None and got[0] == a and got[1] == 1.5e10
    assert T.cache_load(p, 'sig "b"|S=4') == (b, 2.0e9)
    assert T.cache_load(p, "sig-c") is None
    # overwrite keeps one entry per signature; the file is plain JSON
    a2 = dict(a, block=256)
    T.cache_store(p, "sig-a", a2, 1.6e10)
    assert T.cache_load(p, "sig-a") == (a2, 1.6e10)
    doc = json.load(open(p))
    assert doc["format"] == "dpmd_amd autotune cache 1"
    assert sorted(doc["entries"]) == ['sig "b"|S=4', "sig-a"]
    assert doc["entries"]["sig-a"]["block"] == 256


def test_cache_rejects_foreign_or_malformed_files(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"entries": {}}')
    with pytest.raises(L.DpmdError, match="not a dpmd_amd autotune cache"):
        T.cache_load(str(p), "x")
    p.write_text('{"format": "dpmd_amd autotune cache 1", "entries": {"x": {"skin": 0.3}')
    with pytest.raises(L.DpmdError, match="malformed"):
        T.cache_load(str(p), "x")
    p.write_text('{"format": "dpmd_amd autotune cache 1", "entries": {"x": {"skin": 0.3}}}')
    with pytest.raises(L.DpmdError, match="cell_factor"):
        T.cache_load(str(p), "x")
    # a store into a malformed file fails instead of overwriting it
    p.write_text("not json")
    with pytest.raises(L.DpmdError):
        T.cache_store(str(p), "x", dict(skin=0.3, cell_factor=1.0, block=256, max_batch=128, fused_vv=1), 1.0)
    assert p.read_text() == "not json"

"""mtmd_find_topology (the cell-grid topology search on the GPU) against the
oracle's brute-force restatement of the reference's loops: the same bonds,
angles and dihedrals in the same order, with and without the periodic
distance, on water, chain molecules and a dense random packing with many
links per atom."""
import time

import numpy as np
import pytest

from moltwister_amd import workloads as W
from moltwister_amd.engine import find_topology

from test_oracle_topology import waters

pytestmark = pytest.mark.gpu


def _same(P, which, pos, types, box, rules):
    got = find_topology(which, pos, types, box, rules)
    want = P.find_topology(which, pos, types, box, rules)
    assert got.shape == want.shape and np.array_equal(got, want), (which, got.shape, want.shape)
    return got


@pytest.mark.parametrize("pbc", [False, True])
def test_water_box(oracle_mod, pbc):
    s = W.water_box(300, seed=5)
    box = np.array(s.box) if pbc else None
    b = _same(oracle_mod, 1, s.pos, s.types, box, [((0, 1), 1.2)])
    assert b.shape[0] == 600
    a = _same(oracle_mod, 2, s.pos, s.types, box, [((1, 0, 1), 1.2)])
    assert a.shape[0] == 300
    # a longer R0 links atoms of neighbouring molecules too
    _same(oracle_mod, 1, s.pos, s.types, box, [((0, 1), 2.4), ((0, 0), 3.2)])


def test_chain_molecules(oracle_mod):
    s = W.chain_molecules(30, 6, 24.0, seed=4)
    box = np.array(s.box)
    rules2 = [((0, 1), 1.75), ((1, 1), 1.75)]
    _same(oracle_mod, 1, s.pos, s.types, box, rules2)
    _same(oracle_mod, 2, s.pos, s.types, box, [((0, 1, 1), 1.75), ((1, 1, 1), 1.75)])
    d = _same(oracle_mod, 3, s.pos, s.types, box, [((0, 1, 1, 1), 1.75), ((1, 1, 1, 1), 1.75)])
    assert d.shape[0] > 0


@pytest.mark.parametrize("pbc", [False, True])
def test_dense_packing_many_links(oracle_mod, pbc):
    rng = np.random.default_rng(7)
    n, L = 900, 9.0
    pos = rng.uniform(-L / 2, L / 2, (n, 3))
    types = rng.integers(0, 3, n).astype(np.int32)
    box = np.array([L, L, L]) if pbc else None
    _same(oracle_mod, 1, pos, types, box, [((0, 1), 1.3), ((2, 2), 1.1), ((1, 0), 1.2)])
    _same(oracle_mod, 2, pos, types, box, [((0, 1, 2), 1.2), ((2, 2, 2), 1.0)])
    _same(oracle_mod, 3, pos, types, box, [((0, 1, 2, 0), 1.0)])


def test_water_30k_against_brute_force(oracle_mod):
    pos, types, L = waters(10000, spacing=3.2)
    box = np.array([L, L, L])
    t0 = time.perf_counter()
    b = find_topology(1, pos, types, box, [((0, 1), 1.2)])
    a = find_topology(2, pos, types, box, [((1, 0, 1), 1.2)])
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    bo = oracle_mod.find_topology(1, pos, types, box, [((0, 1), 1.2)])
    t_cpu = time.perf_counter() - t0
    print(f"water 30k bond search: gpu (bonds + angles) {t_gpu:.3f} s, brute-force bonds {t_cpu:.3f} s", flush=True)
    assert np.array_equal(b, bo) and b.shape[0] == 20000 and a.shape[0] == 10000

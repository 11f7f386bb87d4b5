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


# ---- the bond-graph criteria: molecule, visible bonds, 1-4 ----
ALL_R, MOL_R, VISIBLE, ONLY14 = 0, 1, 2, 3


def graph_of(n, bonds, symmetric=True):
    g = [[] for _ in range(n)]
    for a, b in bonds:
        g[a].append(int(b))
        if symmetric:
            g[b].append(int(a))
    return g


def _same_g(P, which, pos, types, box, rules, g):
    got = find_topology(which, pos, types, box, rules, g)
    want = P.find_topology(which, pos, types, box, rules, g)
    assert got.shape == want.shape and np.array_equal(got, want), (which, rules, got.shape, want.shape)
    return got


@pytest.mark.parametrize("pbc", [False, True])
def test_graph_criteria_water(oracle_mod, pbc):
    s = W.water_box(300, seed=5)
    box = np.array(s.box) if pbc else None
    g = graph_of(s.n_atoms, s.bonds[:, :2])
    # R0 2.4 reaches into neighbouring molecules: mol_r_less_than must not
    b = _same_g(oracle_mod, 1, s.pos, s.types, box, [((0, 1), 2.4, MOL_R), ((0, 1), 1.2, VISIBLE)], g)
    assert b.shape[0] == 600
    _same_g(oracle_mod, 1, s.pos, s.types, box, [((1, 1), 2.4, MOL_R), ((1, 1), 1.0, ONLY14), ((0, 0), 3.2, ALL_R)], g)
    a = _same_g(oracle_mod, 2, s.pos, s.types, box, [((1, 0, 1), 1.2, MOL_R), ((1, 0, 1), 1.0, VISIBLE)], g)
    assert a.shape[0] == 300


def test_graph_criteria_chains(oracle_mod):
    s = W.chain_molecules(30, 6, 24.0, seed=4)
    box = np.array(s.box)
    g = graph_of(s.n_atoms, s.bonds[:, :2])
    _same_g(oracle_mod, 1, s.pos, s.types, box, [((0, 1), 1.75, MOL_R), ((1, 1), 1.75, VISIBLE), ((0, 1), 1.0, ONLY14),
                                                 ((1, 1), 1.0, ONLY14), ((0, 0), 1.0, ONLY14)], g)
    _same_g(oracle_mod, 2, s.pos, s.types, box, [((0, 1, 1), 1.75, MOL_R), ((1, 1, 1), 1.0, VISIBLE)], g)
    d = _same_g(oracle_mod, 3, s.pos, s.types, box, [((0, 1, 1, 1), 1.75, MOL_R), ((1, 1, 1, 1), 1.75, VISIBLE),
                                                     ((0, 1, 1, 0), 2.6, MOL_R)], g)
    assert d.shape[0] > 0


def _random_molecules(n, L, size, extra, seed, asym=0):
    """Random trees of `size` atoms grown from near neighbours, `extra` ring
    bonds inside each, and `asym` bonds attached at one end only."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-L / 2, L / 2, (n, 3))
    types = rng.integers(0, 3, n).astype(np.int32)
    order = rng.permutation(n)
    edges, oneway = [], []
    for m0 in range(0, n, size):
        mem = order[m0:m0 + size]
        for i in range(1, len(mem)):
            d = np.linalg.norm(pos[mem[:i]] - pos[mem[i]], axis=1)
            edges.append((int(mem[np.argmin(d)]), int(mem[i])))
        for _ in range(extra if len(mem) > 3 else 0):
            a, b = rng.choice(mem, 2, replace=False)
            edges.append((int(a), int(b)))
    for _ in range(asym):
        a, b = rng.choice(n, 2, replace=False)
        oneway.append((int(a), int(b)))
    perm = rng.permutation(len(edges))
    g = graph_of(n, [edges[i] for i in perm])
    for a, b in oneway:
        g[a].append(b)
    return pos, types, g


@pytest.mark.parametrize("pbc", [False, True])
def test_graph_criteria_random_molecules(oracle_mod, pbc):
    # 900 atoms in 9 A: at R0 2.6 an atom has ~90 atoms within R0, above the
    # 64-atom row scan, so the walk-order fallback runs; at 1.1 the rows do
    n, L = 900, 9.0
    pos, types, g = _random_molecules(n, L, 30, 4, seed=11, asym=40)
    box = np.array([L, L, L]) if pbc else None
    _same_g(oracle_mod, 1, pos, types, box, [((0, 1), 1.1, MOL_R), ((2, 2), 2.6, MOL_R), ((1, 0), 1.0, VISIBLE),
                                             ((2, 1), 1.0, ONLY14)], g)
    _same_g(oracle_mod, 2, pos, types, box, [((0, 1, 2), 1.4, MOL_R), ((2, 2, 2), 2.6, MOL_R), ((1, 1, 0), 1.0, VISIBLE)], g)
    _same_g(oracle_mod, 3, pos, types, box, [((0, 1, 2, 0), 1.6, MOL_R), ((0, 1, 1, 2), 1.0, VISIBLE)], g)


def test_graph_criteria_one_large_molecule(oracle_mod):
    # a single 1,500-atom molecule: every walk covers it
    n, L = 1500, 11.0
    pos, types, g = _random_molecules(n, L, n, 25, seed=3)
    box = np.array([L, L, L])
    _same_g(oracle_mod, 1, pos, types, box, [((0, 1), 1.3, MOL_R), ((2, 0), 1.0, ONLY14)], g)
    _same_g(oracle_mod, 2, pos, types, box, [((0, 1, 2), 1.2, MOL_R)], g)


def test_graph_criteria_need_the_graph():
    from moltwister_amd.engine import MtmdError
    pos = np.zeros((2, 3))
    types = np.zeros(2, np.int32)
    with pytest.raises(MtmdError):
        find_topology(1, pos, types, None, [((0, 0), 1.0, VISIBLE)])
    with pytest.raises(MtmdError):   # 1-4 is a bond criterion only
        find_topology(2, pos, types, None, [((0, 0, 0), 1.0, ONLY14)], [[1], [0]])

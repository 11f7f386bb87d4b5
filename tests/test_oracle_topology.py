"""The oracle's topology search (all_r_less_than criterion of
CMolTwisterStateTools::getAllMD{Bonds,Angles,Dihedrals}InSystem,
MolTwisterStateTools.cpp:544-1140, with the duplicate removal of
MolecularSystemTools.cpp:25-200) against topologies known by construction:
isolated water molecules, a butane chain, a bond across the periodic
boundary, and the same pair found by two force fields. The reference's
StateTools need CMolTwisterState (GLUT), so these known answers are the pins."""
import numpy as np

O, H = 0, 1


def waters(n, spacing=6.0, seed=1):
    rng = np.random.default_rng(seed)
    pos, types = [], []
    g = int(np.ceil(n ** (1 / 3)))
    for i in range(n):
        c = np.array([i % g, (i // g) % g, i // (g * g)], float) * spacing
        th = np.radians(104.5)
        R = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        h1 = R @ np.array([0.96, 0.0, 0.0])
        h2 = R @ np.array([0.96 * np.cos(th), 0.96 * np.sin(th), 0.0])
        pos += [c, c + h1, c + h2]
        types += [O, H, H]
    return np.array(pos), np.array(types, np.int32), g * spacing


def test_waters_bonds_and_angles(oracle_mod):
    n = 27
    pos, types, L = waters(n)
    bonds = oracle_mod.find_topology(1, pos, types, None, [((O, H), 1.2)])
    want = np.array([[3 * i, 3 * i + h, 0] for i in range(n) for h in (1, 2)])
    assert np.array_equal(bonds, want)
    # H-O-H found as (H1, O, H2) and (H2, O, H1): the reverse is dropped
    ang = oracle_mod.find_topology(2, pos, types, None, [((H, O, H), 1.2)])
    assert np.array_equal(ang, np.array([[3 * i + 1, 3 * i, 3 * i + 2, 0] for i in range(n)]))
    # a rule listing the pair the other way round finds the same bonds, (H, O) first
    rev = oracle_mod.find_topology(1, pos, types, None, [((H, O), 1.2)])
    assert np.array_equal(rev, np.array([[3 * i + h, 3 * i, 0] for i in range(n) for h in (1, 2)]))


def test_butane_chain(oracle_mod):
    # C1 - C2 - C3 - C4, 1.54 A bonds, trans zig-zag; types CH3 (0) and CH2 (1)
    pos = np.array([[0.0, 0.0, 0.0], [1.54, 0.0, 0.0], [2.05, 1.45, 0.0], [3.59, 1.45, 0.0]])
    types = np.array([0, 1, 1, 0], np.int32)
    rules_b = [((0, 1), 1.75), ((1, 1), 1.75)]
    b = oracle_mod.find_topology(1, pos, types, None, rules_b)
    assert np.array_equal(b, [[0, 1, 0], [3, 2, 0], [1, 2, 1]])
    a = oracle_mod.find_topology(2, pos, types, None, [((0, 1, 1), 1.75)])
    assert np.array_equal(a, [[0, 1, 2, 0], [3, 2, 1, 0]])
    d = oracle_mod.find_topology(3, pos, types, None, [((0, 1, 1, 0), 1.75)])
    assert np.array_equal(d, [[0, 1, 2, 3, 0]])          # (3, 2, 1, 0) is its reverse
    # R0 below the bond length finds nothing
    assert oracle_mod.find_topology(1, pos, types, None, [((0, 1), 1.5)]).shape[0] == 0


def test_bond_across_the_periodic_boundary(oracle_mod):
    L = 10.0
    pos = np.array([[4.7, 0.0, 0.0], [-4.6, 0.1, 0.0], [0.0, 0.0, 0.0]])
    types = np.array([0, 1, 1], np.int32)
    assert oracle_mod.find_topology(1, pos, types, None, [((0, 1), 1.0)]).shape[0] == 0
    b = oracle_mod.find_topology(1, pos, types, np.array([L, L, L]), [((0, 1), 1.0)])
    assert np.array_equal(b, [[0, 1, 0]])


def test_first_force_field_keeps_a_shared_pair(oracle_mod):
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    types = np.array([0, 0], np.int32)
    b = oracle_mod.find_topology(1, pos, types, None, [((0, 0), 1.2), ((0, 0), 1.5)])
    assert np.array_equal(b, [[0, 1, 0]])

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


# ---- the bond-graph criteria (CMDFFBond::EDetCrit 1-3) ----
ALL_R, MOL_R, VISIBLE, ONLY14 = 0, 1, 2, 3


def symmetric(n, edges):
    """Bond graph with each bond attached at both atoms, in the order given
    (what `genbonds` makes: CAtom::attachBondTo on both ends)."""
    g = [[] for _ in range(n)]
    for a, b in edges:
        g[a].append(b)
        g[b].append(a)
    return g


def test_visible_bonds_ignore_distance(oracle_mod):
    # butane-like chain 0-1-2-3 (types 0,1,1,0) with the atoms 10 A apart:
    # only_visible_bonds follows the graph, all_r_less_than finds nothing
    pos = np.array([[0.0, 0, 0], [10.0, 0, 0], [20.0, 0, 0], [30.0, 0, 0]])
    types = np.array([0, 1, 1, 0], np.int32)
    g = symmetric(4, [(0, 1), (1, 2), (2, 3)])
    b = oracle_mod.find_topology(1, pos, types, None, [((0, 1), 1.2, VISIBLE), ((1, 1), 1.2, VISIBLE)], g)
    assert np.array_equal(b, [[0, 1, 0], [3, 2, 0], [1, 2, 1]])
    assert oracle_mod.find_topology(1, pos, types, None, [((0, 1), 1.2, ALL_R)], g).shape[0] == 0
    a = oracle_mod.find_topology(2, pos, types, None, [((0, 1, 1), 1.2, VISIBLE)], g)
    assert np.array_equal(a, [[0, 1, 2, 0], [3, 2, 1, 0]])
    d = oracle_mod.find_topology(3, pos, types, None, [((0, 1, 1, 0), 1.2, VISIBLE)], g)
    assert np.array_equal(d, [[0, 1, 2, 3, 0]])


def test_visible_bonds_follow_the_graph_direction(oracle_mod):
    pos = np.zeros((2, 3))
    types = np.array([0, 1], np.int32)
    g = [[1], []]                      # attached at atom 0 only
    assert np.array_equal(oracle_mod.find_topology(1, pos, types, None, [((0, 1), 1.0, VISIBLE)], g), [[0, 1, 0]])
    assert oracle_mod.find_topology(1, pos, types, None, [((1, 0), 1.0, VISIBLE)], g).shape[0] == 0


def test_visible_angle_naming_an_atom_twice_is_dropped(oracle_mod):
    # water O(0)-H(1), O(0)-H(2): H1 -> O -> H1 is walked, and dropped
    pos = np.array([[0.0, 0, 0], [0.96, 0, 0], [-0.24, 0.93, 0]])
    types = np.array([0, 1, 1], np.int32)
    g = symmetric(3, [(0, 1), (0, 2)])
    a = oracle_mod.find_topology(2, pos, types, None, [((1, 0, 1), 1.2, VISIBLE)], g)
    assert np.array_equal(a, [[1, 0, 2, 0]])


def test_14_bonds(oracle_mod):
    # The walk a -> b -> c -> d only requires d to differ from a, b and c, so
    # it may step back through a (c == a): a's other bonded atoms are then
    # listed too. Each answer below is derived walk by walk.
    pos = np.zeros((4, 3))
    t0 = np.zeros(4, np.int32)
    chain = symmetric(4, [(0, 1), (1, 2), (2, 3)])
    # from 0: 0-1-2-3; from 1: 1-0-1-2, 1-2-1-0; from 2: 2-1-2-3, 2-3-2-1 (= 1-2);
    # from 3: 3-2-1-0 (= 0-3)
    b = oracle_mod.find_topology(1, pos, t0, None, [((0, 0), 1.0, ONLY14)], chain)
    assert np.array_equal(b, [[0, 3, 0], [1, 2, 0], [1, 0, 0], [2, 3, 0]])
    # star 0-{1,2,3}: from 0, 0-1-0-2, 0-1-0-3, 0-2-0-1, ...; from a leaf every
    # walk ends on the centre, which is b
    star = symmetric(4, [(0, 1), (0, 2), (0, 3)])
    b = oracle_mod.find_topology(1, pos, t0, None, [((0, 0), 1.0, ONLY14)], star)
    assert np.array_equal(b, [[0, 2, 0], [0, 3, 0], [0, 1, 0]])
    # the atom of the second type is matched at the far end only
    t1 = np.array([0, 0, 0, 1], np.int32)
    assert np.array_equal(oracle_mod.find_topology(1, pos, t1, None, [((0, 1), 1.0, ONLY14)], chain),
                          [[0, 3, 0], [2, 3, 0]])
    # a pair of atoms has none
    assert oracle_mod.find_topology(1, pos[:2], t0[:2], None, [((0, 0), 1.0, ONLY14)],
                                    symmetric(2, [(0, 1)])).shape[0] == 0


def test_mol_criterion_stays_inside_the_molecule(oracle_mod):
    # two waters; H(1) of the first lies 1.0 A from O(3) of the second
    pos = np.array([[0.0, 0, 0], [0.96, 0, 0], [-0.24, 0.93, 0],
                    [1.96, 0, 0], [2.2, 0.93, 0], [2.92, 0, 0]])
    types = np.array([0, 1, 1, 0, 1, 1], np.int32)
    g = symmetric(6, [(0, 1), (0, 2), (3, 4), (3, 5)])
    allr = oracle_mod.find_topology(1, pos, types, None, [((0, 1), 1.2, ALL_R)], g)
    assert [3, 1, 0] in allr.tolist()
    mol = oracle_mod.find_topology(1, pos, types, None, [((0, 1), 1.2, MOL_R)], g)
    assert np.array_equal(mol, [[0, 1, 0], [0, 2, 0], [3, 4, 0], [3, 5, 0]])
    a = oracle_mod.find_topology(2, pos, types, None, [((1, 0, 1), 1.2, MOL_R)], g)
    assert np.array_equal(a, [[1, 0, 2, 0], [4, 3, 5, 0]])


def test_mol_criterion_walks_depth_first(oracle_mod):
    # atom 0 bonds 1 then 4; 1-2-3 hang below 1. Depth first from 0 visits
    # 0, 1, 2, 3, 4 (breadth first would give 0, 1, 4, 2, 3)
    pos = np.array([[0.0, 0, 0], [0.5, 0, 0], [0.5, 0.5, 0], [0.5, 0.5, 0.5], [-0.5, 0, 0]])
    types = np.array([0, 1, 1, 1, 1], np.int32)
    g = symmetric(5, [(0, 1), (0, 4), (1, 2), (2, 3)])
    b = oracle_mod.find_topology(1, pos, types, None, [((0, 1), 2.0, MOL_R)], g)
    assert np.array_equal(b, [[0, 1, 0], [0, 2, 0], [0, 3, 0], [0, 4, 0]])
    # bond order in the graph decides: 0 bonds 3 first, then 1 and 2
    pos2 = np.array([[0.0, 0, 0], [1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    g2 = [[3, 1, 2], [0], [0], [0]]
    t2 = np.array([0, 1, 1, 1], np.int32)
    b2 = oracle_mod.find_topology(1, pos2, t2, None, [((0, 1), 1.2, MOL_R)], g2)
    assert np.array_equal(b2, [[0, 3, 0], [0, 1, 0], [0, 2, 0]])
    assert np.array_equal(oracle_mod.find_topology(1, pos2, t2, None, [((0, 1), 1.2, ALL_R)], g2),
                          [[0, 1, 0], [0, 2, 0], [0, 3, 0]])

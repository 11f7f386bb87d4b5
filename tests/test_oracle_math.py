"""Pins for the parts of the oracle the reference build cannot exercise
(CVelVerlet/CSimulationBox need CMolTwisterState, whose C3DView translation
unit needs GLUT): the velocity-Verlet step, force cap, PBC wrap, thermo
reductions and the gradient-descent step, checked against closed forms and
invariants written independently here with numpy. Each check is chosen so a
plausible slip (a sign, a missing half-kick, a dropped term, the wrong mass)
fails it.
"""
import numpy as np
import pytest

from moltwister_amd import abi, workloads as W

from systems import cfg

CONV_T, CONV_t, CONV_P = 120.272, 100.0, 16388.2461


@pytest.fixture
def lj(oracle_mod):
    s = W.fcc_lj(5, density=0.5, seed=9)
    rng = np.random.default_rng(2)
    s.pos = s.pos + rng.normal(0, 0.05, s.pos.shape)
    s.mass = 1.0 + rng.random(s.n_atoms)          # unequal masses catch p/m slips
    c = cfg(2.5, 0.3, ensemble=abi.NVE, dt=0.5)
    c.cutoff_force = 0.0                            # cap off unless a test turns it on
    return s, c


def test_one_vv_step_closed_form(oracle_mod, lj):
    """r1 = r0 + (p0 + F0 dt/2) dt/m; p1 = p0 + (F0 + F1) dt/2 (VelVerlet.cpp:74-93)."""
    s, c = lj
    o = oracle_mod.Oracle(s, c)
    r0, p0 = o.get_state()
    F0, _ = o.compute_forces()
    o.step(1)
    r1, p1 = o.get_state()
    F1, _ = o.get_forces()
    dt = c.timestep / CONV_t
    m = s.mass[:, None]
    r1_expect = r0 + (p0 + F0 * (dt / 2)) * (dt / m)
    L = np.array(s.box)
    d = r1 - r1_expect
    d -= L * np.round(d / L)                        # the wrap may move an atom by one box
    assert np.abs(d).max() < 1e-13
    assert np.allclose(p1, p0 + (F0 + F1) * (dt / 2), rtol=0, atol=1e-12)


def test_wrap_into_box(oracle_mod, lj):
    """After a step every coordinate lies in [-L/2, L/2) (SimulationBox.cpp:135-166)."""
    s, c = lj
    o = oracle_mod.Oracle(s, c)
    pos = s.pos.copy()
    pos[0] += np.array(s.box) * 1.7               # far outside: the wrap removes whole boxes
    pos[1] -= np.array(s.box) * 0.6
    o.set_state(pos=pos)
    o.step(1)
    r, _ = o.get_state()
    L = np.array(s.box)
    assert np.all(r >= -L / 2) and np.all(r < L / 2)


def test_force_cap(oracle_mod, lj):
    """cutForces clamps each component to +-Fcut (VelVerlet.cpp:96-122)."""
    s, c = lj
    pos = s.pos.copy()
    pos[1] = pos[0] + np.array([0.6, 0.0, 0.0])    # overlap: |F| >> cap
    c.cutoff_force = 50.0
    o = oracle_mod.Oracle(s, c)
    o.set_state(pos=pos)
    o.step(1)
    F, _ = o.get_forces()
    assert np.abs(F).max() == 50.0
    raw = oracle_mod.Oracle(s, c)
    raw.set_state(pos=o.get_state()[0])
    Fr, _ = raw.compute_forces()                    # compute_forces is uncapped
    assert np.abs(Fr).max() > 50.0
    assert np.array_equal(F, np.clip(Fr, -50.0, 50.0))


def test_momentum_conservation(oracle_mod, lj):
    """Full neighbour lists give sum F = 0 (Newton 3), so NVE keeps sum p."""
    s, c = lj
    o = oracle_mod.Oracle(s, c)
    F, _ = o.compute_forces()
    assert np.abs(F.sum(axis=0)).max() < 1e-10 * np.abs(F).max() * np.sqrt(s.n_atoms)
    _, p0 = o.get_state()
    o.step(200)
    _, p1 = o.get_state()
    assert np.abs(p1.sum(axis=0) - p0.sum(axis=0)).max() < 1e-9


def test_energy_conservation_second_order(oracle_mod):
    """Harmonic dimers (no cutoff crossings): the NVE energy error of velocity
    Verlet is O(dt^2) -- halving dt cuts it about fourfold. (The LJ model of the
    reference is truncated without a shift, so LJ runs also carry the U(rc)
    jumps of pairs crossing the cutoff, which no integrator removes.)"""
    from moltwister_amd.system import System
    rng = np.random.default_rng(5)
    nd = 64
    centres = rng.uniform(-8, 8, (nd, 3))
    d = rng.normal(0, 1, (nd, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pos = np.concatenate([centres - 0.55 * d, centres + 0.55 * d])
    order = np.stack([np.arange(nd), np.arange(nd) + nd], 1).reshape(-1)
    pos = pos[order]
    bonds = np.stack([np.arange(nd) * 2, np.arange(nd) * 2 + 1, np.zeros(nd, int)], 1)
    mass = np.tile([1.0, 2.0], nd)
    mom = rng.normal(0, 1, (2 * nd, 3)) * np.sqrt(mass)[:, None]
    errs = []
    for dt in (0.4, 0.2):
        s = System(pos=pos, mom=mom, mass=mass, charge=np.zeros(2 * nd), types=np.zeros(2 * nd, np.int32),
                   box=(40, 40, 40), nonbonded=[], bond_ff=[(abi.BOND_HARM, (300.0, 1.0))], bonds=bonds)
        c = cfg(3.0, 0.5, ensemble=abi.NVE, dt=dt)
        c.cutoff_force = 0.0
        o = oracle_mod.Oracle(s, c)
        E0 = o.thermo()["Etot"]
        nst = int(round(40 / dt))
        dev = 0.0
        for _ in range(10):
            o.step(nst // 10)
            dev = max(dev, abs(o.thermo()["Etot"] - E0))
        errs.append(dev)
    assert errs[0] < 1e-2 * abs(E0)
    assert 2.5 < errs[0] / errs[1] < 6.0, errs


def test_time_reversibility(oracle_mod, lj):
    s, c = lj
    o = oracle_mod.Oracle(s, c)
    r0, p0 = o.get_state()
    o.step(50)
    r1, p1 = o.get_state()
    o.set_state(pos=r1, mom=-p1)
    o.step(50)
    r2, p2 = o.get_state()
    L = np.array(s.box)
    d = r2 - r0
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-8
    assert np.abs(p2 + p0).max() < 1e-8


def test_thermo_definitions(oracle_mod, lj):
    """T = sum p.p/m/(3N), P = (sum p.p/m + sum F.r)/(3V), K, U sums, units
    (SimulationBox.cpp:168-231, Constants.h:27,32)."""
    s, c = lj
    o = oracle_mod.Oracle(s, c)
    o.step(3)
    r, p = o.get_state()
    F, U = o.get_forces()
    th = o.thermo()
    p2m = np.sum(np.sum(p * p, axis=1) / s.mass)
    V = np.prod(s.box)
    assert th["T"] == pytest.approx(p2m / (3 * s.n_atoms) * CONV_T, rel=1e-13)
    assert th["K"] == pytest.approx(0.5 * p2m, rel=1e-13)
    assert th["U"] == pytest.approx(np.sum(U.astype(np.float64)), rel=1e-13)
    assert th["P"] == pytest.approx((p2m + np.sum(F * r)) / (3 * V) * CONV_P, rel=1e-12)
    assert th["V"] == pytest.approx(V, rel=1e-15)


def test_thermo_closed_forms(oracle_mod):
    """T, K, P, U and V from states whose values are known without the
    reduction formulas: four free atoms with masses 1..4 and momenta chosen so
    sum p.p/m = 8, and an isolated LJ dimer at rest, whose virial sum F.r is
    d f(d) with f(d) = 24 (2/d^13 - 1/d^7) and whose energy is 4 (d^-12 - d^-6).
    A wrong degree count (2N), a dropped 1/m, a virial of the wrong sign or
    without its 1/3, or a missing unit factor each fails one of them."""
    from moltwister_amd.system import System
    L = 20.0
    V = L ** 3
    # free atoms: farther apart than the cutoff, so no forces and U = 0
    pos = [[-6.0, -6.0, -6.0], [6.0, -6.0, -6.0], [-6.0, 6.0, 6.0], [6.0, 6.0, 0.0]]
    mom = [[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 3.0], [2.0, 2.0, 0.0]]     # p.p/m = 1, 2, 3, 2
    s = System(pos=pos, mass=[1.0, 2.0, 3.0, 4.0], charge=[0] * 4, types=[0] * 4, box=(L, L, L),
               nonbonded=[(0, 0, abi.NB_LJ, (1.0, 1.0, 1.0))], mom=mom)
    o = oracle_mod.Oracle(s, cfg(2.5, 0.3, ensemble=abi.NVE))
    o.compute_forces()
    th = o.thermo()
    assert th["T"] == pytest.approx(8.0 / 12.0 * 120.272, rel=1e-14)
    assert th["K"] == pytest.approx(4.0, rel=1e-14)
    assert th["P"] == pytest.approx(8.0 / (3.0 * V) * 16388.2461, rel=1e-14)
    assert th["U"] == 0.0
    assert th["V"] == pytest.approx(V, rel=1e-15)
    # dimer at rest at separation d: T = K = 0, P = d f(d) / (3V), U = LJ(d)
    for d in (1.05, 1.3, 2.0):
        s = System(pos=[[-d / 2, 0.0, 0.0], [d / 2, 0.0, 0.0]], mass=[1.0, 1.0], charge=[0, 0], types=[0, 0],
                   box=(L, L, L), nonbonded=[(0, 0, abi.NB_LJ, (1.0, 1.0, 1.0))], mom=[[0.0] * 3] * 2)
        o = oracle_mod.Oracle(s, cfg(2.5, 0.3, ensemble=abi.NVE))
        o.compute_forces()
        th = o.thermo()
        f = 24.0 * (2.0 / d ** 13 - 1.0 / d ** 7)
        assert th["T"] == 0.0 and th["K"] == 0.0
        # the tabulated force is the analytic one to ~1e-4 of 24 eps/sigma
        assert th["P"] == pytest.approx(d * f / (3.0 * V) * 16388.2461, abs=d * 24 * 2e-4 / (3.0 * V) * 16388.2461)
        assert th["U"] == pytest.approx(4.0 * (d ** -12 - d ** -6), abs=5e-4)   # tabulation error (1.2e-4 at 1.05)


def test_gradient_descent_step(oracle_mod, lj):
    """runOptimization: r += lr F with uncapped forces (MDLoop.cpp:143-173)."""
    s, c = lj
    o = oracle_mod.Oracle(s, c)
    r0, _ = o.get_state()
    F0, _ = o.compute_forces()
    done = o.optimize(1, 1e-4, 0.0, 100)
    r1, _ = o.get_state()
    assert done == 1
    assert np.allclose(r1, r0 + F0 * 1e-4, rtol=0, atol=1e-15)


def test_two_atom_lj_known_answers(oracle_mod):
    """Two LJ atoms: no force at the minimum 2^(1/6) sigma, the analytic
    24 eps (2/r^13 - 1/r^7) elsewhere to the tabulation error, and U(r)
    counted once (only for i > k)."""
    from moltwister_amd.system import System
    for r, tol in ((2.0 ** (1 / 6), 2e-3), (1.1, 2e-4), (1.5, 2e-4), (2.4, 2e-4)):
        s = System(pos=[[0, 0, 0], [r, 0, 0]], mass=[1, 1], charge=[0, 0], types=[0, 0], box=(12, 12, 12),
                   nonbonded=[(0, 0, abi.NB_LJ, (1.0, 1.0, 1.0))])
        o = oracle_mod.Oracle(s, cfg(2.5, 0.3))
        F, U = o.compute_forces()
        f_exact = 24.0 * (2.0 / r ** 13 - 1.0 / r ** 7)
        assert F[1, 0] == pytest.approx(f_exact, abs=tol * 24)
        assert F[0, 0] == pytest.approx(-F[1, 0], rel=1e-15)
        assert U[1] == 0.0                           # atom k holds the pair only when i > k
        assert U[0] == pytest.approx(4.0 * (r ** -12 - r ** -6), abs=1e-4)


def test_two_charge_coulomb(oracle_mod):
    """K q1 q2 / r with K = 1389.32 (Constants.h:40), all-pairs, no cutoff."""
    from moltwister_amd.system import System
    r = 7.3
    s = System(pos=[[0, 0, 0], [r, 0, 0]], mass=[1, 1], charge=[0.5, -0.8], types=[0, 0], box=(30, 30, 30),
               nonbonded=[])
    o = oracle_mod.Oracle(s, cfg(2.5, 0.3, coulomb=abi.COUL_ALLPAIRS))
    F, U = o.compute_forces()
    K = float(np.float32(1389.32) * np.float32(0.5) * np.float32(-0.8))
    assert F[1, 0] == pytest.approx(K / r ** 2, rel=1e-6)
    assert U.sum() == pytest.approx(K / r, rel=1e-6)


def test_harmonic_bond_known_answer(oracle_mod):
    """U = k (r - r0)^2, F = -2k (r - r0) along the bond (MolTwisterMDFFBond_Harm.cpp:90-110)."""
    from moltwister_amd.system import System
    k, r0, r = 500.0, 1.0, 1.1
    s = System(pos=[[0, 0, 0], [r, 0, 0]], mass=[1, 1], charge=[0, 0], types=[0, 0], box=(20, 20, 20),
               nonbonded=[], bond_ff=[(abi.BOND_HARM, (k, r0))], bonds=[[0, 1, 0]])
    o = oracle_mod.Oracle(s, cfg(5.0, 0.5))
    F, U = o.compute_forces()
    assert F[1, 0] == pytest.approx(-2 * k * (r - r0), rel=1e-4)
    assert U.sum() == pytest.approx(k * (r - r0) ** 2, rel=1e-3)


def test_harmonic_angle_known_answer(oracle_mod):
    """Forces of U = k (theta - theta0)^2 on a bent triatomic sum to zero and
    match -dU/dr by central differences of the analytic energy."""
    from moltwister_amd.system import System
    k, th0 = 100.0, 104.5
    th = np.radians(112.0)
    pos = np.array([[1.0, 0, 0], [0, 0, 0], [np.cos(th), np.sin(th), 0.0]])
    s = System(pos=pos, mass=[1, 1, 1], charge=[0, 0, 0], types=[0, 0, 0], box=(20, 20, 20), nonbonded=[],
               angle_ff=[(abi.ANGLE_HARM, (k, th0))], angles=[[0, 1, 2, 0]])
    o = oracle_mod.Oracle(s, cfg(5.0, 0.5))
    F, U = o.compute_forces()

    def energy(p):
        a, b = p[0] - p[1], p[2] - p[1]
        t = np.arccos(a @ b / np.linalg.norm(a) / np.linalg.norm(b))
        return k * (t - np.radians(th0)) ** 2

    h = 1e-6
    for i in range(3):
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, d] += h
            pm[i, d] -= h
            assert F[i, d] == pytest.approx(-(energy(pp) - energy(pm)) / (2 * h), abs=2e-3 * k)
    assert np.abs(F.sum(axis=0)).max() < 1e-6 * k
    assert U.sum() == pytest.approx(energy(pos), rel=2e-3)


def _normals_angle(p):
    """Angle in [0, pi] between the normals r23 x r21 and r32 x r34 of the
    planes (1,2,3) and (2,3,4): pi for a cis torsion, 0 for trans."""
    b1, b2, b3 = p[1] - p[0], p[2] - p[1], p[3] - p[2]
    n1, n2 = np.cross(b2, p[0] - p[1]), np.cross(-b2, b3)
    return np.arccos(n1 @ n2 / np.linalg.norm(n1) / np.linalg.norm(n2))


@pytest.mark.parametrize("kind,prm,U", [
    (abi.DIH_HARM, (3.0, 1, 3), lambda f: 3.0 * (1.0 + np.cos(3 * f))),
    (abi.DIH_FOURIER4T, (1.5, -2.0, 0.7, 0.3),
     lambda f: 0.5 * (1.5 * (1 + np.cos(f)) - 2.0 * (1 - np.cos(2 * f)) + 0.7 * (1 + np.cos(3 * f))
                      + 0.3 * (1 - np.cos(4 * f)))),
])
def test_dihedral_forces_known_answer(oracle_mod, kind, prm, U):
    """Forces of a torsion on four atoms are -dE/dr of the closed-form energy
    E = U(phi) (central differences) to table accuracy, they sum to zero, and
    the energies of the two centres add up to E. phi is the dihedral with cis
    = 0, i.e. pi minus the angle between the plane normals: the profile
    geometry (MolTwisterMDFFDih.cpp:272-320: r1 = (0,1,0), r4 = (sin x,
    cos x, 1)) is cis at x = 0, where the classes' normals are antiparallel."""
    from moltwister_amd.system import System
    pos = np.array([[0.1, 1.2, -0.1], [0.0, 0.0, 0.0], [0.05, -0.1, 1.3], [0.9, 0.6, 1.5]])
    s = System(pos=pos, mass=[1, 1, 1, 1], charge=[0, 0, 0, 0], types=[0, 0, 0, 0], box=(30, 30, 30),
               nonbonded=[], dihedral_ff=[(kind, prm)], dihedrals=[[0, 1, 2, 3, 0]])
    c = cfg(5.0, 0.5)
    c.scale14 = 0.0
    o = oracle_mod.Oracle(s, c)
    F, Ua = o.compute_forces()

    def energy(p):
        return U(np.pi - _normals_angle(p))

    h = 1e-6
    Fmax = 0.0
    for i in range(4):
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, d] += h
            pm[i, d] -= h
            ref = -(energy(pp) - energy(pm)) / (2 * h)
            Fmax = max(Fmax, abs(ref))
            assert F[i, d] == pytest.approx(ref, abs=2e-3 * max(1.0, abs(ref)))
    assert Fmax > 0.5
    assert np.abs(F.sum(axis=0)).max() < 1e-6 * Fmax
    assert float(Ua.sum()) == pytest.approx(energy(pos), rel=1e-4, abs=1e-5)


def test_class2_angle_tables_closed_form(oracle_mod):
    """CMDFFAngle_Class2 (MolTwisterMDFFAngle_Class2.cpp:145-296) at the
    profile's unit bond lengths: U(theta) = K2 d^2 + K3 d^3 + K4 d^4
    + M (1 - r1)(1 - r2) + (N1 (1 - r1') + N2 (1 - r2')) d, d = theta - theta0;
    with only the Ea terms the force column is -dU/dtheta, and Ea with K3 =
    K4 = 0 is the harmonic angle."""
    two_pi = float(np.float32(2 * np.pi))
    q = (112.0, 40.0, -20.0, 10.0, 5.0, 1.1, 1.3, 3.0, 4.0, 1.2, 0.9)
    x, f, e = oracle_mod.profile(2, abi.ANGLE_CLASS2, q, 0.0, two_pi)
    th = x.astype(np.float64)
    sel = (th > 0.05) & (th < np.pi - 0.05)
    d = th - np.radians(q[0])
    U = q[1] * d ** 2 + q[2] * d ** 3 + q[3] * d ** 4 + q[4] * (1 - q[5]) * (1 - q[6]) \
        + (q[7] * (1 - q[9]) + q[8] * (1 - q[10])) * d
    assert np.allclose(e[sel], U[sel], rtol=2e-6, atol=2e-5)
    qa = q[:4] + (0.0,) * 7
    _, fa, _ = oracle_mod.profile(2, abi.ANGLE_CLASS2, qa, 0.0, two_pi)
    dU = 2 * q[1] * d + 3 * q[2] * d ** 2 + 4 * q[3] * d ** 3
    assert np.allclose(fa[sel], -dU[sel], rtol=1e-5, atol=1e-3)
    _, f2, e2 = oracle_mod.profile(2, abi.ANGLE_CLASS2, (104.5, 75.0) + (0.0,) * 9, 0.0, two_pi)
    _, fh, eh = oracle_mod.profile(2, abi.ANGLE_HARM, (75.0, 104.5), 0.0, two_pi)
    assert np.array_equal(f2, fh)
    assert np.allclose(e2, eh, rtol=1e-6, atol=1e-6)


def test_ljbuck_tables(oracle_mod):
    """CMDFFNonBonded_LJBuck: the energy column of its LJ (alpha 1) and
    Buckingham forms, and a force column of zeros (its calcForces zeroes both
    forces at the end, MolTwisterMDFFNonBonded_LJBuck.cpp:250)."""
    for lb, plain in (((0, 1.2, 3.1), (abi.NB_LJ, (1.2, 3.1, 1.0))),
                      ((1, 1000.0, 0.2, 5.0), (abi.NB_BUCK, (1000.0, 0.2, 5.0)))):
        x, f, e = oracle_mod.profile(0, abi.NB_LJBUCK, lb, 0.01, 10.0)
        xp, fp, ep = oracle_mod.profile(0, plain[0], plain[1], 0.01, 10.0)
        assert np.array_equal(x, xp) and np.array_equal(e, ep)
        assert np.all(f == 0.0) and np.any(fp != 0.0)


def test_oracle_threads_give_identical_results(oracle_mod):
    """The oracle's per-atom loops on several host threads (the CPU baseline's
    all-core figure) give the single-thread results bit for bit."""
    from systems import make
    for name in ("binary_lj", "water_300"):
        s, c = make(name)
        o1, o4 = oracle_mod.Oracle(s, c), oracle_mod.Oracle(s, c)
        o4.set_threads(4)
        F1, U1 = o1.compute_forces()
        F4, U4 = o4.compute_forces()
        assert np.array_equal(F1, F4) and np.array_equal(U1, U4)
        n1, l1 = o1.build_neighbors()
        n4, l4 = o4.build_neighbors()
        assert np.array_equal(n1, n4) and np.array_equal(l1, l4)


def test_exclusion_scaling_known_answer(oracle_mod):
    """Which pairs the per-atom bonded lists mark as 1-2, 1-3, 1-4 and same
    molecule (FunctorCalcForce.cpp:95-157, priority 1-2 > 1-3 > 1-4 > 1N),
    from the forces alone: a five-atom chain 0-1-2-3-4 (four bonds, three
    angles, two dihedrals, all of zero strength) and a sixth atom in its own
    molecule, LJ between every pair, the four scale factors all different.
    The force on each atom is sum_j s_ij f_LJ(r_ij) r_ij / r_ij with s = 0.1
    (1-2), 0.25 (1-3), 0.5 (1-4), 0.75 (same molecule otherwise: 0-4) and 1
    (other molecule), computed here analytically."""
    from moltwister_amd.system import System
    pos = np.array([[0.0, 0.0, 0.0], [1.1, 0.0, 0.0], [1.1, 1.1, 0.0], [0.0, 1.1, 0.0], [0.0, 1.1, 1.1],
                    [-0.3, -0.2, 1.4]])
    bonds = [[0, 1, 0], [1, 2, 0], [2, 3, 0], [3, 4, 0]]
    angles = [[0, 1, 2, 0], [1, 2, 3, 0], [2, 3, 4, 0]]
    dih = [[0, 1, 2, 3, 0], [1, 2, 3, 4, 0]]
    s = System(pos=pos, mass=[1.0] * 6, charge=[0.0] * 6, types=[0] * 6, box=(20, 20, 20),
               nonbonded=[(0, 0, abi.NB_LJ, (1.0, 1.0, 1.0))],
               bond_ff=[(abi.BOND_HARM, (0.0, 1.1))], angle_ff=[(abi.ANGLE_HARM, (0.0, 90.0))],
               dihedral_ff=[(abi.DIH_HARM, (0.0, 1, 3))], bonds=bonds, angles=angles, dihedrals=dih)
    c = cfg(2.5, 0.3)
    c.scale12, c.scale13, c.scale14, c.scale1N = 0.1, 0.25, 0.5, 0.75
    o = oracle_mod.Oracle(s, c)
    F, U = o.compute_forces()
    sc = np.ones((6, 6))
    for a, b in ((0, 4),):
        sc[a, b] = sc[b, a] = 0.75
    for a, b in ((0, 3), (1, 4)):
        sc[a, b] = sc[b, a] = 0.5
    for a, b in ((0, 2), (1, 3), (2, 4)):
        sc[a, b] = sc[b, a] = 0.25
    for a, b, _ in bonds:
        sc[a, b] = sc[b, a] = 0.1
    want = np.zeros((6, 3))
    u_want = 0.0
    for i in range(6):
        for j in range(6):
            if i == j:
                continue
            d = pos[i] - pos[j]
            r = np.linalg.norm(d)
            want[i] += sc[i, j] * 24.0 * (2.0 / r ** 13 - 1.0 / r ** 7) * d / r
            if i > j:
                u_want += sc[i, j] * 4.0 * (r ** -12 - r ** -6)
    # the tabulated LJ is the analytic one to ~2e-4 of 24 eps/sigma at r >= 1.1
    assert np.abs(F - want).max() < 24 * 5e-4, np.abs(F - want).max()
    assert U.sum(dtype=np.float64) == pytest.approx(u_want, abs=1e-3)

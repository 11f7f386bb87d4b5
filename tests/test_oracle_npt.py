"""The oracle's Nose-Hoover and MTK integrators (NVT and NPT) pinned by the
equations of motion they discretise.

The reference's step (MDLoop.cpp:124-138: NH_P, NH_T, velocity Verlet,
NH_T, NH_P) is a symmetric Trotter splitting of the Martyna-Tuckerman-Klein
equations with Nose-Hoover chains on the particles and on the barostat
(NHChain.cpp:54-113 with CFctT/CFctP, MDLoop.cpp:29-83; VelVerlet.cpp:57-71,
124-174). For free particles (all pair forces zero) those equations are,
with alpha = 1 + 1/N and u = p_eps/W:

    dp_i/dt    = -(alpha u + p_eta1/Q1) p_i
    deps/dt    = u
    dp_eps/dt  = alpha sum p^2/m - 3 P V0 exp(3 eps) - (p_xi1/Q'1) p_eps
    chain on the particles: G_1 = sum p^2/m - 3 N T,   G_j = p_eta(j-1)^2/Q(j-1) - T
    chain on the barostat:  G'_1 = p_eps^2/W - T,      G'_j = p_xi(j-1)^2/Q'(j-1) - T
    dp_eta_j/dt = G_j - p_eta_j p_eta(j+1)/Q(j+1),   deta_j/dt = p_eta_j/Q_j

They are integrated here with scipy (DOP853, tight tolerances), independently
of the splitting. The oracle's trajectory must converge to that solution with
order 2 in the time step: a dropped term, a wrong factor (alpha, 1/N, 3PV, a
Q) or a misplaced half step changes the limit or the order and fails.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

from moltwister_amd import abi
from moltwister_amd.system import System

from systems import cfg as make_cfg

CONV_T, CONV_t, CONV_PRESS = 120.272, 100.0, 16388.2461


def free_particles(n=8, L=10.0, seed=3):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-L / 2, L / 2, (n, 3))
    mass = 1.0 + np.arange(n, dtype=float)
    mom = rng.normal(0.0, 1.0, (n, 3)) * np.sqrt(mass)[:, None] * 2.0
    # LJ with epsilon = 0: every tabulated force and energy is zero
    return System(pos=pos, mass=mass, charge=np.zeros(n), types=np.zeros(n, np.int32), box=(L, L, L),
                  type_names=["X"], nonbonded=[(0, 0, abi.NB_LJ, (0.0, 1.0, 1.0))], mom=mom)


def settings(ensemble, dt_fs):
    c = make_cfg(2.5, 0.3, ensemble=ensemble, dt=dt_fs, T=300.0)
    c.temperature_relax = 20.0
    c.pressure = 3000.0
    c.pressure_relax = 40.0
    c.temperature_nh_len = c.pressure_nh_len = 3
    c.temperature_respa = c.pressure_respa = 2
    c.cutoff_force = 0.0
    return c


def rhs_factory(s, c, npt):
    n, m = s.n_atoms, s.mass
    M, Mp = c.temperature_nh_len, c.pressure_nh_len if npt else 0
    T = c.temperature / CONV_T
    tau, tauP = c.temperature_relax / CONV_t, c.pressure_relax / CONV_t
    Q = np.array([3 * n * T * tau * tau] + [T * tau * tau] * (M - 1))
    W = T * tauP * tauP
    Qp = np.array([T * tauP * tauP] * Mp)
    P = c.pressure / CONV_PRESS
    V0 = float(np.prod(s.box))
    alpha = 1.0 + 1.0 / n

    def unpack(y):
        p = y[:3 * n].reshape(n, 3)
        i = 3 * n
        eps, peps = y[i], y[i + 1]
        i += 2
        eta, peta = y[i:i + M], y[i + M:i + 2 * M]
        i += 2 * M
        xi, pxi = y[i:i + Mp], y[i + Mp:i + 2 * Mp]
        return p, eps, peps, eta, peta, xi, pxi

    def chain(p_chain, Qc, G1):
        G = np.empty_like(p_chain)
        G[0] = G1
        G[1:] = p_chain[:-1] ** 2 / Qc[:-1] - T
        d = G.copy()
        d[:-1] -= p_chain[:-1] * p_chain[1:] / Qc[1:]
        return d

    def rhs(_, y):
        p, eps, peps, eta, peta, xi, pxi = unpack(y)
        ke2 = float(np.sum(np.sum(p * p, 1) / m))
        u = peps / W if npt else 0.0
        dp = -(alpha * u + peta[0] / Q[0]) * p
        out = [dp.ravel()]
        if npt:
            dpeps = alpha * ke2 - 3.0 * P * V0 * np.exp(3.0 * eps) - (pxi[0] / Qp[0]) * peps
            out.append([u, dpeps])
        else:
            out.append([0.0, 0.0])
        out += [peta / Q, chain(peta, Q, ke2 - 3 * n * T)]
        if npt:
            out += [pxi / Qp, chain(pxi, Qp, peps * peps / W - T)]
        return np.concatenate([np.ravel(np.asarray(x, float)) for x in out])

    return rhs, W, Q, Qp


def start_state(c, n, W, Q, Qp, npt):
    tau, tauP = c.temperature_relax / CONV_t, c.pressure_relax / CONV_t
    eta = np.array([0.1, -0.2, 0.05])
    peta = np.array([0.3, -0.4, 0.2]) * Q / tau
    eps, peps = (0.0, 0.6 * W / tauP) if npt else (0.0, 0.0)
    xi = np.array([0.0, 0.1, -0.1]) if npt else np.zeros(0)
    pxi = np.array([-0.3, 0.2, 0.1]) * Qp / tauP if npt else np.zeros(0)
    return eta, peta, eps, peps, xi, pxi


def oracle_run(P, s, c, npt, t_end):
    o = P.Oracle(s, c)
    rhs, W, Q, Qp = rhs_factory(s, c, npt)
    eta, peta, eps, peps, xi, pxi = start_state(c, s.n_atoms, W, Q, Qp, npt)
    o.set_thermostat(eta, peta)
    if npt:
        o.set_barostat(eps, peps, xi, pxi)
    nsteps = int(round(t_end / (c.timestep / CONV_t)))
    o.step(nsteps)
    _, mom = o.get_state()
    eta1, peta1 = o.get_thermostat()
    parts = [mom.ravel()]
    if npt:
        b = o.get_barostat()
        parts += [[b["eps"], b["p_eps"]], eta1, peta1, b["eta"], b["p_eta"]]
    else:
        parts += [[0.0, 0.0], eta1, peta1]
    return np.concatenate([np.ravel(np.asarray(x, float)) for x in parts])


@pytest.mark.parametrize("ensemble", [abi.NVT, abi.NPT])
def test_oracle_converges_to_mtk_equations_with_order_two(oracle_mod, ensemble):
    npt = ensemble == abi.NPT
    s = free_particles()
    t_end = 0.6                                   # reduced time (60 fs)
    c0 = settings(ensemble, 1.0)
    rhs, W, Q, Qp = rhs_factory(s, c0, npt)
    eta, peta, eps, peps, xi, pxi = start_state(c0, s.n_atoms, W, Q, Qp, npt)
    y0 = np.concatenate([s.mom.ravel(), [eps, peps], eta, peta, xi, pxi])
    sol = solve_ivp(rhs, (0.0, t_end), y0, method="DOP853", rtol=1e-13, atol=1e-13)
    y_exact = sol.y[:, -1]
    scale = np.maximum(np.abs(y_exact), 1e-3 * np.abs(y_exact).max())
    errs = []
    for dt_fs in (1.0, 0.5, 0.25):
        y = oracle_run(oracle_mod, s, settings(ensemble, dt_fs), npt, t_end)
        errs.append(float(np.max(np.abs(y - y_exact) / scale)))
    # the dynamics moved the state well away from the start
    assert np.max(np.abs(y_exact - y0) / scale) > 0.05
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.2 < r1 < 4.8 and 3.2 < r2 < 4.8, errs
    assert errs[2] < 0.03, errs


def test_npt_box_follows_the_volume(oracle_mod):
    """getSimBox (VelVerlet.cpp:199-211): the box scales with (V/V0)^(1/3),
    and the thermo volume is V0 exp(3 eps)."""
    s = free_particles()
    c = settings(abi.NPT, 1.0)
    o = oracle_mod.Oracle(s, c)
    o.set_barostat(0.0, 0.5 * (c.temperature / CONV_T) * (c.pressure_relax / CONV_t), None, None)
    o.step(40)
    b = o.get_barostat()
    V = np.prod(s.box) * np.exp(3.0 * b["eps"])
    assert abs(b["eps"]) > 1e-4
    assert np.allclose(b["box"], np.array(s.box) * (V / np.prod(s.box)) ** (1.0 / 3.0), rtol=1e-14)
    th = o.thermo()
    assert th["V"] == pytest.approx(V, rel=1e-14)
    pos, _ = o.get_state()
    assert np.all(pos >= -b["box"] / 2) and np.all(pos < b["box"] / 2)

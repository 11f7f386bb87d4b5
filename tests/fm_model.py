"""Independent numpy FP64 restatement of the force-matching model (test
helper): LJ 4 eps [(s/r)^12 - (s/r)^6] between atoms of different molecules
within the cutoff, harmonic bonds k (r - r0)^2 and angles k (theta - theta0)^2
(theta0 in degrees), minimum image. The formulas are those of the MDFF
classes (MolTwisterMDFFNonBonded_LJ.cpp:46-60, MolTwisterMDFFBond_Harm.cpp:43-51,
MolTwisterMDFFAngle_Harm.cpp:47-55); forces are minus the gradient of the
energy, taken here by central differences of the energy on tiny systems as a
pin independent of any hand-derived force expression."""
import numpy as np


def molecules(n, bonds):
    par = list(range(n))

    def find(a):
        while par[a] != a:
            par[a] = par[par[a]]
            a = par[a]
        return a
    for a, b, _ in bonds:
        ra, rb = find(a), find(b)
        if ra != rb:
            par[max(ra, rb)] = min(ra, rb)
    return np.array([find(k) for k in range(n)])


def energy(s, pos, params, cutoff):
    L = np.array(s.box)
    mol = molecules(s.n_atoms, s.bonds)
    U = 0.0
    p = 0
    for a, b, _, _ in s.nonbonded:
        eps, sig = params[p], params[p + 1]
        p += 2
        for i in range(s.n_atoms):
            for j in range(i + 1, s.n_atoms):
                if mol[i] == mol[j]:
                    continue
                ti, tj = s.types[i], s.types[j]
                if not ((ti == a and tj == b) or (ti == b and tj == a)):
                    continue
                d = pos[i] - pos[j]
                d -= L * np.round(d / L)
                r = np.sqrt(d @ d)
                if r < cutoff:
                    U += 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)
    nb = len(s.bond_ff)
    for a, b, t in s.bonds:
        k, r0 = params[p + 2 * t], params[p + 2 * t + 1]
        d = pos[a] - pos[b]
        d -= L * np.round(d / L)
        U += k * (np.sqrt(d @ d) - r0) ** 2
    p += 2 * nb
    for a, c, b, t in s.angles:
        k, th0 = params[p + 2 * t], np.radians(params[p + 2 * t + 1])
        u = pos[a] - pos[c]
        v = pos[b] - pos[c]
        u -= L * np.round(u / L)
        v -= L * np.round(v / L)
        cth = np.clip(u @ v / np.sqrt((u @ u) * (v @ v)), -1, 1)
        U += k * (np.arccos(cth) - th0) ** 2
    return U


def forces_fd(s, pos, params, cutoff, h=1e-6):
    F = np.zeros_like(pos)
    for i in range(s.n_atoms):
        for a in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, a] += h
            pm[i, a] -= h
            F[i, a] = -(energy(s, pp, params, cutoff) - energy(s, pm, params, cutoff)) / (2 * h)
    return F

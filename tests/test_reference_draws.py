"""mtmd_reference_draws: the reference's initial random state
(CSimulationBox::initSystem, SimulationBox.cpp:53-75) with glibc's rand().

Pins:
- the generator: its outputs are the C library's own rand() after srand(seed)
  in this process (the container's glibc), for several seeds including 0 and
  one above 2^31;
- the sequence and arithmetic: a restatement of SimulationBox.cpp:98-133,
  NHChain.cpp:115-143 and VelVerlet.cpp:176-183 driven by the C library's
  rand() itself, bit for bit;
- the chain draws: CNHChain::setRandNHPos/setRandNHMom compiled from the
  reference sources (oracle/_ref), interleaved with the momentum draws as
  initSystem interleaves them, bit for bit (skipped without the reference build).
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from moltwister_amd import abi, config
from moltwister_amd.engine import reference_draws
from moltwister_amd.system import System

CONV_T, CONV_t = 120.272, 100.0
RAND_MAX = 2147483647.0

libc = C.CDLL("libc.so.6")
libc.srand.argtypes = [C.c_uint]
libc.rand.restype = C.c_int


def small_system(n=40, immobile=(3, 17)):
    rng = np.random.default_rng(5)
    mobile = np.ones(n, np.uint8)
    mobile[list(immobile)] = 0
    return System(pos=rng.uniform(-5, 5, (n, 3)), mass=rng.uniform(1.0, 40.0, n), charge=np.zeros(n),
                  types=np.zeros(n, np.int32), box=(10.0, 10.0, 10.0), mobile=mobile)


def cfg(mt=4, mp=3):
    c = config.defaults()
    c.temperature, c.temperature_relax, c.temperature_nh_len = 298.0, 20.0, mt
    c.pressure_relax, c.pressure_nh_len = 5000.0, mp
    return c


@pytest.mark.parametrize("seed", [1, 7, 0, 123456789, 3000000000])
def test_generator_is_the_c_library_rand(seed):
    s = small_system(n=6, immobile=())
    d = reference_draws(s, cfg(mt=8, mp=0), seed)
    libc.srand(seed)
    expect = np.array([libc.rand() / RAND_MAX for _ in range(8)])
    assert np.array_equal(d["eta_T"], expect)


def restated(s, c, seed):
    """initSystem's draws, restated from the reference sources, on the C library's rand()"""
    n, mt, mp = s.n_atoms, c.temperature_nh_len, c.pressure_nh_len
    T = c.temperature / CONV_T
    tt, tp = c.temperature_relax / CONV_t, c.pressure_relax / CONV_t
    a = 2.0 / RAND_MAX
    qt = [(float(3 * n) if j == 0 else 1.0) * T * tt * tt for j in range(mt)]
    qp = [(float(1) if j == 0 else 1.0) * T * tp * tp for j in range(mp)]
    libc.srand(seed)
    et = [libc.rand() / RAND_MAX for _ in range(mt)]          # NHChain.cpp:126-132
    ep = [libc.rand() / RAND_MAX for _ in range(mp)]
    p = np.zeros((n, 3))                                     # SimulationBox.cpp:98-133
    for k in range(n):
        for d in range(3):
            v = a * float(libc.rand()) - 1.0
            p[k, d] = 1.0 if v == 0.0 else v
    for k in range(n):
        if not s.mobile[k]:
            p[k] = 0.0
    tsum = 0.0
    for k in range(n):
        tsum += ((p[k, 0] * p[k, 0]) + (p[k, 1] * p[k, 1]) + (p[k, 2] * p[k, 2])) / 1.0   # masses still 1
    scale = math.sqrt(T / (tsum / (3 * float(n))))
    p = p * scale
    pt = []
    for j in range(mt):                                      # NHChain.cpp:134-143
        v = (a * float(libc.rand()) - 1.0) * (qt[j] / tt)
        pt.append(qt[j] / tt if v == 0.0 else v)
    pp = []
    for j in range(mp):
        v = (a * float(libc.rand()) - 1.0) * (qp[j] / tp)
        pp.append(qp[j] / tp if v == 0.0 else v)
    W = T * tp * tp                                          # VelVerlet.cpp:176-183
    pe = (a * float(libc.rand()) - 1.0) * (W / tp)
    pe = W / tp if pe == 0.0 else pe
    return dict(mom=p, eta_T=np.array(et), p_eta_T=np.array(pt), eta_P=np.array(ep), p_eta_P=np.array(pp), p_eps=pe)


@pytest.mark.parametrize("seed", [1, 42])
def test_initial_state_matches_restatement(seed):
    s, c = small_system(), cfg()
    d, r = reference_draws(s, c, seed), restated(s, c, seed)
    for k in ("mom", "eta_T", "p_eta_T", "eta_P", "p_eta_P"):
        assert np.array_equal(d[k], r[k]), k
    assert d["p_eps"] == r["p_eps"]
    assert np.all(d["mom"][[3, 17]] == 0.0)
    # unit masses in the scaling: sum p^2 / (3 n) is the target in reduced units
    assert abs(np.sum(d["mom"] ** 2) / (3 * s.n_atoms) - 298.0 / CONV_T) < 1e-12


REF_LIB = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libmtref.so")


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference build oracle/_ref/libmtref.so not present")
@pytest.mark.parametrize("seed", [1, 9])
def test_chain_draws_are_the_reference_code(seed):
    ref = C.CDLL(REF_LIB)
    dp = C.POINTER(C.c_double)
    ref.ref_nh_draw.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, dp, dp]
    s, c = small_system(), cfg(mt=4, mp=3)
    n = s.n_atoms
    T = c.temperature / CONV_T
    tt, tp = c.temperature_relax / CONV_t, c.pressure_relax / CONV_t
    et, pt, ep, pp = (np.zeros(4), np.zeros(4), np.zeros(3), np.zeros(3))
    ptr = lambda a: a.ctypes.data_as(dp)
    libc.srand(seed)                                  # initSystem's order, SimulationBox.cpp:66-71
    ref.ref_nh_draw(0, 4, n, 3, T, tt, ptr(et), None)
    ref.ref_nh_draw(0, 3, 1, 1, T, tp, ptr(ep), None)
    for _ in range(3 * n):
        libc.rand()
    ref.ref_nh_draw(1, 4, n, 3, T, tt, None, ptr(pt))
    ref.ref_nh_draw(1, 3, 1, 1, T, tp, None, ptr(pp))
    d = reference_draws(s, c, seed)
    assert np.array_equal(d["eta_T"], et) and np.array_equal(d["p_eta_T"], pt)
    assert np.array_equal(d["eta_P"], ep) and np.array_equal(d["p_eta_P"], pp)

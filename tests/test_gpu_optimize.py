"""moldyn optimizeenergy (CMDLoop::runOptimization, MDLoop.cpp:143-173) on the
sorted engines: after the first full build their skin lists are rebuilt only
when an atom moved more than half the skin, while the EXACT engine rebuilds at
rc every step as the reference does. The descent must follow EXACT (FP64 to
rounding, MIXED to its FP32 pair error), with fewer rebuilds than steps."""
import numpy as np
import pytest

from moltwister_amd import abi
from moltwister_amd import workloads as W
from moltwister_amd.engine import Engine

from systems import make

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["binary_lj", "water_300"])
@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_descent_follows_exact(name, engine):
    print(f"case optimize {name} engine={engine}", flush=True)
    s, c = make(name, coulomb=abi.COUL_CUTOFF)
    s2, c2 = make(name, engine=engine, coulomb=abi.COUL_CUTOFF)
    lr = 2e-5 if name == "binary_lj" else 1e-4
    ex, fa = Engine(s, c), Engine(s2, c2)
    ex.optimize(60, lr, 0.0)
    fa.optimize(60, lr, 0.0)
    pe, _ = ex.get_state()
    pf, _ = fa.get_state()
    L = np.array(s.box)
    d = pf - pe
    d -= L * np.rint(d / L)
    tol = 1e-9 if engine == abi.ENGINE_FP64 else 1e-4
    assert np.abs(d).max() < tol, np.abs(d).max()
    # per-atom energies are float sums in every engine (the reference keeps U_k
    # as float) and split differently between the pair's atoms, so the totals
    # agree to float rounding
    ue, uf = ex.thermo()["U"], fa.thermo()["U"]
    assert uf == pytest.approx(ue, rel=1e-6 if engine == abi.ENGINE_FP64 else 1e-5)
    assert np.abs(pf - s.pos).max() > 1e-3                    # the atoms did move
    st = fa.stats()
    assert 1 <= st["rebuilds"] < 60, st
    ex.close()
    fa.close()


def test_descent_on_the_tile_path_with_rebuilds():
    """131,072-atom mixture (16-bit block lists): 40 descent steps with
    conditional rebuilds give the FP64 engine's energies."""
    print("case optimize tile", flush=True)
    from test_gpu_tile import lj_cfg
    s = W.binary_lj_large(131072, tiles=2, seed=5)
    em, ef = Engine(s, lj_cfg(abi.ENGINE_MIXED, 0.3)), Engine(s, lj_cfg(abi.ENGINE_FP64, 0.3))
    em.optimize(40, 2e-5, 0.0)
    ef.optimize(40, 2e-5, 0.0)
    assert em.thermo()["U"] == pytest.approx(ef.thermo()["U"], rel=1e-5)
    rm, rf = em.stats()["rebuilds"], ef.stats()["rebuilds"]
    assert 1 <= rm < 40 and 1 <= rf < 40, (rm, rf)
    em.close()
    ef.close()

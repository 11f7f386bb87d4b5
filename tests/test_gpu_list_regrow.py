"""Neighbour-list growth inside a multi-step call. The reference grows the list
and builds it again when a row overflows (MDFFMatrices.cpp:163-178); the
FP64/MIXED engines find an overflow at the end of a call, grow the lists,
restore the state the call started from and run the call again.

MTMD_LIST_SLACK=0 sizes the lists to the longest row of the latest build
(starting from 16 words), so list builds inside a call overflow as the rows
fluctuate. Such a call must end without error, with a grown capacity, on the
trajectory of the same call run with lists that never overflowed: FP64 to the
rounding of a different summation order, MIXED to its FP32 pair error."""
import copy
import os

import numpy as np
import pytest

from moltwister_amd import abi
from moltwister_amd import workloads as W
from moltwister_amd.engine import Engine

from systems import make

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def liquid():
    s = W.binary_lj_large(16384, tiles=1, seed=3)
    c = make("binary_lj", engine=abi.ENGINE_MIXED, T=120.272, dt=0.5)[1]
    e = Engine(s, c)
    e.optimize(100, 2e-5, 0.0)
    for _ in range(2):
        e.step(100)
        _, mom = e.get_state()
        e.set_state(mom=mom * np.sqrt(120.272 / e.thermo()["T"]))
    s.pos, s.mom = e.get_state()
    e.close()
    return s


def engine(s, eng, ensemble, slack):
    c = make("binary_lj", engine=eng, ensemble=ensemble, T=120.272, dt=0.5)[1]
    old = os.environ.get("MTMD_LIST_SLACK")
    os.environ["MTMD_LIST_SLACK"] = "1" if slack else "0"
    try:
        return Engine(s, c)
    finally:
        if old is None:
            os.environ.pop("MTMD_LIST_SLACK")
        else:
            os.environ["MTMD_LIST_SLACK"] = old


CALL = 25          # steps a call: short enough for the chaos of the liquid to stay near rounding


@pytest.mark.parametrize("eng,ensemble", [(abi.ENGINE_FP64, abi.NVE), (abi.ENGINE_MIXED, abi.NVE),
                                          (abi.ENGINE_MIXED, abi.NVT)])
def test_overflow_inside_a_call_reruns_it(liquid, eng, ensemble):
    """Each call of the tight engine is repeated by a fresh engine with the
    default list margin from the state (and thermostat) the call started
    from; the calls in which the tight lists grew are the ones that matter."""
    print(f"case list regrow engine={eng} ensemble={ensemble}", flush=True)
    tight = engine(liquid, eng, ensemble, False)
    tight.compute_forces()
    cap0 = tight.stats()["capacity"]
    L = np.array(liquid.box)
    checked = 0
    for call in range(40):
        pos0, mom0 = tight.get_state()
        nh0 = tight.get_thermostat()
        tight.step(CALL)
        cap = tight.stats()["capacity"]
        if cap > cap0:
            s = copy.copy(liquid)
            s.pos, s.mom = pos0, mom0
            ref = engine(s, eng, ensemble, True)
            if ensemble == abi.NVT:
                ref.set_thermostat(*nh0)
            ref.step(CALL)
            assert cap <= ref.stats()["capacity"]
            pr, mr = ref.get_state()
            pt, mt = tight.get_state()
            dp = pt - pr
            dp -= L * np.rint(dp / L)
            tol = 1e-9 if eng == abi.ENGINE_FP64 else 1e-4
            print(f"  call {call}: capacity {cap0} -> {cap}, max |dpos| {np.abs(dp).max():.2e}, "
                  f"max |dmom| {np.abs(mt - mr).max() / np.abs(mr).max():.2e}", flush=True)
            assert np.abs(dp).max() < tol, np.abs(dp).max()
            assert np.abs(mt - mr).max() < tol * np.abs(mr).max()
            tr, tt = ref.thermo(), tight.thermo()
            assert abs(tt["Etot"] - tr["Etot"]) <= max(tol, 1e-6) * abs(tr["K"]), (tt["Etot"], tr["Etot"])
            ref.close()
            checked += 1
            if checked == 2:
                break
        cap0 = cap
    assert checked >= 1, "no list build overflowed inside a call"
    assert tight.stats()["steps"] == CALL * (call + 1)
    tight.close()

"""mtmd_get_thermo_noalloc against mtmd_get_thermo through the C ABI: the
same reductions with their partials in the library's fixed device array, so
every field must agree bit for bit. Two calls in a row, the second on a
smaller system, on every engine; the tile path's U_cut reduction on the
131,072-atom mixture; and the refusal without forces."""
import pytest

from moltwister_amd import abi
from moltwister_amd import workloads as W
from moltwister_amd.engine import Engine, MtmdError

from systems import make

pytestmark = pytest.mark.gpu

FIELDS = [k for k, _ in abi.MtmdThermo._fields_]


def _same(a, b):
    for k in FIELDS:
        assert a[k] == b[k], (k, a[k], b[k])


@pytest.mark.parametrize("engine", [abi.ENGINE_EXACT, abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_noalloc_thermo_matches(engine):
    print(f"case noalloc thermo engine={engine}", flush=True)
    # first call: 1500-atom mixture after a few steps; second call: a smaller
    # charged system (300 waters, cutoff Coulomb in U_cut)
    for name, coul in (("binary_lj", abi.COUL_CUTOFF), ("water_300", abi.COUL_CUTOFF)):
        s, c = make(name, engine=engine, coulomb=coul)
        e = Engine(s, c)
        e.step(5)
        got = e.thermo_noalloc()
        want = e.thermo()
        _same(got, want)
        # and again after more steps (the array keeps the last call's contents)
        e.step(3)
        _same(e.thermo_noalloc(), e.thermo())
        e.close()


def test_noalloc_thermo_tile_path():
    print("case noalloc thermo tile path", flush=True)
    from test_gpu_tile import lj_cfg
    s = W.binary_lj_large(131072, tiles=2, seed=5)
    e = Engine(s, lj_cfg(abi.ENGINE_MIXED, 0.3))
    e.compute_forces()
    _same(e.thermo_noalloc(), e.thermo())
    e.close()


def test_noalloc_thermo_needs_forces():
    print("case noalloc thermo without forces", flush=True)
    s, c = make("binary_lj", engine=abi.ENGINE_MIXED)
    e = Engine(s, c)
    with pytest.raises(MtmdError) as ex:
        e.thermo_noalloc()
    assert ex.value.status == abi.MTMD_EINVAL
    e.compute_forces()
    _same(e.thermo_noalloc(), e.thermo())
    e.close()

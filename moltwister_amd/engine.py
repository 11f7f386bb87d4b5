"""Python handle over libmtmd.so (the HIP engine), mirroring the reference's
CSimulationBox/CVelVerlet/CMDLoop operations on one system.

Every call goes through the C ABI of include/mtmd.h; a non-zero status raises
MtmdError with the library's message. There is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref

import numpy as np

from . import abi
from .system import System

# handles still open when the interpreter exits are closed by an atexit hook,
# which runs before the interpreter tears its modules down and before the HIP
# runtime's own exit handlers, instead of from __del__ during that teardown
_open_handles: "weakref.WeakSet" = weakref.WeakSet()


def track(obj) -> None:
    """close obj (which has a close() method) at interpreter exit if still open"""
    _open_handles.add(obj)


@atexit.register
def _close_open_handles() -> None:
    for obj in list(_open_handles):
        try:
            obj.close()
        except Exception:
            pass

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)


class MtmdError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"mtmd status {status}: {msg}")
        self.status = status


def _f64(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i32(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _f32(a):
    return None if a is None else a.ctypes.data_as(_fp)


def profile(which: int, kind: int, params, x0: float, x1: float, n: int = 1000):
    """CMDFF*::calc1DForceProfile/calc1DPotentialProfile through mtmd_profile."""
    lib = abi.load_library()
    p = np.zeros(abi.BONDED_PARAMS)          # 4 used by non-bonded kinds
    p[:len(params)] = params
    x, f, e = (np.zeros(n, np.float32) for _ in range(3))
    rc = lib.mtmd_profile(which, kind, _f64(p), x0, x1, n, _f32(x), _f32(f), _f32(e))
    if rc:
        raise MtmdError(rc, "mtmd_profile")
    return x, f, e


def reference_draws(system: System, cfg: abi.MtmdConfig, seed: int = 1) -> dict:
    """The reference's initial random state (CSimulationBox::initSystem,
    SimulationBox.cpp:53-75, with glibc's rand() after srand(seed)) through
    mtmd_reference_draws: momenta [n, 3], the thermostat chain (eta_T, p_eta_T),
    the barostat chain (eta_P, p_eta_P) and p_eps. Seed 1 is the state a first
    `moldyn run` in a fresh MolTwister process starts from."""
    lib = abi.load_library()
    n, mt, mp = system.n_atoms, cfg.temperature_nh_len, cfg.pressure_nh_len
    mom = np.zeros((n, 3))
    et, pt, ep, pp = np.zeros(max(mt, 1)), np.zeros(max(mt, 1)), np.zeros(max(mp, 1)), np.zeros(max(mp, 1))
    pe = C.c_double(0.0)
    csys = system.to_c()
    rc = lib.mtmd_reference_draws(C.byref(csys), C.byref(cfg), int(seed) & 0xFFFFFFFF, _f64(mom), _f64(et),
                                  _f64(pt), _f64(ep), _f64(pp), C.byref(pe))
    if rc:
        raise MtmdError(rc, "mtmd_reference_draws")
    return dict(mom=mom, eta_T=et[:mt], p_eta_T=pt[:mt], eta_P=ep[:mp], p_eta_P=pp[:mp], p_eps=pe.value)


class Engine:
    def __init__(self, system: System, cfg: abi.MtmdConfig):
        self.lib = abi.load_library()
        self.system = system
        self.cfg = cfg
        self.n = system.n_atoms
        self._csys = system.to_c()
        h = C.c_void_p()
        rc = self.lib.mtmd_create(C.byref(self._csys), C.byref(cfg), C.byref(h))
        if rc:
            raise MtmdError(rc, self.lib.mtmd_last_error(None).decode())
        self.h = h
        track(self)

    def close(self):
        if getattr(self, "h", None):
            self.lib.mtmd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc:
            raise MtmdError(rc, self.lib.mtmd_last_error(self.h).decode())

    def set_state(self, pos=None, mom=None):
        pos = None if pos is None else np.ascontiguousarray(pos, np.float64)
        mom = None if mom is None else np.ascontiguousarray(mom, np.float64)
        self._chk(self.lib.mtmd_set_state(self.h, _f64(pos), _f64(mom)))

    def get_state(self):
        pos, mom = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        self._chk(self.lib.mtmd_get_state(self.h, _f64(pos), _f64(mom)))
        return pos, mom

    def set_thermostat(self, eta, p_eta):
        eta, p_eta = np.ascontiguousarray(eta, np.float64), np.ascontiguousarray(p_eta, np.float64)
        self._chk(self.lib.mtmd_set_thermostat(self.h, _f64(eta), _f64(p_eta)))

    def get_thermostat(self):
        M = self.cfg.temperature_nh_len
        eta, p_eta = np.zeros(M), np.zeros(M)
        self._chk(self.lib.mtmd_get_thermostat(self.h, _f64(eta), _f64(p_eta)))
        return eta, p_eta

    def set_barostat(self, eps, p_eps, eta=None, p_eta=None):
        eta = None if eta is None else np.ascontiguousarray(eta, np.float64)
        p_eta = None if p_eta is None else np.ascontiguousarray(p_eta, np.float64)
        self._chk(self.lib.mtmd_set_barostat(self.h, float(eps), float(p_eps), _f64(eta), _f64(p_eta)))

    def get_barostat(self) -> dict:
        M = self.cfg.pressure_nh_len
        eps, peps = C.c_double(0), C.c_double(0)
        eta, p_eta, box = np.zeros(M), np.zeros(M), np.zeros(3)
        self._chk(self.lib.mtmd_get_barostat(self.h, C.byref(eps), C.byref(peps), _f64(eta), _f64(p_eta),
                                             _f64(box)))
        return dict(eps=eps.value, p_eps=peps.value, eta=eta, p_eta=p_eta, box=box)

    def build_cells(self):
        nc = np.zeros(3, np.int32)
        self._chk(self.lib.mtmd_build_cells(self.h, _i32(nc), None, None, None, None))
        tot = int(np.prod(nc))
        ixyz = np.zeros((self.n, 3), np.int32)
        first, count, lst = np.zeros(tot, np.int32), np.zeros(tot, np.int32), np.zeros(self.n, np.int32)
        self._chk(self.lib.mtmd_build_cells(self.h, _i32(nc), _i32(ixyz), _i32(first), _i32(count), _i32(lst)))
        return dict(nc=nc, ixyz=ixyz, first=first, count=count, list=lst)

    def build_neighbors(self):
        req = C.c_int32(0)
        cnt = np.zeros(self.n, np.int32)
        self._chk(self.lib.mtmd_build_neighbors(self.h, 0, _i32(cnt), None, C.byref(req)))
        w = max(int(req.value), 1)
        lst = np.zeros((self.n, w), np.int32)
        self._chk(self.lib.mtmd_build_neighbors(self.h, w, _i32(cnt), _i32(lst), C.byref(req)))
        return cnt, lst

    def compute_forces(self):
        F, U = np.zeros((self.n, 3)), np.zeros(self.n, np.float32)
        self._chk(self.lib.mtmd_compute_forces(self.h, _f64(F), None, _f32(U)))
        return F, U

    def get_forces(self):
        F, U = np.zeros((self.n, 3)), np.zeros(self.n, np.float32)
        self._chk(self.lib.mtmd_get_forces(self.h, _f64(F), _f32(U)))
        return F, U

    def step(self, n: int = 1):
        self._chk(self.lib.mtmd_step(self.h, int(n)))

    def thermo(self) -> dict:
        t = abi.MtmdThermo()
        self._chk(self.lib.mtmd_get_thermo(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in abi.MtmdThermo._fields_}

    def thermo_noalloc(self) -> dict:
        """mtmd_get_thermo_noalloc: the thermo of the last forces, computed
        without any allocation (partials in the library's fixed device array)."""
        t = abi.MtmdThermo()
        self._chk(self.lib.mtmd_get_thermo_noalloc(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in abi.MtmdThermo._fields_}

    def optimize(self, max_steps: int, lr: float, accuracy: float) -> int:
        done = C.c_int32(0)
        self._chk(self.lib.mtmd_optimize(self.h, max_steps, lr, accuracy, C.byref(done)))
        return int(done.value)

    def stats(self) -> dict:
        r, s, m = C.c_int64(0), C.c_int64(0), C.c_int32(0)
        self._chk(self.lib.mtmd_get_stats(self.h, C.byref(r), C.byref(s), C.byref(m)))
        return dict(rebuilds=r.value, steps=s.value, capacity=m.value)

    def skin_lists(self):
        """The FP64/MIXED engine's neighbour lists of its last build (radius
        cutoff + skin), as (count [n], list [n, max count]) in atom indices."""
        req = C.c_int32(0)
        cnt = np.zeros(self.n, np.int32)
        self._chk(self.lib.mtmd_get_skin_lists(self.h, 0, _i32(cnt), None, C.byref(req)))
        w = max(int(req.value), 1)
        lst = np.zeros((self.n, w), np.int32)
        self._chk(self.lib.mtmd_get_skin_lists(self.h, w, _i32(cnt), _i32(lst), C.byref(req)))
        return cnt, lst

    def bench_kernel(self, which: int, reps: int) -> float:
        """Average ms per launch (HIP events on the engine's stream):
        0 pair forces, 1 list build, 2 MD step, 3 streaming copy."""
        ms = C.c_double(0.0)
        self._chk(self.lib.mtmd_bench_kernel(self.h, which, reps, C.byref(ms)))
        return ms.value

    def synchronize(self):
        self._chk(self.lib.mtmd_synchronize(self.h))

    def append_xyz(self, path: str, t: int):
        """CMDLoop::appendToXYZFile: one frame of the current positions."""
        self._chk(self.lib.mtmd_append_xyz(self.h, path.encode(), int(t)))

    def append_dcd(self, path: str, n_steps: int, stride: int):
        """CMDLoop::appendToDCDFile: header when the file is new, then one frame."""
        self._chk(self.lib.mtmd_append_dcd(self.h, path.encode(), int(n_steps), int(stride)))


def find_topology(which: int, pos, types, box, rules, graph=None) -> np.ndarray:
    """mtmd_find_topology: bonds (which 1), angles (2) or dihedrals (3) of the
    rules [(types tuple, R0[, crit])], rows as mtmd_system's lists (atoms, then
    the rule index); box None for the plain distance; graph: one list of bond
    destinations per atom (needed by the molecule, visible-bond and 1-4
    criteria)."""
    lib = abi.load_library()
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    types = np.ascontiguousarray(types, np.int32)
    arr = abi.topo_rules(rules)
    bs, bd = abi.bond_graph(graph, pos.shape[0])
    b = None if box is None else np.ascontiguousarray(box, np.float64)
    n = C.c_int32(0)
    args = (which, pos.shape[0], _f64(pos), _i32(types), _f64(b), _i32(bs), _i32(bd), len(rules), arr)
    rc = lib.mtmd_find_topology(*args, 0, None, C.byref(n))
    if rc:
        raise MtmdError(rc, lib.mtmd_topology_last_error().decode())
    out = np.zeros((max(n.value, 1), which + 2), np.int32)
    rc = lib.mtmd_find_topology(*args, n.value, _i32(out), C.byref(n))
    if rc:
        raise MtmdError(rc, lib.mtmd_topology_last_error().decode())
    return out[:n.value]


def run(system: System, cfg: abi.MtmdConfig, optimize: bool = False) -> None:
    """CMDSimulator::run / optimize through mtmd_run / mtmd_run_optimize: the
    files named by cfg (thermo table, XYZ, DCD, histograms) are written."""
    lib = abi.load_library()
    cs = system.to_c()
    rc = (lib.mtmd_run_optimize if optimize else lib.mtmd_run)(C.byref(cfg), C.byref(cs), None)
    if rc:
        raise MtmdError(rc, lib.mtmd_last_error(None).decode() or "mtmd_run failed")


class BatchEngine:
    """Many independent systems advanced together (mtmd_batch_*): one
    captured step graph for all replicas, or per-replica launches."""

    def __init__(self, systems: list, cfgs: list):
        self.lib = abi.load_library()
        self.systems = systems
        self.n_systems = len(systems)
        self._csys = [s.to_c() for s in systems]
        arr = (abi.MtmdSystem * self.n_systems)(*self._csys)
        carr = (abi.MtmdConfig * self.n_systems)(*cfgs)
        self._keep = (arr, carr)
        h = C.c_void_p()
        rc = self.lib.mtmd_batch_create(self.n_systems, arr, carr, C.byref(h))
        if rc:
            raise MtmdError(rc, self.lib.mtmd_batch_last_error(None).decode())
        self.h = h
        track(self)

    def _chk(self, rc):
        if rc:
            raise MtmdError(rc, self.lib.mtmd_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mtmd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, n: int = 1):
        self._chk(self.lib.mtmd_batch_step(self.h, int(n)))

    def step_separate(self, n: int = 1):
        self._chk(self.lib.mtmd_batch_step_separate(self.h, int(n)))

    def synchronize(self):
        self._chk(self.lib.mtmd_batch_synchronize(self.h))

    def thermo(self) -> list:
        t = (abi.MtmdThermo * self.n_systems)()
        self._chk(self.lib.mtmd_batch_get_thermo(self.h, t))
        return [{k: getattr(x, k) for k, _ in abi.MtmdThermo._fields_} for x in t]

    def get_state(self, i: int):
        n = self.systems[i].n_atoms
        pos, mom = np.zeros((n, 3)), np.zeros((n, 3))
        self._chk(self.lib.mtmd_batch_get_state(self.h, i, _f64(pos), _f64(mom)))
        return pos, mom

    def get_forces(self, i: int):
        n = self.systems[i].n_atoms
        F, U = np.zeros((n, 3)), np.zeros(n, np.float32)
        self._chk(self.lib.mtmd_batch_get_forces(self.h, i, _f64(F), _f32(U)))
        return F, U

    def autotune(self, reps: int = 3) -> dict:
        f, l = C.c_double(0), C.c_double(0)
        self._chk(self.lib.mtmd_batch_autotune(self.h, reps, C.byref(f), C.byref(l)))
        return dict(force_ms=f.value, list_ms=l.value, **self.info(), **self.tuning())

    def tuning(self) -> dict:
        v = [C.c_int32(0) for _ in range(4)]
        self._chk(self.lib.mtmd_batch_tuning(self.h, *[C.byref(x) for x in v]))
        return dict(persistent=bool(v[0].value), block=v[1].value, lanes=v[2].value, cell_split=v[3].value)

    def info(self) -> dict:
        v = [C.c_int32(0) for _ in range(4)]
        self._chk(self.lib.mtmd_batch_info(self.h, *[C.byref(x) for x in v]))
        return dict(n_systems=v[0].value, n_atoms=v[1].value, force_block=v[2].value, list_block=v[3].value)

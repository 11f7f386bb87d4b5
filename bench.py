#!/usr/bin/env python3
"""Benchmark of the MD hot path on one MI355X: atom-steps/s of the MIXED
engine (FP32 pair forces, FP64 accumulation/integration) on the 1,048,576-atom
binary Lennard-Jones mixture, beside the plain FP64 engine on the same state,
the 30,000-atom flexible-water box, and the reference algorithm on the host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-cpu] [--skip-water]

One step = thermostat/half-kick/drift, displacement check and (when needed)
neighbour-list rebuild, forces, force cap, half-kick, PBC wrap: the body of
MDLoop::runSimulation. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import copy
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "atom-steps/s (1M LJ, 30k water), mixed vs FP64 on same GPU at stated drift"
HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
EQ_ROUNDS, EQ_STEPS = 10, 200      # equilibration of the 1M-atom mixture (NVE rounds rescaled to T* = 1)
DRIFT_STEPS, DRIFT_EVERY = 10000, 1000
WATER_EQ_ROUNDS, WATER_EQ_STEPS, WATER_DRIFT_STEPS = 40, 200, 1000
WATER_NVT_EQ_STEPS = 2000
BUCK = (1000.0, 0.1, 1.0)          # Buckingham variant of the mixture: A, rho, C for every type pair


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def lj_config(engine, ensemble):
    from moltwister_amd import abi, config
    c = config.defaults()
    c.engine = engine
    c.ensemble = ensemble
    c.cutoff_radius = 2.5          # sigma
    c.neigh_shell = 0.3
    c.skin = 0.3
    c.timestep = 0.5               # fs = 0.005 tau
    c.temperature = 120.272        # T* = 1
    c.temperature_relax = 20.0     # fs = 0.2 tau
    c.cutoff_force = 1000.0
    c.coulomb = abi.COUL_CUTOFF
    return c


def water_config(engine):
    from moltwister_amd import abi, config
    c = config.defaults()
    c.engine = engine
    c.ensemble = abi.NVT
    c.cutoff_radius = 10.0
    c.neigh_shell = 1.0
    c.skin = 1.0
    c.timestep = 0.5
    c.temperature = 298.0
    c.temperature_relax = 20.0
    c.coulomb = abi.COUL_CUTOFF
    return c


def rms_rel(F, Fr):
    return float(np.sqrt(np.sum((F - Fr) ** 2) / np.sum(Fr ** 2)))


def profiled_traffic():
    """HBM bytes per launch of the MIXED force kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy reference),
    or None when it is absent."""
    import re
    path = os.path.join(ROOT, "profiles", "traffic_force_mixed_binary_lj_1M.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.]+)", open(path).read())
    return float(m.group(1)) if m else None


def profiled_valu():
    """VALU instructions (wave64) per launch of the MIXED force kernel from the
    committed rocprofv3 SQ_INSTS_VALU summary, or None when it is absent."""
    import re
    path = os.path.join(ROOT, "profiles", "valu_force_mixed_binary_lj_1M.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"valu_insts_per_launch:\s*([0-9.]+)", open(path).read())
    return float(m.group(1)) if m else None


def valu_issue(t_force):
    """The vector-ALU side of the force kernel: a wave64 VALU instruction
    occupies a 16-lane SIMD for 4 cycles, so the profiled instruction count
    over 4 SIMDs x 256 CUs at 2.4 GHz is the issue time the kernel cannot go
    below without fewer instructions (transcendentals take longer still)."""
    insts = profiled_valu()
    if insts is None:
        return None
    t_issue = insts * 4.0 / (4 * 256 * 2.4e9)
    return {"insts_per_launch": insts, "issue_ms": t_issue * 1e3, "issue_frac": t_issue / t_force,
            "source": "profiles/valu_force_mixed_binary_lj_1M.md (rocprofv3 --pmc SQ_INSTS_VALU)"}


def energy_drift(th0, th1):
    """|E(t) - E(0)| / |E(0)| of the total energy, and of Etot - U_cut (the
    energy without the jumps of the unshifted truncated pair potential, which
    the integrator conserves)."""
    e0, e1 = th0["Etot"], th1["Etot"]
    s0, s1 = e0 - th0["U_cut"], e1 - th1["U_cut"]
    return abs(e1 - e0) / abs(e0), abs(s1 - s0) / abs(s0)


def host_threads() -> int:
    """Host threads for the all-core CPU baseline: OMP_NUM_THREADS when set
    (the share of the host this process may use), else the CPU count."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    return int(v) if v.isdigit() and int(v) > 0 else (os.cpu_count() or 1)


def cpu_model() -> str:
    """/proc/cpuinfo's model name with its family / model / stepping (a VM may
    report a short name), the host's logical CPUs and this process's share."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                if k in ("model name", "cpu family", "model", "stepping", "cpu MHz") and k not in info:
                    info[k] = v
    except OSError:
        pass
    name = info.get("model name", "unknown")
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return (f"{name} (family {info.get('cpu family', '?')}, model {info.get('model', '?')}, "
            f"stepping {info.get('stepping', '?')}, {info.get('cpu MHz', '?')} MHz; "
            f"{os.cpu_count()} logical CPUs on the host, {usable} usable by this process)")


def timed_steps(eng, k):
    eng.synchronize()
    t0 = time.perf_counter()
    eng.step(k)
    eng.synchronize()
    return time.perf_counter() - t0


def kernels_only(args, warm, sysA) -> int:
    """Average launch times (HIP events) of the dominant kernels after a short
    warm-up: 50 NVE steps from the relaxed packing."""
    from moltwister_amd import abi
    from moltwister_amd.engine import Engine
    warm.step(50)
    sysA.pos, sysA.mom = warm.get_state()
    warm.close()
    out = {"workload": "1,048,576-atom binary LJ after 100 descent + 50 NVE steps"}
    for name, eng_id in (("mixed", abi.ENGINE_MIXED), ("fp64", abi.ENGINE_FP64)):
        c = lj_config(eng_id, abi.NVE)
        c.block_size = args.block
        e = Engine(sysA, c)
        e.step(5)
        out[name] = {"force_ms": e.bench_kernel(0, 5), "build_ms": e.bench_kernel(1, 3),
                     "copy_ms": e.bench_kernel(3, 5)}
        log(f"bench: {name} {out[name]}")
        e.close()
    print(json.dumps(out))
    return 0


def batch_leg(K: int) -> dict:
    """Configuration 4: 512 LJ replicas of uneven size advanced together
    (persistent kernel, autotuned), timed over max(100, K / 4) steps after 200
    warm-up steps, beside one launch per replica per step."""
    from moltwister_amd import abi, workloads as W
    from moltwister_amd.engine import BatchEngine
    t0 = time.perf_counter()
    systems, temps = W.replicas(512, seed=7)
    cfgs = []
    for t in temps:
        cb = lj_config(abi.ENGINE_MIXED, abi.NVT)
        cb.temperature = t * 120.272
        cfgs.append(cb)
    be = BatchEngine(systems, cfgs)
    tune = be.autotune(3)
    t_setup = time.perf_counter() - t0
    be.step(200)                                  # warm-up
    Kb = max(100, K // 4)
    be.synchronize()
    t1 = time.perf_counter()
    be.step(Kb)
    be.synchronize()
    dt_b = time.perf_counter() - t1
    Ks = 4
    t1 = time.perf_counter()
    be.step_separate(Ks)
    be.synchronize()
    dt_s = time.perf_counter() - t1
    th = be.thermo()
    natoms = sum(x.n_atoms for x in systems)
    batch = {"workload": "512 LJ replicas, N uniform in [256, 4096], rho* in [0.6, 1.0], T* in [0.6, 1.5], "
                         "NH NVT, rc 2.5, skin 0.3, dt 0.005 tau",
             "kernel": "persistent: one launch for all steps of a call, a workgroup per replica"
                       if tune.get("persistent") else "captured step graph of separate kernels",
             "atoms": natoms, "steps": Kb,
             "replica_steps_per_s": 512 * Kb / dt_b, "ms_per_step": dt_b / Kb * 1e3,
             "atom_steps_per_s": natoms * Kb / dt_b,
             "separate_launches": {"steps": Ks, "replica_steps_per_s": 512 * Ks / dt_s,
                                   "ms_per_step": dt_s / Ks * 1e3,
                                   "what": "one launch a replica a step (the cost of 512 separate systems)"},
             "speedup_vs_separate": (512 * Kb / dt_b) / (512 * Ks / dt_s),
             "autotune": tune,
             "T_err_rel_mean": float(np.mean([abs(x["T"] / 120.272 - t) / t for x, t in zip(th, temps)])),
             "setup_s": t_setup}
    be.close()
    log(f"bench: batch {batch}")
    return batch


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-water", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-fm", action="store_true")
    ap.add_argument("--fm-frames", type=int, default=2048)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 512-replica batch leg (for profiler runs); prints its JSON")
    ap.add_argument("--kernels-only", action="store_true",
                    help="short warm-up, then only the kernel timings (force pass, list build, copy) of the "
                         "mixed and FP64 engines; for profiler runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; sharding a system over several GPUs is a "
                                   "design only, not implemented", "n_gpus": args.gpus}))
        return 2

    from moltwister_amd import abi, workloads as W
    from moltwister_amd.engine import Engine

    K, Wm = args.steps, args.warmup
    if args.batch_only:
        print(json.dumps(batch_leg(K)))
        return 0
    t_gen = time.perf_counter()
    sysA = W.binary_lj_large(1048576, seed=3, T_star=1.0)
    n = sysA.n_atoms
    log(f"bench: generated {n} atoms in {time.perf_counter() - t_gen:.1f}s, box {sysA.box[0]:.3f}")

    # equilibration, the same whatever --steps/--warmup say: gradient descent
    # on the random packing (moldyn optimizeenergy), then EQ_ROUNDS rounds of
    # EQ_STEPS NVE steps, each followed by a rescale of the momenta to T* = 1
    # with the reference's initial-temperature scaling (SimulationBox.cpp:124-132).
    # (No NVT here: the reference's Nose-Hoover chain, kept operation for
    # operation, resets p_eta[0] > 15 to 10 whatever N is, which leaves a
    # million-atom system almost unthermostatted.)
    cw = lj_config(abi.ENGINE_MIXED, abi.NVE)
    cw.block_size = args.block
    warm = Engine(sysA, cw)
    warm.optimize(100, 2e-5, 0.0)
    if args.kernels_only:
        return kernels_only(args, warm, sysA)
    t_eq = time.perf_counter()
    for rnd in range(EQ_ROUNDS):
        warm.step(EQ_STEPS)
        pos, mom = warm.get_state()
        thw = warm.thermo()
        mom *= math.sqrt(120.272 / thw["T"])
        warm.set_state(mom=mom)
    log(f"bench: equilibrated: {EQ_ROUNDS} x {EQ_STEPS} NVE steps rescaled to T* = 1 "
        f"(last T before rescale {thw['T']:.2f} K) in {time.perf_counter() - t_eq:.1f}s")
    warm.close()
    sysA.pos, sysA.mom = pos, mom
    pos_eq, mom_eq = pos.copy(), mom.copy()

    # headline: MIXED engine, NVE (configuration 3): W untimed steps, K timed
    cm = lj_config(abi.ENGINE_MIXED, abi.NVE)
    cm.block_size = args.block
    em = Engine(sysA, cm)
    em.step(max(1, Wm))                             # first list build, sort, graph capture
    st0 = em.stats()
    dt_m = timed_steps(em, K)
    st1 = em.stats()
    value = n * K / dt_m
    rebuild_every = K / max(1, st1["rebuilds"] - st0["rebuilds"])
    log(f"bench: mixed {dt_m / K * 1e3:.3f} ms/step, {value:.4g} atom-steps/s, rebuild every {rebuild_every:.1f}")

    # dominant kernel: the pair-force pass, timed with HIP events on the engine's stream
    t_force = em.bench_kernel(0, 50) * 1e-3
    t_copy = em.bench_kernel(3, 50) * 1e-3
    t_build = em.bench_kernel(1, 10) * 1e-3
    rho = n / (sysA.box[0] ** 3)
    r_list = cm.cutoff_radius + cm.skin
    n_full = rho * 4.0 / 3.0 * math.pi * r_list ** 3
    # the MIXED tile path's lists hold 16-bit words (the neighbour's place in
    # its block's staged cells)
    b_force = 16 + 4 + 2 * n_full + 24                  # own position, count, list words, force
    achieved = n * b_force / t_force
    b_step = 184 + 2 * n_full
    b_rb = 32 + 2 * n_full
    step_bytes = b_step + b_rb / rebuild_every
    cap = em.stats()["capacity"]
    copy_bw = 2 * n * 4 * cap / t_copy                  # in-place pass over the list buffer

    # FP64 engine on the same state and the force error of MIXED against it
    pos_m, mom_m = em.get_state()
    F_mixed, U_mixed = em.compute_forces()
    # the same tiled force pass in FP64 arithmetic (the same lists, staging and
    # table; FP64 positions, square root and sums): what precision alone costs
    try:
        t_tile64 = em.bench_kernel(4, 20) * 1e-3
        F_tile64, _ = em.get_forces()
    except Exception as ex:                          # tile mode off, or 32-bit blocks
        log(f"bench: no FP64 tiled pass ({ex})")
        t_tile64, F_tile64 = None, None
    em.close()
    c64 = lj_config(abi.ENGINE_FP64, abi.NVE)
    c64.block_size = args.block
    sysA.pos, sysA.mom = pos_m, mom_m
    e64 = Engine(sysA, c64)
    F64, _ = e64.compute_forces()
    force_rms = rms_rel(F_mixed, F64)
    e64.step(10)
    K64 = max(50, K // 5)
    dt_64 = timed_steps(e64, K64)
    value64 = n * K64 / dt_64
    t_force64 = e64.bench_kernel(0, 10) * 1e-3
    e64.close()
    log(f"bench: fp64 {dt_64 / K64 * 1e3:.3f} ms/step, {value64:.4g} atom-steps/s; force rms rel {force_rms:.2e}")

    # the mixture's Buckingham variant (A = 1000, rho = 0.1, C = 1 for every
    # type pair) at the same positions: the pair passes read tables, so only
    # the table contents change. Its pair potential has no repulsive wall
    # (A e^(-r/rho) / rho stays below 6 C / r^7 at every r), so it is measured
    # as force passes on this state, not as dynamics.
    sysB = copy.copy(sysA)
    sysB.pos, sysB.mom = pos_m, mom_m
    sysB.nonbonded = [(0, 0, abi.NB_BUCK, BUCK), (0, 1, abi.NB_BUCK, BUCK), (1, 1, abi.NB_BUCK, BUCK)]
    buck, F_b = {}, {}
    for eng_id, name in ((abi.ENGINE_MIXED, "mixed"), (abi.ENGINE_FP64, "fp64")):
        eb = Engine(sysB, lj_config(eng_id, abi.NVE))
        F_b[name], _ = eb.compute_forces()
        buck[f"{name}_force_pass_ms"] = eb.bench_kernel(0, 20)
        eb.close()
    buck["force_rms_rel_mixed_vs_fp64"] = rms_rel(F_b["mixed"], F_b["fp64"])
    buck["mixed_vs_fp64_force_pass"] = buck["fp64_force_pass_ms"] / buck["mixed_force_pass_ms"]
    buck["parameters"] = {"A": BUCK[0], "rho": BUCK[1], "C": BUCK[2]}
    log(f"bench: buckingham variant {buck}")

    # NVE drift: DRIFT_STEPS steps of each engine from the same equilibrated
    # state, the energies read every DRIFT_EVERY steps
    drift = {}
    for name, eng_id in (("mixed", abi.ENGINE_MIXED), ("fp64", abi.ENGINE_FP64)):
        sysA.pos, sysA.mom = pos_eq, mom_eq
        cd = lj_config(eng_id, abi.NVE)
        cd.block_size = args.block
        ed = Engine(sysA, cd)
        th0 = ed.thermo()
        worst, worst_s = 0.0, 0.0
        t0 = time.perf_counter()
        for blk in range(DRIFT_STEPS // DRIFT_EVERY):
            ed.step(DRIFT_EVERY)
            d_raw, d_shift = energy_drift(th0, ed.thermo())
            worst, worst_s = max(worst, d_raw), max(worst_s, d_shift)
        th1 = ed.thermo()
        d_raw, d_shift = energy_drift(th0, th1)
        drift[name] = {"steps": DRIFT_STEPS, "drift_rel": d_raw, "drift_shifted_rel": d_shift,
                       "max_over_trace_rel": worst, "max_over_trace_shifted_rel": worst_s,
                       "T_start_K": th0["T"], "T_end_K": th1["T"], "seconds": time.perf_counter() - t0}
        ed.close()
        log(f"bench: {name} NVE drift over {DRIFT_STEPS} steps: {d_raw:.2e} (shifted {d_shift:.2e}), "
            f"T {th0['T']:.1f} -> {th1['T']:.1f} K")

    water = {}
    if not args.skip_water:
        from moltwister_amd.engine import MtmdError
        sysW = W.water_box(10000, seed=2)
        # relax the generated packing first (moldyn optimizeenergy), as a run would
        cwr = water_config(abi.ENGINE_MIXED)
        cwr.coulomb = abi.COUL_ALLPAIRS
        relax = Engine(sysW, cwr)
        relax.optimize(300, 1e-4, 0.0)
        pos_w0, _ = relax.get_state()
        relax.close()
        for coul, mode in ((abi.COUL_ALLPAIRS, "allpairs"), (abi.COUL_CUTOFF, "cutoff10")):
            # thermalise at 298 K in this Coulomb mode: NVE rounds on the MIXED
            # engine, each followed by a rescale of the momenta to 298 K
            sysW.pos = pos_w0
            cth = water_config(abi.ENGINE_MIXED)
            cth.coulomb = coul
            cth.ensemble = abi.NVE
            eth = Engine(sysW, cth)
            for _ in range(WATER_EQ_ROUNDS):
                eth.step(WATER_EQ_STEPS)
                pw, mw = eth.get_state()
                mw *= math.sqrt(298.0 / eth.thermo()["T"])
                eth.set_state(mom=mw)
            eth.close()
            sysW.pos, sysW.mom = pw, mw
            # then the timed run's own Nose-Hoover NVT at 298 K until the
            # temperature has settled; its chain state goes on into the timed runs
            ceq = water_config(abi.ENGINE_MIXED)
            ceq.coulomb = coul
            eeq = Engine(sysW, ceq)
            t_eq = []
            for _ in range(WATER_NVT_EQ_STEPS // 200):
                eeq.step(200)
                t_eq.append(eeq.thermo()["T"])
            sysW.pos, sysW.mom = eeq.get_state()
            nh_w = eeq.get_thermostat()
            eeq.close()
            # the reference chain resets p_eta[0] > 15 to 10 at every call
            # (NHChain.cpp:61-65), which caps its friction at about 15 / Q0 with
            # Q0 = 3 N T tau^2: the chain's first momentum shows when it runs there
            water[f"thermostat_{mode}"] = {"p_eta0_after_equilibration": float(nh_w[1][0]),
                                           "Q0": 3 * sysW.n_atoms * (298.0 / 120.272) * (20.0 / 100.0) ** 2}
            F_w = {}
            for eng_id, ename in ((abi.ENGINE_MIXED, "mixed"), (abi.ENGINE_FP64, "fp64")):
                key = f"{ename}_{mode}"
                cwt = water_config(eng_id)
                cwt.coulomb = coul
                cwt.block_size = args.block
                try:
                    ew = Engine(sysW, cwt)
                    ew.set_thermostat(*nh_w)
                    F_w[ename] = ew.compute_forces()[0]
                    T0 = ew.thermo()["T"]
                    ew.step(20)
                    kw = 200 if eng_id == abi.ENGINE_MIXED else 100
                    dtw = timed_steps(ew, kw)
                    T1 = ew.thermo()["T"]
                    # the temperature over the 200 NVT steps that follow (untimed)
                    t_after = []
                    for _ in range(10):
                        ew.step(20)
                        t_after.append(ew.thermo()["T"])
                    t_mean = float(np.mean(t_after))
                    water[key] = {"atom_steps_per_s": sysW.n_atoms * kw / dtw, "ms_per_step": dtw / kw * 1e3,
                                  "nvt_steps": kw, "T_start_K": T0, "T_end_K": T1,
                                  "T_mean_next_200_K": t_mean, "T_within_10K_of_298": bool(abs(t_mean - 298.0) <= 10.0
                                                                                          and abs(T1 - 298.0) <= 10.0),
                                  "nvt_equilibration_T_K": [round(x, 1) for x in t_eq]}
                    ew.close()
                    # NVE drift from the same thermalised state
                    cwe = water_config(eng_id)
                    cwe.coulomb = coul
                    cwe.ensemble = abi.NVE
                    cwe.block_size = args.block
                    ee = Engine(sysW, cwe)
                    th0 = ee.thermo()
                    ee.step(WATER_DRIFT_STEPS)
                    d_raw, d_shift = energy_drift(th0, ee.thermo())
                    water[key].update({"nve_drift_steps": WATER_DRIFT_STEPS, "nve_drift_rel": d_raw,
                                       "nve_drift_shifted_rel": d_shift})
                    ee.close()
                except MtmdError as ex:
                    water[key] = {"error": str(ex)}
            if len(F_w) == 2:
                water[f"force_rms_rel_mixed_vs_fp64_{mode}"] = rms_rel(F_w["mixed"], F_w["fp64"])
        water["workload"] = ("10,000 flexible 3-site waters (30,000 atoms), 1 g/cm^3, L = 66.88 A; 300 descent "
                             f"steps, then per Coulomb mode {WATER_EQ_ROUNDS} x {WATER_EQ_STEPS} NVE steps each "
                             f"rescaled to 298 K and {WATER_NVT_EQ_STEPS} Nose-Hoover NVT steps at 298 K; timed "
                             "Nose-Hoover NVT at 298 K (tau 20 fs, M 4, n 4), dt 0.5 fs, from that state and chain; "
                             f"NVE drift over {WATER_DRIFT_STEPS} steps from the same state")
        log(f"bench: water {water}")

    # configuration 1: the reference's CPU-runnable case on the GPU engines
    cfg1 = {}
    sys1 = W.fcc_lj(10)
    for name, eng_id in (("exact", abi.ENGINE_EXACT), ("fp64", abi.ENGINE_FP64), ("mixed", abi.ENGINE_MIXED)):
        c1 = lj_config(eng_id, abi.NVE)
        c1.temperature = 86.596
        e1 = Engine(sys1, c1)
        e1.step(10)
        th_a = e1.thermo()
        k1 = 1000
        dt1 = timed_steps(e1, k1)
        th_b = e1.thermo()
        cfg1[name] = {"steps_per_s": k1 / dt1, "atom_steps_per_s": sys1.n_atoms * k1 / dt1,
                      "nve_drift_shifted_rel": energy_drift(th_a, th_b)[1]}
        e1.close()
    cfg1["workload"] = "FCC LJ, 4000 atoms, rho*=0.8442, T*=0.72, rc 2.5, NVE, dt 0.005 tau, 1000 timed steps"
    log(f"bench: configuration 1 {cfg1}")

    batch = {} if args.skip_batch else batch_leg(K)


    fmres = {}
    if not args.skip_fm:
        from moltwister_amd.forcematch import ForceMatch, params_of, fit
        t0 = time.perf_counter()
        n_frames = args.fm_frames
        sysF, frames = W.water_frames(1000, n_frames, 100, seed=4,
                                      progress=lambda f: log(f"bench: fm frames {f}/{n_frames} "
                                                             f"({time.perf_counter() - t0:.1f}s)"))
        t_gen = time.perf_counter() - t0
        fmh = ForceMatch(sysF, 10.0, frames)
        p0 = params_of(sysF)
        truth = p0 * np.random.default_rng(11).uniform(0.9, 1.1, p0.shape)
        fmh.target_from_params(truth)
        fmh.loss_grad(p0)                                   # warm
        reps = 5
        t1 = time.perf_counter()
        for _ in range(reps):
            L32, g32 = fmh.loss_grad(p0, fp64=False)
        t_32 = (time.perf_counter() - t1) / reps
        t1 = time.perf_counter()
        L64, g64 = fmh.loss_grad(p0, fp64=True)
        t_64 = time.perf_counter() - t1
        # central differences of the FP64 loss at h = 1e-6 relative (the stated
        # check) and at h = 1e-5: their difference is the differences' own
        # rounding error, which at 1e-6 is of the order of the bar itself
        def central(k, rel):
            h = rel * abs(p0[k])
            pp, pm = p0.copy(), p0.copy()
            pp[k] += h
            pm[k] -= h
            return (fmh.loss_grad(pp, fp64=True)[0] - fmh.loss_grad(pm, fp64=True)[0]) / (2 * h)
        fd_err = fd_err5 = fd_noise = 0.0
        for k in range(len(p0)):
            fd6, fd5 = central(k, 1e-6), central(k, 1e-5)
            fd_err = max(fd_err, abs(g64[k] - fd6) / abs(fd6))
            fd_err5 = max(fd_err5, abs(g64[k] - fd5) / abs(fd5))
            fd_noise = max(fd_noise, abs(fd6 - fd5) / abs(fd5))
        t1 = time.perf_counter()
        theta, hist = fit(fmh, p0, iters=500, lr=2e-2)
        t_fit = time.perf_counter() - t1
        fmres = {"workload": f"1000 waters (3000 atoms), {n_frames} frames every 100 steps of 298 K NVT, "
                             "targets from parameters scaled by U[0.9, 1.1]",
                 "frames": n_frames, "frame_gen_s": t_gen,
                 "loss_grad_ms_fp32": t_32 * 1e3, "loss_grad_ms_fp64": t_64 * 1e3,
                 "frames_per_s_fp32": n_frames / t_32,
                 "grad_rel_err_fp32_vs_fp64": float(np.max(np.abs(g32 - g64) / np.abs(g64))),
                 "grad_rel_err_fp64_vs_fd": fd_err,
                 "grad_rel_err_fp64_vs_fd_h1e-5": fd_err5,
                 "fd_h1e-6_vs_h1e-5": fd_noise,
                 "adam_500": {"loss_start": hist[0], "loss_end": hist[-1], "seconds": t_fit,
                              "param_rel_err": [float(x) for x in np.abs(theta / truth - 1)]}}
        fmh.close()
        log(f"bench: force matching {fmres}")

    cpu = None
    oracle_check = None
    if not args.skip_cpu:
        from oracle import pyoracle
        c_ref = lj_config(abi.ENGINE_EXACT, abi.NVE)
        threads = host_threads()
        # the MIXED and FP64 forces of the bench state (the timed run's end)
        # against the oracle's, at full size
        sysA.pos, sysA.mom = pos_m, mom_m
        o1m = pyoracle.Oracle(sysA, c_ref)
        o1m.set_threads(threads)
        t0 = time.perf_counter()
        F_orc, U_orc = o1m.compute_forces()
        t_orc = time.perf_counter() - t0
        del o1m
        scale = np.maximum(np.linalg.norm(F_orc, axis=1), np.sqrt(np.mean(np.sum(F_orc ** 2, 1))))
        oracle_check = {"mixed_force_rms_rel": rms_rel(F_mixed, F_orc),
                        "mixed_force_max_rel_per_atom": float(np.max(np.linalg.norm(F_mixed - F_orc, axis=1) / scale)),
                        "fp64_force_max_rel_per_atom": float(np.max(np.linalg.norm(F64 - F_orc, axis=1) / scale)),
                        "mixed_U_rel": abs(float(U_mixed.sum(dtype=np.float64)) - float(U_orc.sum(dtype=np.float64)))
                        / float(np.abs(U_orc).sum(dtype=np.float64)),
                        "oracle_seconds": t_orc, "oracle_threads": threads}
        log(f"bench: 1M state vs oracle {oracle_check}")
        # bounded sample: the 1/8 sub-box (47.8 sigma, 131,072 atoms at the same
        # density) of the timed run's end state, so the per-atom work is that of the workload
        Lh = sysA.box[0] / 2.0
        sel = np.all((pos_m >= -Lh) & (pos_m < 0.0), axis=1)
        from moltwister_amd.system import System
        sub = System(pos=pos_m[sel] + Lh / 2.0, mom=mom_m[sel], mass=sysA.mass[sel], charge=sysA.charge[sel],
                     types=sysA.types[sel], box=(Lh, Lh, Lh), nonbonded=sysA.nonbonded)
        legs = {}
        for nth, nst in ((1, 2), (threads, 8)):
            o = pyoracle.Oracle(sub, c_ref)
            o.set_threads(nth)
            o.compute_forces()
            t0 = time.perf_counter()
            o.step(nst)
            legs[nth] = (sub.n_atoms * nst / (time.perf_counter() - t0), nst)
            del o
        cpu = {"value": legs[1][0], "unit": "atom-steps/s", "cores": 1, "kind": "port", "cpu_model": cpu_model(),
               "sample": f"{legs[1][1]} velocity-Verlet steps of a {sub.n_atoms}-atom sub-box (1/8 of the timed "
                         f"run's 1,048,576-atom end state, same density) with the oracle: the reference algorithm, "
                         f"single thread, lists rebuilt at rc every step, its all-pairs Coulomb pass skipped "
                         f"because every charge is zero; the first force call not counted",
               "all_cores": {"value": legs[threads][0], "unit": "atom-steps/s", "cores": threads, "kind": "port",
                             "sample": f"{legs[threads][1]} steps of the same sub-box, the oracle's per-atom loops "
                                       f"(neighbour rows, forces) on {threads} OpenMP threads, same results"}}
        log(f"bench: cpu oracle {cpu['value']:.4g} atom-steps/s on 1 core, {legs[threads][0]:.4g} on {threads}")
        # MolTwister's own force path (oracle/_ref, built from its sources) on
        # configuration 1, where it runs in seconds: cell list, neighbour list
        # and the force functor including its all-pairs Coulomb loop
        if os.path.exists(pyoracle.REF_LIB):
            c_ref1 = lj_config(abi.ENGINE_EXACT, abi.NVE)
            r1 = pyoracle.Reference(sys1, c_ref1)
            r1.compute_forces()
            t0 = time.perf_counter()
            nrep = 5
            for _ in range(nrep):
                r1.compute_forces()
            t_r1 = (time.perf_counter() - t0) / nrep
            cpu["reference_config1"] = {"value": sys1.n_atoms / t_r1, "unit": "atom-force-evaluations/s",
                                        "cores": 1, "kind": "reference",
                                        "sample": f"{nrep} force calls of MolTwister's CPU functors on the 4000-atom "
                                                  "FCC system (cells, neighbour list, functor with its all-pairs "
                                                  "Coulomb loop); the O(N) integrator is not included"}
            log(f"bench: reference functor {t_r1 * 1e3:.1f} ms per force call on {sys1.n_atoms} atoms")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": dt_m / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "binary LJ 80:20, 1,048,576 atoms, rho*=1.2, rc=2.5 sigma, skin 0.3, NVE, dt=0.005 tau; "
                        "start: a 131,072-atom RSA box (min sep 0.8 sigma) tiled 2x2x2 with per-atom types and "
                        "momenta, 100 descent steps, 10 x 200 NVE steps each rescaled to T*=1, then W untimed steps",
            "engine": "mixed: f32 pair math, f64 accumulation/integration",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": profiled_traffic(),
            "traffic_source": "profiles/traffic_force_mixed_binary_lj_1M.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; "
                              "reads scaled x2.0, writes x1.0 by the in-place list-buffer copy)",
            "kernel": "k_fast_force_tile",
            "kernel_ms": t_force * 1e3,
            "bytes_per_atom": b_force,
            "valu": valu_issue(t_force),
            "limiter": "VALU issue and LDS reads of the staged positions (see valu); the fraction above is of "
                       "the HBM peak, the roofline the kernel's algorithmic bytes are quoted against",
        },
        "cpu_baseline": cpu,
        "fp64": {"value": value64, "ms_per_step": dt_64 / K64 * 1e3, "steps": K64,
                 "force_kernel_ms": t_force64 * 1e3, "mixed_speedup": value / value64,
                 "tiled_fp64_force_pass": None if t_tile64 is None else {
                     "kernel_ms": t_tile64 * 1e3, "vs_mixed_force_pass": t_tile64 / t_force,
                     "vs_fp64_engine_force_pass": t_tile64 / t_force64,
                     "force_rms_rel_vs_fp64_engine": rms_rel(F_tile64, F64),
                     "what": "k_fast_force_tile's force-only pass with FP64 arithmetic (block-frame FP64 positions, "
                             "FP64 rsqrt and sums) on the same 16-bit lists and table: the MIXED/FP64 force-pass "
                             "ratio at equal design"}},
        "precision": {"force_rms_rel_vs_fp64": force_rms,
                      "vs_oracle": oracle_check,
                      "nve_drift_steps": DRIFT_STEPS,
                      "mixed": drift["mixed"], "fp64": drift["fp64"],
                      "note": "drift = |E(t)-E(0)|/|E(0)| after DRIFT_STEPS NVE steps of each engine from the same "
                              "equilibrated state (max_over_trace: the largest value at the 1000-step checkpoints); "
                              "'shifted' uses Etot - U_cut, which has no jumps from pairs crossing the unshifted "
                              "cutoff"},
        "buckingham_variant": buck,
        "step_roofline": {"bytes_per_atom_step": step_bytes, "frac": value * step_bytes / HBM_PEAK,
                          "rebuild_every_steps": rebuild_every, "list_build_ms": t_build * 1e3,
                          "copy_GBps": copy_bw / 1e9},
        "configuration_1": cfg1,
        "water_30k": water,
        "replicas_512": batch,
        "force_matching": fmres,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

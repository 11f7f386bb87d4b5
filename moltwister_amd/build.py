"""Build libmtmd.so in-tree for gfx950 (``python -m moltwister_amd.build``).

Translation units:
  mtmd_exact.hip   reference-exact kernels + integrator/thermostat/reductions,
                   -ffp-contract=off
  mtmd_fast.hip    FP64/MIXED engine kernels (contraction allowed; the exact
                   device helpers disable it locally)
  mtmd_host.cpp    C ABI, set-up and step driver (hipcc, host code)
  mtmd_draws.cpp   the reference's initial random draws (hipcc, host code)
  mtmd_topo.hip    topology search (bonds, angles, dihedrals by distance
                   criteria), -ffp-contract=off
  mtmd_tables.cpp  force/energy profiles, g++ -ffp-contract=off so the float
                   tables are the ones a default x86-64 build of the reference
                   computes
  mtmd_output.cpp  XYZ / DCD trajectory and histogram files of mtmd_run (g++)
  mtmd_gen.c       seeded workload generators (libmtgen.so, host only)
Every HIP compile line carries ``-Xclang -target-feature -Xclang -packed-fp32-ops``
(no packed FP32 instructions).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmtmd.so")
GENLIB = os.path.join(HERE, "libmtgen.so")
ARCH = os.environ.get("MTMD_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
NO_PACKED = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-munsafe-fp-atomics", *NO_PACKED,
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {os.path.basename(cmd[-1])}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    jobs = [
        [HIPCC, *HIP_FLAGS, "-ffp-contract=off", "-c", "-o", os.path.join(OUT, "mtmd_exact.o"),
         os.path.join(SRC, "mtmd_exact.hip")],
        [HIPCC, *HIP_FLAGS, "-c", "-o", os.path.join(OUT, "mtmd_fast.o"), os.path.join(SRC, "mtmd_fast.hip")],
        [HIPCC, *HIP_FLAGS, "-c", "-o", os.path.join(OUT, "mtmd_batch.o"), os.path.join(SRC, "mtmd_batch.hip")],
        [HIPCC, *HIP_FLAGS, "-c", "-o", os.path.join(OUT, "mtmd_fm.o"), os.path.join(SRC, "mtmd_fm.hip")],
        [HIPCC, *HIP_FLAGS, "-ffp-contract=off", "-c", "-o", os.path.join(OUT, "mtmd_topo.o"),
         os.path.join(SRC, "mtmd_topo.hip")],
        [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", "-o", os.path.join(OUT, "mtmd_host.o"),
         os.path.join(SRC, "mtmd_host.cpp")],
        [HIPCC, *HIP_FLAGS, "-ffp-contract=off", "-x", "hip", "-c", "-o", os.path.join(OUT, "mtmd_draws.o"),
         os.path.join(SRC, "mtmd_draws.cpp")],
        [shutil.which("g++") or "g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-c",
         "-o", os.path.join(OUT, "mtmd_tables.o"), os.path.join(SRC, "mtmd_tables.cpp")],
        [shutil.which("g++") or "g++", "-O2", "-fPIC", "-std=c++17", "-Wall", "-c",
         "-o", os.path.join(OUT, "mtmd_output.o"), os.path.join(SRC, "mtmd_output.cpp")],
    ]
    with ThreadPoolExecutor(max_workers=8) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    objs = [os.path.join(OUT, f) for f in ("mtmd_exact.o", "mtmd_fast.o", "mtmd_batch.o", "mtmd_fm.o", "mtmd_topo.o",
                                           "mtmd_host.o", "mtmd_draws.o", "mtmd_tables.o", "mtmd_output.o")]
    # workload generator helpers (host only)
    _run([shutil.which("gcc") or "gcc", "-O2", "-fPIC", "-shared", "-o", GENLIB, os.path.join(SRC, "mtmd_gen.c"), "-lm"])
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

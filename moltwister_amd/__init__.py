"""MI355X-native engine for MolTwister's `moldyn` hot path.

The engine is libmtmd.so (C ABI in include/mtmd.h, HIP kernels for gfx950);
this package is its host-side harness: the `moldyn cfg` configuration mirror,
the system description handed across the C ABI, seeded workload generators and
an Engine wrapper.
"""
from . import abi, config, system, workloads  # noqa: F401

__all__ = ["abi", "config", "system", "workloads"]

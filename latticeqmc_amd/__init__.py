"""latticeqmc_amd: determinant-QMC sweep for the Hubbard model on AMD MI355X (gfx950).

Keeps the reference's call surface (lqmc/__init__.py:9-96): ``HubbardModel``,
``Lattice``, ``Configuration``, ``LatticeQMC``, the tools, ``ParallelProcessManager``,
``SerialProcessManager``, ``measure`` and ``measure_betas``, with the submodules under the
reference's names (``hubbard``, ``lattice``, ``configuration``, ``lqmc``, ``tools``,
``multiprocessing``), so ``import latticeqmc_amd as lqmc`` serves the reference's callers.
The sweep itself runs in hand-written HIP kernels behind the C ABI of include/lqmc_hip.h.
"""
from .lattice import Lattice
from .configuration import Configuration
from .hubbard import HubbardModel
from .tools import (check_params, compute_pole_gf_tau, decompose, fermi_fct, filling, local_gf, local_moment,
                    matsubara_frequencies, pole_gf_tau, reconstruct, binned_mean, jackknife)
from .engine import ChainSpec, SweepEngine, slice_params, wrap_green
from .lqmc import LatticeQMC
from .batched import BatchedDQMC, get_datadir, get_datapath, measure, measure_betas, measure_scan
from .multiprocessing import ParallelProcessManager, SerialProcessManager

__all__ = [
    "Lattice", "Configuration", "HubbardModel", "LatticeQMC", "ChainSpec", "SweepEngine", "wrap_green",
    "BatchedDQMC", "ParallelProcessManager", "SerialProcessManager", "measure", "measure_betas", "measure_scan",
    "get_datadir", "get_datapath", "check_params", "filling", "local_moment", "local_gf", "matsubara_frequencies",
    "fermi_fct", "decompose", "reconstruct", "pole_gf_tau", "compute_pole_gf_tau", "binned_mean", "jackknife",
    "slice_params",
]

"""latticeqmc_amd: determinant-QMC sweep for the Hubbard model on AMD MI355X (gfx950).

Keeps the reference's call surface (lqmc/__init__.py): ``HubbardModel``,
``Lattice``, ``Configuration``, ``LatticeQMC``, ``measure``, ``measure_betas``
and the observable helpers, with the sweep itself running in hand-written HIP
kernels behind the C ABI of include/lqmc_hip.h.
"""
from .lattice import Lattice
from .configuration import Configuration
from .hubbard import HubbardModel
from .tools import check_params, filling, local_moment, local_gf, binned_mean
from .engine import ChainSpec, SweepEngine, slice_params
from .lqmc import LatticeQMC
from .batched import BatchedDQMC, get_datadir, get_datapath, measure, measure_betas, measure_scan

__all__ = [
    "Lattice", "Configuration", "HubbardModel", "LatticeQMC", "ChainSpec", "SweepEngine",
    "BatchedDQMC", "measure", "measure_betas", "measure_scan", "get_datadir", "get_datapath", "check_params", "filling", "local_moment",
    "local_gf", "binned_mean", "slice_params",
]

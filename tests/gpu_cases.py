"""Shared builders for the GPU parity tests (oracle inputs -> engine chains)."""
import numpy as np

from latticeqmc_amd.engine import ChainSpec, SweepEngine
from oracle import dqmc, model


def spec_from(k, beta, L, u, field, seed=0):
    _, lam, ek, eki = model.params(k, beta, L, u)
    return ChainSpec(k, beta, L, u, field, seed=seed, expk=ek, expk_inv=eki, lam=lam)


def square_case(side, L, beta, u, seed):
    k = model.square_k(side, side)
    f = model.random_field(side * side, L, seed)
    return k, f


def masked_equal(acc_gpu, acc_ref, ratio, uni, tol=1e-12):
    tie = np.abs(ratio - uni) < tol
    return bool(np.array_equal(acc_gpu[~tie], acc_ref[~tie])), int(tie.sum())

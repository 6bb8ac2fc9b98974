"""ctypes binding of liblqmc_hip.so (the C ABI in include/lqmc_hip.h).

The library is built in-tree by ``python -m latticeqmc_amd.build``.  There is
no CPU fallback: importing the binding without the library raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblqmc_hip.so")

MODE_DQMC, MODE_REFERENCE, MODE_DET = 0, 1, 2
FP64, FP32, BF16X3 = 0, 1, 2
RNG_REPLAY, RNG_PHILOX = 0, 1

ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "ESTATE", -5: "ENUMERIC"}
FLAG_NAMES = {1: "NaN/Inf ratio", 2: "zero pivot", 4: "replay uniforms exhausted", 8: "NaN/Inf in G",
              16: "flip barrier timeout", 32: "recompute deviation above max_dg"}

EXPORTS = (
    "lqmc_create", "lqmc_set_chain", "lqmc_init", "lqmc_get_ek_split", "lqmc_set_uniforms", "lqmc_sweep",
    "lqmc_sync", "lqmc_use_graph", "lqmc_get_green", "lqmc_get_green_dev", "lqmc_get_recompute_err", "lqmc_get_spin_corr", "lqmc_get_green_tau",
    "lqmc_get_field",
    "lqmc_get_stats", "lqmc_reset_stats", "lqmc_get_trace", "lqmc_get_observables",
    "lqmc_profile", "lqmc_profile_result", "lqmc_wrap_green_workspace", "lqmc_wrap_green",
    "lqmc_wrap_green_static", "lqmc_last_error", "lqmc_destroy",
)


class Options(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("precision", ctypes.c_int32), ("delay", ctypes.c_int32),
                ("n_stab", ctypes.c_int32), ("rng", ctypes.c_int32), ("trace", ctypes.c_int32),
                ("obs_capacity", ctypes.c_int32), ("flip_threads", ctypes.c_int32),
                ("flip_groups", ctypes.c_int32), ("time_displaced", ctypes.c_int32),
                ("max_dg", ctypes.c_double), ("dense_ek", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("attempted", ctypes.c_uint64), ("accepted", ctypes.c_uint64),
                ("sweeps", ctypes.c_uint64), ("measured", ctypes.c_uint64),
                ("recomputes", ctypes.c_uint64), ("max_err", ctypes.c_double),
                ("flags", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class LqmcError(RuntimeError):
    pass


_lib = None


def load(path=None):
    """Load the shared library (once) and declare the prototypes."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m latticeqmc_amd.build` "
                          "(the GPU sweep has no CPU fallback)")
    # One HIP runtime per process.  PyTorch-ROCm ships its own libamdhip64; once torch is
    # imported, the library's HIP symbols bind to that copy, so torch and the engine share one
    # runtime (device pointers and stream handles pass between them).  Loaded before torch, the
    # library would bind to the system runtime instead, and whichever of the two runtimes
    # initialises the GPU second finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    P, I, D, V = ctypes.POINTER, ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    i32p, f64p = P(ctypes.c_int32), P(ctypes.c_double)
    sig = {
        "lqmc_create": (I, [P(V), I, I, i32p, i32p, P(Options)]),
        "lqmc_set_chain": (I, [V, I, f64p, f64p, f64p, D, P(ctypes.c_int8), ctypes.c_uint64]),
        "lqmc_init": (I, [V]),
        "lqmc_get_ek_split": (I, [V, I, i32p, i32p]),
        "lqmc_set_uniforms": (I, [V, I, f64p, ctypes.c_int64]),
        "lqmc_sweep": (I, [V, I, I, V]),
        "lqmc_sync": (I, [V]),
        "lqmc_use_graph": (I, [V, I]),
        "lqmc_get_green": (I, [V, I, f64p, I]),
        "lqmc_get_green_dev": (I, [V, I, V, I, V]),
        "lqmc_get_recompute_err": (I, [V, I, f64p, i32p]),
        "lqmc_get_spin_corr": (I, [V, I, P(ctypes.c_double)]),
        "lqmc_get_green_tau": (I, [V, I, f64p, i32p, i32p]),
        "lqmc_get_field": (I, [V, I, P(ctypes.c_int8)]),
        "lqmc_get_stats": (I, [V, I, P(Stats)]),
        "lqmc_reset_stats": (I, [V]),
        "lqmc_get_trace": (I, [V, I, f64p, P(ctypes.c_uint8)]),
        "lqmc_get_observables": (I, [V, I, f64p, I, P(I)]),
        "lqmc_profile": (I, [V, I]),
        "lqmc_profile_result": (I, [V, f64p, P(ctypes.c_int64), f64p]),
        "lqmc_wrap_green_workspace": (ctypes.c_size_t, [I, I]),
        "lqmc_wrap_green": (I, [I, I, I, V, V, V, V, V, ctypes.c_size_t, V]),
        "lqmc_wrap_green_static": (I, [I, I, I, V, V, V, V, V]),
        "lqmc_last_error": (ctypes.c_char_p, []),
        "lqmc_destroy": (I, [V]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what=""):
    if rc != 0:
        msg = _lib.lqmc_last_error().decode(errors="replace") if _lib else ""
        raise LqmcError(f"{what} failed ({ERRORS.get(rc, rc)}): {msg}")
    return rc

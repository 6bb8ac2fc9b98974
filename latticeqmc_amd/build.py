"""Build liblqmc_hip.so (gfx950) in-tree with hipcc.

    python -m latticeqmc_amd.build            (liblqmc_hip.so)
    python -m latticeqmc_amd.build --probe    (liblqmc_hip_probe.so, flip phase clocks)

Every HIP source is compiled with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops) and linked into one shared
library next to this file; the Python binding loads it from there.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "liblqmc_hip.so")
SOURCES = ["engine.hip", "gemm.hip", "kron.hip", "linalg.hip", "flip.hip", "qr.hip", "wrap.hip"]
ARCH = os.environ.get("LQMC_OFFLOAD_ARCH", "gfx950")


def hipcc():
    cand = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(cand):
        raise RuntimeError("hipcc not found; the HIP toolchain is required to build liblqmc_hip.so")
    return cand


def flags():
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument"]


# Per-source extra flags.  gemm.hip: keep the matrix-core accumulators in VGPRs (the default
# AGPR form copied all 32 accumulator registers between the two files on every k-tile of the
# GEMM loop).
EXTRA = {"gemm.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


PROBE_LIB = os.path.join(HERE, "liblqmc_hip_probe.so")


def build(verbose=False, probe=False):
    """Compile and link the library.  probe=True builds liblqmc_hip_probe.so instead: the same
    sources with the delayed flip's phase clocks compiled in (-DLQMC_FLIP_PROBE), for
    latticeqmc_amd.flip_phases; the product library never carries them."""
    obj_dir = os.path.join(OBJ, "probe") if probe else OBJ
    os.makedirs(obj_dir, exist_ok=True)
    cc = hipcc()
    extra_all = ["-DLQMC_FLIP_PROBE"] if probe else []

    def compile_one(src):
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        cmd = [cc, *flags(), *extra_all, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
        if verbose and r.stderr.strip():
            print(r.stderr[-3000:], file=sys.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    out = PROBE_LIB if probe else LIB
    tmp = out + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True, probe="--probe" in sys.argv[1:]))

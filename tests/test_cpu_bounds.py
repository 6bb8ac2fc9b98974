"""CPU checks of the flip and QR kernels: the HIP sources are compiled for the host with a
small emulation layer (tests/cpu_emu/hip_cpu_shim.h) and run at the sizes the
engine uses -- under AddressSanitizer with every buffer an exact-size allocation
(bounds), and under ThreadSanitizer (cross-wavefront races between barriers, such
as a field entry flipped before every wave has read it)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CLANG = "/opt/rocm/llvm/bin/clang++"


def _build(tmp_path_factory, sanitizer):
    if not os.path.exists(CLANG):
        pytest.skip("host clang++ with sanitizers not available")
    out = str(tmp_path_factory.mktemp("emu") / f"flip_harness_{sanitizer}")
    cmd = [CLANG, "-std=c++20", "-O1", "-g", f"-fsanitize={sanitizer}", "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "cpu_emu", "flip_harness.cpp"),
           "-o", out, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    return _build(tmp_path_factory, "address")


@pytest.fixture(scope="module")
def harness_tsan(tmp_path_factory):
    return _build(tmp_path_factory, "thread")


@pytest.mark.parametrize("args", [
    "4 16 1 256 0 1 0 1",           # rank-1, tiny
    "36 20 32 256 0 2 3 0 40",      # delayed, rejections, compared with rank-1
    "100 60 8 256 0 2 59 1 40",     # ragged last block, measurement slice
    "64 40 8 256 1 2 39 1 30",      # fp32
    "256 80 32 512 0 1 0 0 60",     # the throughput configuration
    "100 6 16 256 0 2 1 0 200 2",   # two workgroups per chain, ragged, all blocks concurrent
    "100 6 32 256 0 1 5 1 200 3",   # three workgroups per chain, k = 32, measurement slice
    "256 3 16 512 0 1 2 1 120 4",   # the throughput form: N = 256, 512 threads, four workgroups
    "64 3 16 256 0 2 2 1 100 4",    # four workgroups of 256 threads, one spin each: rows staged in LDS
    "64 40 8 256 2 2 39 1 30",      # split-BF16 rank-k products (FP32 G), measurement slice
    "64 16 16 512 2 1 15 1 40",     # split-BF16, 16-site blocks on the 512-thread bound
    "100 4 1 256 0 3 3 1 80 3",     # rank-1 on three workgroups per chain, three chains, uneven rows
    "256 2 1 512 0 1 1 1 100 4",    # the plain form at N = 256 on four workgroups
])
def test_flip_kernel_in_bounds(harness, args):
    """In bounds under ASan; delayed runs (9 args) must match the rank-1 kernel on
    the same state and uniforms (FP64 to 1e-11, FP32 to 1e-4, split-BF16 products to 2e-3;
    same field), and the rank-1
    kernel split over several workgroups must match it bit for bit."""
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([harness, *args.split()], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "ok:" in r.stdout, (r.stdout + r.stderr)[-4000:]


@pytest.mark.parametrize("args", [
    "64 16 8 256 0 1 0 1",      # delayed fp64, several waves
    "64 16 8 256 1 1 15 1",     # delayed fp32
    "64 16 16 512 0 1 15 1",    # delayed fp64, 16-site blocks on the 512-thread bound
    "36 20 32 256 0 1 3 0",
    "64 16 1 256 0 1 0 1",      # rank-1
    "64 3 16 256 0 1 2 1 100 2",  # two workgroups per chain: the chain barrier orders G
    "256 2 16 512 0 1 1 1 100 2",  # the throughput form's N = 256 on 512 threads, two workgroups
    "64 3 16 256 0 1 2 1 100 4",   # one spin per workgroup: rows B staged in LDS while wave 0 runs the sites
    "64 3 1 256 0 1 2 1 50 2",     # rank-1 on two workgroups: the barrier after each accepted flip
])
def test_flip_kernel_race_free(harness_tsan, args):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=0 report_signal_unsafe=0")
    r = subprocess.run([harness_tsan, *args.split()], capture_output=True, text=True, env=env, timeout=900)
    assert "ok:" in r.stdout, (r.stdout + r.stderr)[-3000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[:4000]


def _build_qr(tmp_path_factory, sanitizer):
    if not os.path.exists(CLANG):
        pytest.skip("host clang++ with sanitizers not available")
    out = str(tmp_path_factory.mktemp("emu") / f"qr_harness_{sanitizer}")
    cmd = [CLANG, "-std=c++20", "-O1", "-g", f"-fsanitize={sanitizer}", "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "cpu_emu", "qr_harness.cpp"),
           "-o", out, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


@pytest.fixture(scope="module")
def qr_harness(tmp_path_factory):
    return _build_qr(tmp_path_factory, "address")


@pytest.fixture(scope="module")
def qr_harness_tsan(tmp_path_factory):
    return _build_qr(tmp_path_factory, "thread")


@pytest.mark.parametrize("args", ["4 2 1", "37 20 3", "100 40 5", "257 40 6", "64 20 7 near",
                                  "37 20 3 onchip", "100 40 5 onchip", "128 30 2 onchip", "129 30 4 onchip",
                                  "256 40 6 onchip"])
def test_blocked_qr_and_trsm(qr_harness, args):
    """QR, the column-block Q formation and the permuted triangular solve on graded
    matrices: W = Q diag(D) V'' and W P = Q diag(D) R' to 1e-12 column-relative, Q
    orthogonal, R' upper triangular with unit-modulus diagonal, R' Y' = Z to 1e-9,
    all accesses in bounds.  "near" pairs nearly dependent columns, so the panel's
    downdated column norms collapse and take the exact-recompute path.  "onchip" factors
    with the register-resident kernel (N <= 256; above 128 columns two workgroups run at
    once and hand the left half's panels over through memory)."""
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([qr_harness, *args.split()], capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0 and "ok" in r.stdout, (r.stdout + r.stderr)[-3000:]


@pytest.mark.parametrize("args", ["70 20 2", "70 20 2 onchip", "150 30 3 onchip"])
def test_blocked_qr_race_free(qr_harness_tsan, args):
    """No cross-wavefront race in the QR, Q-formation and solve kernels.  (The 8-column
    panels of N > 256 are factored by one wavefront; their multi-wavefront phases are the
    ones checked here, and the 257 case runs under AddressSanitizer above.)"""
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=0 report_signal_unsafe=0")
    r = subprocess.run([qr_harness_tsan, *args.split()], capture_output=True, text=True, env=env, timeout=1500)
    assert "ok" in r.stdout, (r.stdout + r.stderr)[-3000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[:4000]


def _build_kron(tmp_path_factory, sanitizer):
    if not os.path.exists(CLANG):
        pytest.skip("host clang++ with sanitizers not available")
    out = str(tmp_path_factory.mktemp("emu") / f"kron_harness_{sanitizer}")
    cmd = [CLANG, "-std=c++20", "-O1", "-g", f"-fsanitize={sanitizer}", "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "cpu_emu", "kron_harness.cpp"),
           "-o", out, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


@pytest.fixture(scope="module")
def kron_harness(tmp_path_factory):
    return _build_kron(tmp_path_factory, "address")


@pytest.fixture(scope="module")
def kron_harness_tsan(tmp_path_factory):
    return _build_kron(tmp_path_factory, "thread")


@pytest.mark.parametrize("args", ["16 16 6 6 1", "16 16 16 16 6", "3 9 4 16 2", "2 2 16 5 3", "5 7 16 3 4"])
def test_kron_products(kron_harness, args):
    """Left and right products with a Kronecker-factored E = Ex (x) Ey (two tasks of different
    sizes in one launch, ld larger than N in the second, every scale kind) against the dense
    product to 1e-13 relative, the padding columns of C untouched, every access in bounds and
    no read of unwritten LDS (it starts as NaN)."""
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([kron_harness, *args.split()], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "ok:" in r.stdout, (r.stdout + r.stderr)[-3000:]


def test_kron_products_race_free(kron_harness_tsan):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=0 report_signal_unsafe=0")
    r = subprocess.run([kron_harness_tsan, "6", "6", "3", "9", "5"], capture_output=True, text=True, env=env,
                       timeout=900)
    assert "ok:" in r.stdout, (r.stdout + r.stderr)[-3000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[:4000]


def _build_det(tmp_path_factory, sanitizer):
    if not os.path.exists(CLANG):
        pytest.skip("host clang++ with sanitizers not available")
    out = str(tmp_path_factory.mktemp("emu") / f"det_harness_{sanitizer}")
    cmd = [CLANG, "-std=c++20", "-O1", "-g", f"-fsanitize={sanitizer}", "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "cpu_emu", "det_harness.cpp"),
           "-o", out, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


@pytest.mark.parametrize("args", ["5 6 2", "32 1 1"])
def test_det_sweep_in_bounds_and_matches_host(tmp_path_factory, args):
    """The determinant-mode kernel on uneven chains (N - c sites, L - c slices; N = 32 is the
    largest it takes): in bounds, and the field, accept count and measured inv(M(0)) equal a
    plain host restatement on the same Philox uniforms."""
    exe = _build_det(tmp_path_factory, "address")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([exe, *args.split()], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "ok" in r.stdout, (r.stdout + r.stderr)[-3000:]


def test_det_sweep_race_free(tmp_path_factory):
    exe = _build_det(tmp_path_factory, "thread")
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=0 report_signal_unsafe=0")
    r = subprocess.run([exe, "6", "3", "2"], capture_output=True, text=True, env=env, timeout=900)
    assert "ok" in r.stdout, (r.stdout + r.stderr)[-3000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[:4000]

// Slice wrap of equal-time Green's functions outside a sweep context.
//
// For each of `count` matrices (e.g. both spins of a slice, or many chains of one size):
//   G_m <- B_m^-1 G_m B_m,  B_m = diag(v_m) E_K,  B_m^-1 = E_K^-1 diag(1 / v_m),
// the consistent form of the reference's wrap (lqmc/lqmc.py:338-345), as the sweep engine
// applies it after every slice: T = E_K^-1 diag(1/v) G diag(v), then G = T E_K, two
// batched launches of the MFMA GEMM (gemm.hip) with the diagonal factors fused.  The task
// descriptors are written on the device by a setup kernel from the launch arguments, so
// a call copies nothing from the host and can be captured into a graph.
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/lqmc_hip.h"
#include "common.h"

namespace lqmc {
void launch_gemm(int fp32, int cls, const GemmTask* d_tasks, int ntasks, int ntiles, hipStream_t s);
int gemm_class(const GemmTask& t);
int gemm_tiles(int M, int N, int* tiles_n);
void lqmc_set_error(const std::string& msg);
}  // namespace lqmc

using namespace lqmc;

namespace {

// Workspace layout: 2*count GemmTask descriptors, then count N x N temporaries.
size_t wrap_desc_bytes(int count) { return ((sizeof(GemmTask) * 2 * (size_t)count + 255) / 256) * 256; }
size_t wrap_bytes(int n, int count) { return wrap_desc_bytes(count) + sizeof(double) * (size_t)n * n * count; }

// Task m of the first product (m < count) and of the second (count + m).  v holds the
// diagonal of V for matrix m at v + m n; its inverse is fused as the SC_INV k-scale.
__global__ void k_wrap_tasks(GemmTask* tasks, double* tw, int n, int count, const double* expk,
                             const double* expk_inv, const double* v, double* g, int tiles, int tiles_n) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= count) return;
  const size_t n2 = (size_t)n * n;
  GemmTask a{};
  a.M = a.N = a.K = n;
  a.lda = a.ldb = a.ldc = n;
  a.alpha = 1.0;
  a.beta = 0.0;
  a.tiles_n = tiles_n;
  a.tile_base = m * tiles;
  GemmTask b = a;
  // T = E_K^-1 diag(1/v) G diag(v)
  a.A = expk_inv;
  a.B = g + m * n2;
  a.C = tw + m * n2;
  a.ks = v + (size_t)m * n;
  a.ks_kind = SC_INV;
  a.cs = v + (size_t)m * n;
  a.cs_kind = SC_PLAIN;
  // G = T E_K
  b.A = tw + m * n2;
  b.B = expk;
  b.C = g + m * n2;
  tasks[m] = a;
  tasks[count + m] = b;
}

int wrap_check(int n, int count, const double* expk, const double* expk_inv, const double* v, double* g) {
  if (n < 1 || n > 4096 || count < 1 || !expk || !expk_inv || !v || !g) {
    lqmc_set_error("lqmc_wrap_green: bad arguments (1 <= n <= 4096, count >= 1, non-null pointers)");
    return LQMC_EINVAL;
  }
  return LQMC_OK;
}

// Both launches of one wrap; ws is device memory of at least wrap_bytes(n, count) bytes.
int wrap_issue(int n, int count, const double* expk, const double* expk_inv, const double* v, double* g,
               char* ws, hipStream_t s) {
  GemmTask proto{};
  proto.M = proto.N = proto.K = n;
  const int cls = gemm_class(proto);
  int tiles_n = 0;
  const int tiles = gemm_tiles(n, n, &tiles_n);
  GemmTask* tasks = reinterpret_cast<GemmTask*>(ws);
  double* tw = reinterpret_cast<double*>(ws + wrap_desc_bytes(count));
  hipLaunchKernelGGL(k_wrap_tasks, dim3((count + 63) / 64), dim3(64), 0, s, tasks, tw, n, count, expk, expk_inv,
                     v, g, tiles, tiles_n);
  launch_gemm(0, cls, tasks, count, tiles * count, s);
  launch_gemm(0, cls, tasks + count, count, tiles * count, s);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    lqmc_set_error(std::string("lqmc_wrap_green: launch failed: ") + hipGetErrorString(e));
    return LQMC_EHIP;
  }
  return LQMC_OK;
}

// ---- allocation-free variant ------------------------------------------------------------------
// lqmc_wrap_green_static keeps the workspace of lqmc_wrap_green in a file-scope device array of
// WRAP_SCRATCH_BYTES, so a call allocates nothing (for code under graph capture, serving loops
// and harnesses that forbid hipMalloc).  Each GPU holds its own copy of the array until the
// process ends.  In host code the array's name is not a device address: the address of each
// device's copy is looked up once per device with hipGetSymbolAddress and kept in
// g_scratch_ptr, and only that pointer is handed to the kernels.  Every call on a GPU shares
// that GPU's array: calls on one stream are ordered, calls that overlap on one GPU (other
// streams or threads) race.  Nothing in it needs clearing: each call writes its descriptors
// and every temporary entry before reading them.
constexpr size_t WRAP_SCRATCH_BYTES = size_t(16) << 20;
constexpr int WRAP_MAX_DEVICES = 64;
static __device__ __align__(256) char scratch[WRAP_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[WRAP_MAX_DEVICES] = {};

// Device address of `scratch` on `device` (current device on return); null on failure.
char* scratch_on(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return g_scratch_ptr[device];
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    lqmc_set_error(std::string("lqmc_wrap_green_static: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    return nullptr;
  }
  g_scratch_ptr[device] = static_cast<char*>(p);
  return g_scratch_ptr[device];
}

}  // namespace

extern "C" {

size_t lqmc_wrap_green_workspace(int n, int count) {
  if (n < 1 || count < 1) return 0;
  return wrap_bytes(n, count);
}

int lqmc_wrap_green(int device, int n, int count, const double* expk, const double* expk_inv, const double* v,
                    double* g, void* workspace, size_t ws_bytes, void* stream) {
  int rc = wrap_check(n, count, expk, expk_inv, v, g);
  if (rc) return rc;
  if (!workspace || ws_bytes < wrap_bytes(n, count)) {
    lqmc_set_error("lqmc_wrap_green: workspace missing or smaller than lqmc_wrap_green_workspace(n, count)");
    return LQMC_EINVAL;
  }
  if (hipSetDevice(device) != hipSuccess) {
    lqmc_set_error("lqmc_wrap_green: hipSetDevice failed");
    return LQMC_EHIP;
  }
  return wrap_issue(n, count, expk, expk_inv, v, g, static_cast<char*>(workspace), static_cast<hipStream_t>(stream));
}

int lqmc_wrap_green_static(int device, int n, int count, const double* expk, const double* expk_inv,
                           const double* v, double* g, void* stream) {
  int rc = wrap_check(n, count, expk, expk_inv, v, g);
  if (rc) return rc;
  if (wrap_bytes(n, count) > WRAP_SCRATCH_BYTES) {
    lqmc_set_error("lqmc_wrap_green_static: n*n*count*8 bytes plus descriptors exceed the 16 MiB static "
                   "workspace; use lqmc_wrap_green with a workspace of lqmc_wrap_green_workspace(n, count) bytes");
    return LQMC_EINVAL;
  }
  if (device < 0 || device >= WRAP_MAX_DEVICES) {
    lqmc_set_error("lqmc_wrap_green_static: device index out of range");
    return LQMC_EINVAL;
  }
  const hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    lqmc_set_error(std::string("lqmc_wrap_green_static: hipSetDevice failed: ") + hipGetErrorString(e));
    return LQMC_EHIP;
  }
  char* ws = scratch_on(device);
  if (!ws) return LQMC_EHIP;
  return wrap_issue(n, count, expk, expk_inv, v, g, ws, static_cast<hipStream_t>(stream));
}

}  // extern "C"

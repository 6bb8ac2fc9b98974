// FP64 helper kernels: partial-pivot LU solve (the reference mode's inv(I + P)),
// the recompute epilogue, identity fills and the exp(V) tables.
//
// One workgroup per matrix (the batch is chains x spins).  Inner loops map the
// 64 lanes of a wavefront to 64 consecutive columns (coalesced row-major
// access) and the wavefronts of the workgroup to interleaved row groups; the
// per-column dot products are combined through LDS.
#include <algorithm>

#include "common.h"

namespace lqmc {

constexpr int LA_THREADS = 256;
constexpr int LA_WAVES = LA_THREADS / 64;

__device__ inline void set_flag(ChainState* st, int chain, unsigned f) {
  if (st) atomicOr(&st[chain].flags, f);
}

// Block-wide argmax (first index on ties) of (val, idx) pairs.
__device__ inline void block_argmax(double& val, int& idx, double* rv, int* ri) {
  const int tid = threadIdx.x;
  rv[tid] = val;
  ri[tid] = idx;
  __syncthreads();
  for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
    if (tid < s) {
      double v2 = rv[tid + s];
      int i2 = ri[tid + s];
      if (v2 > rv[tid] || (v2 == rv[tid] && i2 < ri[tid])) {
        rv[tid] = v2;
        ri[tid] = i2;
      }
    }
    __syncthreads();
  }
  val = rv[0];
  idx = ri[0];
  __syncthreads();
}

__device__ inline double block_sum(double v, double* rv) {
  const int tid = threadIdx.x;
  rv[tid] = v;
  __syncthreads();
  for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
    if (tid < s) rv[tid] += rv[tid + s];
    __syncthreads();
  }
  double r = rv[0];
  __syncthreads();
  return r;
}

// Gaussian elimination with partial pivoting on [X | RHS] (N x 2N, ld 2N),
// then back substitution; Y overwrites RHS.
__global__ __launch_bounds__(LA_THREADS) void k_solve(const SolveTask* __restrict__ tasks, ChainState* st) {
  const SolveTask tk = tasks[blockIdx.x];
  const int N = tk.N, N2 = 2 * N;
  double* __restrict__ A = tk.aug;
  LQMC_DYNAMIC_LDS(double, sm);
  double* mult = sm;                  // N
  double* rv = mult + N;              // LA_THREADS
  int* ri = reinterpret_cast<int*>(rv + LA_THREADS);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;

  for (size_t e = tid; e < (size_t)N * N; e += blockDim.x) {
    const int r = (int)(e / N), c = (int)(e % N);
    double v;
    if (tk.U) v = tk.U[(size_t)c * N + r] / fmax(tk.D[r], 1.0);
    else v = (r == c) ? 1.0 : 0.0;
    A[(size_t)r * N2 + N + c] = v;
    if (tk.add_identity && r == c) A[(size_t)r * N2 + r] += 1.0;
  }
  __syncthreads();

  for (int j = 0; j < N; ++j) {
    double best = -1.0;
    int bi = N;
    for (int r = j + tid; r < N; r += blockDim.x) {
      double a = fabs(A[(size_t)r * N2 + j]);
      if (a > best) { best = a; bi = r; }
    }
    block_argmax(best, bi, rv, ri);
    const int p = bi;
    if (p != j)
      for (int c = j + tid; c < N2; c += blockDim.x) {
        double a = A[(size_t)j * N2 + c];
        A[(size_t)j * N2 + c] = A[(size_t)p * N2 + c];
        A[(size_t)p * N2 + c] = a;
      }
    __syncthreads();
    const double piv = A[(size_t)j * N2 + j];
    if (piv == 0.0 && tid == 0) set_flag(st, tk.chain, 2u);
    const double ipiv = 1.0 / piv;
    for (int r = j + 1 + tid; r < N; r += blockDim.x) mult[r] = A[(size_t)r * N2 + j] * ipiv;
    __syncthreads();
    for (int c0 = j + 1; c0 < N2; c0 += 64) {
      const int c = c0 + lane;
      if (c < N2) {
        const double top = A[(size_t)j * N2 + c];
        for (int r = j + 1 + wave; r < N; r += LA_WAVES) A[(size_t)r * N2 + c] -= mult[r] * top;
      }
    }
    __syncthreads();
  }
  // back substitution on the RHS block
  for (int j = N - 1; j >= 0; --j) {
    const double ipiv = 1.0 / A[(size_t)j * N2 + j];
    for (int c = tid; c < N; c += blockDim.x) A[(size_t)j * N2 + N + c] *= ipiv;
    __syncthreads();
    for (int c0 = 0; c0 < N; c0 += 64) {
      const int c = c0 + lane;
      if (c < N) {
        const double yj = A[(size_t)j * N2 + N + c];
        for (int r = wave; r < j; r += LA_WAVES) A[(size_t)r * N2 + N + c] -= A[(size_t)r * N2 + j] * yj;
      }
    }
    __syncthreads();
  }
  if (tk.Yout)
    for (size_t e = tid; e < (size_t)N * N; e += blockDim.x) {
      const int r = (int)(e / N), c = (int)(e % N);
      tk.Yout[e] = A[(size_t)r * N2 + N + c];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_finish(const FinishTask* __restrict__ tasks, ChainState* st) {
  const FinishTask tk = tasks[blockIdx.x];
  const size_t n2 = (size_t)tk.N * tk.N;
  T* G = static_cast<T*>(tk.G);
  double m = 0.0;
  bool bad = false;
  for (size_t e = blockIdx.y * blockDim.x + threadIdx.x; e < n2; e += (size_t)gridDim.y * blockDim.x) {
    const double s = tk.Gst[e];
    if (!isfinite(s)) bad = true;
    if (tk.compare) m = fmax(m, fabs((double)G[e] - s));
    G[e] = (T)s;
  }
  __shared__ double rv[256];
  rv[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) rv[threadIdx.x] = fmax(rv[threadIdx.x], rv[threadIdx.x + s]);
    __syncthreads();
  }
  if (bad) set_flag(st, tk.chain, 8u);
  if (threadIdx.x == 0 && tk.compare) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(rv[0]);
    atomicMax(&st[tk.chain].max_err_bits, bits);
    if (tk.slot) atomicMax(tk.slot, bits);
    if (tk.tol > 0.0 && !(rv[0] <= tk.tol)) set_flag(st, tk.chain, 32u);
    if (blockIdx.y == 0) atomicAdd(&st[tk.chain].recomputes, 1ull);
  }
}

// Identity matrices: ptrs[b] -> N[b] x N[b]
__global__ void k_identity(double* const* ptrs, const int* ns) {
  double* p = ptrs[blockIdx.x];
  const int N = ns[blockIdx.x];
  for (size_t e = threadIdx.x; e < (size_t)N * N; e += blockDim.x) p[e] = (e / N == e % N) ? 1.0 : 0.0;
}

// ev/evi tables of every slice from the field.
__global__ void k_expv(const ChainDev* __restrict__ chains) {
  const ChainDev ch = chains[blockIdx.x];
  const int NL = ch.N * ch.L;
  for (int e = threadIdx.x; e < NL; e += blockDim.x) {
    const int s = ch.field[e];
    const double up = s > 0 ? ch.e1m : ch.e1p;   // exp(-lambda*s), sigma = +1
    const double dn = s > 0 ? ch.e1p : ch.e1m;   // sigma = -1
    ch.ev[e] = up;
    ch.ev[NL + e] = dn;
    ch.evi[e] = 1.0 / up;
    ch.evi[NL + e] = 1.0 / dn;
  }
}

// One chain's G in FP64 into caller device memory: src in the form's precision, divided by
// the chain's measured-sweep count when mean != 0 (the accumulated sum -> the mean).
__global__ void k_green_out(double* out, const void* src, int fp32, size_t n, const ChainState* st, int mean) {
  const double m = mean ? (double)(st->measured ? st->measured : 1ull) : 1.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const double v = fp32 ? (double)static_cast<const float*>(src)[e] : static_cast<const double*>(src)[e];
    out[e] = mean ? v / m : v;
  }
}

void launch_green_out(double* out, const void* src, int fp32, size_t n, const ChainState* st, int mean,
                      hipStream_t s) {
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(k_green_out, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, s, out, src, fp32, n, st, mean);
}

void launch_solve(const SolveTask* d, int n, int nmax, ChainState* st, hipStream_t s) {
  if (n <= 0) return;
  size_t sh = sizeof(double) * (nmax + LA_THREADS) + sizeof(int) * LA_THREADS;
  hipLaunchKernelGGL(k_solve, dim3(n), dim3(LA_THREADS), sh, s, d, st);
}

void launch_finish(int fp32, const FinishTask* d, int n, int nmax, ChainState* st, hipStream_t s) {
  if (n <= 0) return;
  int per = (int)(((size_t)nmax * nmax + 256 * 64 - 1) / (256 * 64));
  if (per < 1) per = 1;
  if (per > 32) per = 32;
  if (fp32) hipLaunchKernelGGL(k_finish<float>, dim3(n, per), dim3(256), 0, s, d, st);
  else hipLaunchKernelGGL(k_finish<double>, dim3(n, per), dim3(256), 0, s, d, st);
}

void launch_identity(double* const* ptrs, const int* ns, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_identity, dim3(n), dim3(256), 0, s, ptrs, ns);
}

void launch_expv(const ChainDev* chains, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_expv, dim3(n), dim3(256), 0, s, chains);
}

}  // namespace lqmc

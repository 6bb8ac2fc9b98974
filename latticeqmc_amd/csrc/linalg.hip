// FP64 stabilisation kernels: pivoted Householder QR (UDV step), partial-pivot
// LU solve of the stable Green's-function system, and the recompute epilogue.
//
// One workgroup per matrix (the batch is chains x spins).  Inner loops map the
// 64 lanes of a wavefront to 64 consecutive columns (coalesced row-major
// access) and the wavefronts of the workgroup to interleaved row groups; the
// per-column dot products are combined through LDS.
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

// Apply the Householder reflector (I - tau v v^T), v = vbuf[j..N-1] (vbuf[j] = 1),
// to columns [c_begin, N) of rows [j, N) of X (ld N).  If norms != null, store the
// squared norm of rows j+1..N-1 of each updated column.
__device__ void apply_reflector(double* X, int N, int j, double tau, int c_begin,
                                const double* vbuf, double* part, double* norms) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  for (int c0 = c_begin; c0 < N; c0 += 64) {
    const int c = c0 + lane;
    double w = 0.0;
    if (c < N)
      for (int r = j + wave; r < N; r += LA_WAVES) w += vbuf[r] * X[(size_t)r * N + c];
    part[wave * 64 + lane] = w;
    __syncthreads();
    double wt = 0.0;
#pragma unroll
    for (int q = 0; q < LA_WAVES; ++q) wt += part[q * 64 + lane];
    wt *= tau;
    double ns = 0.0;
    if (c < N)
      for (int r = j + wave; r < N; r += LA_WAVES) {
        double x = X[(size_t)r * N + c] - vbuf[r] * wt;
        X[(size_t)r * N + c] = x;
        if (r > j) ns += x * x;
      }
    __syncthreads();
    if (norms) {
      part[wave * 64 + lane] = ns;
      __syncthreads();
      if (wave == 0 && c < N) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < LA_WAVES; ++q) t += part[q * 64 + lane];
        norms[c] = t;
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(LA_THREADS) void k_qr(const QrTask* __restrict__ tasks, ChainState* st) {
  const QrTask tk = tasks[blockIdx.x];
  const int N = tk.N;
  double* __restrict__ W = tk.W;
  extern __shared__ double sm[];
  double* norms = sm;             // N
  double* vbuf = norms + N;       // N
  double* tau = vbuf + N;         // N
  double* part = tau + N;         // LA_THREADS
  double* rv = part + LA_THREADS; // LA_THREADS
  int* ri = reinterpret_cast<int*>(rv + LA_THREADS);  // LA_THREADS
  int* perm = ri + LA_THREADS;                        // N
  const int tid = threadIdx.x;

  // initial column norms
  for (int c0 = 0; c0 < N; c0 += 64) {
    const int c = c0 + (tid & 63);
    double s = 0.0;
    if (c < N)
      for (int r = tid >> 6; r < N; r += LA_WAVES) {
        double x = W[(size_t)r * N + c];
        s += x * x;
      }
    part[tid] = s;
    __syncthreads();
    if (tid < 64 && c < N) {
      double t = 0.0;
      for (int q = 0; q < LA_WAVES; ++q) t += part[q * 64 + tid];
      norms[c] = t;
      perm[c] = c;
    }
    __syncthreads();
  }

  for (int j = 0; j < N; ++j) {
    // pivot: column of largest remaining norm
    double best = -1.0;
    int bi = N;
    for (int c = j + tid; c < N; c += blockDim.x)
      if (norms[c] > best) { best = norms[c]; bi = c; }
    block_argmax(best, bi, rv, ri);
    const int p = bi;
    if (p != j) {
      for (int r = tid; r < N; r += blockDim.x) {
        double a = W[(size_t)r * N + j];
        W[(size_t)r * N + j] = W[(size_t)r * N + p];
        W[(size_t)r * N + p] = a;
      }
      if (tid == 0) {
        double tn = norms[j]; norms[j] = norms[p]; norms[p] = tn;
        int tp = perm[j]; perm[j] = perm[p]; perm[p] = tp;
      }
    }
    __syncthreads();
    // Householder vector of column j
    double s = 0.0;
    for (int r = j + tid; r < N; r += blockDim.x) {
      double x = W[(size_t)r * N + j];
      s += x * x;
    }
    const double nrm2 = block_sum(s, rv);
    const double alpha = W[(size_t)j * N + j];
    const double nrm = sqrt(nrm2);
    double beta, tj, scal;
    if (nrm == 0.0) {
      beta = 0.0; tj = 0.0; scal = 0.0;
      if (tid == 0) set_flag(st, tk.chain, 2u);
    } else {
      beta = alpha >= 0.0 ? -nrm : nrm;
      tj = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    for (int r = j + 1 + tid; r < N; r += blockDim.x) {
      double v = W[(size_t)r * N + j] * scal;
      vbuf[r] = v;
      W[(size_t)r * N + j] = v;
    }
    if (tid == 0) {
      vbuf[j] = 1.0;
      W[(size_t)j * N + j] = beta;
      tau[j] = tj;
    }
    __syncthreads();
    if (j + 1 < N) apply_reflector(W, N, j, tj, j + 1, vbuf, part, norms);
  }

  // D and T = D^-1 R P^T
  for (int i = tid; i < N; i += blockDim.x) {
    double d = fabs(W[(size_t)i * N + i]);
    tk.D[i] = d > 0.0 ? d : 1e-300;
  }
  __syncthreads();
  for (size_t e = tid; e < (size_t)N * N; e += blockDim.x) {
    const int i = (int)(e / N), c = (int)(e % N);
    const double d = tk.D[i];
    tk.T[(size_t)i * N + perm[c]] = c >= i ? W[e] / d : 0.0;
  }
  // Q = H_0 H_1 ... H_{N-1} (backward accumulation)
  double* Q = tk.Q;
  for (size_t e = tid; e < (size_t)N * N; e += blockDim.x) Q[e] = (e / N == e % N) ? 1.0 : 0.0;
  __syncthreads();
  for (int j = N - 1; j >= 0; --j) {
    for (int r = j + 1 + tid; r < N; r += blockDim.x) vbuf[r] = W[(size_t)r * N + j];
    if (tid == 0) vbuf[j] = 1.0;
    __syncthreads();
    if (tau[j] != 0.0) apply_reflector(Q, N, j, tau[j], j, vbuf, part, nullptr);
  }
}

// Gaussian elimination with partial pivoting on [X | RHS] (N x 2N, ld 2N),
// then back substitution; Y overwrites RHS.
__global__ __launch_bounds__(LA_THREADS) void k_solve(const SolveTask* __restrict__ tasks, ChainState* st) {
  const SolveTask tk = tasks[blockIdx.x];
  const int N = tk.N, N2 = 2 * N;
  double* __restrict__ A = tk.aug;
  extern __shared__ double sm[];
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
    atomicMax(&st[tk.chain].max_err_bits, (unsigned long long)__double_as_longlong(rv[0]));
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

void launch_qr(const QrTask* d, int n, int nmax, ChainState* st, hipStream_t s) {
  if (n <= 0) return;
  size_t sh = sizeof(double) * (3 * nmax + 2 * LA_THREADS) + sizeof(int) * (LA_THREADS + nmax);
  hipLaunchKernelGGL(k_qr, dim3(n), dim3(LA_THREADS), sh, s, d, st);
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

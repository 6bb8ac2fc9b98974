// Blocked FP64 factorisations for the UDV stabilisation (one workgroup per matrix).
//
// k_qr_blocked: pre-pivoted Householder QR, W P = Q R with the columns of W
//   sorted once by decreasing norm (pre-pivoting, Bai, Lee, Li, Xu 2011), then an
//   unpivoted blocked factorisation: each panel of nb columns is factored in LDS
//   and applied to the trailing matrix in compact-WY form
//   (I - V T V^T)^T A = A - V T^T (V^T A), two passes over the trailing matrix
//   per panel instead of per column.  Outputs Q (explicit, backward blocked
//   accumulation), D = |diag R| and V'' = D^-1 R P^T written over W, so that
//   W = Q diag(D) V''.  The factor lives in the scratch matrix A.
// k_trsm_perm: solves V'' Y = Z for the V'' of k_qr_blocked (R' = D^-1 R is
//   upper triangular with unit-modulus diagonal, V''[i][perm[c]] = R'[i][c]),
//   block back substitution on 16-row blocks.
#include "common.h"

namespace lqmc {

constexpr int QR_THREADS = 1024;
constexpr int QR_WAVES = QR_THREADS / 64;
constexpr int TRSM_B = 16;

__device__ inline double qr_wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline double qr_block_sum(double v, double* red) {
  v = qr_wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int q = 0; q < QR_WAVES; ++q) r += red[q];
  __syncthreads();
  return r;
}

__device__ inline int qr_panel_width(int N) { return N > 256 ? 8 : 16; }

// T (w x w, ld nb) of the compact-WY form of the panel's reflectors (V explicit in pnl).
__device__ void qr_build_t(const double* pnl, int nb, int r0, int N, int w, const double* tau, double* Tm,
                           double* S) {
  for (int e = threadIdx.x; e < w * w; e += QR_THREADS) {
    const int p = e / w, q = e % w;
    if (p < q) {
      double s = 0.0;
      for (int r = r0; r < N; ++r) s += pnl[r * nb + p] * pnl[r * nb + q];
      S[p * 64 + q] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < w; ++i) {
      Tm[i * nb + i] = tau[i];
      for (int p = 0; p < i; ++p) {
        double acc = 0.0;
        for (int p2 = p; p2 < i; ++p2) acc += Tm[p * nb + p2] * S[p2 * 64 + i];
        Tm[p * nb + i] = -tau[i] * acc;
      }
      for (int p = i + 1; p < w; ++p) Tm[p * nb + i] = 0.0;
    }
  }
  __syncthreads();
}

typedef double qr_v4d __attribute__((ext_vector_type(4)));

__device__ inline qr_v4d qr_mma(double a, double b, qr_v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// X[r0:N, c_begin:N] <- (I - V T' V^T) X[r0:N, c_begin:N] on the FP64 matrix cores,
// V = pnl (explicit reflectors, rows r0..N-1, w <= 16 columns), T' = T^T (trans) or T.
// One wavefront per 16-column tile:  W = V^T X (K = rows), W' = T' W, X -= V W'.
// In the 16x16x4 FP64 layout the accumulator element v of lane l is row l/16 + 4v,
// column l%16 -- exactly the B operand lane l supplies for k-step v -- so W and W'
// chain through registers.  Ends with a workgroup barrier.
__device__ void qr_apply_block(double* X, int N, int r0, int c_begin, const double* pnl, int nb, int w,
                               const double* Tm, bool trans) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nct = (N - c_begin + 15) >> 4;
  for (int ct = wave; ct < nct; ct += QR_WAVES) {
    const int c = c_begin + ct * 16 + li;
    qr_v4d wt = {0.0, 0.0, 0.0, 0.0};
    for (int kk = r0; kk < N; kk += 4) {
      const int r = kk + lk;
      const double a = (r < N && li < w) ? pnl[r * nb + li] : 0.0;
      const double b = (r < N && c < N) ? X[(size_t)r * N + c] : 0.0;
      wt = qr_mma(a, b, wt);
    }
    qr_v4d w2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int p = 4 * s + lk;
      const double a = (li < w && p < w) ? (trans ? Tm[p * nb + li] : Tm[li * nb + p]) : 0.0;
      w2 = qr_mma(a, wt[s], w2);
    }
    for (int rt = r0; rt < N; rt += 16) {
      qr_v4d acc;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = rt + lk + 4 * v;
        acc[v] = (row < N && c < N) ? X[(size_t)row * N + c] : 0.0;
      }
      const int ra = rt + li;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int q = 4 * s + lk;
        const double a = (ra < N && q < w) ? -pnl[ra * nb + q] : 0.0;
        acc = qr_mma(a, w2[s], acc);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = rt + lk + 4 * v;
        if (row < N && c < N) X[(size_t)row * N + c] = acc[v];
      }
    }
  }
  __syncthreads();
}

// Orders LDS traffic between the lanes of one wavefront.
__device__ inline void qr_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Householder factorisation of the panel (rows j0..N-1, w <= NB columns, in LDS) by one
// wavefront, with panel-local column pivoting: at step q the remaining panel column of
// largest norm over rows j..N-1 is swapped into place (with its R entries above the
// panel in A and its pivot index).  The norms of all panel columns and the reflector's
// dot products with them are accumulated in one pass each and reduced across lanes
// together.  On return pnl holds R (on/above the diagonal) and the reflectors below it.
template <int NB>
__device__ void qr_panel_wave(double* pnl, double* A, int N, int j0, int w, int* perm, double* taus,
                              bool& zero_pivot) {
  const int lane = threadIdx.x & 63;
  for (int q = 0; q < w; ++q) {
    const int j = j0 + q;
    double nrm[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) nrm[c] = 0.0;
    for (int r = j + lane; r < N; r += 64) {
      const double* row = pnl + r * NB;
#pragma unroll
      for (int c = 0; c < NB; ++c) nrm[c] += row[c] * row[c];
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) nrm[c] = qr_wave_sum(nrm[c]);
    int p = q;
    double best = -1.0;
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (c >= q && c < w && nrm[c] > best) { best = nrm[c]; p = c; }
    if (p != q) {
      for (int r = j0 + lane; r < N; r += 64) {
        const double t0 = pnl[r * NB + q];
        pnl[r * NB + q] = pnl[r * NB + p];
        pnl[r * NB + p] = t0;
      }
      for (int r = lane; r < j0; r += 64) {
        const double t0 = A[(size_t)r * N + j0 + q];
        A[(size_t)r * N + j0 + q] = A[(size_t)r * N + j0 + p];
        A[(size_t)r * N + j0 + p] = t0;
      }
      if (lane == 0) {
        const int tp = perm[j0 + q];
        perm[j0 + q] = perm[j0 + p];
        perm[j0 + p] = tp;
      }
    }
    qr_wave_sync();
    const double alpha = pnl[j * NB + q];
    const double nrmv = sqrt(best > 0.0 ? best : 0.0);
    double beta, tau, scal;
    if (nrmv == 0.0) {
      beta = 0.0; tau = 0.0; scal = 0.0;
      zero_pivot = true;
    } else {
      beta = alpha >= 0.0 ? -nrmv : nrmv;
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    for (int r = j + 1 + lane; r < N; r += 64) pnl[r * NB + q] *= scal;
    qr_wave_sync();
    double d[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) d[c] = 0.0;
    for (int r = j + lane; r < N; r += 64) {
      const double v = r == j ? 1.0 : pnl[r * NB + q];
      const double* row = pnl + r * NB;
#pragma unroll
      for (int c = 0; c < NB; ++c) d[c] += v * row[c];
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) d[c] = qr_wave_sum(d[c]) * tau;
    for (int r = j + lane; r < N; r += 64) {
      const double v = r == j ? 1.0 : pnl[r * NB + q];
      double* row = pnl + r * NB;
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c > q && c < w) row[c] -= v * d[c];
    }
    qr_wave_sync();
    if (lane == 0) {
      pnl[j * NB + q] = beta;
      taus[j] = tau;
    }
    qr_wave_sync();
  }
}

__global__ __launch_bounds__(QR_THREADS) void k_qr_blocked(const QrTask* __restrict__ tasks, ChainState* st) {
  const QrTask tk = tasks[blockIdx.x];
  const int N = tk.N;
  const int nb = qr_panel_width(N);
  double* __restrict__ W = tk.W;
  double* __restrict__ A = tk.A;
  double* __restrict__ Q = tk.Q;
  LQMC_DYNAMIC_LDS(double, sm);
  double* pnl = sm;                    // [N][nb]
  double* Tm = pnl + (size_t)N * nb;   // [nb][nb]
  double* Wt = Tm + nb * nb;           // [nb][64] (also S)
  double* red = Wt + nb * 64;          // [64]
  double* norms = red + 64;            // [N]
  double* taus = norms + N;            // [N]
  int* perm = reinterpret_cast<int*>(taus + N);  // [N]
  const int tid = threadIdx.x;
  bool zero_pivot = false, bad = false;

  // pre-pivoting: order columns by decreasing norm (ties by index; NaN last)
  for (int c = tid; c < N; c += QR_THREADS) {
    double s = 0.0;
    for (int r = 0; r < N; ++r) {
      const double x = W[(size_t)r * N + c];
      s += x * x;
    }
    if (!(s >= 0.0)) { s = -1.0; bad = true; }
    norms[c] = s;
  }
  __syncthreads();
  for (int c = tid; c < N; c += QR_THREADS) {
    const double nc = norms[c];
    int rank = 0;
    for (int c2 = 0; c2 < N; ++c2) {
      const double n2 = norms[c2];
      rank += (n2 > nc) || (n2 == nc && c2 < c);
    }
    perm[rank] = c;
  }
  __syncthreads();
  for (size_t e = tid; e < (size_t)N * N; e += QR_THREADS) {
    const int r = (int)(e / N), j = (int)(e % N);
    A[e] = W[(size_t)r * N + perm[j]];
  }
  __syncthreads();

  for (int j0 = 0; j0 < N; j0 += nb) {
    const int w = min(nb, N - j0);
    for (int e = tid; e < (N - j0) * nb; e += QR_THREADS) {
      const int r = j0 + e / nb, q = e % nb;
      pnl[r * nb + q] = q < w ? A[(size_t)r * N + j0 + q] : 0.0;
    }
    __syncthreads();
    if ((tid >> 6) == 0) {
      if (nb == 16) qr_panel_wave<16>(pnl, A, N, j0, w, perm, taus, zero_pivot);
      else qr_panel_wave<8>(pnl, A, N, j0, w, perm, taus, zero_pivot);
    }
    __syncthreads();
    // R (upper part) and V (below the diagonal) back to A; pnl becomes the explicit V
    for (int e = tid; e < (N - j0) * w; e += QR_THREADS) {
      const int r = j0 + e / w, q = e % w;
      A[(size_t)r * N + j0 + q] = pnl[r * nb + q];
    }
    __syncthreads();
    for (int e = tid; e < w * w; e += QR_THREADS) {
      const int rr = e / w, q = e % w;
      if (rr < q) pnl[(j0 + rr) * nb + q] = 0.0;
      else if (rr == q) pnl[(j0 + rr) * nb + q] = 1.0;
    }
    __syncthreads();
    if (j0 + w < N) {
      qr_build_t(pnl, nb, j0, N, w, taus + j0, Tm, Wt);
      qr_apply_block(A, N, j0, j0 + w, pnl, nb, w, Tm, true);
    }
  }

  // D = |diag R|, V'' = D^-1 R P^T over W
  for (int i = tid; i < N; i += QR_THREADS) {
    const double d = fabs(A[(size_t)i * N + i]);
    tk.D[i] = d > 0.0 ? d : 1e-300;
    if (tk.perm) tk.perm[i] = perm[i];
  }
  __syncthreads();
  for (size_t e = tid; e < (size_t)N * N; e += QR_THREADS) {
    const int i = (int)(e / N), c = (int)(e % N);
    W[(size_t)i * N + perm[c]] = c >= i ? A[e] / tk.D[i] : 0.0;
  }
  // Q = H_0 ... H_{N-1}: backward over panels, Q[j0:, j0:] <- (I - V T V^T) Q[j0:, j0:]
  for (size_t e = tid; e < (size_t)N * N; e += QR_THREADS) Q[e] = (e / N == e % N) ? 1.0 : 0.0;
  __syncthreads();
  for (int j0 = ((N - 1) / nb) * nb; j0 >= 0; j0 -= nb) {
    const int w = min(nb, N - j0);
    for (int e = tid; e < (N - j0) * nb; e += QR_THREADS) {
      const int r = j0 + e / nb, q = e % nb;
      const int rr = r - j0;
      double v = 0.0;
      if (q < w) v = rr < q ? 0.0 : (rr == q ? 1.0 : A[(size_t)r * N + j0 + q]);
      pnl[r * nb + q] = v;
    }
    __syncthreads();
    qr_build_t(pnl, nb, j0, N, w, taus + j0, Tm, Wt);
    qr_apply_block(Q, N, j0, j0, pnl, nb, w, Tm, false);
  }
  if (tid == 0 && st) {
    if (zero_pivot) atomicOr(&st[tk.chain].flags, 2u);
    if (bad) atomicOr(&st[tk.chain].flags, 8u);
  }
  if (bad && st) atomicOr(&st[tk.chain].flags, 8u);
}

__global__ __launch_bounds__(QR_THREADS) void k_trsm_perm(const TrsmTask* __restrict__ tasks, ChainState* st) {
  const TrsmTask tk = tasks[blockIdx.x];
  const int N = tk.N;
  const double* __restrict__ V = tk.V;
  const int* __restrict__ perm = tk.perm;
  double* __restrict__ Z = tk.Z;
  LQMC_DYNAMIC_LDS(double, sm);
  double* blk = sm;                                   // [TRSM_B][64]
  double* rb = blk + TRSM_B * 64;                     // [TRSM_B][TRSM_B] diagonal block of R'
  int* pm = reinterpret_cast<int*>(rb + TRSM_B * TRSM_B);  // [N]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int c = tid; c < N; c += QR_THREADS) pm[c] = perm[c];
  __syncthreads();
  bool zero = false;
  for (int i0 = ((N - 1) / TRSM_B) * TRSM_B; i0 >= 0; i0 -= TRSM_B) {
    const int b = min(TRSM_B, N - i0);
    for (int e = tid; e < b * b; e += QR_THREADS) {
      const int ii = e / b, kk = e % b;
      rb[ii * TRSM_B + kk] = V[(size_t)(i0 + ii) * N + pm[i0 + kk]];
    }
    for (int cc = 0; cc < N; cc += 64) {
      const int c = cc + lane;
      // rows of this block owned by the wave: ii = wave + 4m
      double acc[TRSM_B / QR_WAVES];
#pragma unroll
      for (int m = 0; m < TRSM_B / QR_WAVES; ++m) {
        const int ii = wave + QR_WAVES * m;
        acc[m] = (ii < b && c < N) ? Z[(size_t)(i0 + ii) * N + c] : 0.0;
      }
      if (c < N)
        for (int k = i0 + b; k < N; ++k) {
          const double zk = Z[(size_t)k * N + c];
          const int pk = pm[k];
#pragma unroll
          for (int m = 0; m < TRSM_B / QR_WAVES; ++m) {
            const int ii = wave + QR_WAVES * m;
            if (ii < b) acc[m] -= V[(size_t)(i0 + ii) * N + pk] * zk;
          }
        }
#pragma unroll
      for (int m = 0; m < TRSM_B / QR_WAVES; ++m) {
        const int ii = wave + QR_WAVES * m;
        if (ii < b) blk[ii * 64 + lane] = acc[m];
      }
      __syncthreads();
      if (wave == 0) {
        for (int ii = b - 1; ii >= 0; --ii) {
          double s = blk[ii * 64 + lane];
          for (int kk = ii + 1; kk < b; ++kk) s -= rb[ii * TRSM_B + kk] * blk[kk * 64 + lane];
          const double d = rb[ii * TRSM_B + ii];
          if (d == 0.0) zero = true;
          blk[ii * 64 + lane] = s / d;
        }
      }
      __syncthreads();
      if (c < N)
        for (int ii = wave; ii < b; ii += QR_WAVES) Z[(size_t)(i0 + ii) * N + c] = blk[ii * 64 + lane];
      __syncthreads();
    }
  }
  for (size_t e = tid; e < (size_t)N * N; e += QR_THREADS) {
    const int r = (int)(e / N), c = (int)(e % N);
    tk.Y[(size_t)pm[r] * N + c] = Z[e];
  }
  if (zero && st) atomicOr(&st[tk.chain].flags, 2u);
}

size_t qr_blocked_shmem(int nmax) {
  const size_t nb = nmax > 256 ? 8 : 16;
  return sizeof(double) * (nmax * nb + nb * nb + nb * 64 + 64 + 2 * (size_t)nmax) + sizeof(int) * nmax;
}

size_t trsm_shmem(int nmax) {
  return sizeof(double) * (TRSM_B * 64 + TRSM_B * TRSM_B) + sizeof(int) * nmax;
}

void launch_qr_blocked(const QrTask* d, int n, int nmax, ChainState* st, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_qr_blocked, dim3(n), dim3(QR_THREADS), qr_blocked_shmem(nmax), s, d, st);
}

void launch_trsm(const TrsmTask* d, int n, int nmax, ChainState* st, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_trsm_perm, dim3(n), dim3(QR_THREADS), trsm_shmem(nmax), s, d, st);
}

}  // namespace lqmc

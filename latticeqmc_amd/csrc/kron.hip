// Products with the kinetic propagator E_K = exp(-dtau K) of a lattice whose hopping matrix is a
// Kronecker sum, K = Kx (x) I + I (x) Ky (the square and cubic lattices with uniform hopping per
// direction).  Then E_K = Ex (x) Ey exactly, and with the site index i = a n2 + b a column x of
// the operand, read as the n1 x n2 array M[a][b] = x[a n2 + b], maps to Ex M Ey^T.  A product
// E X (N x N, N = n1 n2) costs 2 N^2 (n1 + n2) flops instead of 2 N^3, and each launch reads
// and writes its operands once: at N = 256 a memory-bound pass in place of a 256^3 GEMM.
//
// One 256-thread workgroup per 16 columns of X (left product E X) or 16 rows of X (right
// product X E, handled as E^T X^T).  The tile goes to LDS as V[a'][b'][j] (the k index split
// into its two lattice coordinates), the four wavefronts contract b' with Ey on the matrix
// cores (16x16x4 FP64, one 16 x 16 x 16 product per a'), then a' with Ex (one per b), and
// write the result with the diagonal scales of the task applied.
#include "common.h"

namespace lqmc {

typedef double kr_v4d __attribute__((ext_vector_type(4)));

constexpr int KR_T = 16;             // columns (left) or rows (right) of X per workgroup
constexpr int KR_BS = 17;            // LDS stride of b' (odd: transposed writes spread over banks)
constexpr int KR_AS = 16 * KR_BS;    // LDS stride of a'

__device__ inline kr_v4d kr_mma(double a, double b, kr_v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// RIGHT: C = diag(rs) X diag(ks) E diag(cs), computed as the left product with E^T = Ex^T (x) Ey^T
// applied to the rows of X.
template <bool RIGHT>
__global__ __launch_bounds__(256) void k_kron(const KronTask* __restrict__ tasks, int ntasks) {
  const int tile = blockIdx.x;
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].tile_base <= tile) lo = mid; else hi = mid - 1;
  }
  const KronTask tk = tasks[lo];
  const int N = tk.N, n1 = tk.n1, n2 = tk.n2;
  const int t0 = (tile - tk.tile_base) * KR_T;  // first column (left) or row (right) of the tile
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, li = lane & 15, lk = lane >> 4;
  LQMC_DYNAMIC_LDS(double, vs);  // [16 a'][16 b' (stride KR_BS)][16 j]
  const LQMC_GLOBAL double* X = gptr(tk.X);

  // 1. V[a'][b'][j] = ks[k] X[k][t0 + j] (left) or ks[k] X[t0 + j][k] (right), k = a' n2 + b';
  //    zero outside a' < n1, b' < n2, t0 + j < N.  The fastest index of e runs along memory:
  //    j for the left product (columns of a row), b' for the right one (along a row).
  {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = tid + 256 * u;
      const int j = RIGHT ? e >> 8 : e & 15;
      const int bp = RIGHT ? e & 15 : (e >> 4) & 15;
      const int ap = RIGHT ? (e >> 4) & 15 : e >> 8;
      const bool ok = ap < n1 && bp < n2 && t0 + j < N;
      const int k = ap * n2 + bp;
      const size_t ix = RIGHT ? (size_t)(t0 + j) * tk.ldx + k : (size_t)k * tk.ldx + t0 + j;
      v[u] = ldg0(X, ix, ok);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = tid + 256 * u;
      const int j = RIGHT ? e >> 8 : e & 15;
      const int bp = RIGHT ? e & 15 : (e >> 4) & 15;
      const int ap = RIGHT ? (e >> 4) & 15 : e >> 8;
      const bool ok = ap < n1 && bp < n2 && t0 + j < N;
      const double s = ok ? scale_val(tk.ks_kind, tk.ks, ap * n2 + bp) : 1.0;
      vs[ap * KR_AS + bp * KR_BS + j] = v[u] * s;
    }
  }
  // the A operands: lane holds F[row li][k 4s + lk] of F = Ey (stage 1) and Ex (stage 2), or
  // their transposes for the right product; zero past n2 / n1
  double ay[4], ax[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int kk = 4 * s + lk;
    ay[s] = tk.Ey[RIGHT ? kk * 16 + li : li * 16 + kk];
    ax[s] = tk.Ex[RIGHT ? kk * 16 + li : li * 16 + kk];
  }
  __syncthreads();

  // 2. T[a'][b][j] = sum_b' Ey[b][b'] V[a'][b'][j], over this wavefront's a' (in place: each
  //    a' block is read and written by one wavefront only)
  for (int ap = wave; ap < n1; ap += 4) {
    kr_v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = kr_mma(ay[s], vs[ap * KR_AS + (4 * s + lk) * KR_BS + li], acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) vs[ap * KR_AS + (lk + 4 * r) * KR_BS + li] = acc[r];
  }
  __syncthreads();

  // 3. Y[a][b][j] = sum_a' Ex[a][a'] T[a'][b][j], over this wavefront's b (blocks a' >= n1 hold
  //    zeros from step 1)
  for (int b = wave; b < n2; b += 4) {
    kr_v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = kr_mma(ax[s], vs[(4 * s + lk) * KR_AS + b * KR_BS + li], acc);
    if (!RIGHT) {
      // C[a n2 + b][t0 + j]: 16 consecutive columns per row
      const int c = t0 + li;
      if (c < N) {
        const double cv = scale_val(tk.cs_kind, tk.cs, c);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int a = lk + 4 * r;
          if (a < n1) {
            const int row = a * n2 + b;
            tk.C[(size_t)row * tk.ldc + c] = scale_val(tk.rs_kind, tk.rs, row) * acc[r] * cv;
          }
        }
      }
    } else {
      // back to LDS (this wavefront's b only), written out along rows below
#pragma unroll
      for (int r = 0; r < 4; ++r) vs[(lk + 4 * r) * KR_AS + b * KR_BS + li] = acc[r];
    }
  }
  if (RIGHT) {
    __syncthreads();
    // C[t0 + j][a n2 + b], b fastest: along a row of C
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
      const int e = tid + 256 * u;
      const int j = e >> 8, a = (e >> 4) & 15, b = e & 15;
      const int row = t0 + j, col = a * n2 + b;
      if (a < n1 && b < n2 && row < N)
        tk.C[(size_t)row * tk.ldc + col] =
            scale_val(tk.rs_kind, tk.rs, row) * vs[a * KR_AS + b * KR_BS + j] * scale_val(tk.cs_kind, tk.cs, col);
    }
  }
}

size_t kron_shmem() { return sizeof(double) * 16 * KR_AS; }

void launch_kron(int right, const KronTask* d_tasks, int ntasks, int ntiles, hipStream_t s) {
  if (ntasks <= 0 || ntiles <= 0) return;
  if (right) hipLaunchKernelGGL(k_kron<true>, dim3(ntiles), dim3(256), kron_shmem(), s, d_tasks, ntasks);
  else hipLaunchKernelGGL(k_kron<false>, dim3(ntiles), dim3(256), kron_shmem(), s, d_tasks, ntasks);
}

int kron_tiles(int N) { return (N + KR_T - 1) / KR_T; }

}  // namespace lqmc

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

#include <algorithm>

namespace lqmc {

typedef double kr_v4d __attribute__((ext_vector_type(4)));

constexpr int KR_T = 16;             // columns (left) or rows (right) of X per workgroup
constexpr int KR_BS = 17;            // LDS stride of b' (odd: transposed writes spread over banks)
constexpr int KR_AS = 16 * KR_BS;    // LDS stride of a'

__device__ inline kr_v4d kr_mma(double a, double b, kr_v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Index of the task holding tile `tile` (tasks sorted by tile_base).
__device__ inline int kr_task(const KronTask* __restrict__ tasks, int ntasks, int tile) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].tile_base <= tile) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Element u of this thread's share of tile `tile`'s operand: V[a'][b'][j] with e = tid + 256 u,
// the fastest index of e running along memory (j for the left product, b' for the right one).
template <bool RIGHT>
__device__ inline void kr_coords(int e, int& ap, int& bp, int& j) {
  j = RIGHT ? e >> 8 : e & 15;
  bp = RIGHT ? e & 15 : (e >> 4) & 15;
  ap = RIGHT ? (e >> 4) & 15 : e >> 8;
}

// A diagonal scale entry from its stored value (the kinds of scale_val).
__device__ inline double kr_scale_of(int kind, double v) {
  switch (kind) {
    case SC_PLAIN: return v;
    case SC_INV: return 1.0 / v;
    case SC_INV_BIG: return 1.0 / fmax(v, 1.0);
    case SC_SMALL: return fmin(v, 1.0);
    default: return 1.0;
  }
}

// The tile's operand share (v) and entry tid of the task's three scale vectors (raw).
template <bool RIGHT, bool F16>
__device__ inline void kr_load(const KronTask& tk, int t0, double (&v)[16], double (&sc)[3]) {
  const int tid = threadIdx.x;
  sc[0] = tk.ks_kind != SC_NONE && tid < tk.N ? tk.ks[tid] : 1.0;
  sc[1] = tk.rs_kind != SC_NONE && tid < tk.N ? tk.rs[tid] : 1.0;
  sc[2] = tk.cs_kind != SC_NONE && tid < tk.N ? tk.cs[tid] : 1.0;
  const LQMC_GLOBAL double* X = gptr(tk.X);
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    int ap, bp, j;
    kr_coords<RIGHT>((int)threadIdx.x + 256 * u, ap, bp, j);
    const bool ok = F16 || (ap < tk.n1 && bp < tk.n2 && t0 + j < tk.N);
    const int k = ap * tk.n2 + bp;
    const size_t ix = RIGHT ? (size_t)(t0 + j) * tk.ldx + k : (size_t)k * tk.ldx + t0 + j;
    if constexpr (F16) v[u] = X[ix];
    else v[u] = ldg0(X, ix, ok);
  }
}

// RIGHT: C = diag(rs) X diag(ks) E diag(cs), computed as the left product with E^T = Ex^T (x) Ey^T
// applied to the rows of X.  Each workgroup walks tiles blockIdx.x, + gridDim.x, ...; the next
// tile's operand loads are issued before the current tile's contractions, so they are in flight
// during them.
template <bool RIGHT, bool F16>
__global__ __launch_bounds__(256, 4) void k_kron(const KronTask* __restrict__ tasks, int ntasks, int ntiles) {
  int tile = blockIdx.x;
  if (tile >= ntiles) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, li = lane & 15, lk = lane >> 4;
  LQMC_DYNAMIC_LDS(double, vs);  // [16 a'][16 b' (stride KR_BS)][16 j], then ks, rs, cs [256] each
  double* const ksl = vs + 16 * KR_AS;
  double* const rsl = ksl + 256;
  double* const csl = rsl + 256;
  KronTask tk = tasks[kr_task(tasks, ntasks, tile)];
  int t0 = (tile - tk.tile_base) * KR_T;  // first column (left) or row (right) of the tile
  double v[16], sc[3];
  kr_load<RIGHT, F16>(tk, t0, v, sc);
  for (;;) {
    // F16: every task of the launch is 16 x 16 (N = 256): no tile reaches past the lattice
    const int N = F16 ? 256 : tk.N, n1 = F16 ? 16 : tk.n1, n2 = F16 ? 16 : tk.n2;
    // 1. V[a'][b'][j] = X[k][t0 + j] (left) or X[t0 + j][k] (right), k = a' n2 + b'; zero outside
    //    a' < n1, b' < n2, t0 + j < N.  The scales to LDS (ks is applied to the stage-1 operands).
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      int ap, bp, j;
      kr_coords<RIGHT>(tid + 256 * u, ap, bp, j);
      vs[ap * KR_AS + bp * KR_BS + j] = v[u];
    }
    ksl[tid] = kr_scale_of(tk.ks_kind, sc[0]);
    rsl[tid] = kr_scale_of(tk.rs_kind, sc[1]);
    csl[tid] = kr_scale_of(tk.cs_kind, sc[2]);
    // the A operands: lane holds F[row li][k 4s + lk] of F = Ey (stage 1) and Ex (stage 2), or
    // their transposes for the right product; zero past n2 / n1
    double ay[4], ax[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int kk = 4 * s + lk;
      ay[s] = tk.Ey[RIGHT ? kk * 16 + li : li * 16 + kk];
      ax[s] = tk.Ex[RIGHT ? kk * 16 + li : li * 16 + kk];
    }
    // the next tile's operand, in flight during the contractions below
    const int nt = tile + (int)gridDim.x;
    const bool more = nt < ntiles;
    KronTask tn = tk;
    int n0 = 0;
    if (more) {
      tn = tasks[kr_task(tasks, ntasks, nt)];
      n0 = (nt - tn.tile_base) * KR_T;
      kr_load<RIGHT, F16>(tn, n0, v, sc);
    }
    __syncthreads();

    // 2. T[a'][b][j] = sum_b' Ey[b][b'] ks[k] V[a'][b'][j] for this wavefront's a' = wave + 8h + 4q
    //    (in place: each a' block is read and written by one wavefront only; blocks past n1
    //    hold zeros and stay zero)
    // two a' blocks at a time (register budget of four workgroups per compute unit)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      kr_v4d acc[2];
      double bv[2][4];
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int ap = wave + 8 * h + 4 * q, bp = 4 * s + lk;
          const double kf = F16 || (ap < n1 && bp < n2) ? ksl[ap * n2 + bp] : 0.0;
          bv[q][s] = vs[ap * KR_AS + bp * KR_BS + li] * kf;
        }
#pragma unroll
      for (int q = 0; q < 2; ++q) acc[q] = kr_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int q = 0; q < 2; ++q) acc[q] = kr_mma(ay[s], bv[q][s], acc[q]);
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) vs[(wave + 8 * h + 4 * q) * KR_AS + (lk + 4 * r) * KR_BS + li] = acc[q][r];
    }
    __syncthreads();

    // 3. Y[a][b][j] = sum_a' Ex[a][a'] T[a'][b][j] for this wavefront's b = wave + 8h + 4q
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      kr_v4d acc[2];
      double bv[2][4];
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int s = 0; s < 4; ++s) bv[q][s] = vs[(4 * s + lk) * KR_AS + (wave + 8 * h + 4 * q) * KR_BS + li];
#pragma unroll
      for (int q = 0; q < 2; ++q) acc[q] = kr_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int q = 0; q < 2; ++q) acc[q] = kr_mma(ax[s], bv[q][s], acc[q]);
      if (!RIGHT) {
        // C[a n2 + b][t0 + j]: 16 consecutive columns per row
        const int c = t0 + li;
        if (F16 || c < N) {
          const double cv = csl[c];
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const int b = wave + 8 * h + 4 * q;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int a = lk + 4 * r;
              if (F16 || (a < n1 && b < n2)) {
                const int row = a * n2 + b;
                tk.C[(size_t)row * tk.ldc + c] = rsl[row] * acc[q][r] * cv;
              }
            }
          }
        }
      } else {
        // back to LDS (this wavefront's b only), written out along rows below
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
          for (int r = 0; r < 4; ++r) vs[(lk + 4 * r) * KR_AS + (wave + 8 * h + 4 * q) * KR_BS + li] = acc[q][r];
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
        if (F16 || (a < n1 && b < n2 && row < N))
          tk.C[(size_t)row * tk.ldc + col] =
              rsl[row] * vs[a * KR_AS + b * KR_BS + j] * csl[col];
      }
    }
    if (!more) break;
    __syncthreads();  // every read of this tile's LDS is done before the next tile is staged
    tile = nt;
    tk = tn;
    t0 = n0;
  }
}

size_t kron_shmem() { return sizeof(double) * (16 * KR_AS + 3 * 256); }  // 40 KiB: four per CU

// Grid of a launch over ntiles tiles: four workgroups per compute unit (what the LDS of one
// allows), each walking ntiles / grid tiles.
int kron_grid(int ntiles, int cus) { return std::max(1, std::min(ntiles, 4 * std::max(cus, 1))); }

// all16: every task of the launch has n1 = n2 = 16.
void launch_kron(int right, int all16, const KronTask* d_tasks, int ntasks, int ntiles, int cus, hipStream_t s) {
  if (ntasks <= 0 || ntiles <= 0) return;
  const int grid = kron_grid(ntiles, cus);
  if (right && all16) hipLaunchKernelGGL((k_kron<true, true>), dim3(grid), dim3(256), kron_shmem(), s, d_tasks, ntasks, ntiles);
  else if (right) hipLaunchKernelGGL((k_kron<true, false>), dim3(grid), dim3(256), kron_shmem(), s, d_tasks, ntasks, ntiles);
  else if (all16) hipLaunchKernelGGL((k_kron<false, true>), dim3(grid), dim3(256), kron_shmem(), s, d_tasks, ntasks, ntiles);
  else hipLaunchKernelGGL((k_kron<false, false>), dim3(grid), dim3(256), kron_shmem(), s, d_tasks, ntasks, ntiles);
}

int kron_tiles(int N) { return (N + KR_T - 1) / KR_T; }

}  // namespace lqmc

// Batched GEMM on CDNA4 matrix cores for the slice products and wraps.
//
// One launch covers a heterogeneous list of tasks (chains of different N, both
// spins, several product kinds); each 256-thread workgroup computes one 64x64
// tile of one task.  Diagonal factors of the DQMC algebra (V = exp(-s*lambda*h),
// the UDV D splits) are fused: diag(rs) and diag(cs) in the epilogue, diag(ks)
// on the A tile as it is staged in LDS.  Matrix-core ops: v_mfma_f64_16x16x4f64
// (FP64) and v_mfma_f32_16x16x4f32 (FP32), 2x2 of them per wavefront per k-step.
#include "common.h"

namespace lqmc {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
// Accumulator layout of the 16x16 output block: lane holds column lane%16 and
// four rows, row(lane, r).  FP64: lane/16 + 4r;  FP32: 4*(lane/16) + r.
template <> struct Mfma<double> {
  using acc_t = v4d;
  __device__ static inline acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
  using acc_t = v4f;
  __device__ static inline acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return ((lane >> 4) << 2) + r; }
};

constexpr int BM = 64, BN = 64, BK = 16, PAD = 4;
constexpr int GEMM_KMAX = 1024;  // longest inner dimension (the chains' N <= 1024)

// One 64x64 tile of C = alpha diag(rs) op(A) diag(ks) op(B) diag(cs) + beta C for one
// transpose combination (TA, TB known at compile time, so every operand load is one load
// through a computed index).  The k-scaling factors of the whole K range are formed once, at
// the start, into LDS and applied as the A tile is stored; the row and column factors of the
// epilogue are formed once per lane.
//  FULL: every tile of the task lies inside C and K is a multiple of BK, so the operand
//  loads need no bounds checks.
template <typename T, int TA, int TB, bool FULL>
__device__ void gemm_tile(const GemmTask& tk, int m0, int n0, T (*As)[BM + PAD], T (*Bs)[BN + PAD],
                          double* ksv) {
  using acc_t = typename Mfma<T>::acc_t;
  const int M = tk.M, N = tk.N, K = tk.K;
  const LQMC_GLOBAL T* __restrict__ A = gptr(static_cast<const T*>(tk.A));
  const LQMC_GLOBAL T* __restrict__ B = gptr(static_cast<const T*>(tk.B));
  T* __restrict__ C = static_cast<T*>(tk.C);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const bool kscale = tk.ks_kind != SC_NONE;

  // global->register staging coordinates
  // A: non-trans: row m = tid/4, k = (tid%4)*4 + q ; trans: k = tid/16, m = (tid%16)*4 + q
  // B: non-trans: k = tid/16, n = (tid%16)*4 + q ; trans: n = tid/4, k = (tid%4)*4 + q
  T ra[4], rb[4];
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int m = TA ? m0 + ((tid & 15) << 2) + q : m0 + (tid >> 2);
      const int k = TA ? k0 + (tid >> 4) : k0 + ((tid & 3) << 2) + q;
      const size_t ia = TA ? (size_t)k * tk.lda + m : (size_t)m * tk.lda + k;
      ra[q] = (FULL || (m < M && k < K)) ? A[ia] : T(0);
      const int n = TB ? n0 + (tid >> 2) : n0 + ((tid & 15) << 2) + q;
      const int kb = TB ? k0 + ((tid & 3) << 2) + q : k0 + (tid >> 4);
      const size_t ib = TB ? (size_t)n * tk.ldb + kb : (size_t)kb * tk.ldb + n;
      rb[q] = (FULL || (n < N && kb < K)) ? B[ib] : T(0);
    }
  };
  auto store_tiles = [&](int k0) {
    // ksv holds the k-scaling factors of the whole K range (formed before the first barrier)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int ka = TA ? (tid >> 4) : ((tid & 3) << 2) + q;
      const T av = kscale ? ra[q] * T(ksv[k0 + ka]) : ra[q];
      if (!TA) As[((tid & 3) << 2) + q][tid >> 2] = av;
      else As[tid >> 4][((tid & 15) << 2) + q] = av;
      if (!TB) Bs[tid >> 4][((tid & 15) << 2) + q] = rb[q];
      else Bs[((tid & 3) << 2) + q][tid >> 2] = rb[q];
    }
  };

  acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  const int nk = (K + BK - 1) / BK;
  load_tiles(0);
  if (kscale)
    for (int k = tid; k < nk * BK; k += 256) ksv[k] = k < K ? scale_val(tk.ks_kind, tk.ks, k) : 1.0;
  __syncthreads();
  store_tiles(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) load_tiles((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      T a0 = As[kr][wm * 32 + (lane & 15)];
      T a1 = As[kr][wm * 32 + 16 + (lane & 15)];
      T b0 = Bs[kr][wn * 32 + (lane & 15)];
      T b1 = Bs[kr][wn * 32 + 16 + (lane & 15)];
      acc[0][0] = Mfma<T>::mma(a0, b0, acc[0][0]);
      acc[0][1] = Mfma<T>::mma(a0, b1, acc[0][1]);
      acc[1][0] = Mfma<T>::mma(a1, b0, acc[1][0]);
      acc[1][1] = Mfma<T>::mma(a1, b1, acc[1][1]);
    }
    __syncthreads();  // every read of As and Bs for this tile is done
    if (kt + 1 < nk) {
      store_tiles((kt + 1) * BK);
      __syncthreads();
    }
  }

  // epilogue: lane holds rows Mfma<T>::row(lane, r), column lane%16 of each 16x16 block
  const bool plain = tk.rs_kind == SC_NONE && tk.cs_kind == SC_NONE && tk.alpha == 1.0 && tk.beta == 0.0;
  double rsv[2][4], csv[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int col = n0 + wn * 32 + j * 16 + (lane & 15);
    csv[j] = plain || col >= N ? 1.0 : scale_val(tk.cs_kind, tk.cs, col);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = m0 + wm * 32 + i * 16 + Mfma<T>::row(lane, r);
      rsv[i][r] = plain || row >= M ? 1.0 : tk.alpha * scale_val(tk.rs_kind, tk.rs, row);
    }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn * 32 + j * 16 + (lane & 15);
      if (col >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 32 + i * 16 + Mfma<T>::row(lane, r);
        if (row >= M) continue;
        T* p = C + (size_t)row * tk.ldc + col;
        if (plain) {
          *p = acc[i][j][r];
        } else {
          double v = rsv[i][r] * (double)acc[i][j][r] * csv[j];
          if (tk.beta != 0.0) v += tk.beta * (double)(*p);
          *p = (T)v;
        }
      }
    }
}

// One kernel per (transpose case, full tiles) class: a launch holds tasks of one class
// only, so each kernel keeps the registers of a single tile body.
template <typename T, int TA, int TB, bool FULL>
__global__ __launch_bounds__(256) void k_gemm(const GemmTask* __restrict__ tasks, int ntasks) {
  // the tiles of one product (sharing the operands' row and column panels) on one XCD
  const int tile = xcd_grouped_block();
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].tile_base <= tile) lo = mid; else hi = mid - 1;
  }
  const GemmTask tk = tasks[lo];
  const int local = tile - tk.tile_base;
  const int m0 = (local / tk.tiles_n) * BM;
  const int n0 = (local % tk.tiles_n) * BN;
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];
  __shared__ double ksv[GEMM_KMAX];
  gemm_tile<T, TA, TB, FULL>(tk, m0, n0, As, Bs, ksv);
}

int gemm_class(const GemmTask& t) {
  const int full = t.M % BM == 0 && t.N % BN == 0 && t.K % BK == 0;
  return (t.transA ? 4 : 0) + (t.transB ? 2 : 0) + full;
}

template <typename T>
static void launch_gemm_t(int cls, const GemmTask* d_tasks, int ntasks, int ntiles, hipStream_t s) {
  switch (cls) {
#define LQMC_GEMM_CASE(C, TA, TB, F) \
    case C: hipLaunchKernelGGL((k_gemm<T, TA, TB, F>), dim3(ntiles), dim3(256), 0, s, d_tasks, ntasks); break;
    LQMC_GEMM_CASE(0, 0, 0, false)
    LQMC_GEMM_CASE(1, 0, 0, true)
    LQMC_GEMM_CASE(2, 0, 1, false)
    LQMC_GEMM_CASE(3, 0, 1, true)
    LQMC_GEMM_CASE(4, 1, 0, false)
    LQMC_GEMM_CASE(5, 1, 0, true)
    LQMC_GEMM_CASE(6, 1, 1, false)
    LQMC_GEMM_CASE(7, 1, 1, true)
#undef LQMC_GEMM_CASE
    default: break;
  }
}

void launch_gemm(int fp32, int cls, const GemmTask* d_tasks, int ntasks, int ntiles, hipStream_t s) {
  if (ntasks <= 0 || ntiles <= 0) return;
  if (fp32) launch_gemm_t<float>(cls, d_tasks, ntasks, ntiles, s);
  else launch_gemm_t<double>(cls, d_tasks, ntasks, ntiles, s);
}

int gemm_tiles(int M, int N, int* tiles_n) {
  *tiles_n = (N + BN - 1) / BN;
  return ((M + BM - 1) / BM) * (*tiles_n);
}

}  // namespace lqmc

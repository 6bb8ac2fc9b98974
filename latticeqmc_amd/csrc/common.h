// Device-side descriptors shared by the engine (host) and the kernels.
#pragma once
#ifndef LQMC_CPU_EMU
#include <hip/hip_runtime.h>
// Dynamic LDS of a kernel (redirected to a host buffer by the CPU bounds-check build).
#define LQMC_DYNAMIC_LDS(type, name) extern __shared__ type name[]
// Global address space: loads through such pointers are global_load (vector-memory
// counter only) rather than flat loads, which also hold the LDS counter.
#define LQMC_GLOBAL __attribute__((address_space(1)))
// The value of x (in an SGPR: "s", a VGPR: "v") passed through an empty asm: expressions
// formed from it afterwards are new to the compiler and are not merged with earlier ones.
#define LQMC_OPAQUE(x, reg) asm volatile("" : "+" reg(x))
#endif
#include <stdint.h>

template <typename T>
__device__ inline LQMC_GLOBAL T* gptr(T* p) {
  return (LQMC_GLOBAL T*)p;
}

// Block index regrouped by XCD.  Workgroups are dispatched to the eight XCDs round-robin
// (block b on XCD b % 8), each XCD with its own L2.  The returned index runs consecutively
// over the blocks of one XCD, so neighbouring indices (the tiles or column blocks of one
// matrix, which read the same operand panels) share an L2; blocks past the last multiple of
// 8 keep their index.
__device__ inline int xcd_grouped_block() {
  const int q8 = (int)gridDim.x >> 3, b = (int)blockIdx.x;
  return b < (q8 << 3) ? (b & 7) * q8 + (b >> 3) : b;
}

// p[ok ? i : 0], then 0 where !ok: the load is unconditional (no branch, so loads of
// a loop body can all be in flight together) and always addresses valid memory.
template <typename T>
__device__ inline T ld0(const T* p, size_t i, bool ok) {
  const T v = p[ok ? i : 0];
  return ok ? v : T(0);
}
template <typename T>
__device__ inline T ldg0(const LQMC_GLOBAL T* p, size_t i, bool ok) {
  const T v = p[ok ? i : 0];
  return ok ? v : T(0);
}

namespace lqmc {

// Static description of one Markov chain (read-only on the device).
struct ChainDev {
  int N, L;
  int mode;            // 0 dqmc, 1 reference
  int pad0;
  double lam;
  double e1m, e1p;     // exp(-lambda), exp(+lambda): V entries for sigma*s = +1 / -1
  double e2p, e2m;     // exp(+2 lambda), exp(-2 lambda)
  uint64_t seed;
  int8_t* field;       // [L][N]
  double* ev;          // [2][L][N] exp(-sigma*lambda*s)
  double* evi;         // [2][L][N] 1/ev
  void* G;             // [2][N][N] in the form's precision
  double* Gacc;        // [2][N][N] sum of measured G
  double* Glast;       // [2][N][N] last measured G
  double* Sacc;        // [N][N] sum over measured sweeps of <S^z_i S^z_j>, S^z = n_up - n_dn
  double* hop;         // [N][N] off-diagonal part of K
  const double* expk;  // [N][N] E_K = expm(-dtau K)
  double* uni;         // replay uniforms [n_replay][L][N]
  double* tr_ratio;    // [L][N] trace
  uint8_t* tr_acc;     // [L][N] trace
  double* obs;         // [obs_cap][3]
  void* dws;           // delayed-update workspace: Z tiles of both spins in register order
  int obs_cap;
  int pad1;
};

// Mutable per-chain counters (device).
struct ChainState {
  unsigned long long sweep;        // completed sweeps (global sweep index)
  unsigned long long replay_base;  // sweep index of uni[0]
  unsigned long long replay_count; // sweeps held in uni
  unsigned long long attempted;
  unsigned long long accepted;
  unsigned long long measured;
  unsigned long long recomputes;
  unsigned long long max_err_bits; // non-negative double as ordered bits
  unsigned int flags;
  unsigned int pad;
};

// Diagonal-scale kinds applied to GEMM rows (rs), inner index (ks) or columns (cs).
enum ScaleKind : int { SC_NONE = 0, SC_PLAIN = 1, SC_INV = 2, SC_INV_BIG = 3, SC_SMALL = 4 };

// Entry idx of a diagonal scale of the given kind (1 for SC_NONE).
__device__ inline double scale_val(int kind, const double* s, int idx) {
  if (kind == SC_NONE) return 1.0;
  double v = s[idx];
  switch (kind) {
    case SC_PLAIN: return v;
    case SC_INV: return 1.0 / v;
    case SC_INV_BIG: return 1.0 / fmax(v, 1.0);
    case SC_SMALL: return fmin(v, 1.0);
    default: return 1.0;
  }
}

// C = alpha * diag(rs) * op(A) * diag(ks) * op(B) * diag(cs) + beta * C
struct GemmTask {
  const void* A;
  const void* B;
  void* C;
  const double* rs;
  const double* ks;
  const double* cs;
  int M, N, K;
  int lda, ldb, ldc;
  int transA, transB;
  int rs_kind, ks_kind, cs_kind;
  int tiles_n;      // tiles along N
  int tile_base;    // first tile index of this task in the launch
  int pad;
  double alpha, beta;
};

// C = diag(rs) E diag(ks) X diag(cs) (left) or C = diag(rs) X diag(ks) E diag(cs) (right) for
// a Kronecker-factored E = Ex (x) Ey (site i = a n2 + b; Ex n1 x n1, Ey n2 x n2, n1, n2 <= 16,
// each stored as a 16 x 16 row-major block, zero outside).  X and C are N x N, N = n1 n2.
struct KronTask {
  const double* X;
  double* C;
  const double* Ex;
  const double* Ey;
  const double* rs;
  const double* ks;
  const double* cs;
  int N, n1, n2;
  int ldx, ldc;
  int rs_kind, ks_kind, cs_kind;
  int tile_base;  // first 16-column (left) or 16-row (right) tile of this task in the launch
};

// Pre-pivoted blocked Householder QR: W = Q diag(D) V'' with V'' written over W.
struct QrTask {
  double* W;   // in: matrix; out: V'' = D^-1 R P^T, or R' = D^-1 R when perm is requested
  double* Q;   // out: orthogonal factor
  double* D;   // out: |diag R|
  double* A;   // scratch N x N (holds R and the reflectors)
  int* perm;   // optional out: pivot order (W then holds R' unpermuted)
  double* Tt;  // scratch: compact-WY T of every panel, 16 * (N + 16) doubles
  int N;
  int chain;
};

// Solve R' Y' = Z, Y[perm[r]] = Y'[r], for the R' and pivot order of a QrTask (Z destroyed).
struct TrsmTask {
  const double* V;
  const int* perm;
  double* Z;
  double* Y;
  int N;
  int chain;
};

// Solve [X | RHS] (N x 2N, ld 2N) in place; RHS built from U (transposed,
// scaled by 1/max(D,1)) or the identity; optional copy of Y to Yout (ld N).
struct SolveTask {
  double* aug;
  const double* U;   // null -> RHS = I
  const double* D;
  double* Yout;      // optional
  int N;
  int add_identity;  // X += I before solving (reference-mode M = I + P)
  int chain;
  int pad;
};

// Compare the propagated G (form precision) with the recomputed Gst (f64),
// record the max deviation (per chain, and per recompute point in slot), raise the
// precision flag above tol, then overwrite G by Gst.
struct FinishTask {
  void* G;
  const double* Gst;
  unsigned long long* slot;  // max deviation at this recompute point (ordered bits), or null
  double tol;                // > 0: flag 32 when the deviation exceeds it
  int N;
  int chain;
  int compare;
  int pad;
};

// Arguments of the flip kernels for step t (slice l = L-1-t of each chain).
struct FlipArgs {
  const ChainDev* chains;
  ChainState* st;
  int t;
  int measure;
  int rng;
  int trace;
  int delay;
  int wpc;        // workgroups per chain of the delayed kernel (0 or 1: one)
  unsigned* bar;  // [nchains][4] barrier counters and XCC bits (wpc > 1), zero between launches
  int nchains;
  int bf3;        // delayed FP32 form: rank-k products as three split-BF16 matrix-core products
  int yrows;      // delayed flip, k = 16, one spin per workgroup: the block's rows of G and its Z tiles
                  // kept in LDS (flip_delayed_stage_rows; flip_delayed_rows_shmem more bytes of LDS)
};

}  // namespace lqmc

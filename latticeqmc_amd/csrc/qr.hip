// Blocked FP64 factorisations for the UDV stabilisation.
//
// k_qr_blocked (one workgroup per matrix): pre-pivoted Householder QR, W P = Q R
//   with the columns of W sorted once by decreasing norm (pre-pivoting, Bai, Lee,
//   Li, Xu 2011) and panel-local column pivoting: each panel of nb columns is
//   factored in LDS by one wavefront and applied to the trailing matrix in
//   compact-WY form (I - V T V^T)^T A = A - V T^T (V^T A) on the matrix cores.
//   Outputs D = |diag R| and V'' = D^-1 R P^T written over W (or R' = D^-1 R and
//   the pivot order when perm is requested), the reflectors below the diagonal of
//   the scratch matrix A and the T factor of every panel in Tt.
// k_qr_formq (one workgroup per 64-column block of Q): Q[:, c0:c1] = H_0 ... H_{c1-1}
//   E[:, c0:c1] by backward blocked accumulation from A and Tt; the column blocks of
//   Q are independent, so the explicit Q is formed by many workgroups per matrix.
// k_trsm_perm (one workgroup per 64-column block): solves R' Y' = Z by block back
//   substitution on 16-row blocks (off-diagonal updates on the matrix cores) and
//   scatters Y[perm[r]] = Y'[r].
#include <algorithm>

#include "common.h"

namespace lqmc {

constexpr int QR_THREADS = 1024;
constexpr int QR_WAVES = QR_THREADS / 64;
constexpr int QF_COLS = 64;                 // columns per workgroup in k_qr_formq / k_trsm_perm
constexpr int QF_THREADS = 64 * QF_COLS / 16;
constexpr int TRSM_B = 16;

__device__ inline double qr_wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline int qr_panel_width(int N) { return N > 256 ? 8 : 16; }

typedef double qr_v4d __attribute__((ext_vector_type(4)));

__device__ inline qr_v4d qr_mma(double a, double b, qr_v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Keeps the instruction scheduler from moving code across this point: the loads issued
// before it are all in flight before the arithmetic after it starts (left to itself the
// scheduler interleaves each LDS load with its use and waits on every one).
__device__ inline void qr_sched_fence() { __builtin_amdgcn_sched_barrier(0); }

// X[r0:N, c_begin:c_end] <- (I - V T' V^T) X[r0:N, c_begin:c_end] on the FP64 matrix
// cores, V = pnl (explicit reflectors, rows r0..N-1, w <= 16 columns, row stride ps),
// T' = T^T (trans) or T (w x w, leading dimension nb).  One wavefront per 16-column tile
// (NW wavefronts in the workgroup): W = V^T X (K = rows), W' = T' W, X -= V W'.  In the
// 16x16x4 FP64 layout the accumulator element v of lane l is row l/16 + 4v, column
// l%16 -- exactly the B operand lane l supplies for k-step v -- so W and W' chain
// through registers.  A tile is one wavefront's serial walk down the rows, so its loads
// are unconditional (clamped: zero beyond row N, which adds nothing to the products) and
// issued in batches ahead of the products -- KB k-steps of V^T X, RB 16-row tiles of the
// update -- to keep many global loads in flight instead of waiting on each.  That pays
// where a few wavefronts per workgroup share a compute unit with other workgroups (Q
// formation: KB = 8, RB = 4); with 16 wavefronts of one factorisation all streaming tiles
// the compute unit's memory path is the limit, and row batches only add register pressure
// (KB = 8, RB = 1).  Ends with a workgroup barrier.
template <int NW, int KB, int RB>
__device__ void qr_apply_block(double* X, int N, int r0, int c_begin, int c_end, const double* pnl, int ps,
                               int w, const double* Tm, int nb, bool trans) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nct = (c_end - c_begin + 15) >> 4;
  LQMC_GLOBAL double* Xg = gptr(X);
  for (int ct = wave; ct < nct; ct += NW) {
    const int c = c_begin + ct * 16 + li;
    const bool cok = c < c_end;
    qr_v4d wt = {0.0, 0.0, 0.0, 0.0};
    for (int kk = r0; kk < N; kk += 4 * KB) {
      double a[KB], b[KB];
#pragma unroll
      for (int u = 0; u < KB; ++u) {
        const int r = kk + 4 * u + lk;
        a[u] = ld0(pnl, (size_t)r * ps + li, r < N && li < w);
        b[u] = ldg0(Xg, (size_t)r * N + c, r < N && cok);
      }
      qr_sched_fence();
#pragma unroll
      for (int u = 0; u < KB; ++u) wt = qr_mma(a[u], b[u], wt);
    }
    qr_v4d w2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int p = 4 * s + lk;
      const double a = ld0(Tm, trans ? (size_t)p * nb + li : (size_t)li * nb + p, li < w && p < w);
      w2 = qr_mma(a, wt[s], w2);
    }
    for (int rt = r0; rt < N; rt += 16 * RB) {
      qr_v4d acc[RB];
      double a[RB][4];
#pragma unroll
      for (int t = 0; t < RB; ++t) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = rt + 16 * t + lk + 4 * v;
          acc[t][v] = ldg0(Xg, (size_t)row * N + c, row < N && cok);
        }
        const int ra = rt + 16 * t + li;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int q = 4 * s + lk;
          a[t][s] = -ld0(pnl, (size_t)ra * ps + q, ra < N && q < w);
        }
      }
      qr_sched_fence();
#pragma unroll
      for (int t = 0; t < RB; ++t) {
#pragma unroll
        for (int s = 0; s < 4; ++s) acc[t] = qr_mma(a[t][s], w2[s], acc[t]);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = rt + 16 * t + lk + 4 * v;
          if (row < N && cok) Xg[(size_t)row * N + c] = acc[t][v];
        }
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

// Lane exchange inside 16-lane rows through the DPP network (a VALU operation, no LDS
// round trip); CTRL is the GFX9 dpp_ctrl encoding, applied to both halves of a double.
template <int CTRL>
__device__ inline double qr_dpp(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Column sums over the 64 lanes of a wavefront: lane l holds NB (8 or 16) partial sums
// v[c]; on return the NB totals are in the LDS scratch red[0..NB).  Two halving
// exchanges across the 16-lane rows (offset 32: lanes with bit 5 set keep the upper half
// of their values and trade the lower half with their partner; then offset 16 on the
// remaining half) leave NB/4 columns per lane, (NB/2) b5 + (NB/4) b4 + i; each is then
// summed over its 16-lane row by four DPP steps (swap within pairs, within quads, rotate
// the row by 4, by 8).  Reading both halves into values before the conditional keeps v
// in registers (a conditional over the elements would select their addresses).
// The two lane-half exchanges as gfx950 permlane swaps (VALU, no LDS round trip): after
// qr_swap<32>(a, b), lanes 0..31 of a + b hold a summed over lanes l and l + 32 and lanes
// 32..63 hold b summed likewise; qr_swap<16> does the same for 16-lane rows (0 with 1, 2 with 3).
template <int HALF>
__device__ inline void qr_swap(double& a, double& b) {
  const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
  const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
  unsigned alo = (unsigned)ua, ahi = (unsigned)(ua >> 32), blo = (unsigned)ub, bhi = (unsigned)(ub >> 32);
  if constexpr (HALF == 32) {
    const auto l = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
    alo = l[0]; blo = l[1]; ahi = h[0]; bhi = h[1];
  } else {
    const auto l = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
    alo = l[0]; blo = l[1]; ahi = h[0]; bhi = h[1];
  }
  a = __longlong_as_double((long long)(((unsigned long long)ahi << 32) | alo));
  b = __longlong_as_double((long long)(((unsigned long long)bhi << 32) | blo));
}

template <int NB>
__device__ inline void qr_wave_colsum(double (&v)[NB], double* red, int lane) {
  static_assert(NB == 8 || NB == 16, "panel width");
  // column i and i + NB/2 halve over lanes l, l + 32 (then i, i + NB/4 over rows of 16 lanes):
  // the same pairs of operands as an exchange with lane l ^ 32 (l ^ 16), so the same sums
#pragma unroll
  for (int i = 0; i < NB / 2; ++i) {
    double a = v[i], b = v[NB / 2 + i];
    qr_swap<32>(a, b);
    v[i] = a + b;
  }
#pragma unroll
  for (int i = 0; i < NB / 4; ++i) {
    double a = v[i], b = v[NB / 4 + i];
    qr_swap<16>(a, b);
    v[i] = a + b;
  }
#pragma unroll
  for (int i = 0; i < NB / 4; ++i) {
    double x = v[i];
    x += qr_dpp<0xB1>(x);   // quad_perm [1,0,3,2]
    x += qr_dpp<0x4E>(x);   // quad_perm [2,3,0,1]
    x += qr_dpp<0x124>(x);  // row_ror:4
    x += qr_dpp<0x128>(x);  // row_ror:8
    v[i] = x;
  }
  qr_wave_sync();  // earlier readers of red are done
  if ((lane & 15) == 0) {
#pragma unroll
    for (int i = 0; i < NB / 4; ++i) red[(NB / 2) * (lane >> 5) + (NB / 4) * ((lane >> 4) & 1) + i] = v[i];
  }
  qr_wave_sync();
}

// Column sums over rows r0..N-1 of the panel, reduced over the wavefront into red[0..NB):
// the squares of every column (SQ) or the products of column q with every column.
template <int NB, bool SQ>
__device__ inline void qr_panel_gram(const double* pnl, int N, int r0, int q, double* red, int lane) {
  constexpr int PS = NB + 1;
  double part[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) part[c] = 0.0;
  for (int r = r0 + lane; r < N; r += 64) {
    const double* row = pnl + r * PS;
    const double x = SQ ? 1.0 : row[q];
#pragma unroll
    for (int c0 = 0; c0 < NB; c0 += 8) {  // batches of 8 loads in flight
      if (!SQ && c0 + 8 <= q) continue;   // columns before the pivot are not needed
      double rv[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) rv[c] = row[c0 + c];
      qr_sched_fence();
#pragma unroll
      for (int c = 0; c < 8; ++c) part[c0 + c] += (SQ ? rv[c] : x) * rv[c];
    }
  }
  qr_wave_colsum<NB>(part, red, lane);
}

// Householder factorisation of the panel (rows j0..N-1, w <= NB columns, in LDS with row
// stride NB + 1) by one wavefront, with panel-local column pivoting: at step q the
// remaining panel column of largest norm over rows j..N-1 is swapped into place (and its
// pivot index with it).  One wavefront reduction per column: y_c = x^T col_c over rows
// j..N-1 for the pivot column x and every panel column c gives the exact |x|^2 = y_q for the
// reflector and, with u = x - beta e_j, the products u^T col_c = y_c - beta col_c[j] of the
// update.  The pivot choice uses the column norms downdated after each step,
// |col_c|^2 -= R[j][c]^2 (Householder reflections keep column norms), with the norms of
// every column recomputed when one of them has lost all but sqrt(eps) of its value since
// its last exact computation (the safeguard of LAPACK's xLAQP2).  The column swaps are
// recorded in swp[] for the R entries above the panel, which the caller permutes
// afterwards.  nrm, org: NB-entry LDS scratch (current and reference squared norms); flg:
// one int.  On return pnl holds R (on/above the diagonal) and the reflectors below it.
template <int NB>
__device__ void qr_panel_wave(double* pnl, int N, int j0, int w, int* perm, int* swp, double* taus,
                              double* red, double* nrm, double* org, int* flg, bool& zero_pivot) {
  const int lane = threadIdx.x & 63;
  constexpr int PS = NB + 1;  // padded row stride: lanes walking rows spread over the banks
  constexpr double kTol = 1.4901161193847656e-08;  // sqrt(eps), eps = 2^-52
  qr_panel_gram<NB, true>(pnl, N, j0, 0, red, lane);
  if (lane < NB) nrm[lane] = org[lane] = red[lane];
  qr_wave_sync();
  for (int q = 0; q < w; ++q) {
    const int j = j0 + q;
    int p = q;
    double best = -1.0;
#pragma unroll
    for (int c = 0; c < NB; ++c) {  // static trip count: the NB loads are issued together
      const double nc = nrm[c];
      if (c >= q && c < w && nc > best) { best = nc; p = c; }
    }
    qr_wave_sync();  // every lane has read nrm before lane 0 swaps entries
    if (p != q) {
      for (int r = j0 + lane; r < N; r += 64) {
        const double t0 = pnl[r * PS + q];
        pnl[r * PS + q] = pnl[r * PS + p];
        pnl[r * PS + p] = t0;
      }
      if (lane == 0) {
        const int tp = perm[j0 + q];
        perm[j0 + q] = perm[j0 + p];
        perm[j0 + p] = tp;
        const double tn = nrm[q], to = org[q];
        nrm[q] = nrm[p]; org[q] = org[p];
        nrm[p] = tn; org[p] = to;
      }
    }
    if (lane == 0) { swp[q] = p; *flg = 0; }
    qr_wave_sync();
    qr_panel_gram<NB, false>(pnl, N, j, q, red, lane);
    const double alpha = pnl[j * PS + q];
    const double xx = red[q];
    const double nrmv = sqrt(xx > 0.0 ? xx : 0.0);
    double beta, tau, scal;
    if (nrmv == 0.0) {
      beta = 0.0; tau = 0.0; scal = 0.0;
      zero_pivot = true;
    } else {
      beta = alpha >= 0.0 ? -nrmv : nrmv;
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    // dp_c = tau v^T col_c with v = u / (alpha - beta); zero outside (q, w)
    double dp[NB];
    {
      const double* rowj = pnl + j * PS;
      const double ts = tau * scal;
#pragma unroll
      for (int c = 0; c < NB; ++c) dp[c] = (c > q && c < w) ? ts * (red[c] - beta * rowj[c]) : 0.0;
    }
    qr_wave_sync();  // row j read by every lane before it is updated
    // update, with the reflector below the diagonal scaled into column q
    for (int r = j + lane; r < N; r += 64) {
      double* row = pnl + r * PS;
      const double v = r == j ? 1.0 : row[q] * scal;
#pragma unroll
      for (int c0 = 0; c0 < NB; c0 += 8) {
        if (c0 + 8 <= q) continue;  // columns before the pivot do not change
        double rv[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) rv[c] = row[c0 + c];
        qr_sched_fence();
#pragma unroll
        for (int c = 0; c < 8; ++c)
          row[c0 + c] = (c0 + c == q && r > j) ? v : rv[c] - v * dp[c0 + c];  // dp = 0 off (q, w)
      }
    }
    qr_wave_sync();
    if (lane == 0) {
      pnl[j * PS + q] = beta;
      taus[j] = tau;
    }
    // downdate the norms of the remaining columns over rows j+1..N-1
    if (lane > q && lane < w) {
      const double rjc = pnl[j * PS + lane];
      const double nn = nrm[lane] - rjc * rjc;
      nrm[lane] = nn;
      if (!(nn > kTol * org[lane])) *flg = 1;
    }
    qr_wave_sync();
    if (q + 1 < w && *flg) {  // uniform: every lane reads the same flag
      qr_panel_gram<NB, true>(pnl, N, j + 1, 0, red, lane);
      if (lane > q && lane < w) nrm[lane] = org[lane] = red[lane];
      qr_wave_sync();
    }
  }
}

// T (w x w, leading dimension nb) of the compact-WY form I - V T V^T = H_0 ... H_{w-1}
// of a panel's reflectors (V explicit in pnl, rows r0..N-1, row stride ps).  The wavefront forms
// S = V^T V on the matrix cores (A and B operands coincide), then lane p computes row p
// of T in registers from the recurrence T[p][p] = tau_p,
// T[p][i] = -tau_i * sum_{p <= q < i} T[p][q] S[q][i]  (rows are independent).
// S is a 16 x 16 LDS scratch.  Called by one whole wavefront; ends with a wavefront sync.
__device__ void qr_build_t(const double* pnl, int ps, int nb, int r0, int N, int w, const double* tau,
                           double* Tm, double* S) {
  {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    qr_v4d acc = {0.0, 0.0, 0.0, 0.0};
    for (int kk = r0; kk < N; kk += 4) {
      const int r = kk + lk;
      const double v = ld0(pnl, (size_t)r * ps + li, r < N && li < w);
      acc = qr_mma(v, v, acc);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) S[(lk + 4 * v) * 16 + li] = acc[v];
    qr_wave_sync();
    if (lane < w) {
      const int p = lane;
      double t[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < i; ++q)
          if (q >= p) s += t[q] * S[q * 16 + i];
        t[i] = (i < p || i >= w) ? 0.0 : (i == p ? tau[p] : -tau[i] * s);
      }
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (i < w) Tm[p * nb + i] = t[i];
    }
  }
  qr_wave_sync();
}

__global__ __launch_bounds__(QR_THREADS) void k_qr_blocked(const QrTask* __restrict__ tasks, ChainState* st) {
  const QrTask tk = tasks[blockIdx.x];
  const int N = tk.N;
  const int nb = qr_panel_width(N);
  const int ps = nb + 1;               // panel row stride (padded)
  double* __restrict__ W = tk.W;
  double* __restrict__ A = tk.A;
  LQMC_DYNAMIC_LDS(double, sm);
  double* pnl = sm;                    // [N][ps]
  double* Tm = pnl + (size_t)N * ps;   // [nb][nb]
  double* Ssc = Tm + nb * nb;          // [16][16]
  double* red = Ssc + 256;             // [16]
  double* pnrm = red + 16;             // [16] panel column norms (squared), current
  double* porg = pnrm + 16;            // [16] and at their last exact computation
  double* norms = porg + 16;           // [N]
  double* taus = norms + N;            // [N]
  int* perm = reinterpret_cast<int*>(taus + N);  // [N]
  int* swp = perm + N;                 // [16] column swaps of the current panel
  int* fp = swp + 16;                  // [16] resulting panel column order
  int* pflg = fp + 16;                 // [1] norm recompute flag of the panel
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  bool zero_pivot = false, bad = false;

  // pre-pivoting: order columns by decreasing norm (ties by index; NaN last).  Lane =
  // column within a 64-column tile, the wavefronts of a tile split the rows; partial
  // sums meet in the (not yet used) panel buffer.
  {
    const int nct = (N + 63) >> 6;                       // 64-column tiles
    const int ng = QR_WAVES > nct ? QR_WAVES / nct : 1;  // row groups per tile
    for (int t = wave; t < nct * ng; t += QR_WAVES) {
      const int ct = t % nct, g = t / nct;
      const int c = ct * 64 + lane;
      double s = 0.0;
      if (c < N)
        for (int r0 = g; r0 < N; r0 += 8 * ng) {  // 8 loads in flight, summed in row order
          double x[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int r = r0 + u * ng;
            x[u] = ldg0(gptr(W), (size_t)r * N + c, r < N);
          }
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (r0 + u * ng < N) s += x[u] * x[u];
        }
      if (c < N) pnl[g * N + c] = s;
    }
    __syncthreads();
    for (int c = tid; c < N; c += QR_THREADS) {
      double s = 0.0;
      for (int q = 0; q < ng; ++q) s += pnl[q * N + c];
      if (!(s >= 0.0)) { s = -1.0; bad = true; }
      norms[c] = s;
    }
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
  const size_t n2 = (size_t)N * N;
  for (size_t e0 = tid; e0 < n2; e0 += 8 * QR_THREADS) {  // 8 loads in flight per thread
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const size_t e = e0 + (size_t)u * QR_THREADS;
      const bool in = e < n2;
      v[u] = ldg0(gptr(W), in ? (e / N) * N + perm[e % N] : 0, in);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const size_t e = e0 + (size_t)u * QR_THREADS;
      if (e < n2) A[e] = v[u];
    }
  }
  __syncthreads();

  for (int j0 = 0; j0 < N; j0 += nb) {
    const int w = min(nb, N - j0);
    for (int e = tid; e < (N - j0) * nb; e += QR_THREADS) {
      const int r = j0 + e / nb, q = e % nb;
      pnl[r * ps + q] = q < w ? A[(size_t)r * N + j0 + q] : 0.0;
    }
    __syncthreads();
    if (wave == 0) {
      if (nb == 16) qr_panel_wave<16>(pnl, N, j0, w, perm, swp, taus, red, pnrm, porg, pflg, zero_pivot);
      else qr_panel_wave<8>(pnl, N, j0, w, perm, swp, taus, red, pnrm, porg, pflg, zero_pivot);
      if (lane == 0) {
        for (int q = 0; q < nb; ++q) fp[q] = q;
        for (int q = 0; q < w; ++q) {
          const int t0 = fp[q];
          fp[q] = fp[swp[q]];
          fp[swp[q]] = t0;
        }
      }
    }
    __syncthreads();
    // R (upper part) and V (below the diagonal) back to A; the R entries above the panel
    // follow the panel's column swaps; pnl becomes the explicit V
    for (int e = tid; e < (N - j0) * w; e += QR_THREADS) {
      const int r = j0 + e / w, q = e % w;
      A[(size_t)r * N + j0 + q] = pnl[r * ps + q];
    }
    for (int r = tid; r < j0; r += QR_THREADS) {
      double* Ar = A + (size_t)r * N + j0;
      double t[16];
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q < w) t[q] = Ar[fp[q]];
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q < w) Ar[q] = t[q];
    }
    __syncthreads();
    for (int e = tid; e < w * w; e += QR_THREADS) {
      const int rr = e / w, q = e % w;
      if (rr < q) pnl[(j0 + rr) * ps + q] = 0.0;
      else if (rr == q) pnl[(j0 + rr) * ps + q] = 1.0;
    }
    __syncthreads();
    if (wave == 0) qr_build_t(pnl, ps, nb, j0, N, w, taus + j0, Tm, Ssc);
    __syncthreads();
    double* tout = tk.Tt + (size_t)(j0 / nb) * nb * nb;
    for (int e = tid; e < nb * nb; e += QR_THREADS) {
      const int p = e / nb, q = e % nb;
      tout[e] = (p < w && q < w) ? Tm[e] : 0.0;
    }
    if (j0 + w < N) qr_apply_block<QR_WAVES, 8, 1>(A, N, j0, j0 + w, N, pnl, ps, w, Tm, nb, true);
  }

  // D = |diag R|; V'' = D^-1 R P^T over W, or R' = D^-1 R (unpermuted) with the pivot order
  for (int i = tid; i < N; i += QR_THREADS) {
    const double d = fabs(A[(size_t)i * N + i]);
    norms[i] = d > 0.0 ? d : 1e-300;
    tk.D[i] = norms[i];
    if (tk.perm) tk.perm[i] = perm[i];
  }
  __syncthreads();
  for (size_t e0 = tid; e0 < n2; e0 += 8 * QR_THREADS) {  // 8 loads in flight per thread
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const size_t e = e0 + (size_t)u * QR_THREADS;
      v[u] = ldg0(gptr(A), e, e < n2 && (int)(e % N) >= (int)(e / N));
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const size_t e = e0 + (size_t)u * QR_THREADS;
      if (e >= n2) continue;
      const int i = (int)(e / N), c = (int)(e % N);
      const double y = c >= i ? v[u] / norms[i] : 0.0;
      if (tk.perm) W[e] = y;
      else W[(size_t)i * N + perm[c]] = y;
    }
  }
  if (st) {
    if (tid == 0 && zero_pivot) atomicOr(&st[tk.chain].flags, 2u);
    if (bad) atomicOr(&st[tk.chain].flags, 8u);  // any thread whose columns had a NaN norm
  }
}

// Q[:, c0:c1] of the factorisation left in A / Tt by k_qr_blocked.  Reflector H_j acts on
// rows >= j only, so the identity columns c0..c1-1 are touched by panels j0 <= c1-1, and
// panel j0 changes rows and columns >= j0 only.
__global__ __launch_bounds__(QF_THREADS) void k_qr_formq(const QrTask* __restrict__ tasks, int nblk) {
  const int gb = xcd_grouped_block();  // a matrix's column blocks on one XCD
  const QrTask tk = tasks[gb / nblk];
  const int N = tk.N;
  const int c0 = (gb % nblk) * QF_COLS;
  if (c0 >= N) return;
  const int c1 = min(N, c0 + QF_COLS), cw = c1 - c0;
  const int nb = qr_panel_width(N);
  const int ps = nb + 1;
  const double* __restrict__ A = tk.A;
  double* __restrict__ Q = tk.Q;
  LQMC_DYNAMIC_LDS(double, sm);
  double* pnl = sm;                    // [N][ps]
  double* Tm = pnl + (size_t)N * ps;   // [nb][nb]
  const int tid = threadIdx.x;
  for (int e = tid; e < N * cw; e += QF_THREADS) {
    const int r = e / cw, c = c0 + e % cw;
    Q[(size_t)r * N + c] = r == c ? 1.0 : 0.0;
  }
  for (int j0 = ((c1 - 1) / nb) * nb; j0 >= 0; j0 -= nb) {
    const int w = min(nb, N - j0);
    const int tot = (N - j0) * nb;
    for (int e0 = tid; e0 < tot; e0 += 8 * QF_THREADS) {  // 8 loads in flight per thread
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * QF_THREADS;
        const int q = e % nb, rr = e / nb;
        const bool in = e < tot && q < w;
        const double x = ldg0(gptr(A), (size_t)(j0 + rr) * N + j0 + q, in && rr > q);
        v[u] = (in && rr == q) ? 1.0 : x;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * QF_THREADS;
        if (e < tot) pnl[(j0 + e / nb) * ps + e % nb] = v[u];
      }
    }
    const double* tin = tk.Tt + (size_t)(j0 / nb) * nb * nb;
    for (int e = tid; e < nb * nb; e += QF_THREADS) Tm[e] = tin[e];
    __syncthreads();
    qr_apply_block<QF_THREADS / 64, 8, 4>(Q, N, j0, max(c0, j0), c1, pnl, ps, w, Tm, nb, false);
  }
}

// R' Y' = Z on the columns c0..c1-1 (R' upper triangular, unit-modulus diagonal), then
// Y[perm[r]][c] = Y'[r][c].  Per 16-row block from the bottom: wavefront t forms
// Z[blk, tile t] - R'[blk, below] Y'[below, tile t] on the matrix cores, then one lane per
// column solves the 16 x 16 diagonal block.
__global__ __launch_bounds__(QF_THREADS) void k_trsm_perm(const TrsmTask* __restrict__ tasks, int nblk,
                                                          ChainState* st) {
  const int gb = xcd_grouped_block();  // a matrix's column blocks on one XCD
  const TrsmTask tk = tasks[gb / nblk];
  const int N = tk.N;
  const int c0 = (gb % nblk) * QF_COLS;
  if (c0 >= N) return;
  const int c1 = min(N, c0 + QF_COLS);
  const double* __restrict__ R = tk.V;
  double* __restrict__ Z = tk.Z;
  const LQMC_GLOBAL double* Rg = gptr(R);
  const LQMC_GLOBAL double* Zg = gptr(Z);
  LQMC_DYNAMIC_LDS(double, sm);
  double* blk = sm;                     // [TRSM_B][QF_COLS]
  double* rb = blk + TRSM_B * QF_COLS;  // [TRSM_B][TRSM_B]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int c = c0 + wave * 16 + li;
  const bool cok = c < c1;
  bool zero = false;
  for (int i0 = ((N - 1) / TRSM_B) * TRSM_B; i0 >= 0; i0 -= TRSM_B) {
    const int b = min(TRSM_B, N - i0);
    for (int e = tid; e < TRSM_B * TRSM_B; e += QF_THREADS) {
      const int ii = e / TRSM_B, kk = e % TRSM_B;
      rb[e] = (ii < b && kk < b) ? R[(size_t)(i0 + ii) * N + i0 + kk] : 0.0;
    }
    qr_v4d acc;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = lk + 4 * v;
      acc[v] = (row < b && cok) ? Z[(size_t)(i0 + row) * N + c] : 0.0;
    }
    const int ra = i0 + li;
    constexpr int KB = 8;  // k-steps whose loads are issued together (zero beyond N)
    for (int k = i0 + b; k < N; k += 4 * KB) {
      double a[KB], y[KB];
#pragma unroll
      for (int u = 0; u < KB; ++u) {
        const int kr = k + 4 * u + lk;
        a[u] = -ldg0(Rg, (size_t)ra * N + kr, li < b && kr < N);
        y[u] = ldg0(Zg, (size_t)kr * N + c, kr < N && cok);
      }
      qr_sched_fence();
#pragma unroll
      for (int u = 0; u < KB; ++u) acc = qr_mma(a[u], y[u], acc);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) blk[(lk + 4 * v) * QF_COLS + wave * 16 + li] = acc[v];
    __syncthreads();
    if (tid < QF_COLS) {
      for (int ii = b - 1; ii >= 0; --ii) {
        double s = blk[ii * QF_COLS + tid];
        for (int kk = ii + 1; kk < b; ++kk) s -= rb[ii * TRSM_B + kk] * blk[kk * QF_COLS + tid];
        const double d = rb[ii * TRSM_B + ii];
        if (d == 0.0) zero = true;
        blk[ii * QF_COLS + tid] = s / d;
      }
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = lk + 4 * v;
      if (row < b && cok) Z[(size_t)(i0 + row) * N + c] = blk[row * QF_COLS + wave * 16 + li];
    }
    __syncthreads();
  }
  const int* __restrict__ perm = tk.perm;
  for (int e0 = tid; e0 < N * QF_COLS; e0 += 8 * QF_THREADS) {  // 8 loads in flight per thread
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * QF_THREADS;
      const int r = e / QF_COLS, cc = c0 + e % QF_COLS;
      v[u] = ldg0(Zg, (size_t)r * N + cc, e < N * QF_COLS && cc < c1);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * QF_THREADS;
      const int r = e / QF_COLS, cc = c0 + e % QF_COLS;
      if (e < N * QF_COLS && cc < c1) tk.Y[(size_t)perm[r] * N + cc] = v[u];
    }
  }
  if (zero && st) atomicOr(&st[tk.chain].flags, 2u);
}

// ---- on-chip factorisation (N <= 256) ----------------------------------------------------------
// k_qr_onchip: the same pre-pivoted Householder QR with the whole matrix held in registers.
// The pre-pivoted columns are split into halves of 128, one 512-thread workgroup per half
// (two for 128 < N <= 256, one below), and each wavefront holds one 16-column strip: 16 row
// tiles in the 16x16x4 FP64 accumulator layout (row 16t + l/16 + 4v, column l%16 of lane l),
// 64 doubles per lane.  Panel P is the strip of wavefront P%8 in half P/8.  A wavefront
//   1. applies Q_P^T = I - V T^T V^T of every panel P left of its strip, from LDS, on the
//      matrix cores: W = V^T X, W' = T^T W, X -= V W' (the accumulator layout of X is the
//      operand layout of these products, so nothing is transposed);
//   2. factors its own panel: the strip goes through LDS into a row-per-lane layout (rows
//      j0 + lane + 64i), one wavefront reduction per column (Householder, no pivoting beyond
//      the pre-pivot: the columns of the UDV chain are graded, and sorting them by norm once
//      is what keeps the factorisation accurate), then V and T go to LDS for the others,
//      the reflectors to A, T to Tt, |diag R| to D, and the strip's columns of D^-1 R (final:
//      later panels touch only columns right of it) to W;
//   3. only meets the workgroup barriers of the remaining panels (its registers are free).
// The right half receives the left half's panels (reflectors in A, T in Tt, D) through global
// memory behind a per-matrix release counter, hand[2 task] = panels published; the left half
// writes W (which both halves read at the start) only after the right half has released
// hand[2 task + 1] = 1.  The two workgroups of a matrix get block indices of equal residue
// mod 8 (one XCD under round-robin dispatch), and a launch holds at most as many workgroups
// as the device has compute units, so both are resident; every wait is still bounded and
// raises flag 16 on expiry.  Outputs as k_qr_blocked (D, V'' or R' and the pivot order,
// reflectors below the diagonal of A, T per panel in Tt), so k_qr_formq forms Q the same way.
constexpr int QO_THREADS = 512;
constexpr int QO_NMAX = 256;
constexpr int QO_PS = 17;                 // LDS row stride of the strip and panel buffers
constexpr unsigned QO_SPINS = 1u << 22;   // polls of a cross-workgroup wait

typedef double qo_acc_t __attribute__((ext_vector_type(4)));

__device__ inline double qo_readlane(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Bounded wait of one thread until *c > after (acquire at device scope); on expiry flag 16.
__device__ inline void qo_wait(unsigned* c, unsigned after, ChainState* st, int chain, bool& broken) {
  if (!broken) {
    unsigned spins = 0;
    while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= after) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins >= QO_SPINS) {
        broken = true;
        if (st) atomicOr(&st[chain].flags, 16u);
        break;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Q_P^T applied to one strip held in registers (rows j0 = 16 P .. 16 NS - 1; the panel's V,
// explicit with unit diagonal and zeros above it, in vbuf; T in tbuf).
__device__ inline void qo_apply(qo_acc_t (&acc)[16], const double* vbuf, const double* tbuf, int NS, int P) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  qr_v4d wt = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    if (t < P || t >= NS) continue;
    double a[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) a[v] = vbuf[(16 * t + lk + 4 * v) * QO_PS + li];
#pragma unroll
    for (int v = 0; v < 4; ++v) wt = qr_mma(a[v], acc[t][v], wt);
  }
  qr_v4d w2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 4; ++s) w2 = qr_mma(tbuf[(4 * s + lk) * 16 + li], wt[s], w2);
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    if (t < P || t >= NS) continue;
    double a[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) a[s] = -vbuf[(16 * t + li) * QO_PS + 4 * s + lk];
#pragma unroll
    for (int s = 0; s < 4; ++s) acc[t] = qr_mma(a[s], w2[s], acc[t]);
  }
}

// Q formation for N <= 256 with the column block held in registers: one 256-thread workgroup
// per 64-column block of Q, wavefront w holding the 16 columns c0 + 16 w .. in the accumulator
// layout of the FP64 matrix cores (all rows).  From the identity, the panels' block reflectors
// H_P = I - V T V^T are applied from the last to the first (Q = H_0 ... H_{NS-1}): each panel's
// V (explicit, unit diagonal, from the reflectors below the diagonal of A) and T (stored
// transposed, since qo_apply applies I - V T_buf^T V^T) go to LDS once per workgroup, and a
// wavefront skips the panels right of its columns, which leave them unchanged.  Q is written
// once, at the end (the blocked k_qr_formq reads and writes its column block per panel).
__global__ __launch_bounds__(QF_THREADS, 2) void k_qr_formq_onchip(const QrTask* __restrict__ tasks, int nblk) {
  const int gb = xcd_grouped_block();  // a matrix's column blocks on one XCD
  const QrTask tk = tasks[gb / nblk];
  const int N = tk.N;
  const int c0 = (gb % nblk) * QF_COLS;
  if (c0 >= N) return;
  const int c1 = min(N, c0 + QF_COLS);
  const int NS = (N + 15) >> 4;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  LQMC_DYNAMIC_LDS(double, sm);
  double* vbuf = sm;                      // [256][QO_PS] V of the panel
  double* tbuf = vbuf + QO_NMAX * QO_PS;  // [16][16] T^T of the panel
  const int cs = (c0 >> 4) + wave;        // my strip's column tile
  const int c = (cs << 4) + li;           // my lane's column
  qo_acc_t acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[t][v] = (16 * t + lk + 4 * v) == c ? 1.0 : 0.0;
  for (int P = (c1 - 1) >> 4; P >= 0; --P) {
    const int j0 = P << 4, w = min(16, N - j0);
    __syncthreads();  // every wavefront is done with the previous panel's buffers
    for (int e = tid; e < (QO_NMAX - j0) * 16; e += QF_THREADS) {
      const int r = j0 + (e >> 4), cc = e & 15, jc = j0 + cc;
      const bool in = r < N && cc < w;
      const double a = ldg0(gptr(tk.A), (size_t)r * N + jc, in && r > jc);
      vbuf[r * QO_PS + cc] = in ? (r > jc ? a : (r == jc ? 1.0 : 0.0)) : 0.0;
    }
    for (int e = tid; e < 256; e += QF_THREADS) {
      const int i = e >> 4, j = e & 15;
      tbuf[e] = (i < w && j < w) ? tk.Tt[(size_t)P * 256 + j * 16 + i] : 0.0;
    }
    __syncthreads();
    if (cs >= P && cs < NS) qo_apply(acc, vbuf, tbuf, NS, P);
  }
  if (c < c1)
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * t + lk + 4 * v;
        if (r < N) tk.Q[(size_t)r * N + c] = acc[t][v];
      }
}

size_t qr_formq_onchip_shmem() { return sizeof(double) * ((size_t)QO_NMAX * QO_PS + 256); }

// R' Y' = Z, Y[perm[r]] = Y'[r] for N <= 256 with the solution held in registers: one
// 256-thread workgroup per 64-column block, wavefront w holding Y'[:, c0 + 16 w ..] in the
// accumulator layout (row 16 t + l/16 + 4 v, column l%16 of lane l).  Row blocks of 16 from the
// bottom: the block's band of R' goes to LDS, each wavefront forms Z[blk] - R'[blk, below]
// Y'[below] on the matrix cores from its registers (the accumulator layout of Y' is the operand
// layout of the product), the 16 x 16 diagonal systems are solved one lane per column through
// LDS, and the block joins the registers.  Y is written once, permuted, at the end (the blocked
// k_trsm_perm re-reads the solution below each block from global memory).
constexpr int TO_LD = QO_NMAX + 1;  // LDS row stride of the R' band

__global__ __launch_bounds__(QF_THREADS, 2) void k_trsm_onchip(const TrsmTask* __restrict__ tasks, int nblk,
                                                               ChainState* st) {
  const int gb = xcd_grouped_block();  // a matrix's column blocks on one XCD
  const TrsmTask tk = tasks[gb / nblk];
  const int N = tk.N;
  const int c0 = (gb % nblk) * QF_COLS;
  if (c0 >= N) return;
  const int c1 = min(N, c0 + QF_COLS);
  const int NS = (N + 15) >> 4;
  const LQMC_GLOBAL double* Rg = gptr(tk.V);
  const LQMC_GLOBAL double* Zg = gptr(tk.Z);
  LQMC_DYNAMIC_LDS(double, sm);
  double* band = sm;                 // [16][TO_LD] rows i0 .. i0+15 of R' (columns i0 .. N-1)
  double* blk = band + 16 * TO_LD;   // [16][QF_COLS] the diagonal block's right-hand sides
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int c = c0 + wave * 16 + li;
  const bool cok = c < c1;
  bool zero = false;
  qo_acc_t y[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) y[t] = qo_acc_t{0.0, 0.0, 0.0, 0.0};
  for (int i = NS - 1; i >= 0; --i) {
    // the band's columns i0 .. 16 NS - 1, zero past N (the products below read whole 16-column
    // tiles, whose rows of Y' past N are zero: stale LDS there could hold a NaN)
    const int i0 = i << 4, b = min(16, N - i0), nc = 16 * NS - i0;
    __syncthreads();  // the previous block's reads of band and blk are done
    for (int e = tid; e < 16 * nc; e += QF_THREADS) {
      const int ii = e / nc, cc = i0 + e % nc;
      band[ii * TO_LD + cc] = ldg0(Rg, (size_t)(i0 + ii) * N + cc, ii < b && cc < N);
    }
    qr_v4d acc;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = lk + 4 * v;
      acc[v] = ldg0(Zg, (size_t)(i0 + row) * N + c, row < b && cok);
    }
    __syncthreads();
    // acc -= R'[blk, j] Y'_j for the blocks j below (A lane l: R'[i0 + l%16][16 j + l/16 + 4 v])
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      if (j <= i || j >= NS) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) acc = qr_mma(-band[li * TO_LD + 16 * j + lk + 4 * v], y[j][v], acc);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) blk[(lk + 4 * v) * QF_COLS + wave * 16 + li] = acc[v];
    __syncthreads();
    if (tid < QF_COLS) {
      for (int ii = b - 1; ii >= 0; --ii) {
        double sv = blk[ii * QF_COLS + tid];
        for (int kk = ii + 1; kk < b; ++kk) sv -= band[ii * TO_LD + i0 + kk] * blk[kk * QF_COLS + tid];
        const double d = band[ii * TO_LD + i0 + ii];
        if (d == 0.0) zero = true;
        blk[ii * QF_COLS + tid] = sv / d;
      }
    }
    __syncthreads();
    qr_v4d sol;
#pragma unroll
    for (int v = 0; v < 4; ++v) sol[v] = blk[(lk + 4 * v) * QF_COLS + wave * 16 + li];
#pragma unroll
    for (int t = 0; t < 16; ++t)
      if (t == i) y[t] = sol;
  }
  if (cok)
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * t + lk + 4 * v;
        if (r < N) tk.Y[(size_t)tk.perm[r] * N + c] = y[t][v];
      }
  if (zero && st) atomicOr(&st[tk.chain].flags, 2u);
}

size_t trsm_onchip_shmem() { return sizeof(double) * (16 * (size_t)TO_LD + 16 * QF_COLS); }

struct QoShared {
  double *sbuf, *vbuf, *tbuf, *ssc, *red, *dbuf, *taus;
  const int* perm;
};

// Step 2 for panel P (rows j0..N-1, w columns) of the calling wavefront's strip.  hand: the
// matrix's two hand-over counters, or null when one workgroup holds the whole matrix;
// left: the calling workgroup is the left half of a split matrix.
__device__ void qo_factor(qo_acc_t (&acc)[16], const QrTask& tk, const QoShared& sh, int N, int P, int w,
                          unsigned* hand, bool left, ChainState* st, bool& broken, bool& zero_pivot) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int j0 = P << 4;
  double* sbuf = sh.sbuf;
  // the whole strip to LDS; the registers are not used again
#pragma unroll
  for (int t = 0; t < 16; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) sbuf[(16 * t + lk + 4 * v) * QO_PS + li] = acc[t][v];
  qr_wave_sync();
  double x[4][16];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = j0 + lane + 64 * i;
    const bool ok = r < N;
#pragma unroll
    for (int c = 0; c < 16; ++c) x[i][c] = ld0(sbuf, (size_t)r * QO_PS + c, ok);
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    if (q < w) {
      const int j = j0 + q;
      // s_c = x^T col_c over rows >= j for the column x = col_q (s_q = |x|^2), into red
      {
        double part[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) part[c] = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = j0 + lane + 64 * i;
          const double xv = r >= j ? x[i][q] : 0.0;  // rows >= N hold zeros
#pragma unroll
          for (int c = q; c < 16; ++c) part[c] += xv * x[i][c];
        }
        qr_wave_colsum<16>(part, sh.red, lane);
      }
      const double alpha = qo_readlane(x[0][q], q);  // row j is lane q's first row
      const double xx = sh.red[q];
      const double nrmv = sqrt(xx > 0.0 ? xx : 0.0);
      double beta, tau, scal;
      if (nrmv == 0.0) {
        beta = 0.0; tau = 0.0; scal = 0.0;
        zero_pivot = true;
      } else {
        beta = alpha >= 0.0 ? -nrmv : nrmv;
        tau = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      const double ts = tau * scal;
      double v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = j0 + lane + 64 * i;
        v[i] = r > j ? x[i][q] * scal : (r == j ? 1.0 : 0.0);
      }
      // column c -= v (tau v^T col_c), tau v^T col_c = ts (s_c - beta col_c[j]), v = u / (alpha - beta)
#pragma unroll
      for (int c = q + 1; c < 16; ++c) {
        const double dpc = c < w ? ts * (sh.red[c] - beta * qo_readlane(x[0][c], q)) : 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i][c] = x[i][c] - v[i] * dpc;
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = j0 + lane + 64 * i;
        x[i][q] = r > j ? v[i] : (r == j ? beta : x[i][q]);
      }
      qr_wave_sync();  // red read by every lane before the next reduction writes it
      if (lane == 0) {
        sh.taus[q] = tau;
        const double d = fabs(beta);
        sh.dbuf[j] = d > 0.0 ? d : 1e-300;
      }
    }
  }
  // V (explicit) to vbuf, reflectors to A, R to sbuf
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = j0 + lane + 64 * i;
    if (r >= QO_NMAX) continue;
    const bool in = r < N;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int jc = j0 + c;
      const bool col = c < w;
      sh.vbuf[r * QO_PS + c] = (in && col) ? (r > jc ? x[i][c] : (r == jc ? 1.0 : 0.0)) : 0.0;
      if (in && col && r <= jc) sbuf[r * QO_PS + c] = x[i][c];
      if (in && col && r > jc) tk.A[(size_t)r * N + jc] = x[i][c];
    }
  }
  qr_wave_sync();
  qr_build_t(sh.vbuf, QO_PS, 16, j0, N, w, sh.taus, sh.tbuf, sh.ssc);
  double* tout = tk.Tt + (size_t)P * 256;
  for (int e = lane; e < 256; e += 64) tout[e] = ((e >> 4) < w && (e & 15) < w) ? sh.tbuf[e] : 0.0;
  if (lane < w) tk.D[j0 + lane] = sh.dbuf[j0 + lane];
  if (hand && left) {
    // hand the panel (reflectors in A, T in Tt, D) to the right half
    __builtin_amdgcn_s_waitcnt(0);
    qr_wave_sync();
    if (lane == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      atomicExch(hand, (unsigned)(P + 1));
      qo_wait(hand + 1, 0u, st, tk.chain, broken);  // the right half has read W
    }
    qr_wave_sync();
  }
}

// R of strip S (rows 0..16S+15, upper triangle of the panel block) from sbuf back to the
// accumulator layout; zero below.  Runs after the panel's release barrier, while the next
// wavefront applies the panel and before it overwrites sbuf.
__device__ inline void qo_keep_r(qo_acc_t (&acc)[16], const QoShared& sh, int S) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int jc = 16 * S + li;
#pragma unroll
  for (int t = 0; t < 16; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * t + lk + 4 * v;
      acc[t][v] = ld0(sh.sbuf, (size_t)r * QO_PS + li, r <= jc);
    }
}

__global__ __launch_bounds__(QO_THREADS) void k_qr_onchip(const QrTask* __restrict__ tasks, int ntasks, int G,
                                                          unsigned* hand, ChainState* st) {
  const int b = blockIdx.x, slot = b >> 3;
  const int ti = (b & 7) + 8 * (slot / G), part = slot % G;
  if (ti >= ntasks) return;
  const QrTask tk = tasks[ti];
  const int N = tk.N;
  const int NS = (N + 15) >> 4;   // strips, and row tiles
  const int Gt = (N + 127) >> 7;  // workgroups this matrix uses
  if (part >= Gt) return;
  unsigned* hd = Gt > 1 ? hand + 2 * ti : nullptr;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  LQMC_DYNAMIC_LDS(double, sm);
  QoShared sh;
  sh.sbuf = sm;                             // [256][17] strip transpose buffer, pre-pivot partials
  sh.vbuf = sh.sbuf + QO_NMAX * QO_PS;      // [256][17] explicit V of the current panel
  sh.tbuf = sh.vbuf + QO_NMAX * QO_PS;      // [16][16] T of the current panel
  sh.ssc = sh.tbuf + 256;                   // [16][16] S = V^T V scratch
  sh.red = sh.ssc + 256;                    // [16] wavefront column sums
  sh.dbuf = sh.red + 16;                    // [256] |diag R|
  sh.taus = sh.dbuf + QO_NMAX;              // [16] reflector coefficients of the panel
  double* norms = sh.taus + 16;             // [256] pre-pivot column norms
  int* perm = reinterpret_cast<int*>(norms + QO_NMAX);  // [256] pre-pivot order
  sh.perm = perm;
  bool zero_pivot = false, bad = false, broken = false;

  // pre-pivoting: column norms (partial sums over row groups, combined in group order, so
  // both workgroups of a matrix derive the same order), then decreasing norm, ties by index,
  // NaN last
  {
    const int ng = QO_THREADS / N;  // >= 2 row groups
    if (tid < ng * N) {
      const int c = tid % N, g = tid / N;
      double s = 0.0;
      for (int r0 = g; r0 < N; r0 += 8 * ng) {
        double xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int r = r0 + u * ng;
          xv[u] = ldg0(gptr(tk.W), (size_t)r * N + c, r < N);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) s += xv[u] * xv[u];
      }
      sh.sbuf[g * N + c] = s;
    }
    __syncthreads();
    for (int c = tid; c < N; c += QO_THREADS) {
      double s = 0.0;
      for (int g = 0; g < ng; ++g) s += sh.sbuf[g * N + c];
      if (!(s >= 0.0)) { s = -1.0; bad = true; }
      norms[c] = s;
    }
    for (int i = tid; i < QO_NMAX; i += QO_THREADS) sh.dbuf[i] = 1.0;
    __syncthreads();
    for (int c = tid; c < N; c += QO_THREADS) {
      const double nc = norms[c];
      int rank = 0;
      for (int c2 = 0; c2 < N; ++c2) {
        const double n2 = norms[c2];
        rank += (n2 > nc) || (n2 == nc && c2 < c);
      }
      perm[rank] = c;
    }
    __syncthreads();
  }
  // my strip
  const int S = part * 8 + wave;
  const bool have = S < NS;
  qo_acc_t acc[16];
  {
    const int c = 16 * S + li;
    const bool cok = have && c < N;
    const int pc = perm[cok ? c : 0];
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * t + lk + 4 * v;
        acc[t][v] = ldg0(gptr(tk.W), (size_t)r * N + pc, cok && r < N);
      }
  }
  // every read of W is complete before any workgroup of the matrix writes it
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (hd && part == 1 && tid == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    atomicExch(hd + 1, 1u);
  }
  const int P0 = part * 8, P1 = min(NS, P0 + 8);  // the panels of my half
  const int Sm = have ? S : P1;
  // 1. panels left of my strip (received from the left half, or factored by a wavefront of
  //    my half before the first barrier of their iteration)
  for (int P = 0; P < Sm; ++P) {
    if (P < P0) {
      const int j0 = P << 4, w = min(16, N - j0);
      if (tid == 0) qo_wait(hd, (unsigned)P, st, tk.chain, broken);
      __syncthreads();
      for (int e = tid; e < (QO_NMAX - j0) * 16; e += QO_THREADS) {
        const int r = j0 + (e >> 4), c = e & 15, jc = j0 + c;
        const bool in = r < N && c < w;
        const double a = ldg0(gptr(tk.A), (size_t)r * N + jc, in && r > jc);
        sh.vbuf[r * QO_PS + c] = in ? (r > jc ? a : (r == jc ? 1.0 : 0.0)) : 0.0;
      }
      for (int e = tid; e < 256; e += QO_THREADS) sh.tbuf[e] = tk.Tt[(size_t)P * 256 + e];
      if (tid < w) sh.dbuf[j0 + tid] = tk.D[j0 + tid];
    }
    __syncthreads();
    if (have) qo_apply(acc, sh.vbuf, sh.tbuf, NS, P);
    __syncthreads();
  }
  // 2. my panel, 3. the panels right of it
  if (have) {
    qo_factor(acc, tk, sh, N, S, min(16, N - 16 * S), hd, part == 0, st, broken, zero_pivot);
    __syncthreads();
    qo_keep_r(acc, sh, S);
    __syncthreads();
    for (int P = S + 1; P < P1; ++P) {
      __syncthreads();
      __syncthreads();
    }
  }
  __syncthreads();
  if (hd && tid == 0) atomicExch(hd + (part == 1 ? 0 : 1), 0u);  // counters back to zero for the next launch
  if (part == 0 && tk.perm)
    for (int i = tid; i < N; i += QO_THREADS) tk.perm[i] = perm[i];
  // 4. D^-1 R of my strip over W: V'' = D^-1 R P^T (columns to their original places), or
  //    R' = D^-1 R with the pivot order.  Off the panel chain: every |diag R| this strip
  //    needs (rows <= its last column) is in dbuf by now.
  if (have) {
    // N and the lane's row made opaque, so the compiler forms these 64 row offsets and
    // bounds here instead of keeping those of the strip load live all kernel (128 VGPRs)
    int Nw = N, lkw = lk;
    LQMC_OPAQUE(Nw, "s");
    LQMC_OPAQUE(lkw, "v");
    const int jc = 16 * S + li;
    if (jc < Nw) {
      const int oc = tk.perm ? jc : perm[jc];
#pragma unroll
      for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = 16 * t + lkw + 4 * v;
          if (r < Nw) tk.W[(size_t)r * Nw + oc] = r <= jc ? acc[t][v] / sh.dbuf[r] : 0.0;
        }
    }
  }
  if (st) {
    if (zero_pivot && lane == 0) atomicOr(&st[tk.chain].flags, 2u);
    if (bad) atomicOr(&st[tk.chain].flags, 8u);
  }
}

size_t qr_onchip_shmem() {
  return sizeof(double) * (2 * (size_t)QO_NMAX * QO_PS + 256 + 256 + 16 + QO_NMAX + 16 + QO_NMAX) +
         sizeof(int) * QO_NMAX;
}

size_t qr_blocked_shmem(int nmax) {
  const size_t nb = nmax > 256 ? 8 : 16;
  return sizeof(double) * (nmax * (nb + 1) + nb * nb + 256 + 48 + 2 * (size_t)nmax) + sizeof(int) * (nmax + 33);
}

size_t qr_formq_shmem(int nmax) {
  const size_t nb = nmax > 256 ? 8 : 16;
  return sizeof(double) * (nmax * (nb + 1) + nb * nb);
}

size_t trsm_shmem(int) { return sizeof(double) * (TRSM_B * QF_COLS + TRSM_B * TRSM_B); }

// QR of n matrices (explicit Q formed).  N <= 256: the on-chip factorisation, at most
// max_wg workgroups per launch (the compute units, so the two workgroups of a matrix are
// resident together) with hand[2 n] (zero between launches) for the hand-over between them;
// larger N: the blocked factorisation with the trailing matrix in global memory.
void launch_qr_blocked(const QrTask* d, int n, int nmax, ChainState* st, unsigned* hand, int max_wg, hipStream_t s) {
  if (n <= 0) return;
  if (nmax <= QO_NMAX) {
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute((const void*)k_qr_onchip, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)qr_onchip_shmem());
      attr = true;
    }
    const int G = (nmax + 127) >> 7;
    const int chunk = G > 1 ? std::max(8, (max_wg / G) / 8 * 8) : n;
    for (int c0 = 0; c0 < n; c0 += chunk) {
      const int cnt = std::min(chunk, n - c0);
      const int grid = 8 * G * ((cnt + 7) / 8);
      hipLaunchKernelGGL(k_qr_onchip, dim3(grid), dim3(QO_THREADS), qr_onchip_shmem(), s, d + c0, cnt, G, hand, st);
    }
  } else {
    hipLaunchKernelGGL(k_qr_blocked, dim3(n), dim3(QR_THREADS), qr_blocked_shmem(nmax), s, d, st);
  }
  const int nblk = (nmax + QF_COLS - 1) / QF_COLS;
  if (nmax <= QO_NMAX)
    hipLaunchKernelGGL(k_qr_formq_onchip, dim3(n * nblk), dim3(QF_THREADS), qr_formq_onchip_shmem(), s, d, nblk);
  else
    hipLaunchKernelGGL(k_qr_formq, dim3(n * nblk), dim3(QF_THREADS), qr_formq_shmem(nmax), s, d, nblk);
}

void launch_trsm(const TrsmTask* d, int n, int nmax, ChainState* st, hipStream_t s) {
  if (n <= 0) return;
  const int nblk = (nmax + QF_COLS - 1) / QF_COLS;
  if (nmax <= QO_NMAX)
    hipLaunchKernelGGL(k_trsm_onchip, dim3(n * nblk), dim3(QF_THREADS), trsm_onchip_shmem(), s, d, nblk, st);
  else
    hipLaunchKernelGGL(k_trsm_perm, dim3(n * nblk), dim3(QF_THREADS), trsm_shmem(nmax), s, d, nblk, st);
}

}  // namespace lqmc

// Metropolis flip phase of one time slice for every chain (one workgroup per chain).
//
// The N attempts of a slice are sequential (lqmc/lqmc.py:309-333): each reads
// G_ii of both spins, forms the ratio, compares with its uniform and, when
// accepted, updates G and flips the field entry.
//  - k_flip_rank1: Sherman-Morrison rank-1 update of the N x N matrices per
//    accepted flip (Mode D textbook coefficients, or the reference's Mode R
//    coefficients of lqmc.py:320-331).
//  - k_flip_delayed: submatrix update.  Sites are taken in blocks of k; the
//    serial accept/reject only touches the k x k diagonal block (kept in LDS,
//    FP64), then the accepted set A applies to the whole G as one rank-|A|
//    product G -= G[:,A] Gamma^-1 (E_A - G[A,:]), Gamma = diag(1/alpha) + (I-G)[A,A],
//    on the matrix cores.
// Both end the slice by refreshing the exp(V) scale tables of the slice and, on
// slice 0, by the measurement (accumulated G and per-sweep observables).
#include "common.h"

#pragma clang fp contract(off)

namespace lqmc {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Value of lane l (wavefront-uniform) broadcast through scalar registers.
__device__ inline double readlane_f64(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// ---- Philox4x32-10 (Salmon et al., SC'11) --------------------------------------------------
__device__ inline double philox_uniform(uint64_t seed, uint64_t q) {
  uint32_t c0 = (uint32_t)q, c1 = (uint32_t)(q >> 32), c2 = 0u, c3 = 0x51D3u;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  const double hi = (double)(c0 >> 5), lo = (double)(c1 >> 6);
  return (hi * 67108864.0 + lo) * (1.0 / 9007199254740992.0);
}

__device__ inline double draw_uniform(const ChainDev& ch, const ChainState* st, int chain, int rng,
                                      unsigned long long sweep, int t, int i, bool& underrun) {
  const unsigned long long NL = (unsigned long long)ch.N * ch.L;
  const unsigned long long within = (unsigned long long)t * ch.N + i;
  if (rng == 1) return philox_uniform(ch.seed, sweep * NL + within);
  const unsigned long long rel = sweep - st->replay_base;
  if (sweep < st->replay_base || rel >= st->replay_count) {
    underrun = true;
    return 2.0;  // rejects every ratio < 2; flagged
  }
  return ch.uni[rel * NL + within];
}

// ---- shared epilogue ---------------------------------------------------------------------------
__device__ inline double wg_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Measurement of one sweep from the equal-time G of both spins (lqmc.py:371-372 plus the
// per-sweep estimators): accumulate G, keep it as the last measured G, accumulate
// <S^z_i S^z_j>, and append density, double occupancy and kinetic energy per site to the
// chain's observable ring.  All threads of the workgroup take part (reductions).
template <typename T>
__device__ void measure_equal_time(const ChainDev& ch, ChainState* st, const T* G, double* red) {
  const int N = ch.N;
  {
    const size_t n2 = (size_t)N * N;
    for (size_t e = threadIdx.x; e < 2 * n2; e += blockDim.x) {
      const double g = (double)G[e];
      ch.Gacc[e] += g;
      ch.Glast[e] = g;
    }
    // <S^z_i S^z_j> by Wick's theorem for the sampled field, S^z = n_up - n_dn:
    // m_i m_j + sum_s (delta_ij - G_s,ji) G_s,ij with m_i = G_dn,ii - G_up,ii
    for (size_t e = threadIdx.x; e < n2; e += blockDim.x) {
      const size_t i = e / N, j = e % N;
      const double mi = (double)G[n2 + i * N + i] - (double)G[i * N + i];
      const double mj = (double)G[n2 + j * N + j] - (double)G[j * N + j];
      double v = mi * mj;
      for (int sp = 0; sp < 2; ++sp) {
        const double gij = (double)G[sp * n2 + i * N + j], gji = (double)G[sp * n2 + j * N + i];
        v += ((i == j ? 1.0 : 0.0) - gji) * gij;
      }
      ch.Sacc[e] += v;
    }
    double dens = 0.0, docc = 0.0, ek = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      const double nu = 1.0 - (double)G[(size_t)i * N + i];
      const double nd = 1.0 - (double)G[n2 + (size_t)i * N + i];
      dens += nu + nd;
      docc += nu * nd;
    }
    for (size_t e = threadIdx.x; e < n2; e += blockDim.x) {
      const size_t i = e / N, j = e % N;
      const double h = ch.hop[e];
      if (h != 0.0) ek -= h * ((double)G[j * N + i] + (double)G[n2 + j * N + i]);
    }
    dens = wg_sum(dens, red);
    docc = wg_sum(docc, red);
    ek = wg_sum(ek, red);
    if (threadIdx.x == 0) {
      const unsigned long long k = st->measured;
      double* o = ch.obs + (size_t)(k % (unsigned long long)ch.obs_cap) * 3;
      o[0] = dens / N;
      o[1] = docc / N;
      o[2] = ek / N;
      st->measured = k + 1;
    }
  }
}

template <typename T>
__device__ void slice_epilogue(const FlipArgs& a, const ChainDev& ch, int c, int l, const T* G,
                               unsigned long long nacc, bool underrun, bool bad, double* red) {
  const int N = ch.N;
  const int NL = N * ch.L;
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int s = ch.field[(size_t)l * N + i];
    const double e1m = ch.e1m, e1p = ch.e1p;  // values, so the conditionals select values
    const double up = s > 0 ? e1m : e1p;
    const double dn = s > 0 ? e1p : e1m;
    ch.ev[(size_t)l * N + i] = up;
    ch.ev[(size_t)NL + (size_t)l * N + i] = dn;
    ch.evi[(size_t)l * N + i] = 1.0 / up;
    ch.evi[(size_t)NL + (size_t)l * N + i] = 1.0 / dn;
  }
  if (threadIdx.x == 0) {
    atomicAdd(&a.st[c].attempted, (unsigned long long)N);
    atomicAdd(&a.st[c].accepted, nacc);
    if (underrun) atomicOr(&a.st[c].flags, 4u);
    if (bad) atomicOr(&a.st[c].flags, 1u);
  }
  if (l == 0 && a.measure) measure_equal_time<T>(ch, &a.st[c], G, red);
  __syncthreads();
  if (l == 0 && threadIdx.x == 0) a.st[c].sweep += 1ull;
}

// ---- rank-1 form --------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(1024) void k_flip_rank1(FlipArgs a) {
  const int c = blockIdx.x;
  const ChainDev ch = a.chains[c];
  if (a.t >= ch.L) return;
  const int N = ch.N;
  const int l = ch.L - 1 - a.t;
  const size_t n2 = (size_t)N * N;
  T* __restrict__ G = static_cast<T*>(ch.G);
  int8_t* fld = ch.field + (size_t)l * N;
  LQMC_DYNAMIC_LDS(double, sm);
  double* xv = sm;          // [2][N] column (Mode D) / e vector (Mode R)
  double* yv = sm + 2 * N;  // [2][N] row (Mode D) / c vector (Mode R)
  double* red = sm + 4 * N; // blockDim
  __shared__ double coef[2];
  const unsigned long long sweep = a.st[c].sweep;
  unsigned long long nacc = 0;
  bool underrun = false, bad = false;
  const bool ref = ch.mode == 1;

  for (int i = 0; i < N; ++i) {
    const int s = fld[i];
    const double gu = (double)G[(size_t)i * N + i];
    const double gd = (double)G[n2 + (size_t)i * N + i];
    const double e2p = ch.e2p, e2m = ch.e2m;
    const double ep = s > 0 ? e2p : e2m;  // exp(+2 lambda s)
    const double em = s > 0 ? e2m : e2p;  // exp(-2 lambda s)
    const double d_up = 1.0 + (1.0 - gu) * (ep - 1.0);
    const double d_dn = 1.0 + (1.0 - gd) * (em - 1.0);
    const double ratio = d_up * d_dn;
    const double u = draw_uniform(ch, &a.st[c], c, a.rng, sweep, a.t, i, underrun);
    const bool acc = u <= ratio;
    if (!isfinite(ratio)) bad = true;
    if (a.trace && threadIdx.x == 0) {
      ch.tr_ratio[(size_t)a.t * N + i] = ratio;
      ch.tr_acc[(size_t)a.t * N + i] = acc ? 1 : 0;
    }
    if (!acc) continue;
    for (int j = threadIdx.x; j < 2 * N; j += blockDim.x) {
      const int sp = j / N, jj = j % N;
      const T* Gs = G + sp * n2;
      xv[j] = (double)Gs[(size_t)jj * N + i];
      yv[j] = (double)Gs[(size_t)i * N + jj];
    }
    __syncthreads();
    if (ref) {
      // c = -(e^{-+arg}-1) G[i,:]; c_i += (e^{-+arg}-1); e = G[:,i]/(1+c_i)   (lqmc.py:320-327)
      const double m1[2] = {em - 1.0, ep - 1.0};
      for (int j = threadIdx.x; j < 2 * N; j += blockDim.x) {
        const int sp = j / N;
        yv[j] = -m1[sp] * yv[j];
      }
      __syncthreads();
      if (threadIdx.x < 2) yv[threadIdx.x * N + i] = yv[threadIdx.x * N + i] + m1[threadIdx.x];
      __syncthreads();
      for (int j = threadIdx.x; j < 2 * N; j += blockDim.x) {
        const int sp = j / N;
        xv[j] = xv[j] / (1.0 + yv[sp * N + i]);
      }
    } else {
      // y = e_i - G[i,:], coefficient alpha/(1 + alpha(1 - G_ii))
      if (threadIdx.x < 2) {
        const double al = threadIdx.x == 0 ? ep - 1.0 : em - 1.0;
        const double de = threadIdx.x == 0 ? d_up : d_dn;
        coef[threadIdx.x] = al / de;
      }
      for (int j = threadIdx.x; j < 2 * N; j += blockDim.x) {
        const int jj = j % N;
        double y = -yv[j];
        if (jj == i) y = y + 1.0;
        yv[j] = y;
      }
    }
    __syncthreads();
    {
      // RB rows per wavefront at a time, every load of a batch issued before its arithmetic
      constexpr int RB = 4;
      const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
      for (int row0 = (threadIdx.x >> 6) * RB; row0 < 2 * N; row0 += nw * RB) {
        for (int k0 = 0; k0 < N; k0 += 256) {
          double gv[RB][4];
#pragma unroll
          for (int b = 0; b < RB; ++b) {
            const int row = row0 + b;
            const bool rok = row < 2 * N;
            const LQMC_GLOBAL T* Gr = gptr(G + (rok ? (size_t)row * N : 0));
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int k = k0 + 64 * q + lane;
              gv[b][q] = (double)ldg0(Gr, (size_t)k, rok && k < N);
            }
          }
#pragma unroll
          for (int b = 0; b < RB; ++b) {
            const int row = row0 + b;
            if (row >= 2 * N) continue;
            const int sp = row >= N ? 1 : 0;
            const int j = row - sp * N;
            const double xj = xv[sp * N + j];
            const double cj = coef[sp];
            T* Gr = G + sp * n2 + (size_t)j * N;
            const double* y = yv + sp * N;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int k = k0 + 64 * q + lane;
              if (k >= N) continue;
              const double p = ref ? xj * y[k] : cj * (xj * y[k]);
              Gr[k] = (T)(gv[b][q] - p);
            }
          }
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) fld[i] = (int8_t)(-s);
    ++nacc;
    __syncthreads();
  }
  slice_epilogue<T>(a, ch, c, l, G, nacc, underrun, bad, red);
}

// ---- delayed (submatrix) form ---------------------------------------------------------------------
template <typename T> struct MfmaOp;
// 16x16 accumulator rows held by a lane: FP64 lane/16 + 4r, FP32 4*(lane/16) + r.
template <> struct MfmaOp<double> {
  using acc_t = v4d;
  __device__ static inline acc_t mma(double x, double y, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct MfmaOp<float> {
  using acc_t = v4f;
  __device__ static inline acc_t mma(float x, float y, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return ((lane >> 4) << 2) + r; }
};

// ---- split-BF16 products -----------------------------------------------------------------------
// x = x_hi + x_lo with x_hi = bf16(x) (round to nearest even) and x_lo = bf16(x - x_hi); a
// product a b is taken as a_hi b_hi + a_hi b_lo + a_lo b_hi (the a_lo b_lo term, ~2^-16 of it,
// dropped), each an exact BF16 product summed in FP32 on the matrix cores.
typedef short v4s __attribute__((ext_vector_type(4)));

__device__ inline unsigned short bf16_rne(float x) {
  const unsigned u = __float_as_uint(x);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// The four FP32 values of one lane (k = 4 (lane / 16) + j of a 16x16x16 operand) as hi and lo halves.
__device__ inline void bf16_split4(const float (&x)[4], v4s& hi, v4s& lo) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const unsigned short h = bf16_rne(x[j]);
    hi[j] = (short)h;
    lo[j] = (short)bf16_rne(x[j] - bf16_to_f32(h));
  }
}

// acc += A B over k = 16 as three BF16 products (16x16x16, FP32 accumulation); A lane l holds
// A[l % 16][4 (l / 16) + j], B lane l holds B[4 (l / 16) + j][l % 16], j = 0..3.
__device__ inline v4f mma_bf16x3(const v4s& ah, const v4s& al, const v4s& bh, const v4s& bl, v4f acc) {
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc, 0, 0, 0);
  return acc;
}

// Orders LDS traffic between the lanes of one wavefront (the serial phase runs in wave 0).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane (l & ~3) + src of each lane quad, for a double: a DPP quad_perm broadcast (a VALU
// operation, no LDS round trip).  src is 0..3 and a constant once the caller's loop unrolls.
template <int CTRL>
__device__ inline double quad_dpp(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline double quad_bcast(double v, int src) {
  switch (src) {
    case 0: return quad_dpp<0x00>(v);
    case 1: return quad_dpp<0x55>(v);
    case 2: return quad_dpp<0xAA>(v);
    default: return quad_dpp<0xFF>(v);
  }
}

// Phase clocks of the delayed flip (built only with -DLQMC_FLIP_PROBE, as the separate
// liblqmc_hip_probe.so of latticeqmc_amd.flip_phases): thread 0 of each workgroup sums the
// wall-clock ticks (100 MHz) of each phase of its blocks and, with tracing on, writes them over
// the slice's trace ratios at t N + 16 w + p (p < 14; 14: blocks, 15: total) when the launch
// ends; the per-site trace is not written.  Phases: 0 chain barrier, 1 uniforms and sync,
// 2 wave 0's diagonal-block loads, 3 sites, 4 wait for the other waves, 12 tables,
// 13 inverses, 5 scatter, 6 Z tiles, 7 strips and copies, 8 tail.
#ifdef LQMC_FLIP_PROBE
#define LQMC_PROBE(p) probe_mark(p)
#else
#define LQMC_PROBE(p) ((void)0)
#endif

// Per block of k sites (k a multiple of 4):
//  1. the block's field entries and uniforms go to LDS;
//  2. wave 0 runs the k attempts on the k x k diagonal blocks of both spins, held in
//     its registers (rank-1 Sherman-Morrison on k x k), recording for accepted j the
//     block parts of x_j = c_j G^(j)[:, a_j] and y_j = e_{a_j} - G^(j)[a_j, :]^T;
//  3. with U[p][j] = y_p(a_j), L[p][j] = x_p(a_j) (p < j), the full factors are
//     X = G0[:, A] C (I + U C)^-1 and Y = (E_A - G0[A, :]^T) (I - L)^-1; four waves form
//     the unit-triangular inverses and scatter Mx = C (I + U C)^-1 and
//     My = (I - L)^-1 onto the block's rows (k x m, zero on rejected sites);
//  4. G <- G0 + P Z with P = -G0[:, B] Mx and Z = My^T (E_B - G0[B, :]) on the matrix
//     cores (= G0 - sum_j x_j y_j^T), form_z and update_strip below.
// Rank-m update of both spins' G by an accepted block of sites B = b0 .. b0+kb-1 (form_z,
// then update_strip over the strips):
//   G_s += P_s Z_s,  P_s = -G0_s[:, B] Mx_s (N x m),  Z_s = My_s^T (E_B - G0_s[B, :]) (m x N),
// with Mx_s, My_s the k x m factors scattered on the block rows (LDS, leading dimension K).
// In the 16x16x4 matrix-core layout, lane l's accumulator element v of a 16 x 16 product
// is exactly what lane l supplies as the A (or B) operand of k-step v of a following
// product, with k = l/16 + 4v (FP64) or 4(l/16) + v (FP32) -- any consistent k order sums
// the same terms.  So
//  phase 1 (all wavefronts): the Z_s tiles (16 accepts x 16 columns) are formed on the
//    matrix cores and stored to the workspace in that register order; lane l later reads
//    back its own element at ((tile * 4 + v) * 64 + l), a coalesced load;
//  phase 2 (after a workgroup barrier): each wavefront owns whole 16-row strips of G_s,
//    forms P_s^T for its strip (in operand order, kept in registers) and streams the
//    strip's 16 x 16 tiles through G += P Z, RU tiles at a time with every load of a group
//    issued before its first matrix-core op.
// Phase 1 reads rows B of G0 before any strip is updated; phase 2 reads G0[strip, B] and
// updates only its own strip.
// spmask: bit s set = the tiles of spin s are formed (a workgroup forms only the spins of the
// strips it updates).
// LROWS: the workgroup forms one spin's tiles and rows B of that spin are in LDS (Yl, leading
// dimension N), read in place of G's rows.  ZL: the tiles are stored to LDS (zl, tile t - tbase
// of the workgroup's spin, same register order) instead of the workspace.
template <typename T, int K, bool LROWS = false, bool ZL = false>
__device__ void form_z(const T* G, size_t n2, int N, int Np, int b0, int kb, int m, const double* My, T* zws,
                       int spmask = 3, const T* Yl = nullptr, T* zl = nullptr) {
  using acc_t = typename MfmaOp<T>::acc_t;
  constexpr int JT = (K + 15) / 16;  // 16-wide tiles over the accepts
  constexpr int KS = K / 4;           // k-steps of one tile
  // tiles per wavefront whose loads are issued together (ZL: one spin, 16 tiles over 8 waves)
  constexpr int TB = K >= 32 || ZL ? 2 : 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nct = Np >> 4;
  const int jtu = (m + 15) >> 4;
  const int tbase = spmask == 2 ? nct * JT : 0;            // first tile formed
  const int ntile = tbase + __builtin_popcount(spmask & 3) * nct * JT;  // one past the last
  const LQMC_GLOBAL T* Gg = gptr(G);
  LQMC_GLOBAL T* Zg = gptr(zws);
  for (int t0 = tbase + wave; t0 < ntile; t0 += TB * nw) {
    T xv[TB][KS], yv[TB][KS];
    // every load of the group first (clamped addresses, masked values), then the products
#pragma unroll
    for (int u = 0; u < TB; ++u) {
      const int t = t0 + u * nw;
      const int sp = (t / (nct * JT)) & 1, ct = (t / JT) % nct, jt = t % JT;
      const bool ok = t < ntile && jt < jtu;
      const double* Mys = My + sp * K * K;
      const LQMC_GLOBAL T* Gs = Gg + sp * n2;
      const int c = (ct << 4) + li, a = (jt << 4) + li;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int q = 4 * ks + lk;
        xv[u][ks] = (T)ld0(Mys, (size_t)q * K + a, ok && q < kb && a < m);
        T y;
        if constexpr (LROWS) y = -(ok && q < kb && c < N ? Yl[q * N + c] : T(0));
        else y = -ldg0(Gs, (size_t)(b0 + q) * N + c, ok && q < kb && c < N);
        if (ok && q < kb && c == b0 + q) y = y + T(1);
        yv[u][ks] = y;
      }
    }
#pragma unroll
    for (int u = 0; u < TB; ++u) {
      const int t = t0 + u * nw;
      if (t >= ntile || t % JT >= jtu) continue;
      acc_t z = acc_t{0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) z = MfmaOp<T>::mma(xv[u][ks], yv[u][ks], z);
      if constexpr (ZL) {
        T* zo = zl + (size_t)(t - tbase) * 256 + lane;
#pragma unroll
        for (int v = 0; v < 4; ++v) zo[v * 64] = z[v];
      } else {
        LQMC_GLOBAL T* zo = Zg + (size_t)t * 256 + lane;
#pragma unroll
        for (int v = 0; v < 4; ++v) zo[v * 64] = z[v];
      }
    }
  }
}

// Phase 2 for one 16-row strip rs (spin rs / nct, rows (rs % nct) * 16 ..) of G, by one
// wavefront: the column-tile groups cg, cg + ng, ... of RU tiles each (default: all).  P is
// formed from G0[strip rows, B], or from the 16 x 16 copy pin of it when given (a strip
// whose column groups are shared by several wavefronts: one of them may already have
// updated the tile holding columns B).
// FULL: N is a multiple of 16 RU (every strip row and column-tile group lies inside G) and
// K = 16 (one tile of accepts, m >= 1), so the tile loads and stores of the column loop need no
// bounds checks; the products are the same.  ZL: the Z tiles are read from LDS (zl, the layout
// form_z<..., ZL> stores).
template <typename T, int K, int RU, bool BF3 = false, bool FULL = false, bool ZL = false>
__device__ void update_strip(T* G, size_t n2, int N, int Np, int b0, int kb, int m, const double* Mx,
                             const T* zws, int rs, int cg = 0, int ng = 1, const double* pin = nullptr,
                             const T* zl = nullptr) {
  static_assert(!FULL || K == 16, "full-tile form: one tile of accepts");
  using acc_t = typename MfmaOp<T>::acc_t;
  constexpr int JT = (K + 15) / 16;
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int nct = Np >> 4;
  const int jtu = (m + 15) >> 4;
  LQMC_GLOBAL T* Gg = gptr(G);
  const LQMC_GLOBAL T* Zg = gptr(zws);
  const int sp = rs / nct, r0 = (rs % nct) << 4;
  const double* Mxs = Mx + sp * K * K;
  LQMC_GLOBAL T* Gs = Gg + sp * n2;
  const int r = r0 + li;
  acc_t pt[JT];
#pragma unroll
  for (int jt = 0; jt < JT; ++jt) {
    pt[jt] = acc_t{0, 0, 0, 0};
    if (jt < jtu) {
      const int a = (jt << 4) + li;
      for (int kk = 0; kk < kb; kk += 4) {
        const int q = kk + lk;
        const T x = (T)ld0(Mxs, (size_t)q * K + a, q < kb && a < m);
        const T y = pin ? (T)ld0(pin, (size_t)li * 16 + q, r < N && q < kb)
                        : ldg0(Gs, (size_t)r * N + b0 + q, r < N && q < kb);
        pt[jt] = MfmaOp<T>::mma(x, y, pt[jt]);
      }
      pt[jt] = -pt[jt];
    }
  }
  for (int c0 = cg * RU; c0 < nct; c0 += RU * ng) {
    acc_t acc[RU];
    T zv[RU][JT][4];
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int ct = c0 + u;
      const int col = (ct << 4) + li;
      const bool ok = ct < nct && col < N;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + MfmaOp<T>::row(lane, v);
        if constexpr (FULL) acc[u][v] = Gs[(size_t)row * N + col];
        else acc[u][v] = ldg0(Gs, (size_t)row * N + col, ok && row < N);
      }
#pragma unroll
      for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if constexpr (ZL) {
            const int zli = ((ct * JT + jt) * 4 + v) * 64 + lane;
            if constexpr (FULL) zv[u][jt][v] = zl[zli];
            else zv[u][jt][v] = ct < nct && jt < jtu ? zl[zli] : T(0);
          } else {
            const size_t zi = ((size_t)((sp * nct + ct) * JT + jt) * 4 + v) * 64 + lane;
            if constexpr (FULL) zv[u][jt][v] = Zg[zi];
            else zv[u][jt][v] = ldg0(Zg, zi, ct < nct && jt < jtu);
          }
        }
    }
    if constexpr (BF3) {
      // FP32 operands: lane l's P^T accumulator elements v = 0..3 are P[l % 16][4 (l / 16) + v] and
      // its Z elements are Z[4 (l / 16) + v][l % 16] -- the k = 16 operand layout of the BF16 product
#pragma unroll
      for (int jt = 0; jt < JT; ++jt)
        if (jt < jtu) {
          const float pv[4] = {(float)pt[jt][0], (float)pt[jt][1], (float)pt[jt][2], (float)pt[jt][3]};
          v4s ph, pl;
          bf16_split4(pv, ph, pl);
#pragma unroll
          for (int u = 0; u < RU; ++u) {
            const float zz[4] = {(float)zv[u][jt][0], (float)zv[u][jt][1], (float)zv[u][jt][2], (float)zv[u][jt][3]};
            v4s zh, zl;
            bf16_split4(zz, zh, zl);
            acc[u] = mma_bf16x3(ph, pl, zh, zl, acc[u]);
          }
        }
    } else {
#pragma unroll
      for (int u = 0; u < RU; ++u)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
          if (jt < jtu) {
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u] = MfmaOp<T>::mma(pt[jt][v], zv[u][jt][v], acc[u]);
          }
    }
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int ct = c0 + u;
      const int col = (ct << 4) + li;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + MfmaOp<T>::row(lane, v);
        if (FULL || (ct < nct && row < N && col < N)) Gs[(size_t)row * N + col] = acc[u][v];
      }
    }
  }
}

// ---- several workgroups per chain ------------------------------------------------------------
// With W > 1 workgroups on one chain (blocks of 16 or 32 sites, so a block is whole 16-row
// strips), every workgroup runs the serial phase, the inverses and the Z tiles of each block
// itself (from the same G rows and uniforms, so with identical results), and the 2 ceil(N/16)
// strips of both spins are shared out over the W x (waves per workgroup) wavefronts.  Between
// blocks the W workgroups meet at a per-chain barrier.  The strips holding the block's own
// rows B are read by every workgroup during the block (serial phase and Z), so their owner
// defers their update until after the next barrier, when no workgroup reads them any more
// (later blocks read only their own rows B' and each strip owner its own rows).
//
// Rendezvous of the W workgroups of one chain, through counters in global memory:
// bar[0] counts arrivals within one launch (the n-th rendezvous waits for n W), bar[1] counts
// exits (the last workgroup to exit clears bar[0..2], so each launch starts from zero), bar[2]
// holds one bit per XCC that runs a workgroup of the chain, set at the start.  A wait is
// bounded by FLIP_BARRIER_SPINS polls; on expiry the workgroup records flag 16, stops waiting
// for the rest of the launch and still arrives, so every wave exits.
//
// A barrier publishes each workgroup's G stores to the others.  At device scope that is an L2
// write-back before arriving and an L2 invalidate after the wait: the XCDs have separate L2s.
// The block map gives a chain's workgroups block indices of equal residue mod 8, which is one
// XCD under round-robin dispatch; workgroups on one XCD share its L2, so it suffices that every
// wave's own stores have reached the L2 (vmcnt 0) before the arrival and that the acquire
// invalidates L1.  The first barrier of a launch is always at device scope; after it every
// workgroup reads the XCC bits and keeps the L2-local form for the rest of the launch only if
// there is a single bit.
constexpr unsigned FLIP_BARRIER_SPINS = 1u << 22;

// XCC that runs the calling wavefront (hardware register XCC_ID, bits 3:0).
__device__ inline unsigned xcc_id() {
  return (unsigned)__builtin_amdgcn_s_getreg(20 | (0 << 6) | ((4 - 1) << 11)) & 15u;
}

struct ChainSync {
  unsigned* bar;
  ChainState* st;
  int* flag;       // one int of LDS: the L2-local decision, broadcast to the workgroup
  unsigned W, n;   // workgroups per chain, rendezvous so far in this launch
  bool broken, local, decided;
};

__device__ inline ChainSync chain_sync_begin(unsigned* bar, ChainState* st, int W, int* flag) {
  if (W > 1 && threadIdx.x == 0) atomicOr(bar + 2, 1u << xcc_id());
  return ChainSync{bar, st, flag, (unsigned)W, 0u, false, false, false};
}

__device__ inline void chain_wait(ChainSync& cs, unsigned target) {
  if (cs.broken) return;
  unsigned spins = 0;
  while (__hip_atomic_load(cs.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(1);
    if (++spins >= FLIP_BARRIER_SPINS) {
      cs.broken = true;
      atomicOr(&cs.st->flags, 16u);
      break;
    }
  }
}

__device__ inline void chain_barrier(ChainSync& cs) {
  const unsigned target = ++cs.n * cs.W;
  if (cs.local) __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are in the shared L2
  __syncthreads();
  if (threadIdx.x == 0) {
    // device scope: release the workgroup's G stores (ordered before this point by the
    // barrier above), arrive, wait, then acquire
    if (!cs.local) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    atomicAdd(cs.bar, 1u);
    chain_wait(cs, target);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (!cs.decided) {  // every workgroup set its XCC bit before arriving here
      const unsigned m = __hip_atomic_load(cs.bar + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *cs.flag = !cs.broken && m != 0 && (m & (m - 1)) == 0;
    }
  }
  __syncthreads();
  if (!cs.decided) {
    cs.local = *cs.flag != 0;
    cs.decided = true;
  }
}

// The reads-before-writes half of a step: every workgroup of the chain has finished reading the
// state it decides and updates from (its loads have returned: their values are in LDS or were
// consumed) before any of them writes.  No memory is published, so no fences; the counter is
// the barrier counter, whose targets it shares.
__device__ inline void chain_sync_reads(ChainSync& cs) {
  const unsigned target = ++cs.n * cs.W;
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(cs.bar, 1u);
    chain_wait(cs, target);
  }
  __syncthreads();
}

__device__ inline void chain_exit(ChainSync& cs) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(cs.bar + 1, 1u) == cs.W - 1) {
      atomicExch(cs.bar, 0u);
      atomicExch(cs.bar + 1, 0u);
      atomicExch(cs.bar + 2, 0u);
    }
  }
}

// Workgroup size bound of the delayed kernel: blocks of 16 or more sites keep their
// k x k work in registers that only fit the 256-VGPR budget of a 512-thread workgroup.
constexpr int flip_delayed_max_threads(int k) { return k >= 16 ? 512 : 1024; }

// Block index -> (chain, workgroup of the chain).  W > 1: the W workgroups of a chain get block
// indices of equal residue mod 8, i.e. the same XCD under round-robin dispatch, so the rows
// they share stay in one L2; blocks past the last chain return at once.
__device__ inline void flip_block_map(const FlipArgs& a, int& c, int& w) {
  const int W = a.wpc > 1 ? a.wpc : 1;
  if (W == 1) {
    c = blockIdx.x;
    w = 0;
    return;
  }
  const int b = blockIdx.x, slot = b >> 3;
  c = (b & 7) + 8 * (slot / W);
  w = slot % W;
}

template <typename T, int K, bool BF3 = false>
__global__ __launch_bounds__(flip_delayed_max_threads(K)) void k_flip_delayed(FlipArgs a) {
  static_assert(K >= 4 && K <= 32 && K % 4 == 0, "block size");
  int c, w;
  flip_block_map(a, c, w);
  const int W = a.wpc > 1 ? a.wpc : 1;
  if (W > 1 && c >= a.nchains) return;
  const ChainDev ch = a.chains[c];
  if (a.t >= ch.L) return;
  const int N = ch.N;
  const int l = ch.L - 1 - a.t;
  const int Np = (N + 15) & ~15;
  const int nct = Np >> 4;
  const size_t n2 = (size_t)N * N;
  T* __restrict__ G = static_cast<T*>(ch.G);
  int8_t* fld = ch.field + (size_t)l * N;
  // workspace: the Z tiles of both spins in register order (2 x K16 x Np, form precision),
  // one per workgroup of the chain
  T* wsT = static_cast<T*>(ch.dws) + (size_t)w * 2 * ((K + 15) & ~15) * Np;
  unsigned* bar = a.bar + 4 * c;

  LQMC_DYNAMIC_LDS(double, sm);
  double* red = sm;                        // blockDim (epilogue reductions)
  double* Xsb = red + blockDim.x;          // [2][K rows][K accepts]: x_j block parts, later Mx scattered
  double* Ysb = Xsb + 2 * K * K;           // [2][K][K]: y_j block parts, later My scattered
  double* cf = Ysb + 2 * K * K;            // [2][K] coefficients of the accepts
  double* ub = cf + 2 * K;                 // [K] uniforms of the block
  int* aidx = reinterpret_cast<int*>(ub + K);  // [K] block-local site of each accept
  int* sb = aidx + K;                          // [K] field entries of the block
  int* m_sh = sb + K;                          // [4] accept count of the block; [1] barrier flag
  double* Mxp = reinterpret_cast<double*>(m_sh + 4);  // [2][K][K] Mx of the previous block (W > 1)
  double* Pin = Mxp + 2 * K * K;  // [2][16][16] G0[rows, columns] of the deferred strips (K = 16)
  T* Yl = reinterpret_cast<T*>(Pin + (K == 16 ? 512 : 0));  // [16][N] rows B of the workgroup's spin (a.yrows)
  T* Zl = Yl + (K == 16 && a.yrows ? 16 * N : 0);  // [Np / 16][256] Z tiles of the workgroup's spin (a.yrows)
  // [16][16] G0[rows of the next block, columns B] and the count of waves done with that strip
  double* PinN = reinterpret_cast<double*>(Zl + (K == 16 && a.yrows ? 16 * Np : 0));
  unsigned* ecnt = reinterpret_cast<unsigned*>(PinN + (K == 16 && a.yrows ? 256 : 0));
  int8_t* fl = reinterpret_cast<int8_t*>(ecnt + (K == 16 && a.yrows ? 4 : 0));  // [N] the slice's field

  const unsigned long long sweep = a.st[c].sweep;
  unsigned long long nacc = 0;
  bool underrun = false, bad = false;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int gw = w * nw + wave, tw = W * nw;  // wavefront index within the chain, and count
  // the spins whose strips this workgroup updates (strip rs belongs to workgroup (rs % tw) / nw);
  // only their Z tiles are formed
  int spmask = 0;
  for (int rs = 0; rs < 2 * nct; ++rs)
    if ((rs % tw) / nw == w) spmask |= 1 << (rs / nct);
  // whole strips and column-tile groups of 4 tiles: the strip updates need no bounds checks
  const bool full = K == 16 && N % 64 == 0;
  // one spin per workgroup (W > 1): rows B of that spin are staged in LDS by the helper waves
  // (below) while wave 0 runs the serial phase, and the Z tiles read them there.  Rows B are
  // final after the block's chain barrier and stay untouched until the next one (the strips
  // holding them are deferred), so the staged rows equal what the Z tiles would read from G.
  // With whole tiles (full) the Z tiles themselves stay in LDS too (Zl): the strips of the
  // block and the deferred strips, which run before the next block's Z tiles, read them there.
  const bool lrows = K == 16 && a.yrows && W > 1 && (spmask == 1 || spmask == 2);
  // Early arrival (one spin per workgroup, whole tiles, eight waves, one strip per wave of the
  // chain, L2-local barriers): the only rows another workgroup reads in the next block are that
  // block's rows, so a workgroup arrives at the chain barrier as soon as those rows are final --
  // at once after its Z tiles if it does not hold their strip, else after the two waves idle in
  // its strip phase (the ones whose strips are this block's and the next block's rows) have
  // updated that strip first and their stores are in L2.  The barrier at the next block's start
  // then only waits.
  const bool early = lrows && full && nw == 8 && tw == 2 * nct;
  bool arrived = false;
  for (int i = threadIdx.x; i < N; i += blockDim.x) fl[i] = fld[i];
  __syncthreads();
  bool pending = false, prepped = false;
  int pb0 = 0, pkb = 0, pm = 0;
#ifdef LQMC_FLIP_PROBE
  unsigned long long ph[16] = {0};
  unsigned long long tprev = wall_clock64();
  const unsigned long long tstart = tprev;
  int nblk = 0;
  auto probe_mark = [&](int p) {
    const unsigned long long tt = wall_clock64();
    ph[p] += tt - tprev;
    tprev = tt;
  };
#endif
  ChainSync cs = chain_sync_begin(bar, &a.st[c], W, m_sh + 1);
  // The deferred strips of the previous block (its own rows, after the chain barrier), with
  // that block's Mx kept in Mxp and its Z tiles still in the workspace.  Each is applied by
  // its owner workgroup; while wave 0 runs the next serial phase the strips go to waves 1 ..
  // Each strip's column-tile groups are shared out over those waves, so a strip is not
  // left to one wavefront's latency.
  // For K = 16 (one strip per spin) each strip's column-tile groups are shared out over
  // those waves, with P formed from the copy Pin; for K = 32 a wave takes whole strips.
  // While wave 0 runs the serial phase, the other work of the block goes to the waves that
  // do not share its SIMD (wave w runs on SIMD w % 4): helper i is wave i + i / 3 + 1, i.e.
  // waves 1, 2, 3, 5, 6, 7 of eight (nw - (nw + 3) / 4 helpers).
  auto helper_wave = [](int i) { return i + i / 3 + 1; };
  auto deferred_strips = [&](bool serial_busy) {
    constexpr int RUd = K >= 16 ? 4 : 2;
    const int s0 = pb0 >> 4, s1 = (pb0 + pkb - 1) >> 4;
    const int w0 = serial_busy ? 1 : 0, nwv = nw - w0;
    const int nhw = nw - (nw + 3) / 4;  // helper waves
    if constexpr (K == 16) {
      const int ngr = (nct + RUd - 1) / RUd;  // column-tile groups per strip
      int item = 0;                           // (strip, group) items of this workgroup
      for (int sp = 0; sp < 2; ++sp) {
        const int rs = sp * nct + s0;
        if ((rs % tw) / nw != w) continue;
        for (int g = 0; g < ngr; ++g, ++item)
          if (wave == (serial_busy ? helper_wave(item % nhw) : w0 + item % nwv))
            if (full && lrows) update_strip<T, K, RUd, BF3, K == 16, true>(G, n2, N, Np, pb0, pkb, pm, Mxp, wsT, rs, g, ngr, Pin + sp * 256, Zl);
            else if (full) update_strip<T, K, RUd, BF3, K == 16>(G, n2, N, Np, pb0, pkb, pm, Mxp, wsT, rs, g, ngr, Pin + sp * 256);
            else update_strip<T, K, RUd, BF3>(G, n2, N, Np, pb0, pkb, pm, Mxp, wsT, rs, g, ngr, Pin + sp * 256);
      }
    } else {
      int i = 0;
      for (int sp = 0; sp < 2; ++sp)
        for (int st = s0; st <= s1; ++st, ++i) {
          const int rs = sp * nct + st, own = rs % tw;
          if (own / nw != w) continue;
          const int dw = serial_busy ? 1 + i % (nw - 1) : own % nw;
          if (wave == dw) update_strip<T, K, RUd, BF3>(G, n2, N, Np, pb0, pkb, pm, Mxp, wsT, rs);  // K = 32
        }
    }
  };

  for (int b0 = 0; b0 < N; b0 += K) {
    const int kb = min(K, N - b0);
    if constexpr (K >= 16) {
      if (pending) {
        if (arrived) {
          // arrival made during the previous block: wait, then acquire
          const unsigned target = ++cs.n * cs.W;
          if (threadIdx.x == 0) {
            chain_wait(cs, target);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          }
          __syncthreads();
        } else {
          chain_barrier(cs);
        }
      }
      arrived = false;
    }
    LQMC_PROBE(0);
#ifdef LQMC_FLIP_PROBE
    ++nblk;
#endif
    // the block's field entries and uniforms, unless wave 4 drew them during the previous
    // block's inverses
    if (!prepped && threadIdx.x < kb) {
      const int i = b0 + threadIdx.x;
      sb[threadIdx.x] = fl[i];
      bool und = false;
      ub[threadIdx.x] = draw_uniform(ch, &a.st[c], c, a.rng, sweep, a.t, i, und);
      if (und) underrun = true;
    }
    prepped = false;
    __syncthreads();
    LQMC_PROBE(1);
    if (wave == 0) {
      // The k x k blocks of both spins live in wave 0's registers: lane
      // (sp, h, q) = (lane / (K H), lane / K % H, lane % K) holds rows h R .. h R + R-1
      // of column q of spin sp (H lanes per column, R = K / H rows each).  The loop over
      // the block's sites stays rolled; row entries are picked by select chains.
      constexpr int H = (64 / (2 * K)) < K ? (64 / (2 * K)) : K;
      constexpr int R = K / H;
      const int sp = lane / (K * H) < 2 ? lane / (K * H) : 1;
      const int h = (lane / K) % H;
      const int q = lane % K;
      const bool live = lane < 2 * K * H && q < kb;
      double col[R];
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const int p = h * R + i;
        col[i] = (live && p < kb) ? (double)G[sp * n2 + (size_t)(b0 + p) * N + b0 + q] : 0.0;
      }
      // the block's field entries and uniforms, lane jj holding site jj (picked by readlane)
      const int sb_l = sb[lane < kb ? lane : 0];
      const double ub_l = ub[lane < kb ? lane : 0];
#ifdef LQMC_FLIP_PROBE
      __builtin_amdgcn_s_waitcnt(0);
      LQMC_PROBE(2);
#endif
      // the chain's exp(+-2 lambda) held in VGPRs: as scalars the compiler keeps them with a
      // 64-byte block of the chain record that it spills and reloads at every site
      double e2pv = ch.e2p, e2mv = ch.e2m;
      LQMC_OPAQUE(e2pv, "v");
      LQMC_OPAQUE(e2mv, "v");
      // One site jj (hj = jj / R) with cj = G(current)[jj][q] of my spin on the lanes with
      // h == hj.  For R <= 8 the sites of one hj are unrolled, so cj is a register of col;
      // otherwise it is picked by a select chain.
      int m = 0;
      auto site = [&](const int jj, const int hj, const double cj) {
        const double gu = readlane_f64(cj, hj * K + jj);
        const double gd = readlane_f64(cj, K * H + hj * K + jj);
        const int s = __builtin_amdgcn_readlane(sb_l, jj);
        const double e2p = e2pv, e2m = e2mv;
        const double ep = s > 0 ? e2p : e2m;
        const double em = s > 0 ? e2m : e2p;
        const double d_up = 1.0 + (1.0 - gu) * (ep - 1.0);
        const double d_dn = 1.0 + (1.0 - gd) * (em - 1.0);
        const double ratio = d_up * d_dn;
        const bool acc = readlane_f64(ub_l, jj) <= ratio;
        if (!isfinite(ratio)) bad = true;
#ifdef LQMC_FLIP_PROBE
        if (false) {
#else
        // to LDS (red is free until the epilogue), written out below; the lane test is recomputed
        // from a VGPR copy (kept as a lane mask it is spilled and reloaded at every site)
        int lv = lane;
        LQMC_OPAQUE(lv, "v");
        if (a.trace && lv == 0) {
#endif
          red[jj] = ratio;
          red[K + jj] = acc ? 1.0 : 0.0;
        }
        // The exchanges and the coefficient do not depend on the decision: they are
        // issued before it, so their latencies overlap.  My spin's coefficient is
        // c = (e^{+-2 lambda s} - 1) / d (one division per lane); y = e_jj - G[jj, :]^T has
        // column q's entry on lane (sp, hj, q); my rows of x = column jj of my spin come
        // from lane src.
        const int src = sp * K * H + h * K + jj;
        double yq = -__shfl(cj, sp * K * H + hj * K + q, 64);
        double xp[R];
#pragma unroll
        for (int i = 0; i < R; ++i) xp[i] = __shfl(col[i], src, 64);
        // the spin test is recomputed from a VGPR copy at every site (kept as a lane mask it is
        // spilled and reloaded)
        int spv = sp;
        LQMC_OPAQUE(spv, "v");
        const double num = spv == 0 ? ep - 1.0 : em - 1.0;
        const double cfs = num / (spv == 0 ? d_up : d_dn);
        if (acc) {
          if (q == jj) yq = yq + 1.0;
          // G <- G - c x y^T on the k x k block.  Entries of rows >= kb and of lanes past
          // the block are never read (the pivots and exchanges come from lanes q = jj < kb,
          // rows < kb), so the update runs unmasked.
          const double cy = cfs * yq;
#pragma unroll
          for (int i = 0; i < R; ++i) col[i] = fma(-cy, xp[i], col[i]);
          if (lane == src) {  // rows >= kb of the block parts are never read
#pragma unroll
            for (int i = 0; i < R; ++i) Xsb[sp * K * K + (h * R + i) * K + m] = cfs * xp[i];
          }
          if (live && h == 0) Ysb[sp * K * K + q * K + m] = yq;
          if (lane == 0) {
            cf[m] = cfs;
            aidx[m] = jj;
            fl[b0 + jj] = (int8_t)(-s);
          }
          if (lane == K * H) cf[K + m] = cfs;
          ++m;
        }
      };
      if constexpr (R <= 8) {
#pragma unroll 1
        for (int hj = 0; hj * R < kb; ++hj) {
#pragma unroll
          for (int ij = 0; ij < R; ++ij)
            if (hj * R + ij < kb) site(hj * R + ij, hj, col[ij]);
        }
      } else {
#pragma unroll 1
        for (int jj = 0; jj < kb; ++jj) {
          double cj = 0.0;
#pragma unroll
          for (int i = 0; i < R; ++i)
            if (i == jj % R) cj = col[i];
          site(jj, jj / R, cj);
        }
      }
      if (lane == 0) m_sh[0] = m;
      LQMC_PROBE(3);
    } else if (K >= 16) {
      if (lrows) {
        const LQMC_GLOBAL T* Gs = gptr(G) + (spmask >> 1) * n2 + (size_t)b0 * N;
        const int hi = wave - 1 - (wave - 1) / 4;  // helper index of this wave
        const int nt = 64 * (nw - (nw + 3) / 4), tt = 64 * hi + lane, tot = kb * N;
        if (wave % 4 != 0)
          for (int e0 = tt; e0 < tot; e0 += 8 * nt) {
            T v[8];  // every load of the batch first
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = ldg0(Gs, (size_t)(e0 + u * nt), e0 + u * nt < tot);
#pragma unroll
            for (int u = 0; u < 8; ++u)
              if (e0 + u * nt < tot) Yl[e0 + u * nt] = v[u];
          }
      }
      if (pending) deferred_strips(true);
    }
    __syncthreads();
#ifdef LQMC_FLIP_PROBE
    if (false) {
#else
    if (a.trace && w == 0 && threadIdx.x < kb) {
#endif
      ch.tr_ratio[(size_t)a.t * N + b0 + threadIdx.x] = red[threadIdx.x];
      ch.tr_acc[(size_t)a.t * N + b0 + threadIdx.x] = red[K + threadIdx.x] != 0.0 ? 1 : 0;
    }
    pending = false;
    const int m = m_sh[0];
    LQMC_PROBE(4);
    if (m == 0) continue;
    nacc += m;
    // The coefficients of the two unit upper-triangular systems per spin, in accept
    // order (r < p < m, else 0): upper T[r][p] = y_r(a_p) c_p, lower T[r][p] = -c_r x_r(a_p).
    // They replace the block parts in place: matrix (spin s, lower) at Xsb + s 2K^2 + lower K^2.
    constexpr int TE = 4 * K * K;
    constexpr int TPT = (TE + 255) / 256;  // entries per thread (at least 256 threads)
    {
      double tv[TPT];
#pragma unroll
      for (int i = 0; i < TPT; ++i) {
        const int e = threadIdx.x + i * blockDim.x;
        double v = 0.0;
        if (e < TE) {
          const int ts = e / (2 * K * K), tl = (e / (K * K)) & 1, r = (e / K) % K, p = e % K;
          if (r < p && p < m) {
            const int ap = aidx[p];
            v = tl ? -Xsb[ts * K * K + ap * K + r] : Ysb[ts * K * K + ap * K + r] * cf[ts * K + p];
          }
        }
        tv[i] = v;
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < TPT; ++i) {
        const int e = threadIdx.x + i * blockDim.x;
        if (e < TE) Xsb[e] = tv[i];
      }
    }
    __syncthreads();
    LQMC_PROBE(12);
    // unit upper-triangular inverses, one per wave: waves 0/1 spin 0 (I+UC)^-1 / (I-L)^-1,
    // waves 2/3 spin 1, then scattered onto the block rows:
    // Mx[aidx[r]][j] = c_r inv[r][j], My[aidx[r]][j] = inv[r][j] (zero on rejected sites).
    const int isp = wave >> 1;
    const bool lower = (wave & 1) != 0;
    if constexpr (K <= 16) {
      // Lane 4 j + g holds rows r = g + 4 i (i < K / 4) of column j.  Column-oriented back
      // substitution: once row p of column j is final (on lane 4 j + p % 4) it is broadcast
      // through the lane quad (DPP) and every row r < p takes its term -T[r][p] inv[p] in one
      // fused multiply-add.  The table is strictly upper triangular with exact zeros
      // elsewhere, so rows r >= p take zero terms and need no mask.  Every table entry a
      // lane uses is read into registers first: the substitution waits on no LDS read.
      constexpr int NI = K / 4;
      const int j = lane >> 2, g = lane & 3;
      double iv[NI];
      if (wave < 4) {
        const double* Tw = Xsb + isp * 2 * K * K + (lower ? K * K : 0);
        double tr[NI][K];  // tr[i][p] = T[g + 4 i][p], for p > 4 i
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
          for (int p = 4 * i + 1; p < K; ++p) tr[i][p] = Tw[(g + 4 * i) * K + p];
#pragma unroll
        for (int i = 0; i < NI; ++i) iv[i] = g + 4 * i == j ? 1.0 : 0.0;
#pragma unroll
        for (int p = K - 1; p >= 1; --p) {
          const double ip = quad_bcast(iv[p / 4], p % 4);
#pragma unroll
          for (int i = 0; i < NI; ++i)
            if (4 * i < p) iv[i] = fma(-tr[i][p], ip, iv[i]);
        }
      } else if (early && wave == 5 && b0 + K < N && cs.decided && cs.local) {
        // the next block's strip of this spin, if this workgroup holds it: G0[next rows, B]
        const int rsn = (spmask >> 1) * nct + (b0 >> 4) + 1;
        if ((rsn % tw) / nw == w) {
          for (int e = lane; e < 256; e += 64)
            PinN[e] = (double)G[(spmask >> 1) * n2 + (size_t)(b0 + K + (e >> 4)) * N + b0 + (e & 15)];
          if (lane == 0) ecnt[0] = 0u;
        }
      } else if (nw > 4 && wave == 4 && b0 + K < N) {
        // wave 4, idle here: the next block's field entries (not changed by this block) and
        // uniforms, read by wave 0 after this block's barriers
        const int i = b0 + K + lane;
        if (lane < min(K, N - b0 - K)) {
          sb[lane] = fl[i];
          bool und = false;
          ub[lane] = draw_uniform(ch, &a.st[c], c, a.rng, sweep, a.t, i, und);
          if (und) underrun = true;
        }
      }
      prepped = nw > 4 && b0 + K < N;
      __syncthreads();
      LQMC_PROBE(13);
      if (wave < 4) {
        const bool live = j < m;
        double* out = (lower ? Ysb : Xsb) + isp * K * K;
        const double* cs = cf + isp * K;
        int ar[NI];
        double cr[NI];
#pragma unroll
        for (int i = 0; i < NI; ++i) {
          ar[i] = aidx[g + 4 * i];  // past the accepts: read, not used
          cr[i] = cs[g + 4 * i];
        }
        // rows qq = g + 4 i of column j zeroed first, then the accepted rows stored over them
        // by (other lanes of) the same wavefront after a wavefront barrier
#pragma unroll
        for (int i = 0; i < NI; ++i)
          if (live && g + 4 * i < kb) out[(g + 4 * i) * K + j] = 0.0;
        wave_sync();
#pragma unroll
        for (int i = 0; i < NI; ++i)
          if (live && g + 4 * i <= j) out[ar[i] * K + j] = lower ? iv[i] : cr[i] * iv[i];
      }
    } else {
      // lane j forms column j by back substitution in registers (the loop over rows stays
      // rolled; inv is written through a select chain)
      double inv[K];
      const int j = lane;
      if (wave < 4) {
        const double* Tw = Xsb + isp * 2 * K * K + (lower ? K * K : 0);
#pragma unroll
        for (int p = 0; p < K; ++p) inv[p] = p == j ? 1.0 : 0.0;
        // row r of the table is read while row r+1 is being used (one LDS wait per row)
        double trow[K];
#pragma unroll
        for (int p = 0; p < K; ++p) trow[p] = Tw[(K - 2) * K + p];
#pragma unroll 1
        for (int r = K - 2; r >= 0; --r) {
          double tnext[K];
          const int rn = r > 0 ? r - 1 : 0;
#pragma unroll
          for (int p = 0; p < K; ++p) tnext[p] = Tw[rn * K + p];
          // four interleaved partial sums shorten the dependent FMA chain
          double v4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int p = 0; p < K; ++p) v4[p & 3] -= trow[p] * inv[p];
          const double v = (v4[0] + v4[1]) + (v4[2] + v4[3]);
#pragma unroll
          for (int p = 0; p < K; ++p)
            if (p == r && r < j) inv[p] = v;
#pragma unroll
          for (int p = 0; p < K; ++p) trow[p] = tnext[p];
        }
      }
      __syncthreads();
      if (wave < 4 && j < m) {
        double* out = (lower ? Ysb : Xsb) + isp * K * K;
        const double* cs = cf + isp * K;
        for (int qq = 0; qq < kb; ++qq) out[qq * K + j] = 0.0;
#pragma unroll
        for (int r = 0; r < K; ++r)
          if (r <= j) out[aidx[r] * K + j] = lower ? inv[r] : cs[r] * inv[r];
      }
    }
    __syncthreads();
    LQMC_PROBE(5);
    if (lrows && full) form_z<T, K, true, true>(G, n2, N, Np, b0, kb, m, Ysb, wsT, spmask, Yl, Zl);
    else if (lrows) form_z<T, K, true>(G, n2, N, Np, b0, kb, m, Ysb, wsT, spmask, Yl);
    else form_z<T, K>(G, n2, N, Np, b0, kb, m, Ysb, wsT, spmask);
    __syncthreads();
    LQMC_PROBE(6);
    {
      // W > 1: the strips of this block's rows wait for the next barrier (see above)
      const bool defer = K >= 16 && W > 1;
      const int s0 = defer ? b0 >> 4 : nct, s1 = defer ? (b0 + kb - 1) >> 4 : -1;
      int rs_skip = -1;  // the next block's strip, when updated first (below)
      if (K == 16 && early && b0 + K < N && cs.decided && cs.local) {
        const int rsn = (spmask >> 1) * nct + (b0 >> 4) + 1, rsb = rsn - 1;
        if ((rsn % tw) / nw != w) {
          if (threadIdx.x == 0) atomicAdd(cs.bar, 1u);  // no row of the next block here
        } else {
          // its column-tile groups split between the wave of the next block's strip and, when
          // this workgroup holds it, the wave of this block's (deferred) strip; P from PinN
          const int vn = rsn % nw, vb = (rsb % tw) / nw == w ? rsb % nw : vn;
          const int ngr = nct / 4, half = vb == vn ? ngr : ngr / 2;
          if (wave == vn || wave == vb) {
            const int g0 = wave == vn ? 0 : half, g1 = wave == vn ? half : ngr;
            for (int g = g0; g < g1; ++g)
              update_strip<T, K, 4, BF3, K == 16, true>(G, n2, N, Np, b0, kb, m, Xsb, wsT, rsn, g, ngr, PinN, Zl);
            __builtin_amdgcn_s_waitcnt(0);  // this wave's stores of the strip are in the shared L2
            wave_sync();
            if (lane == 0 && atomicAdd(ecnt, 1u) == (vb == vn ? 0u : 1u)) atomicAdd(cs.bar, 1u);
          }
          rs_skip = rsn;
        }
        arrived = true;
      }
      for (int rs = gw; rs < 2 * nct; rs += tw) {
        const int st = rs % nct;
        if ((st < s0 || st > s1) && rs != rs_skip) {
          if (full && lrows) update_strip<T, K, (K >= 16 ? 4 : 2), BF3, K == 16, true>(G, n2, N, Np, b0, kb, m, Xsb, wsT, rs, 0, 1, nullptr, Zl);
          else if (full) update_strip<T, K, (K >= 16 ? 4 : 2), BF3, K == 16>(G, n2, N, Np, b0, kb, m, Xsb, wsT, rs);
          else update_strip<T, K, (K >= 16 ? 4 : 2), BF3>(G, n2, N, Np, b0, kb, m, Xsb, wsT, rs);
        }
      }
    }
    if (K >= 16 && W > 1) {
      pending = true;
      pb0 = b0;
      pkb = kb;
      pm = m;
      for (int e = threadIdx.x; e < 2 * K * K; e += blockDim.x) Mxp[e] = Xsb[e];
      if constexpr (K == 16) {
        // rows B x columns B of both spins (untouched until the deferred update); with the
        // rows staged, the workgroup's spin is copied from LDS (the only spin its deferred
        // strips use) and no global read sits before the next barrier
        if (lrows) {
          const int sp0 = spmask >> 1;
          for (int e = threadIdx.x; e < 512; e += blockDim.x) {
            const int sp = e >> 8, i = (e >> 4) & 15, q = e & 15;
            const bool ok = sp == sp0 && b0 + i < N && q < kb;
            Pin[e] = ok ? (double)Yl[i * N + b0 + q] : 0.0;
          }
        } else {
          for (int e = threadIdx.x; e < 512; e += blockDim.x) {
            const int sp = e >> 8, i = (e >> 4) & 15, q = e & 15;
            const bool ok = b0 + i < N && q < kb;
            Pin[e] = ok ? (double)G[sp * n2 + (size_t)(b0 + i) * N + b0 + q] : 0.0;
          }
        }
      }
    }
    __syncthreads();
    LQMC_PROBE(7);
  }
  if constexpr (K >= 16) {
    if (pending) {
      chain_barrier(cs);
      deferred_strips(false);
      __syncthreads();
    }
  }
  // On slice 0 every workgroup has read the sweep counter and (when measuring) every strip is
  // final before workgroup 0 measures and advances the counter.
  if (W > 1 && l == 0) chain_barrier(cs);
  if (w == 0) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) fld[i] = fl[i];
    slice_epilogue<T>(a, ch, c, l, G, nacc, underrun, bad, red);
  }
#ifdef LQMC_FLIP_PROBE
  LQMC_PROBE(8);
  if (a.trace && threadIdx.x == 0 && 16 * (w + 1) <= N) {
    for (int p = 0; p < 14; ++p) ch.tr_ratio[(size_t)a.t * N + w * 16 + p] = (double)ph[p];
    ch.tr_ratio[(size_t)a.t * N + w * 16 + 14] = (double)nblk;
    ch.tr_ratio[(size_t)a.t * N + w * 16 + 15] = (double)(wall_clock64() - tstart);
  }
#endif
  if (W > 1) chain_exit(cs);
}

// ---- determinant mode ------------------------------------------------------------------------
// The reference's brute-force sweep (lqmc.py:236-299), one workgroup per chain, for lattices
// of at most DET_NMAX sites (a cross-check of the fast sweeps on tiny lattices):
//   M_s(l0) = I + prod over l in (l0-1, ..., 0, L-1, ..., l0) of E_K V_s(l)   (lqmc.py:156-185),
//   for l = L-1 .. 0, i = 0 .. N-1: flip s_il, new = det M_up(l) det M_dn(l),
//   accept iff u <= new / old (then old = new), else flip back;
//   measurement: G_s = M_s(0)^-1 after the sweep.
// The products and both LU factorisations (partial pivoting, first maximal |pivot| like
// LAPACK's getrf) stay in LDS; wavefront s factors spin s.  The starting determinant is
// recomputed from the field at the start of every sweep (the reference carries it from the
// previous accepted flip; both are det M(0) up to rounding).
constexpr int DET_NMAX = 32;

struct DetArgs {
  const ChainDev* chains;
  ChainState* st;
  int measure, rng, trace, pad;
};

// P_s <- prod over l in order(l0) of E_K V_s(l), both spins, all threads.  Buffer b (N x N) of
// the ping-pong set starts at pp0 + b N^2; spin s uses buffers s and 2 + s.  Returns the
// index (0 or 1) of the half holding the result.
__device__ int det_product(const ChainDev& ch, const int8_t* fld, const double* ek, double* pp0, int l0) {
  const int N = ch.N, L = ch.L;
  const size_t n2 = (size_t)N * N;
  for (size_t e = threadIdx.x; e < 2 * n2; e += blockDim.x) {
    const int sp = (int)(e / n2);
    const size_t r = (e % n2) / N, cc = (e % n2) % N;
    pp0[sp * n2 + r * N + cc] = r == cc ? 1.0 : 0.0;
  }
  __syncthreads();
  int cur = 0;
  const double e1m = ch.e1m, e1p = ch.e1p;
  for (int q = 0; q < L; ++q) {
    const int l = ((l0 - 1 - q) % L + L) % L;  // l0-1, ..., 0, L-1, ..., l0
    for (size_t e = threadIdx.x; e < 2 * n2; e += blockDim.x) {
      const int sp = (int)(e / n2);
      const int r = (int)((e % n2) / N), cc = (int)((e % n2) % N);
      const double* P = pp0 + (size_t)(cur * 2 + sp) * n2;
      double acc = 0.0;
      for (int k = 0; k < N; ++k) acc += P[r * N + k] * ek[k * N + cc];
      const int sv = fld[l * N + cc];
      // V_s(l) = exp(-s lambda s_l,c): sigma * s = +1 -> e^-lambda
      const double v = (sp == 0) == (sv > 0) ? e1m : e1p;
      pp0[(size_t)((1 - cur) * 2 + sp) * n2 + r * N + cc] = acc * v;
    }
    __syncthreads();
    cur = 1 - cur;
  }
  return cur;
}

// In-place LU with partial pivoting of the N x N matrix A (wavefront-wide, lanes over rows);
// returns det(A), records the row permutation in piv.
__device__ double det_lu_wave(double* A, int N, int* piv) {
  const int lane = threadIdx.x & 63;
  double det = 1.0;
  for (int k = 0; k < N; ++k) {
    // |A[r][k]| over rows r >= k, NaN counted as +inf so that the order is total and every
    // lane ends the butterfly on the same row; lanes outside [k, N) offer -1
    double best = -1.0;
    if (lane >= k && lane < N) {
      const double v = fabs(A[lane * N + k]);
      best = v == v ? v : INFINITY;
    }
    int bi = lane;
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    const int p = min(max(__builtin_amdgcn_readlane(bi, 0), k), N - 1);
    if (lane == 0) piv[k] = p;
    if (p != k) {
      for (int cc = lane; cc < N; cc += 64) {
        const double t = A[k * N + cc];
        A[k * N + cc] = A[p * N + cc];
        A[p * N + cc] = t;
      }
      det = -det;
    }
    wave_sync();
    const double d = A[k * N + k];
    det *= d;
    if (lane > k && lane < N && d != 0.0) {
      const double lr = A[lane * N + k] / d;
      A[lane * N + k] = lr;
      for (int cc = k + 1; cc < N; ++cc) A[lane * N + cc] -= lr * A[k * N + cc];
    }
    wave_sync();
  }
  return det;
}

__global__ __launch_bounds__(256) void k_det_sweep(DetArgs a) {
  const int c = blockIdx.x;
  const ChainDev ch = a.chains[c];
  const int N = ch.N, L = ch.L;
  const size_t n2 = (size_t)N * N;
  LQMC_DYNAMIC_LDS(double, sm);
  double* ek = sm;                       // [N][N]
  double* pp0 = ek + n2;                 // [4][N][N] ping-pong products / LU / inverse
  double* red = ek + 5 * n2;             // [blockDim]
  double* dets = red + blockDim.x;       // [2]
  int* piv = reinterpret_cast<int*>(dets + 2);  // [2][DET_NMAX]
  int8_t* fld = reinterpret_cast<int8_t*>(piv + 2 * DET_NMAX);  // [L][N]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (size_t e = threadIdx.x; e < n2; e += blockDim.x) ek[e] = ch.expk[e];
  for (int e = threadIdx.x; e < N * L; e += blockDim.x) fld[e] = ch.field[e];
  __syncthreads();

  // determinants of I + P for both spins (wave s, spin s) -> dets[s]
  auto both_dets = [&](int l0) {
    const int cur = det_product(ch, fld, ek, pp0, l0);
    if (wave < 2) {
      double* A = pp0 + (size_t)(cur * 2 + wave) * n2;
      for (int r = lane; r < N; r += 64) A[r * N + r] += 1.0;
      wave_sync();
      const double d = det_lu_wave(A, N, piv + wave * DET_NMAX);
      if (lane == 0) dets[wave] = d;
    }
    __syncthreads();
    return cur;
  };

  const unsigned long long sweep = a.st[c].sweep;
  bool underrun = false, bad = false;
  unsigned long long nacc = 0;
  // det M(l) does not depend on l (det(I + AB) = det(I + BA) for the cyclic shifts), so the
  // reference's running determinant equals det M(0) of the current field at every sweep start
  both_dets(0);
  double old_det = dets[0] * dets[1];
  for (int t = 0; t < L; ++t) {
    const int l = L - 1 - t;
    for (int i = 0; i < N; ++i) {
      if (threadIdx.x == 0) fld[l * N + i] = (int8_t)(-fld[l * N + i]);
      __syncthreads();
      both_dets(l);
      const double new_det = dets[0] * dets[1];
      const double ratio = new_det / old_det;
      const double u = draw_uniform(ch, &a.st[c], c, a.rng, sweep, t, i, underrun);
      const bool acc = u <= ratio;
      if (!isfinite(ratio)) bad = true;
      if (a.trace && threadIdx.x == 0) {
        ch.tr_ratio[(size_t)t * N + i] = ratio;
        ch.tr_acc[(size_t)t * N + i] = acc ? 1 : 0;
      }
      if (acc) {
        old_det = new_det;
        ++nacc;
      } else if (threadIdx.x == 0) {
        fld[l * N + i] = (int8_t)(-fld[l * N + i]);
      }
      __syncthreads();
    }
  }
  for (int e = threadIdx.x; e < N * L; e += blockDim.x) ch.field[e] = fld[e];
  if (threadIdx.x == 0) {
    atomicAdd(&a.st[c].attempted, (unsigned long long)N * L);
    atomicAdd(&a.st[c].accepted, nacc);
    if (underrun) atomicOr(&a.st[c].flags, 4u);
    if (bad) atomicOr(&a.st[c].flags, 1u);
  }
  if (a.measure) {
    // G_s = (I + P_s(0))^-1 from the LU: lane j of wave s solves column j
    const int cur = both_dets(0);
    double* G = static_cast<double*>(ch.G);
    if (wave < 2) {
      const double* A = pp0 + (size_t)(cur * 2 + wave) * n2;
      const int* pv = piv + wave * DET_NMAX;
      double* X = pp0 + (size_t)((1 - cur) * 2 + wave) * n2;  // scratch [N][N], X[r][j]
      const int j = lane;
      if (j < N) {
        for (int r = 0; r < N; ++r) X[r * N + j] = r == j ? 1.0 : 0.0;
        for (int k = 0; k < N; ++k)  // the row swaps of the factorisation
          if (pv[k] != k) {
            const double t0 = X[k * N + j];
            X[k * N + j] = X[pv[k] * N + j];
            X[pv[k] * N + j] = t0;
          }
        for (int r = 1; r < N; ++r) {  // unit lower
          double v = X[r * N + j];
          for (int k = 0; k < r; ++k) v -= A[r * N + k] * X[k * N + j];
          X[r * N + j] = v;
        }
        for (int r = N - 1; r >= 0; --r) {  // upper
          double v = X[r * N + j];
          for (int k = r + 1; k < N; ++k) v -= A[r * N + k] * X[k * N + j];
          X[r * N + j] = v / A[r * N + r];
        }
        for (int r = 0; r < N; ++r) G[(size_t)wave * n2 + (size_t)r * N + j] = X[r * N + j];
      }
    }
    __syncthreads();
    measure_equal_time<double>(ch, &a.st[c], G, red);
  }
  __syncthreads();
  if (threadIdx.x == 0) a.st[c].sweep += 1ull;
}

// ---- rank-1 form on several workgroups per chain -----------------------------------------------
// The plain sweep with the compute-unit budget of the delayed form: W workgroups per chain,
// workgroup w owning rows [w N / W, (w + 1) N / W) of both spins' G.  Per site every workgroup
// reads G_ii of both spins (current: a chain barrier ends every accepted flip), and forms the same
// ratio and decision from the same uniform.  On acceptance each workgroup takes row i of both
// spins into LDS and the column entries G[r, i] of its own rows, waits until every workgroup has
// done its reads of this state (the diagonals of the rejected sites before i included: a
// workgroup ahead must not overwrite them), and applies the update of k_flip_rank1,
// G -= c x (e_i - G[i, :])^T (the same operations in the same order, so W = 1 and W > 1 agree
// bit for bit), to its rows except row i, which every workgroup reads during the step: its
// owner keeps the updated row in LDS and writes it after the barrier that ends the step.
template <typename T>
__global__ __launch_bounds__(1024) void k_flip_rank1_split(FlipArgs a) {
  int c, w;
  flip_block_map(a, c, w);
  const int W = a.wpc;
  if (c >= a.nchains) return;
  const ChainDev ch = a.chains[c];
  if (a.t >= ch.L) return;
  const int N = ch.N;
  const int l = ch.L - 1 - a.t;
  const size_t n2 = (size_t)N * N;
  T* __restrict__ G = static_cast<T*>(ch.G);
  int8_t* fld = ch.field + (size_t)l * N;
  const int r0 = (int)((long)N * w / W), r1 = (int)((long)N * (w + 1) / W), nr = r1 - r0;
  LQMC_DYNAMIC_LDS(double, sm);
  double* grow = sm;               // [2][N] row i of both spins before the update
  double* prow = grow + 2 * N;     // [2][N] the owner's updated row of the pending site
  double* xv = prow + 2 * N;       // [2][nr] G[r, i] of my rows
  double* red = xv + 2 * nr;       // blockDim (epilogue)
  double* coef = red + blockDim.x;  // [2] update coefficients of both spins
  int* bflag = reinterpret_cast<int*>(coef + 2);  // [2] the barrier's flag
  int8_t* fl = reinterpret_cast<int8_t*>(bflag + 2);  // [N] the slice's field
  ChainSync cs = chain_sync_begin(a.bar + 4 * c, &a.st[c], W, bflag);
  bool underrun = false, bad = false;
  const unsigned long long sweep = a.st[c].sweep;
  unsigned long long nacc = 0;
  int pend = -1;  // site whose updated row waits in prow (uniform)
  for (int i = threadIdx.x; i < N; i += blockDim.x) fl[i] = fld[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int i = 0; i < N; ++i) {
    const int s = fl[i];
    const double gu = (double)G[(size_t)i * N + i];
    const double gd = (double)G[n2 + (size_t)i * N + i];
    const double e2p = ch.e2p, e2m = ch.e2m;
    const double ep = s > 0 ? e2p : e2m;
    const double em = s > 0 ? e2m : e2p;
    const double d_up = 1.0 + (1.0 - gu) * (ep - 1.0);
    const double d_dn = 1.0 + (1.0 - gd) * (em - 1.0);
    const double ratio = d_up * d_dn;
    const double u = draw_uniform(ch, &a.st[c], c, a.rng, sweep, a.t, i, underrun);
    const bool acc = u <= ratio;
    if (!isfinite(ratio)) bad = true;
    if (a.trace && threadIdx.x == 0 && w == 0) {
      ch.tr_ratio[(size_t)a.t * N + i] = ratio;
      ch.tr_acc[(size_t)a.t * N + i] = acc ? 1 : 0;
    }
    if (!acc) continue;
    // the pending row of the previous accepted site: nobody reads it any more
    if (pend >= r0 && pend < r1)
      for (int j = threadIdx.x; j < 2 * N; j += blockDim.x) {
        const int sp = j / N, jj = j % N;
        G[sp * n2 + (size_t)pend * N + jj] = (T)prow[j];
      }
    __syncthreads();
    for (int j = threadIdx.x; j < 2 * N; j += blockDim.x) {
      const int sp = j / N, jj = j % N;
      grow[j] = (double)G[sp * n2 + (size_t)i * N + jj];
    }
    for (int j = threadIdx.x; j < 2 * nr; j += blockDim.x) {
      const int sp = j / nr, r = r0 + j % nr;
      xv[j] = (double)G[sp * n2 + (size_t)r * N + i];
    }
    if (threadIdx.x < 2) {
      const double al = threadIdx.x == 0 ? ep - 1.0 : em - 1.0;
      const double de = threadIdx.x == 0 ? d_up : d_dn;
      coef[threadIdx.x] = al / de;
    }
    chain_sync_reads(cs);
    // G[r, :] -= c (x_r (e_i - G[i, :])): each wavefront takes RB rows at a time, lanes over
    // 256-column chunks, with every load of a batch issued before its arithmetic (the rows
    // stream from memory; one load at a time would leave the update latency-bound)
    {
      constexpr int RB = 4;
      for (int row0 = (threadIdx.x >> 6) * RB; row0 < 2 * nr; row0 += nw * RB) {
        for (int k0 = 0; k0 < N; k0 += 256) {
          double gv[RB][4];
#pragma unroll
          for (int b = 0; b < RB; ++b) {
            const int row = row0 + b;
            const bool rok = row < 2 * nr;
            const int sp = rok ? row / nr : 0, r = r0 + (rok ? row % nr : 0);
            const LQMC_GLOBAL T* Gr = gptr(G + sp * n2 + (size_t)r * N);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int k = k0 + 64 * q + lane;
              gv[b][q] = (double)ldg0(Gr, (size_t)k, rok && r != i && k < N);
            }
          }
#pragma unroll
          for (int b = 0; b < RB; ++b) {
            const int row = row0 + b;
            if (row >= 2 * nr) continue;
            const int sp = row / nr, r = r0 + row % nr;
            const double xj = xv[row];
            const double cj = coef[sp];
            const double* gi = grow + sp * N;
            T* Gr = G + sp * n2 + (size_t)r * N;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int k = k0 + 64 * q + lane;
              if (k >= N) continue;
              double y = -gi[k];
              if (k == i) y = y + 1.0;
              if (r == i) prow[sp * N + k] = gi[k] - cj * (xj * y);
              else Gr[k] = (T)(gv[b][q] - cj * (xj * y));
            }
          }
        }
      }
    }
    pend = i;
    if (threadIdx.x == 0) fl[i] = (int8_t)(-s);
    ++nacc;
    chain_barrier(cs);
  }
  if (pend >= r0 && pend < r1) {
    for (int j = threadIdx.x; j < 2 * N; j += blockDim.x) {
      const int sp = j / N, jj = j % N;
      G[sp * n2 + (size_t)pend * N + jj] = (T)prow[j];
    }
  }
  // every row final (and the sweep counter read by every workgroup) before workgroup 0 measures
  if (l == 0) chain_barrier(cs);
  if (w == 0) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) fld[i] = fl[i];
    slice_epilogue<T>(a, ch, c, l, G, nacc, underrun, bad, red);
  }
  chain_exit(cs);
}

size_t flip_rank1_split_shmem(int nmax, int threads, int groups) {
  const size_t nr = ((size_t)nmax + groups - 1) / groups;
  return sizeof(double) * (4 * (size_t)nmax + 2 * nr + threads + 3) + (size_t)nmax;
}

size_t det_sweep_shmem(int nmax, int lmax) {
  return sizeof(double) * (5 * (size_t)nmax * nmax + 256 + 2) + sizeof(int) * 2 * DET_NMAX + (size_t)nmax * lmax;
}

void launch_det_sweep(const ChainDev* chains, ChainState* st, int nchains, int measure, int rng, int trace,
                      size_t shmem, hipStream_t s) {
  DetArgs a{chains, st, measure, rng, trace, 0};
  hipLaunchKernelGGL(k_det_sweep, dim3(nchains), dim3(256), shmem, s, a);
}

int flip_delayed_threads_limit(int k) { return flip_delayed_max_threads(k); }

size_t flip_delayed_shmem(int nmax, int k, int threads, int fp32) {
  (void)fp32;
  return sizeof(double) * (threads + 6 * (size_t)k * k + 3 * (size_t)k + (k == 16 ? 512 : 0)) +
         sizeof(int) * (2 * k + 4) + (size_t)nmax;
}

// Whether the delayed flip can stage each block's rows and Z tiles in LDS (FlipArgs::yrows):
// 16-site blocks on several workgroups per chain, N <= 256.  The extra LDS is 16 x nmax
// doubles of rows, 16 x Np doubles of Z tiles and a 16 x 16 copy of the next block's G0[rows, B]
// with its counter (at most 66 KB), added to flip_delayed_shmem.
int flip_delayed_stage_rows(int nmax, int k, int groups) { return k == 16 && groups > 1 && nmax <= 256 ? 1 : 0; }
size_t flip_delayed_rows_shmem(int nmax, int k, int groups) {
  return flip_delayed_stage_rows(nmax, k, groups)
             ? sizeof(double) * (16 * ((size_t)nmax + ((nmax + 15) & ~15)) + 256) + 4 * sizeof(int) : 0;
}

// Bytes of the per-chain delayed-update workspace (the Z tiles of both spins, one set per
// workgroup of the chain).
size_t flip_delayed_workspace(int n, int k, int fp32, int groups) {
  const size_t Np = (n + 15) & ~15;
  const size_t K16 = (k + 15) & ~15;
  return 2 * K16 * Np * (fp32 ? 4 : 8) * (size_t)(groups > 1 ? groups : 1);
}

// Grid of the delayed kernel: W workgroups per chain, chains padded to a multiple of 8.
int flip_grid(int nchains, int groups) {
  return groups > 1 ? 8 * groups * ((nchains + 7) / 8) : nchains;
}

void launch_flip(int fp32, int delayed, const FlipArgs& a, int nchains, int threads, size_t shmem,
                 hipStream_t s) {
  if (delayed) {
    const int grid = flip_grid(nchains, a.wpc);
#define LQMC_DELAYED(KK)                                                                                    \
  case KK:                                                                                                  \
    if (fp32 && a.bf3) hipLaunchKernelGGL((k_flip_delayed<float, KK, true>), dim3(grid), dim3(threads), shmem, s, a); \
    else if (fp32) hipLaunchKernelGGL((k_flip_delayed<float, KK>), dim3(grid), dim3(threads), shmem, s, a); \
    else hipLaunchKernelGGL((k_flip_delayed<double, KK>), dim3(grid), dim3(threads), shmem, s, a);         \
    break;
    switch (a.delay) {
      LQMC_DELAYED(4)
      LQMC_DELAYED(8)
      LQMC_DELAYED(16)
      LQMC_DELAYED(32)
      default: break;
    }
#undef LQMC_DELAYED
  } else if (a.wpc > 1) {
    const int grid = flip_grid(nchains, a.wpc);
    if (fp32) hipLaunchKernelGGL(k_flip_rank1_split<float>, dim3(grid), dim3(threads), shmem, s, a);
    else hipLaunchKernelGGL(k_flip_rank1_split<double>, dim3(grid), dim3(threads), shmem, s, a);
  } else {
    if (fp32) hipLaunchKernelGGL(k_flip_rank1<float>, dim3(nchains), dim3(threads), shmem, s, a);
    else hipLaunchKernelGGL(k_flip_rank1<double>, dim3(nchains), dim3(threads), shmem, s, a);
  }
}

}  // namespace lqmc

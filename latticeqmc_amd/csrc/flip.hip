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

template <typename T>
__device__ void slice_epilogue(const FlipArgs& a, const ChainDev& ch, int c, int l, const T* G,
                               unsigned long long nacc, bool underrun, bool bad, double* red) {
  const int N = ch.N;
  const int NL = N * ch.L;
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int s = ch.field[(size_t)l * N + i];
    const double up = s > 0 ? ch.e1m : ch.e1p;
    const double dn = s > 0 ? ch.e1p : ch.e1m;
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
  if (l == 0 && a.measure) {
    const size_t n2 = (size_t)N * N;
    for (size_t e = threadIdx.x; e < 2 * n2; e += blockDim.x) {
      const double g = (double)G[e];
      ch.Gacc[e] += g;
      ch.Glast[e] = g;
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
      const unsigned long long k = a.st[c].measured;
      double* o = ch.obs + (size_t)(k % (unsigned long long)ch.obs_cap) * 3;
      o[0] = dens / N;
      o[1] = docc / N;
      o[2] = ek / N;
      a.st[c].measured = k + 1;
    }
  }
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
  extern __shared__ double sm[];
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
    const double ep = s > 0 ? ch.e2p : ch.e2m;  // exp(+2 lambda s)
    const double em = s > 0 ? ch.e2m : ch.e2p;  // exp(-2 lambda s)
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
    for (size_t e = threadIdx.x; e < 2 * n2; e += blockDim.x) {
      const int sp = (int)(e / n2);
      const size_t r = e % n2;
      const int j = (int)(r / N), k = (int)(r % N);
      double p;
      if (ref) p = xv[sp * N + j] * yv[sp * N + k];
      else p = coef[sp] * (xv[sp * N + j] * yv[sp * N + k]);
      G[e] = (T)((double)G[e] - p);
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

// G (N x N, ld N) += P (Np x K4, ld pld) * Z (K4 x Np, ld Np); Np = N rounded up to 16,
// K4 = rank rounded up to 4 with zero padding in P and Z.
template <typename T>
__device__ void rank_update(T* G, int N, int Np, const T* P, int pld, const T* Z, int K4) {
  using acc_t = typename MfmaOp<T>::acc_t;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int tn = Np >> 4;
  const int ntile = tn * tn;
  for (int tile = wave; tile < ntile; tile += nw) {
    const int r0 = (tile / tn) << 4, c0 = (tile % tn) << 4;
    const int col = c0 + (lane & 15);
    acc_t acc;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = r0 + MfmaOp<T>::row(lane, q);
      acc[q] = (row < N && col < N) ? G[(size_t)row * N + col] : T(0);
    }
    for (int kk = 0; kk < K4; kk += 4) {
      const T x = P[(size_t)(r0 + (lane & 15)) * pld + kk + (lane >> 4)];
      const T y = Z[(size_t)(kk + (lane >> 4)) * Np + col];
      acc = MfmaOp<T>::mma(x, y, acc);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = r0 + MfmaOp<T>::row(lane, q);
      if (row < N && col < N) G[(size_t)row * N + col] = acc[q];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(1024) void k_flip_delayed(FlipArgs a) {
  const int c = blockIdx.x;
  const ChainDev ch = a.chains[c];
  if (a.t >= ch.L) return;
  const int N = ch.N;
  const int l = ch.L - 1 - a.t;
  const int K = a.delay;
  const int Np = (N + 15) & ~15;
  const int pld = ((K + 3) & ~3) + 1;  // P row stride >= padded rank, +1 against bank conflicts
  const size_t n2 = (size_t)N * N;
  T* __restrict__ G = static_cast<T*>(ch.G);
  int8_t* fld = ch.field + (size_t)l * N;

  extern __shared__ double sm[];
  double* red = sm;                        // blockDim
  double* alpha = red + blockDim.x;        // [2][K]
  double* gam = alpha + 2 * K;             // [K][2K]
  int* aidx = reinterpret_cast<int*>(gam + 2 * K * K);  // [K]
  double* uni = reinterpret_cast<double*>(aidx + ((K + 1) & ~1));  // union region
  // serial phase view
  double* small = uni;                     // [2][K][K]
  double* xs = small + 2 * K * K;          // [2][K]
  double* ys = xs + 2 * K;                 // [2][K]
  // update phase view
  T* P = reinterpret_cast<T*>(uni);        // [Np][pld]
  T* Z = P + (size_t)Np * pld;             // [K4][Np]
  __shared__ double coef[2];
  __shared__ int piv_s;

  const unsigned long long sweep = a.st[c].sweep;
  unsigned long long nacc = 0;
  bool underrun = false, bad = false;

  for (int b0 = 0; b0 < N; b0 += K) {
    const int kb = min(K, N - b0);
    for (int e = threadIdx.x; e < 2 * kb * kb; e += blockDim.x) {
      const int sp = e / (kb * kb), r = e % (kb * kb);
      const int p = r / kb, q = r % kb;
      small[sp * K * K + p * K + q] = (double)G[sp * n2 + (size_t)(b0 + p) * N + b0 + q];
    }
    __syncthreads();
    int m = 0;
    for (int jj = 0; jj < kb; ++jj) {
      const int i = b0 + jj;
      const int s = fld[i];
      const double gu = small[jj * K + jj];
      const double gd = small[K * K + jj * K + jj];
      const double ep = s > 0 ? ch.e2p : ch.e2m;
      const double em = s > 0 ? ch.e2m : ch.e2p;
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
      for (int e = threadIdx.x; e < 2 * kb; e += blockDim.x) {
        const int sp = e / kb, q = e % kb;
        xs[sp * K + q] = small[sp * K * K + q * K + jj];
        double y = -small[sp * K * K + jj * K + q];
        if (q == jj) y = y + 1.0;
        ys[sp * K + q] = y;
      }
      if (threadIdx.x < 2) {
        const double al = threadIdx.x == 0 ? ep - 1.0 : em - 1.0;
        const double de = threadIdx.x == 0 ? d_up : d_dn;
        coef[threadIdx.x] = al / de;
        alpha[threadIdx.x * K + m] = al;
      }
      if (threadIdx.x == 0) {
        aidx[m] = i;
        fld[i] = (int8_t)(-s);
      }
      __syncthreads();
      for (int e = threadIdx.x; e < 2 * kb * kb; e += blockDim.x) {
        const int sp = e / (kb * kb), r = e % (kb * kb);
        const int p = r / kb, q = r % kb;
        small[sp * K * K + p * K + q] -= coef[sp] * (xs[sp * K + p] * ys[sp * K + q]);
      }
      __syncthreads();
      ++m;
    }
    if (m == 0) continue;
    nacc += m;
    const int K4 = (m + 3) & ~3;
    for (int sp = 0; sp < 2; ++sp) {
      T* Gs = G + sp * n2;
      // Gamma = diag(1/alpha) + I - G[A,A], augmented with I for Gauss-Jordan
      for (int e = threadIdx.x; e < m * 2 * m; e += blockDim.x) {
        const int p = e / (2 * m), q = e % (2 * m);
        double v;
        if (q < m) {
          v = -(double)Gs[(size_t)aidx[p] * N + aidx[q]];
          if (p == q) v += 1.0 + 1.0 / alpha[sp * K + p];
        } else {
          v = (q - m == p) ? 1.0 : 0.0;
        }
        gam[p * 2 * K + q] = v;
      }
      __syncthreads();
      for (int j = 0; j < m; ++j) {
        if (threadIdx.x == 0) {
          int pv = j;
          double best = fabs(gam[j * 2 * K + j]);
          for (int r = j + 1; r < m; ++r) {
            const double v = fabs(gam[r * 2 * K + j]);
            if (v > best) { best = v; pv = r; }
          }
          piv_s = pv;
        }
        __syncthreads();
        const int pv = piv_s;
        if (pv != j)
          for (int q = threadIdx.x; q < 2 * m; q += blockDim.x) {
            const double t0 = gam[j * 2 * K + q];
            gam[j * 2 * K + q] = gam[pv * 2 * K + q];
            gam[pv * 2 * K + q] = t0;
          }
        __syncthreads();
        const double ip = 1.0 / gam[j * 2 * K + j];
        __syncthreads();
        for (int q = threadIdx.x; q < 2 * m; q += blockDim.x) gam[j * 2 * K + q] *= ip;
        __syncthreads();
        for (int e = threadIdx.x; e < m * 2 * m; e += blockDim.x) {
          const int r = e / (2 * m), q = e % (2 * m);
          if (r != j) {
            const double f = gam[r * 2 * K + j];
            if (q != j) gam[r * 2 * K + q] -= f * gam[j * 2 * K + q];
          }
        }
        __syncthreads();
        for (int r = threadIdx.x; r < m; r += blockDim.x)
          if (r != j) gam[r * 2 * K + j] = 0.0;
        __syncthreads();
      }
      // P = -G[:,A] Gamma^-1  (Np x K4), Z = E_A - G[A,:]  (K4 x Np)
      for (int e = threadIdx.x; e < Np * K4; e += blockDim.x) {
        const int r = e / K4, q = e % K4;
        double v = 0.0;
        if (r < N && q < m)
          for (int p = 0; p < m; ++p) v += (double)Gs[(size_t)r * N + aidx[p]] * gam[p * 2 * K + m + q];
        P[(size_t)r * pld + q] = (T)(-v);
      }
      for (int e = threadIdx.x; e < K4 * Np; e += blockDim.x) {
        const int p = e / Np, col = e % Np;
        double v = 0.0;
        if (p < m && col < N) {
          v = -(double)Gs[(size_t)aidx[p] * N + col];
          if (col == aidx[p]) v += 1.0;
        }
        Z[(size_t)p * Np + col] = (T)v;
      }
      __syncthreads();
      rank_update<T>(Gs, N, Np, P, pld, Z, K4);
      __syncthreads();
    }
  }
  slice_epilogue<T>(a, ch, c, l, G, nacc, underrun, bad, red);
}

size_t flip_delayed_shmem(int nmax, int k, int threads, int fp32) {
  const int Np = (nmax + 15) & ~15;
  const int K4 = (k + 3) & ~3;
  const size_t esz = fp32 ? 4 : 8;
  size_t head = sizeof(double) * (threads + 2 * k + 2 * (size_t)k * k) + sizeof(int) * ((k + 1) & ~1);
  size_t serial = sizeof(double) * (2 * (size_t)k * k + 4 * k);
  size_t update = esz * ((size_t)Np * (K4 + 1) + (size_t)K4 * Np);
  return head + (serial > update ? serial : update);
}

// Allow dynamic LDS beyond the default per-kernel limit (gfx950 has 160 KB per CU).
int flip_prepare(size_t shmem) {
  hipError_t e = hipSuccess;
  const int b = (int)shmem;
  hipError_t r;
  r = hipFuncSetAttribute((const void*)k_flip_delayed<float>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  if (r != hipSuccess) e = r;
  r = hipFuncSetAttribute((const void*)k_flip_delayed<double>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  if (r != hipSuccess) e = r;
  r = hipFuncSetAttribute((const void*)k_flip_rank1<float>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  if (r != hipSuccess) e = r;
  r = hipFuncSetAttribute((const void*)k_flip_rank1<double>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  if (r != hipSuccess) e = r;
  return (int)e;
}

void launch_flip(int fp32, int delayed, const FlipArgs& a, int nchains, int threads, size_t shmem,
                 hipStream_t s) {
  if (delayed) {
    if (fp32) hipLaunchKernelGGL(k_flip_delayed<float>, dim3(nchains), dim3(threads), shmem, s, a);
    else hipLaunchKernelGGL(k_flip_delayed<double>, dim3(nchains), dim3(threads), shmem, s, a);
  } else {
    if (fp32) hipLaunchKernelGGL(k_flip_rank1<float>, dim3(nchains), dim3(threads), shmem, s, a);
    else hipLaunchKernelGGL(k_flip_rank1<double>, dim3(nchains), dim3(threads), shmem, s, a);
  }
}

}  // namespace lqmc

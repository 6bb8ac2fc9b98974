// Bounds check of the determinant-mode sweep kernel in a CPU build (AddressSanitizer or
// ThreadSanitizer), with uneven chains (N - c sites, L - c slices) and a plain host restatement
// of the same sweep (lqmc.py:236-299: flip, det M(l) of both spins, accept u <= ratio, revert;
// then G = inv(M(0))) on the same Philox uniforms: fields identical, G within 1e-10.
// usage: det_harness N L nchains
#include "hip_cpu_shim.h"
#include "../../latticeqmc_amd/csrc/flip.hip"

#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

using namespace lqmc;

template <typename T>
static T* dalloc(size_t n) { return static_cast<T*>(std::calloc(n ? n : 1, sizeof(T))); }

// M(l0) = I + prod over l in (l0-1, ..., 0, L-1, ..., l0) of E_K diag(exp(-sigma lambda s_l))
static std::vector<double> host_m(const std::vector<int8_t>& f, const std::vector<double>& ek, int N, int L,
                                  double lam, int l0, int sigma) {
  std::vector<double> P(N * N, 0.0), T(N * N);
  for (int r = 0; r < N; ++r) P[r * N + r] = 1.0;
  for (int q = 0; q < L; ++q) {
    const int l = ((l0 - 1 - q) % L + L) % L;
    for (int r = 0; r < N; ++r)
      for (int c = 0; c < N; ++c) {
        double a = 0.0;
        for (int k = 0; k < N; ++k) a += P[r * N + k] * ek[k * N + c];
        T[r * N + c] = a * std::exp(-sigma * lam * f[l * N + c]);
      }
    P.swap(T);
  }
  for (int r = 0; r < N; ++r) P[r * N + r] += 1.0;
  return P;
}

// determinant by Gaussian elimination with partial pivoting; optionally the inverse
static double host_det(std::vector<double> A, int N, std::vector<double>* inv = nullptr) {
  std::vector<double> X(N * N, 0.0);
  for (int r = 0; r < N; ++r) X[r * N + r] = 1.0;
  double det = 1.0;
  for (int k = 0; k < N; ++k) {
    int p = k;
    for (int r = k + 1; r < N; ++r) if (std::fabs(A[r * N + k]) > std::fabs(A[p * N + k])) p = r;
    if (p != k) {
      for (int c = 0; c < N; ++c) { std::swap(A[k * N + c], A[p * N + c]); std::swap(X[k * N + c], X[p * N + c]); }
      det = -det;
    }
    const double d = A[k * N + k];
    det *= d;
    for (int c = 0; c < N; ++c) { A[k * N + c] /= d; X[k * N + c] /= d; }
    for (int r = 0; r < N; ++r) {
      if (r == k) continue;
      const double m = A[r * N + k];
      for (int c = 0; c < N; ++c) { A[r * N + c] -= m * A[k * N + c]; X[r * N + c] -= m * X[k * N + c]; }
    }
  }
  if (inv) *inv = X;
  return det;
}

int main(int argc, char** argv) {
  if (argc < 4) { std::fprintf(stderr, "usage: N L nchains\n"); return 2; }
  const int NM = atoi(argv[1]), LM = atoi(argv[2]), nch = atoi(argv[3]);
  if (NM > DET_NMAX || NM - nch + 1 < 1 || LM - nch + 1 < 1) { std::fprintf(stderr, "bad sizes\n"); return 2; }
  std::mt19937_64 rng(3);
  std::vector<ChainDev> chains(nch);
  std::vector<ChainState> st(nch);
  std::memset(st.data(), 0, sizeof(ChainState) * nch);
  std::vector<std::vector<int8_t>> f0(nch);
  std::vector<std::vector<double>> ekv(nch);
  const double lam = 0.55;
  for (int c = 0; c < nch; ++c) {
    const int N = NM - c, L = LM - c;
    const size_t n2 = (size_t)N * N;
    ChainDev& d = chains[c];
    std::memset(&d, 0, sizeof d);
    d.N = N; d.L = L; d.mode = 0;
    d.lam = lam; d.e1m = std::exp(-lam); d.e1p = std::exp(lam); d.e2p = std::exp(2 * lam); d.e2m = std::exp(-2 * lam);
    d.seed = 77 + c;
    d.field = dalloc<int8_t>((size_t)N * L);
    f0[c].resize((size_t)N * L);
    for (size_t e = 0; e < (size_t)N * L; ++e) f0[c][e] = d.field[e] = (rng() & 1) ? 1 : -1;
    // E_K of a ring with hopping 0.1 to first order, plus a small chemical-potential shift
    double* ek = dalloc<double>(n2);
    ekv[c].assign(n2, 0.0);
    for (int i = 0; i < N; ++i) {
      ek[i * N + i] = 1.02;
      if (N > 1) { ek[i * N + (i + 1) % N] += 0.1; ek[((i + 1) % N) * N + i] += 0.1; }
    }
    for (size_t e = 0; e < n2; ++e) ekv[c][e] = ek[e];
    d.expk = ek;
    d.G = dalloc<double>(2 * n2);
    d.Gacc = dalloc<double>(2 * n2);
    d.Glast = dalloc<double>(2 * n2);
    d.Sacc = dalloc<double>(n2);
    d.hop = dalloc<double>(n2);
    d.tr_ratio = dalloc<double>((size_t)N * L);
    d.tr_acc = dalloc<uint8_t>((size_t)N * L);
    d.obs_cap = 4;
    d.obs = dalloc<double>(3 * 4);
  }
  const size_t lds = det_sweep_shmem(NM, LM);
  std::printf("det N<=%d L<=%d chains=%d lds=%zu\n", NM, LM, nch, lds);
  std::fflush(stdout);
  DetArgs a{chains.data(), st.data(), 1, 1, 1, 0};
  emu_launch(nch, 256, lds, [&] { k_det_sweep(a); });

  double worst = 0.0;
  int fdiff = 0;
  unsigned long long acc_dev = 0, acc_host = 0;
  for (int c = 0; c < nch; ++c) {
    const int N = NM - c, L = LM - c;
    const size_t n2 = (size_t)N * N;
    std::vector<int8_t> f = f0[c];
    double old_det = host_det(host_m(f, ekv[c], N, L, lam, 0, +1), N) * host_det(host_m(f, ekv[c], N, L, lam, 0, -1), N);
    for (int t = 0; t < L; ++t) {
      const int l = L - 1 - t;
      for (int i = 0; i < N; ++i) {
        f[l * N + i] = (int8_t)-f[l * N + i];
        const double nd = host_det(host_m(f, ekv[c], N, L, lam, l, +1), N) * host_det(host_m(f, ekv[c], N, L, lam, l, -1), N);
        const double u = philox_uniform(chains[c].seed, (unsigned long long)t * N + i);
        if (u <= nd / old_det) { old_det = nd; ++acc_host; }
        else f[l * N + i] = (int8_t)-f[l * N + i];
      }
    }
    for (size_t e = 0; e < (size_t)N * L; ++e) fdiff += f[e] != chains[c].field[e];
    for (int s = 0; s < 2; ++s) {
      std::vector<double> g;
      host_det(host_m(f, ekv[c], N, L, lam, 0, s == 0 ? +1 : -1), N, &g);
      const double* gd = static_cast<double*>(chains[c].G) + s * n2;
      for (size_t e = 0; e < n2; ++e) {
        const double dg = std::fabs(g[e] - gd[e]);
        if (!(dg <= worst)) worst = dg;  // a NaN is kept (std::max would drop it)
      }
    }
    acc_dev += st[c].accepted;
    if (st[c].flags) { std::printf("FAIL: flags %u on chain %d\n", st[c].flags, c); return 1; }
  }
  std::printf("vs host: max|dG| %.3e, field differences %d, accepted %llu / %llu\n", worst, fdiff, acc_dev, acc_host);
  if (fdiff || !(worst <= 1e-10) || acc_dev != acc_host) { std::printf("FAIL\n"); return 1; }
  std::printf("ok\n");
  return 0;
}

// Numerical and bounds check of the Kronecker-factored E_K products in a CPU build: random
// factors Ex (n1 x n1), Ey (n2 x n2), E = Ex (x) Ey formed densely, and the left product
// C = diag(rs) E diag(ks) X diag(cs) and the right product C = diag(rs) X diag(ks) E diag(cs)
// of k_kron against the dense product, for a launch holding both given sizes as two tasks
// (ld of X and C larger than N in the second).
// usage: kron_harness n1 n2 n1b n2b seed
#include "hip_cpu_shim.h"
#include "../../latticeqmc_amd/csrc/kron.hip"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>

using namespace lqmc;

struct Case {
  int n1, n2, N, ldx, ldc;
  std::vector<double> ex, ey, X, C, rs, ks, cs;
};

int main(int argc, char** argv) {
  if (argc < 6) return 2;
  std::mt19937_64 rng(atoi(argv[5]));
  std::normal_distribution<double> g(0.0, 1.0);
  std::uniform_real_distribution<double> u(0.5, 2.0);
  double worst = 0.0;
  bool finite = true;
  for (int right = 0; right < 2; ++right) {
    std::vector<Case> cs(2);
    for (int t = 0; t < 2; ++t) {
      Case& c = cs[t];
      c.n1 = atoi(argv[1 + 2 * t]);
      c.n2 = atoi(argv[2 + 2 * t]);
      c.N = c.n1 * c.n2;
      c.ldx = c.N + 3 * t;
      c.ldc = c.N + 5 * t;
      c.ex.assign(256, 0.0);
      c.ey.assign(256, 0.0);
      for (int a = 0; a < c.n1; ++a) for (int b = 0; b < c.n1; ++b) c.ex[a * 16 + b] = g(rng);
      for (int a = 0; a < c.n2; ++a) for (int b = 0; b < c.n2; ++b) c.ey[a * 16 + b] = g(rng);
      c.X.resize((size_t)c.N * c.ldx);
      for (auto& v : c.X) v = g(rng);
      c.C.assign((size_t)c.N * c.ldc, 7.0);  // the padding columns must stay untouched
      c.rs.resize(c.N); c.ks.resize(c.N); c.cs.resize(c.N);
      for (int i = 0; i < c.N; ++i) { c.rs[i] = u(rng); c.ks[i] = u(rng); c.cs[i] = u(rng); }
    }
    std::vector<KronTask> tasks(2);
    int base = 0;
    for (int t = 0; t < 2; ++t) {
      Case& c = cs[t];
      KronTask& k = tasks[t];
      k.X = c.X.data(); k.C = c.C.data(); k.Ex = c.ex.data(); k.Ey = c.ey.data();
      k.rs = c.rs.data(); k.ks = c.ks.data(); k.cs = c.cs.data();
      // plain, inverse and clamped kinds on the three scales
      k.rs_kind = t ? SC_INV : SC_PLAIN; k.ks_kind = t ? SC_SMALL : SC_PLAIN; k.cs_kind = t ? SC_INV_BIG : SC_NONE;
      k.N = c.N; k.n1 = c.n1; k.n2 = c.n2; k.ldx = c.ldx; k.ldc = c.ldc;
      k.tile_base = base;
      base += kron_tiles(c.N);
    }
    // a grid smaller than the tile count: workgroups walk several tiles, across the two tasks
    const int grid = kron_grid(base, (base + 11) / 12);
    // the form without bounds checks when both tasks are 16 x 16
    const bool all16 = cs[0].n1 == 16 && cs[0].n2 == 16 && cs[1].n1 == 16 && cs[1].n2 == 16;
    if (right && all16) emu_launch(grid, 256, kron_shmem(), [&] { k_kron<true, true>(tasks.data(), 2, base); });
    else if (right) emu_launch(grid, 256, kron_shmem(), [&] { k_kron<true, false>(tasks.data(), 2, base); });
    else if (all16) emu_launch(grid, 256, kron_shmem(), [&] { k_kron<false, true>(tasks.data(), 2, base); });
    else emu_launch(grid, 256, kron_shmem(), [&] { k_kron<false, false>(tasks.data(), 2, base); });
    for (int t = 0; t < 2; ++t) {
      Case& c = cs[t];
      const KronTask& k = tasks[t];
      auto sc = [](int kind, const std::vector<double>& s, int i) {
        const double v = s[i];
        switch (kind) {
          case SC_PLAIN: return v;
          case SC_INV: return 1.0 / v;
          case SC_INV_BIG: return 1.0 / std::max(v, 1.0);
          case SC_SMALL: return std::min(v, 1.0);
          default: return 1.0;
        }
      };
      auto E = [&](int i, int j) { return c.ex[(i / c.n2) * 16 + j / c.n2] * c.ey[(i % c.n2) * 16 + j % c.n2]; };
      double num = 0.0, den = 0.0;
      for (int i = 0; i < c.N; ++i)
        for (int j = 0; j < c.N; ++j) {
          double s = 0.0;
          for (int q = 0; q < c.N; ++q)
            s += right ? c.X[(size_t)i * c.ldx + q] * sc(k.ks_kind, c.ks, q) * E(q, j)
                       : E(i, q) * sc(k.ks_kind, c.ks, q) * c.X[(size_t)q * c.ldx + j];
          const double ref = sc(k.rs_kind, c.rs, i) * s * sc(k.cs_kind, c.cs, j);
          const double got = c.C[(size_t)i * c.ldc + j];
          if (!std::isfinite(got)) finite = false;
          num = std::max(num, std::fabs(got - ref));
          den = std::max(den, std::fabs(ref));
        }
      for (int i = 0; i < c.N; ++i)
        for (int j = c.N; j < c.ldc; ++j)
          if (c.C[(size_t)i * c.ldc + j] != 7.0) finite = false;
      const double rel = num / den;
      printf("%s N=%d (%d x %d): max rel err %.3e\n", right ? "right" : "left", c.N, c.n1, c.n2, rel);
      worst = std::max(worst, rel);
    }
  }
  if (!finite || !(worst < 1e-13)) {
    printf("FAIL: worst %.3e finite/padding %d\n", worst, (int)finite);
    return 1;
  }
  printf("ok: worst %.3e\n", worst);
  return 0;
}

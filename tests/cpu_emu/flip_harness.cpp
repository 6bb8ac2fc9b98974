// Bounds check of the flip kernels in a CPU build (AddressSanitizer): every
// device buffer is a separate exact-size allocation and the dynamic LDS is an
// exact-size heap block, so any out-of-range access aborts with a report.
// usage: flip_harness N L delay threads fp32 nchains t measure [gscale [groups]]
// fp32: 0 FP64, 1 FP32, 2 FP32 with the rank-k products as three split-BF16 products
// groups > 1 runs the delayed kernel with that many workgroups per chain, all blocks at once.
#include "hip_cpu_shim.h"
#include "../../latticeqmc_amd/csrc/flip.hip"

#include <cstdio>
#include <cstdlib>
#include <random>

using namespace lqmc;

template <typename T>
static T* dalloc(size_t n) { return static_cast<T*>(std::calloc(n ? n : 1, sizeof(T))); }

int main(int argc, char** argv) {
  if (argc < 9) { std::fprintf(stderr, "usage: N L delay threads fp32 nchains t measure\n"); return 2; }
  const int N = atoi(argv[1]), L = atoi(argv[2]), K = atoi(argv[3]), threads = atoi(argv[4]);
  const int fp32 = atoi(argv[5]), nch = atoi(argv[6]), t = atoi(argv[7]), measure = atoi(argv[8]);
  const size_t n2 = (size_t)N * N;
  std::mt19937_64 rng(1);
  std::uniform_real_distribution<double> U(-0.3, 0.3);
  const double gscale = argc > 9 ? atof(argv[9]) : 1.0;  // off-diagonal scale (larger -> more rejections)
  const int groups = argc > 10 ? atoi(argv[10]) : 1;
  if (groups > 1 && K > 1 && K < 16) { std::fprintf(stderr, "groups > 1 needs delay 1, 16 or 32\n"); return 2; }
  std::vector<ChainDev> chains(nch);
  std::vector<ChainState> st(nch);
  for (int c = 0; c < nch; ++c) {
    ChainDev& d = chains[c];
    std::memset(&d, 0, sizeof d);
    d.N = N; d.L = L; d.mode = 0;
    d.lam = 0.6; d.e1m = std::exp(-0.6); d.e1p = std::exp(0.6); d.e2p = std::exp(1.2); d.e2m = std::exp(-1.2);
    d.seed = 1234 + c;
    d.field = dalloc<int8_t>((size_t)N * L);
    for (size_t e = 0; e < (size_t)N * L; ++e) d.field[e] = (rng() & 1) ? 1 : -1;
    d.ev = dalloc<double>(2 * (size_t)N * L);
    d.evi = dalloc<double>(2 * (size_t)N * L);
    if (fp32) {
      float* g = dalloc<float>(2 * n2);
      for (size_t e = 0; e < 2 * n2; ++e) g[e] = (float)(gscale * U(rng) / N + ((e % n2) / N == (e % n2) % N ? 0.5 : 0.0));
      d.G = g;
    } else {
      double* g = dalloc<double>(2 * n2);
      for (size_t e = 0; e < 2 * n2; ++e) g[e] = gscale * U(rng) / N + ((e % n2) / N == (e % n2) % N ? 0.5 : 0.0);
      d.G = g;
    }
    d.Gacc = dalloc<double>(2 * n2);
    d.Glast = dalloc<double>(2 * n2);
    d.Sacc = dalloc<double>(n2);
    d.hop = dalloc<double>(n2);
    d.uni = nullptr;
    d.tr_ratio = dalloc<double>((size_t)N * L);
    d.tr_acc = dalloc<uint8_t>((size_t)N * L);
    d.obs_cap = 8;
    d.obs = dalloc<double>(3 * 8);
    if (K > 1) {
      d.dws = dalloc<char>(flip_delayed_workspace(N, K, fp32, groups));
    }
  }
  std::memset(st.data(), 0, sizeof(ChainState) * nch);
  std::vector<unsigned> bar(4 * nch, 0u);
  FlipArgs a{chains.data(), st.data(), t, measure, 1, 1, K, groups, bar.data(), nch, fp32 == 2 ? 1 : 0,
              K > 1 ? flip_delayed_stage_rows(N, K, groups) : 0};
  if (K > 1 && (threads > flip_delayed_threads_limit(K) || threads < 256)) {
    std::fprintf(stderr, "threads above the delayed kernel's launch bound %d\n", flip_delayed_threads_limit(K));
    return 2;
  }
  const size_t lds = K > 1 ? flip_delayed_shmem(N, K, threads, fp32) + flip_delayed_rows_shmem(N, K, groups)
                   : groups > 1 ? flip_rank1_split_shmem(N, threads, groups)
                                : sizeof(double) * (4 * (size_t)N + threads);
  std::printf("N=%d L=%d K=%d threads=%d fp32=%d chains=%d t=%d groups=%d lds=%zu\n", N, L, K, threads, fp32, nch, t,
              groups, lds);
  std::fflush(stdout);
  if (K > 1) {
#define RUNK(KK)                                                                          \
  case KK:                                                                                \
    if (groups > 1) {                                                                     \
      if (fp32 == 2) emu_launch_concurrent(flip_grid(nch, groups), threads, lds, [&] { k_flip_delayed<float, KK, true>(a); }); \
      else if (fp32) emu_launch_concurrent(flip_grid(nch, groups), threads, lds, [&] { k_flip_delayed<float, KK>(a); }); \
      else emu_launch_concurrent(flip_grid(nch, groups), threads, lds, [&] { k_flip_delayed<double, KK>(a); });    \
    } else if (fp32 == 2) emu_launch(nch, threads, lds, [&] { k_flip_delayed<float, KK, true>(a); }); \
    else if (fp32) emu_launch(nch, threads, lds, [&] { k_flip_delayed<float, KK>(a); }); \
    else emu_launch(nch, threads, lds, [&] { k_flip_delayed<double, KK>(a); });          \
    break;
    switch (K) { RUNK(4) RUNK(8) RUNK(16) RUNK(32) default: std::printf("unsupported K\n"); return 2; }
  } else if (groups > 1) {
    if (fp32) emu_launch_concurrent(flip_grid(nch, groups), threads, lds, [&] { k_flip_rank1_split<float>(a); });
    else emu_launch_concurrent(flip_grid(nch, groups), threads, lds, [&] { k_flip_rank1_split<double>(a); });
  } else {
    if (fp32) emu_launch(nch, threads, lds, [&] { k_flip_rank1<float>(a); });
    else emu_launch(nch, threads, lds, [&] { k_flip_rank1<double>(a); });
  }
  unsigned long long acc = 0;
  for (auto& s : st) acc += s.accepted;
  for (int c = 0; c < nch; ++c)
    if (st[c].flags) { std::printf("FAIL: flags %u on chain %d\n", st[c].flags, c); return 1; }
  // cross-check against the rank-1 kernel on a copy of the initial state (same uniforms); the
  // rank-1 kernel split over several workgroups must agree with it bit for bit
  if ((K > 1 || groups > 1) && argc > 9) {
    std::vector<ChainDev> ref(nch);
    std::vector<ChainState> st2(nch);
    std::memset(st2.data(), 0, sizeof(ChainState) * nch);
    std::mt19937_64 rng2(1);
    double worst = 0.0;
    int fdiff = 0;
    for (int c = 0; c < nch; ++c) {
      ref[c] = chains[c];
      ChainDev& d = ref[c];
      d.field = dalloc<int8_t>((size_t)N * L);
      for (size_t e = 0; e < (size_t)N * L; ++e) d.field[e] = (rng2() & 1) ? 1 : -1;
      double* g = dalloc<double>(2 * n2);
      for (size_t e = 0; e < 2 * n2; ++e) g[e] = gscale * U(rng2) / N + ((e % n2) / N == (e % n2) % N ? 0.5 : 0.0);
      d.G = g;
      d.Gacc = dalloc<double>(2 * n2); d.Glast = dalloc<double>(2 * n2); d.Sacc = dalloc<double>(n2);
      d.ev = dalloc<double>(2 * (size_t)N * L); d.evi = dalloc<double>(2 * (size_t)N * L);
      d.tr_ratio = dalloc<double>((size_t)N * L); d.tr_acc = dalloc<uint8_t>((size_t)N * L);
      d.obs = dalloc<double>(3 * 8);
    }
    FlipArgs b{ref.data(), st2.data(), t, measure, 1, 1, 1, 0};
    emu_launch(nch, threads, sizeof(double) * (4 * (size_t)N + threads), [&] { k_flip_rank1<double>(b); });
    for (int c = 0; c < nch; ++c) {
      const double* g1 = static_cast<double*>(ref[c].G);
      for (size_t e = 0; e < 2 * n2; ++e) {
        const double g0 = fp32 ? (double)static_cast<float*>(chains[c].G)[e] : static_cast<double*>(chains[c].G)[e];
        const double dg = std::fabs(g0 - g1[e]);
        if (!(dg <= worst)) worst = dg;  // a NaN is kept (std::max would drop it)
      }
      for (size_t e = 0; e < (size_t)N * L; ++e) fdiff += chains[c].field[e] != ref[c].field[e];
    }
    std::printf("vs rank-1: max|dG| %.3e, field differences %d\n", worst, fdiff);
    if (fdiff || !(worst <= (K == 1 ? 0.0 : fp32 == 2 ? 2e-3 : fp32 ? 1e-4 : 1e-11))) { std::printf("FAIL\n"); return 1; }
  }
  for (int c = 0; c < nch; ++c)
    if (st[c].flags) { std::printf("FAIL: flags %u on chain %d\n", st[c].flags, c); return 1; }
  for (unsigned v : bar)
    if (v) { std::printf("FAIL: barrier counters not reset\n"); return 1; }
  std::printf("ok: accepted %llu of %llu, flags %u\n", acc, (unsigned long long)nch * N, st[0].flags);
  return 0;
}

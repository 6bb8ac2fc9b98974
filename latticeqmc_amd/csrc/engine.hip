// Host side of the sweep engine: device arenas, the per-sweep launch program,
// hipGraph replay, profiling and the C ABI declared in include/lqmc_hip.h.
//
// A sweep is a fixed sequence of batched launches (all chains, both spins):
//   Mode D ("dqmc", lqmc/lqmc.py:301-347 made consistent and stabilised)
//     left stack:  for each slice j bottom-up, W = diag(v_j) E_K W; at the end of
//                  every block of n_s slices a pivoted QR gives S_m = U D V
//     for t = 0..L-1 (slice l = L-1-t):
//        flip(l)                          one workgroup per chain
//        wrap  G <- E_K^-1 V^-1 G V E_K    two GEMMs (form precision)
//        right product W <- E_K V W       one FP64 GEMM
//        every n_s slices: QR(W), V_R update, stable solve for G(l-1),
//        compare with the wrapped G and replace it
//   Mode R ("reference", literal lqmc/lqmc.py:301-347)
//     P = prod E_K V(l) (FP64 GEMM chain), G = inv(I + P) (LU), then
//     for t: flip(l); G <- V(l-1) E_K G E_K^-1 V(l-1)^-1
// Descriptors of every launch are built once and kept in device memory, so a
// sweep issues no host->device copies and can be captured into a graph.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lqmc_hip.h"
#include "common.h"

namespace lqmc {
void launch_gemm(int fp32, int cls, const GemmTask* d_tasks, int ntasks, int ntiles, hipStream_t s);
void launch_kron(int right, int all16, const KronTask* d_tasks, int ntasks, int ntiles, int cus, hipStream_t s);
int kron_tiles(int N);
int gemm_class(const GemmTask& t);
int gemm_tiles(int M, int N, int* tiles_n);
void launch_qr_blocked(const QrTask* d, int n, int nmax, ChainState* st, unsigned* pub, int max_wg, hipStream_t s);
void launch_trsm(const TrsmTask* d, int n, int nmax, ChainState* st, hipStream_t s);
void launch_solve(const SolveTask* d, int n, int nmax, ChainState* st, hipStream_t s);
void launch_finish(int fp32, const FinishTask* d, int n, int nmax, ChainState* st, hipStream_t s);
void launch_identity(double* const* ptrs, const int* ns, int n, hipStream_t s);
void launch_expv(const ChainDev* chains, int n, hipStream_t s);
void launch_green_out(double* out, const void* src, int fp32, size_t n, const ChainState* st, int mean,
                      hipStream_t s);
void launch_flip(int fp32, int delayed, const FlipArgs& a, int nchains, int threads, size_t shmem,
                 hipStream_t s);
size_t flip_delayed_shmem(int nmax, int k, int threads, int fp32);
size_t flip_delayed_workspace(int n, int k, int fp32, int groups);
int flip_delayed_stage_rows(int nmax, int k, int groups);
size_t flip_delayed_rows_shmem(int nmax, int k, int groups);
size_t flip_rank1_split_shmem(int nmax, int threads, int groups);
int flip_grid(int nchains, int groups);
int flip_delayed_threads_limit(int k);
size_t qr_blocked_shmem(int nmax);
size_t det_sweep_shmem(int nmax, int lmax);
void launch_det_sweep(const ChainDev* chains, ChainState* st, int nchains, int measure, int rng, int trace,
                      size_t shmem, hipStream_t s);
}  // namespace lqmc

using namespace lqmc;

static thread_local std::string g_err;

namespace lqmc {
// The thread-local error message of the C ABI, for entry points in other source files.
void lqmc_set_error(const std::string& msg) { g_err = msg; }
}  // namespace lqmc

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      return fail(LQMC_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));       \
  } while (0)

namespace {

enum OpKind { OP_GEMM = 0, OP_QR, OP_SOLVE, OP_FINISH, OP_FLIP, OP_EXPV, OP_IDENT, OP_TRSM, OP_DET, OP_KRON };
enum ProfKind { PK_FLIP = 0, PK_GEMM = 1, PK_QR = 2, PK_SOLVE = 3, PK_KRON = 4 };

struct Op {
  int kind;
  int fp32;
  size_t off;   // byte offset into the descriptor buffer
  int count;
  int ntiles;
  int nmax;
  int t;
  double work;  // algorithmic work units (flips for OP_FLIP, flops for OP_GEMM)
  int meas_only = 0;  // skipped in sweeps that do not measure
};

struct Chain {
  int N = 0, L = 0, nb = 0;
  bool set = false;
  double lam = 0;
  uint64_t seed = 0;
  // device buffers (carved from the arena)
  int8_t* field = nullptr;
  double *ev = nullptr, *evi = nullptr;
  void* G = nullptr;
  double *Gacc = nullptr, *Glast = nullptr, *Gst = nullptr, *hop = nullptr, *Sacc = nullptr;
  double *tr_ratio = nullptr, *obs = nullptr;
  uint8_t* tr_acc = nullptr;
  double *expk = nullptr, *expki = nullptr;
  float *expk32 = nullptr, *expki32 = nullptr;
  // E_K = Ex (x) Ey and E_K^-1 = Exi (x) Eyi when the lattice factors (kn1 > 0): the four
  // 16 x 16 blocks Ex, Ey, Exi, Eyi at kr, zero-padded
  int kn1 = 0, kn2 = 0;
  double* kr = nullptr;
  double *Wa[2] = {}, *Wb[2] = {};
  double *LU[2] = {}, *LD[2] = {}, *LV[2] = {};
  double *RU[2] = {}, *RD[2] = {}, *RV[2][2] = {};
  double *Tq[2] = {}, *Tt[2] = {}, *aug[2] = {};
  double *Dx[2] = {}, *Zb[2] = {};
  int* px[2] = {};
  unsigned long long* rec = nullptr;    // [nb] max |G_wrapped - G_recomputed| per recompute point
  double* Gtd = nullptr;                // [nb][2][N][N] sums of G(tau, 0) (time_displaced)
  double *Ztd[2] = {}, *Ytd[2] = {};   // its right-hand side and solution
  double *eye = nullptr, *ones = nullptr;
  void* Tw[2] = {};
  void* dws = nullptr;    // delayed-update operands
  double* uni = nullptr;  // separate allocation (replay)
  int64_t uni_cap = 0;
};

struct Prof {
  int kind = -1;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  std::vector<double> work;
};

}  // namespace

struct lqmc_ctx {
  int device = 0;
  int nchains = 0;
  lqmc_options opt{};
  hipStream_t stream = nullptr;
  hipEvent_t done_ev = nullptr;  // end of the last sweep issued on a caller stream
  hipEvent_t aux_ev = nullptr;   // context-stream point a caller stream waits for
  double max_dg = 1e-3;          // precision-flag threshold of the recompute check (0: off)
  std::vector<Chain> ch;
  char* arena = nullptr;
  size_t arena_bytes = 0;
  ChainDev* d_chains = nullptr;
  ChainState* d_state = nullptr;
  std::vector<ChainDev> h_chains;
  // program
  bool built = false, inited = false;
  std::vector<Op> init_ops, sweep_ops;
  char* d_desc = nullptr;
  size_t desc_bytes = 0;
  std::vector<char> h_desc;
  double** d_eye_ptrs = nullptr;
  int* d_eye_ns = nullptr;
  int nmax = 0, lmax = 0;
  int flip_threads = 256;
  int flip_groups = 1;           // workgroups per chain of the delayed flip
  int flip_yrows = 0;            // delayed flip keeps block rows and Z tiles in LDS (FlipArgs::yrows)
  unsigned* d_bar = nullptr;     // [nchains][4] their barrier counters and XCC bits
  unsigned* d_qrpub = nullptr;   // [4 nchains] hand-over counters of the on-chip QR (two per matrix)
  int cus = 0;                   // compute units of the device
  size_t flip_shmem = 0;
  size_t det_shmem = 0;
  // graphs
  bool use_graph = false;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  hipGraph_t graph[2] = {nullptr, nullptr};
  Prof prof;
};

// ------------------------------------------------------------------------------------------
// arena layout
static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

template <typename T>
static void carve(char* base, size_t& off, T*& p, size_t count) {
  off = align_up(off, 256);
  p = base ? reinterpret_cast<T*>(base + off) : nullptr;
  off += sizeof(T) * count;
}

static void layout_chain(const lqmc_ctx* x, Chain& c, char* base, size_t& off) {
  const size_t N = c.N, L = c.L, n2 = N * N;
  const bool fp32 = x->opt.precision != LQMC_FP64;
  const bool dq = x->opt.mode == LQMC_MODE_DQMC;
  const size_t esz = fp32 ? 4 : 8;
  carve(base, off, c.field, N * L);
  carve(base, off, c.ev, 2 * N * L);
  carve(base, off, c.evi, 2 * N * L);
  char* g = nullptr;
  carve(base, off, g, 2 * n2 * esz);
  c.G = g;
  carve(base, off, c.Gacc, 2 * n2);
  carve(base, off, c.Glast, 2 * n2);
  carve(base, off, c.Sacc, n2);
  carve(base, off, c.Gst, 2 * n2);
  carve(base, off, c.hop, n2);
  carve(base, off, c.obs, 3 * (size_t)x->opt.obs_capacity);
  if (x->opt.trace) {
    carve(base, off, c.tr_ratio, N * L);
    carve(base, off, c.tr_acc, N * L);
  }
  carve(base, off, c.expk, n2);
  carve(base, off, c.expki, n2);
  if (fp32) {
    carve(base, off, c.expk32, n2);
    carve(base, off, c.expki32, n2);
  }
  carve(base, off, c.kr, 4 * 256);
  carve(base, off, c.eye, n2);
  carve(base, off, c.ones, N);
  if (x->opt.delay > 1) {
    char* w = nullptr;
    carve(base, off, w, flip_delayed_workspace(c.N, x->opt.delay, fp32 ? 1 : 0, x->flip_groups));
    c.dws = w;
  }
  for (int s = 0; s < 2; ++s) {
    carve(base, off, c.Wa[s], n2);
    carve(base, off, c.Wb[s], n2);
    carve(base, off, c.aug[s], 2 * n2);
    char* tw = nullptr;
    carve(base, off, tw, n2 * esz);
    c.Tw[s] = tw;
    if (dq) {
      carve(base, off, c.LU[s], (size_t)c.nb * n2);
      carve(base, off, c.LD[s], (size_t)c.nb * N);
      carve(base, off, c.LV[s], (size_t)c.nb * n2);
      carve(base, off, c.RU[s], n2);
      carve(base, off, c.RD[s], N);
      carve(base, off, c.RV[s][0], n2);
      carve(base, off, c.RV[s][1], n2);
      carve(base, off, c.Tq[s], n2);
      carve(base, off, c.Tt[s], 16 * (N + 16));
      carve(base, off, c.Dx[s], N);
      carve(base, off, c.Zb[s], n2);
      carve(base, off, c.px[s], N);
      if (x->opt.time_displaced) {
        carve(base, off, c.Ztd[s], n2);
        carve(base, off, c.Ytd[s], n2);
      }
    }
  }
  if (dq && x->opt.time_displaced) carve(base, off, c.Gtd, (size_t)c.nb * 2 * n2);
  if (dq) carve(base, off, c.rec, (size_t)c.nb);
}

// ------------------------------------------------------------------------------------------
// program builder
namespace {

struct Builder {
  lqmc_ctx* x;
  std::vector<char>& desc;
  explicit Builder(lqmc_ctx* ctx) : x(ctx), desc(ctx->h_desc) {}

  template <typename T>
  size_t push(const std::vector<T>& v) {
    size_t off = align_up(desc.size(), 256);
    desc.resize(off + sizeof(T) * v.size());
    std::memcpy(desc.data() + off, v.data(), sizeof(T) * v.size());
    return off;
  }

  // One launch per kernel class (transpose case, full tiles) present among the tasks; the
  // tasks of one call are independent, so the split launches may run in any order.
  // An FP64 product with a chain's E_K or E_K^-1 as the untransposed left (right) operand,
  // with alpha 1 and beta 0, becomes a left (right) Kronecker task when that chain's E_K factors.
  bool as_kron(const GemmTask& t, int fp32, KronTask& k, int& right) const {
    if (fp32 || t.alpha != 1.0 || t.beta != 0.0 || t.M != t.N || t.N != t.K) return false;
    for (const Chain& c : x->ch) {
      if (c.kn1 == 0 || c.N != t.M) continue;
      const bool la = !t.transA && (t.A == c.expk || t.A == c.expki) && !t.transB;
      const bool rb = !t.transB && (t.B == c.expk || t.B == c.expki) && !t.transA;
      if (!la && !rb) continue;
      right = la ? 0 : 1;
      const bool inv = la ? t.A == c.expki : t.B == c.expki;
      k = KronTask{};
      k.X = static_cast<const double*>(la ? t.B : t.A);
      k.C = static_cast<double*>(t.C);
      k.Ex = c.kr + (inv ? 512 : 0);
      k.Ey = c.kr + (inv ? 768 : 256);
      k.rs = t.rs; k.ks = t.ks; k.cs = t.cs;
      k.rs_kind = t.rs_kind; k.ks_kind = t.ks_kind; k.cs_kind = t.cs_kind;
      k.N = c.N; k.n1 = c.kn1; k.n2 = c.kn2;
      k.ldx = la ? t.ldb : t.lda;
      k.ldc = t.ldc;
      return true;
    }
    return false;
  }

  void gemm(std::vector<Op>& ops, std::vector<GemmTask>& tasks, int fp32) {
    if (tasks.empty()) return;
    {
      std::vector<KronTask> kt[2];
      std::vector<GemmTask> rest;
      for (auto& t : tasks) {
        KronTask k;
        int right = 0;
        if (as_kron(t, fp32, k, right)) kt[right].push_back(k);
        else rest.push_back(t);
      }
      for (int right = 0; right < 2; ++right) {
        if (kt[right].empty()) continue;
        int base = 0, all16 = 1;
        double flops = 0;
        for (auto& k : kt[right]) {
          k.tile_base = base;
          base += kron_tiles(k.N);
          flops += 2.0 * k.N * (double)k.N * (k.n1 + k.n2);
          all16 &= k.n1 == 16 && k.n2 == 16;
        }
        // fp32 field: whether every task is 16 x 16 (the kernel's form without bounds checks)
        ops.push_back(Op{OP_KRON, all16, push(kt[right]), (int)kt[right].size(), base, right, 0, flops});
      }
      tasks.swap(rest);
      if (tasks.empty()) return;
    }
    for (int cls = 0; cls < 8; ++cls) {
      std::vector<GemmTask> part;
      for (auto& t : tasks)
        if (gemm_class(t) == cls) part.push_back(t);
      if (part.empty()) continue;
      int base = 0;
      double flops = 0;
      for (auto& t : part) {
        int tn;
        int nt = gemm_tiles(t.M, t.N, &tn);
        t.tiles_n = tn;
        t.tile_base = base;
        base += nt;
        flops += 2.0 * t.M * t.N * t.K;
      }
      ops.push_back(Op{OP_GEMM, fp32, push(part), (int)part.size(), base, cls, 0, flops});
    }
    tasks.clear();
  }
  void qr(std::vector<Op>& ops, std::vector<QrTask>& tasks) {
    if (tasks.empty()) return;
    int nm = 0;
    double fl = 0;
    for (auto& t : tasks) { nm = std::max(nm, t.N); fl += 8.0 / 3.0 * t.N * (double)t.N * t.N; }
    ops.push_back(Op{OP_QR, 0, push(tasks), (int)tasks.size(), 0, nm, 0, fl});
    tasks.clear();
  }
  void solve(std::vector<Op>& ops, std::vector<SolveTask>& tasks) {
    if (tasks.empty()) return;
    int nm = 0;
    double fl = 0;
    for (auto& t : tasks) { nm = std::max(nm, t.N); fl += 8.0 / 3.0 * t.N * (double)t.N * t.N; }
    ops.push_back(Op{OP_SOLVE, 0, push(tasks), (int)tasks.size(), 0, nm, 0, fl});
    tasks.clear();
  }
  void trsm(std::vector<Op>& ops, std::vector<TrsmTask>& tasks) {
    if (tasks.empty()) return;
    int nm = 0;
    double fl = 0;
    for (auto& t : tasks) { nm = std::max(nm, t.N); fl += (double)t.N * t.N * t.N; }
    ops.push_back(Op{OP_TRSM, 0, push(tasks), (int)tasks.size(), 0, nm, 0, fl});
    tasks.clear();
  }
  void finish(std::vector<Op>& ops, std::vector<FinishTask>& tasks, int fp32) {
    if (tasks.empty()) return;
    int nm = 0;
    for (auto& t : tasks) nm = std::max(nm, t.N);
    ops.push_back(Op{OP_FINISH, fp32, push(tasks), (int)tasks.size(), 0, nm, 0, 0});
    tasks.clear();
  }
};

GemmTask mk(const void* A, int lda, int tA, const void* B, int ldb, int tB, void* C, int ldc, int N) {
  GemmTask g{};
  g.A = A; g.B = B; g.C = C;
  g.M = g.N = g.K = N;
  g.lda = lda; g.ldb = ldb; g.ldc = ldc;
  g.transA = tA; g.transB = tB;
  g.alpha = 1.0; g.beta = 0.0;
  return g;
}

}  // namespace

static void build_program(lqmc_ctx* x) {
  x->h_desc.clear();
  x->init_ops.clear();
  x->sweep_ops.clear();
  Builder b(x);
  const bool fp32 = x->opt.precision != LQMC_FP64;
  const bool ref = x->opt.mode == LQMC_MODE_REFERENCE;
  const int ns = x->opt.n_stab;
  const int nc = x->nchains;
  std::vector<GemmTask> g1, g2;
  std::vector<QrTask> q;
  std::vector<SolveTask> sv;
  std::vector<TrsmTask> ts;
  std::vector<FinishTask> fin;

  // ---- init: scale tables and identity matrices ----
  x->init_ops.push_back(Op{OP_EXPV, 0, 0, nc, 0, 0, 0, 0});
  x->init_ops.push_back(Op{OP_IDENT, 0, 0, nc, 0, 0, 0, 0});

  auto flipop = [&](std::vector<Op>& ops, int t) {
    double w = 0;
    for (auto& c : x->ch) if (t < c.L) w += c.N;
    ops.push_back(Op{OP_FLIP, fp32, 0, nc, 0, 0, t, w});
  };
  auto evp = [&](Chain& c, int s, int l) { return c.ev + (size_t)s * c.N * c.L + (size_t)l * c.N; };
  auto evip = [&](Chain& c, int s, int l) { return c.evi + (size_t)s * c.N * c.L + (size_t)l * c.N; };
  auto Gp = [&](Chain& c, int s) -> void* {
    return static_cast<char*>(c.G) + (size_t)s * c.N * c.N * (fp32 ? 4 : 8);
  };

  if (x->opt.mode == LQMC_MODE_DET) {
    // ---- determinant mode: one brute-force sweep kernel per sweep ----
    double w = 0;
    for (auto& c : x->ch) w += (double)c.N * c.L;
    x->sweep_ops.push_back(Op{OP_DET, 0, 0, nc, 0, 0, 0, w});
  } else if (ref) {
    // ---- Mode R sweep: P = prod_{l=L-1..0} E_K V(l), G = inv(I + P) ----
    std::vector<int> cur(nc * 2, 0);  // 0 -> Wa, 1 -> Wb holds the running product
    for (int m = 0; m < x->lmax; ++m) {
      for (int ci = 0; ci < nc; ++ci) {
        Chain& c = x->ch[ci];
        if (m >= c.L) continue;
        const int l = c.L - 1 - m;
        for (int s = 0; s < 2; ++s) {
          double* out = (m == c.L - 1) ? c.aug[s] : ((m % 2 == 0) ? c.Wa[s] : c.Wb[s]);
          int ldc = (m == c.L - 1) ? 2 * c.N : c.N;
          GemmTask g;
          if (m == 0) {
            g = mk(c.expk, c.N, 0, c.eye, c.N, 0, out, ldc, c.N);
          } else {
            double* in = ((m - 1) % 2 == 0) ? c.Wa[s] : c.Wb[s];
            g = mk(in, c.N, 0, c.expk, c.N, 0, out, ldc, c.N);
          }
          g.cs = evp(c, s, l);
          g.cs_kind = SC_PLAIN;
          g1.push_back(g);
        }
      }
      b.gemm(x->sweep_ops, g1, 0);
    }
    for (int ci = 0; ci < nc; ++ci) {
      Chain& c = x->ch[ci];
      for (int s = 0; s < 2; ++s) {
        SolveTask t{c.aug[s], nullptr, nullptr, c.Gst + (size_t)s * c.N * c.N, c.N, 1, ci, 0};
        sv.push_back(t);
        fin.push_back(FinishTask{Gp(c, s), c.Gst + (size_t)s * c.N * c.N, nullptr, 0.0, c.N, ci, 0, 0});
      }
    }
    b.solve(x->sweep_ops, sv);
    b.finish(x->sweep_ops, fin, fp32);
    for (int t = 0; t < x->lmax; ++t) {
      flipop(x->sweep_ops, t);
      for (int ci = 0; ci < nc; ++ci) {
        Chain& c = x->ch[ci];
        if (t >= c.L - 1) continue;  // no wrap after slice 0
        const int l = c.L - 1 - t;
        for (int s = 0; s < 2; ++s) {
          g1.push_back(mk(c.expk, c.N, 0, Gp(c, s), c.N, 0, c.Tw[s], c.N, c.N));
          GemmTask g = mk(c.Tw[s], c.N, 0, c.expki, c.N, 0, Gp(c, s), c.N, c.N);
          g.rs = evp(c, s, l - 1); g.rs_kind = SC_PLAIN;
          g.cs = evip(c, s, l - 1); g.cs_kind = SC_PLAIN;
          g2.push_back(g);
        }
      }
      b.gemm(x->sweep_ops, g1, 0);
      b.gemm(x->sweep_ops, g2, 0);
    }
  } else {
    // ---- Mode D ----
    // left stack: blocks m = nb-1 .. m_min, slices lo..hi (hi = L-1-m*ns)
    auto left_stack = [&](std::vector<Op>& ops, int m_min) {
      // per chain-spin: which buffer holds the running product
      std::vector<double*> wcur(nc * 2, nullptr);
      for (int j = 0; j < x->lmax; ++j) {
        std::vector<int> ends;
        for (int ci = 0; ci < nc; ++ci) {
          Chain& c = x->ch[ci];
          const int top = c.L - 1 - m_min * ns;  // last slice in the stack
          if (j > top) continue;
          const int m = (c.L - 1 - j) / ns;       // block containing slice j
          const int lo = std::max(c.L - (m + 1) * ns, 0);
          const int hi = c.L - 1 - m * ns;
          for (int s = 0; s < 2; ++s) {
            double*& w = wcur[ci * 2 + s];
            double* out = (j == lo || w == c.Wb[s]) ? c.Wa[s] : c.Wb[s];
            const bool bottom = (m == c.nb - 1);
            const double* in = (j == lo) ? (bottom ? c.eye : c.LU[s] + (size_t)(m + 1) * c.N * c.N) : w;
            GemmTask g = mk(c.expk, c.N, 0, in, c.N, 0, out, c.N, c.N);
            g.rs = evp(c, s, j); g.rs_kind = SC_PLAIN;
            if (j == lo && !bottom) { g.cs = c.LD[s] + (size_t)(m + 1) * c.N; g.cs_kind = SC_PLAIN; }
            g1.push_back(g);
            w = out;
          }
          if (j == hi) ends.push_back(ci);
        }
        b.gemm(ops, g1, 0);
        for (int ci : ends) {
          Chain& c = x->ch[ci];
          const int m = (c.L - 1 - j) / ns;
          const bool bottom = (m == c.nb - 1);
          for (int s = 0; s < 2; ++s) {
            const size_t n2 = (size_t)c.N * c.N;
            q.push_back(QrTask{wcur[ci * 2 + s], c.LU[s] + m * n2, c.LD[s] + (size_t)m * c.N, c.Tq[s], nullptr, c.Tt[s],
                               c.N, ci});
            const double* vin = bottom ? c.eye : c.LV[s] + (m + 1) * n2;
            g2.push_back(mk(wcur[ci * 2 + s], c.N, 0, vin, c.N, 0, c.LV[s] + m * n2, c.N, c.N));
          }
        }
        b.qr(ops, q);
        b.gemm(ops, g2, 0);
      }
    };
    // stable G from the left entry (or identity) and the right factors:
    //   X = Db_L^-1 (U_L^T U_R) Db_R^-1 + Ds_L (V_L V_R^T) Ds_R ;  X = Qx Dx Vx   (QR)
    //   Z = Dx^-1 Qx^T Db_L^-1 U_L^T ;  Vx Y = Z ;  G = U_R Db_R^-1 Y
    // td >= 0 (time_displaced, sweep only): also G(tau, 0) = U_R Db_R^-1 X^-1 Ds_L V_L from the
    // same QR of X, added to the sums of stabilisation point td in measured sweeps:
    //   Z' = Dx^-1 Qx^T Ds_L V_L ;  Vx Y' = Z' ;  Gtd[td] += U_R Db_R^-1 Y'
    // slot >= 0: the recompute point's index along the sweep (per-point deviation record)
    auto recompute = [&](std::vector<Op>& ops, std::vector<int>& which, std::vector<int>& mleft,
                         std::vector<int>& rvsel, bool right_identity, int compare, int td, int slot) {
      for (size_t k = 0; k < which.size(); ++k) {
        Chain& c = x->ch[which[k]];
        const size_t n2 = (size_t)c.N * c.N;
        for (int s = 0; s < 2; ++s) {
          const bool lid = mleft[k] < 0;
          const double* UL = lid ? c.eye : c.LU[s] + mleft[k] * n2;
          const double* DL = lid ? c.ones : c.LD[s] + (size_t)mleft[k] * c.N;
          const double* VL = lid ? c.eye : c.LV[s] + mleft[k] * n2;
          const double* UR = right_identity ? c.eye : c.RU[s];
          const double* DR = right_identity ? c.ones : c.RD[s];
          const double* VR = right_identity ? c.eye : c.RV[s][rvsel[k]];
          GemmTask a = mk(UL, c.N, 1, UR, c.N, 0, c.aug[s], c.N, c.N);
          a.rs = DL; a.rs_kind = SC_INV_BIG; a.cs = DR; a.cs_kind = SC_INV_BIG;
          g1.push_back(a);
          GemmTask bb = mk(VL, c.N, 0, VR, c.N, 1, c.aug[s], c.N, c.N);
          bb.rs = DL; bb.rs_kind = SC_SMALL; bb.cs = DR; bb.cs_kind = SC_SMALL; bb.beta = 1.0;
          g2.push_back(bb);
          q.push_back(QrTask{c.aug[s], c.aug[s] + n2, c.Dx[s], c.Tq[s], c.px[s], c.Tt[s], c.N, which[k]});
        }
      }
      b.gemm(ops, g1, 0);
      b.gemm(ops, g2, 0);
      b.qr(ops, q);
      for (size_t k = 0; k < which.size(); ++k) {
        Chain& c = x->ch[which[k]];
        const size_t n2 = (size_t)c.N * c.N;
        for (int s = 0; s < 2; ++s) {
          const bool lid = mleft[k] < 0;
          const double* UL = lid ? c.eye : c.LU[s] + mleft[k] * n2;
          const double* DL = lid ? c.ones : c.LD[s] + (size_t)mleft[k] * c.N;
          GemmTask z = mk(c.aug[s] + n2, c.N, 1, UL, c.N, 1, c.Zb[s], c.N, c.N);
          z.rs = c.Dx[s]; z.rs_kind = SC_INV;
          z.ks = DL; z.ks_kind = SC_INV_BIG;
          g1.push_back(z);
          ts.push_back(TrsmTask{c.aug[s], c.px[s], c.Zb[s], c.aug[s] + n2, c.N, which[k]});
        }
      }
      b.gemm(ops, g1, 0);
      if (td >= 0) {  // before the equal-time solve overwrites Qx
        const size_t from = ops.size();
        std::vector<GemmTask> gz, gg;
        std::vector<TrsmTask> tz;
        for (size_t k = 0; k < which.size(); ++k) {
          Chain& c = x->ch[which[k]];
          const size_t n2 = (size_t)c.N * c.N;
          for (int s = 0; s < 2; ++s) {
            const bool lid = mleft[k] < 0;
            const double* DL = lid ? c.ones : c.LD[s] + (size_t)mleft[k] * c.N;
            const double* VL = lid ? c.eye : c.LV[s] + mleft[k] * n2;
            const double* UR = right_identity ? c.eye : c.RU[s];
            const double* DR = right_identity ? c.ones : c.RD[s];
            GemmTask z = mk(c.aug[s] + n2, c.N, 1, VL, c.N, 0, c.Ztd[s], c.N, c.N);
            z.rs = c.Dx[s]; z.rs_kind = SC_INV;
            z.ks = DL; z.ks_kind = SC_SMALL;
            gz.push_back(z);
            tz.push_back(TrsmTask{c.aug[s], c.px[s], c.Ztd[s], c.Ytd[s], c.N, which[k]});
            GemmTask g = mk(UR, c.N, 0, c.Ytd[s], c.N, 0, c.Gtd + ((size_t)td * 2 + s) * n2, c.N, c.N);
            g.ks = DR; g.ks_kind = SC_INV_BIG; g.beta = 1.0;
            gg.push_back(g);
          }
        }
        b.gemm(ops, gz, 0);
        b.trsm(ops, tz);
        b.gemm(ops, gg, 0);
        for (size_t i = from; i < ops.size(); ++i) ops[i].meas_only = 1;
      }
      b.trsm(ops, ts);
      for (size_t k = 0; k < which.size(); ++k) {
        Chain& c = x->ch[which[k]];
        const size_t n2 = (size_t)c.N * c.N;
        for (int s = 0; s < 2; ++s) {
          const double* UR = right_identity ? c.eye : c.RU[s];
          const double* DR = right_identity ? c.ones : c.RD[s];
          GemmTask g = mk(UR, c.N, 0, c.aug[s] + n2, c.N, 0, c.Gst + s * n2, c.N, c.N);
          g.ks = DR; g.ks_kind = SC_INV_BIG;
          g1.push_back(g);
          fin.push_back(FinishTask{Gp(c, s), c.Gst + s * n2, slot >= 0 ? c.rec + slot : nullptr, x->max_dg,
                                   c.N, which[k], compare, 0});
        }
      }
      b.gemm(ops, g1, 0);
      b.finish(ops, fin, fp32);
    };

    // init: full stack (m >= 0) and G(L-1) = (I + S_0)^-1
    left_stack(x->init_ops, 0);
    {
      std::vector<int> which, ml, rv;
      for (int ci = 0; ci < nc; ++ci) { which.push_back(ci); ml.push_back(0); rv.push_back(0); }
      recompute(x->init_ops, which, ml, rv, true, 0, -1, -1);
    }
    // sweep
    left_stack(x->sweep_ops, 1);
    std::vector<double*> wcur(nc * 2, nullptr);
    std::vector<int> rvcur(nc, -1);  // -1: right chain is the identity
    for (int t = 0; t < x->lmax; ++t) {
      flipop(x->sweep_ops, t);
      for (int ci = 0; ci < nc; ++ci) {
        Chain& c = x->ch[ci];
        if (t >= c.L) continue;
        const int l = c.L - 1 - t;
        for (int s = 0; s < 2; ++s) {
          const void* ekinv = fp32 ? (const void*)c.expki32 : (const void*)c.expki;
          const void* ek = fp32 ? (const void*)c.expk32 : (const void*)c.expk;
          GemmTask w1 = mk(ekinv, c.N, 0, Gp(c, s), c.N, 0, c.Tw[s], c.N, c.N);
          w1.ks = evip(c, s, l); w1.ks_kind = SC_PLAIN;
          w1.cs = evp(c, s, l); w1.cs_kind = SC_PLAIN;
          g1.push_back(w1);
          g2.push_back(mk(c.Tw[s], c.N, 0, ek, c.N, 0, Gp(c, s), c.N, c.N));
        }
      }
      // right product W = E_K V(l) In; FP64 and independent of the wrap, so in FP64
      // sweeps its tasks join the first wrap launch
      if (fp32) b.gemm(x->sweep_ops, g1, fp32);
      std::vector<int> bnd, ml, rvsel;
      for (int ci = 0; ci < nc; ++ci) {
        Chain& c = x->ch[ci];
        if (t >= c.L) continue;
        const int l = c.L - 1 - t;
        const bool first = (t % ns == 0);
        for (int s = 0; s < 2; ++s) {
          double*& w = wcur[ci * 2 + s];
          double* out = (first || w == c.Wb[s]) ? c.Wa[s] : c.Wb[s];
          const bool rid = rvcur[ci] < 0;
          const double* in = first ? (rid ? c.eye : c.RU[s]) : w;
          GemmTask g = mk(c.expk, c.N, 0, in, c.N, 0, out, c.N, c.N);
          g.ks = evp(c, s, l); g.ks_kind = SC_PLAIN;
          if (first && !rid) { g.cs = c.RD[s]; g.cs_kind = SC_PLAIN; }
          g1.push_back(g);
          w = out;
        }
        if ((t + 1) % ns == 0 || t == c.L - 1) bnd.push_back(ci);
      }
      b.gemm(x->sweep_ops, g1, 0);
      b.gemm(x->sweep_ops, g2, fp32);
      if (bnd.empty()) continue;
      for (int ci : bnd) {
        Chain& c = x->ch[ci];
        const int nxt = rvcur[ci] < 0 ? 0 : 1 - rvcur[ci];
        for (int s = 0; s < 2; ++s) {
          q.push_back(QrTask{wcur[ci * 2 + s], c.RU[s], c.RD[s], c.Tq[s], nullptr, c.Tt[s], c.N, ci});
          const double* vin = rvcur[ci] < 0 ? c.eye : c.RV[s][rvcur[ci]];
          g1.push_back(mk(wcur[ci * 2 + s], c.N, 0, vin, c.N, 0, c.RV[s][nxt], c.N, c.N));
        }
        rvcur[ci] = nxt;
        ml.push_back(t == c.L - 1 ? -1 : (t + 1) / ns);
        rvsel.push_back(nxt);
      }
      b.qr(x->sweep_ops, q);
      b.gemm(x->sweep_ops, g1, 0);
      recompute(x->sweep_ops, bnd, ml, rvsel, false, 1, x->opt.time_displaced ? t / ns : -1, t / ns);
      for (int ci : bnd)
        if (t == x->ch[ci].L - 1) rvcur[ci] = -1;
    }
  }
  x->built = true;
}

// ------------------------------------------------------------------------------------------
// executor
static void run_ops(lqmc_ctx* x, const std::vector<Op>& ops, int measure, hipStream_t s, bool profile) {
  const bool fp32 = x->opt.precision != LQMC_FP64;
  for (const Op& op : ops) {
    if (op.meas_only && !measure) continue;
    int pk = -1;
    if (op.kind == OP_FLIP || op.kind == OP_DET) pk = PK_FLIP;
    else if (op.kind == OP_GEMM) pk = PK_GEMM;
    else if (op.kind == OP_QR) pk = PK_QR;
    else if (op.kind == OP_SOLVE || op.kind == OP_TRSM) pk = PK_SOLVE;
    else if (op.kind == OP_KRON) pk = PK_KRON;
    const bool rec = profile && pk == x->prof.kind;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rec) {
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
      (void)hipEventRecord(e0, s);
    }
    const char* d = x->d_desc + op.off;
    switch (op.kind) {
      case OP_GEMM: launch_gemm(op.fp32, op.nmax, reinterpret_cast<const GemmTask*>(d), op.count, op.ntiles, s); break;
      case OP_KRON: launch_kron(op.nmax, op.fp32, reinterpret_cast<const KronTask*>(d), op.count, op.ntiles, x->cus, s); break;
      case OP_QR:
        launch_qr_blocked(reinterpret_cast<const QrTask*>(d), op.count, op.nmax, x->d_state, x->d_qrpub, x->cus, s);
        break;
      case OP_TRSM: launch_trsm(reinterpret_cast<const TrsmTask*>(d), op.count, op.nmax, x->d_state, s); break;
      case OP_SOLVE: launch_solve(reinterpret_cast<const SolveTask*>(d), op.count, op.nmax, x->d_state, s); break;
      case OP_FINISH: launch_finish(op.fp32, reinterpret_cast<const FinishTask*>(d), op.count, op.nmax, x->d_state, s); break;
      case OP_EXPV: launch_expv(x->d_chains, x->nchains, s); break;
      case OP_IDENT: launch_identity(x->d_eye_ptrs, x->d_eye_ns, x->nchains, s); break;
      case OP_DET:
        launch_det_sweep(x->d_chains, x->d_state, x->nchains, measure, x->opt.rng, x->opt.trace, x->det_shmem, s);
        break;
      case OP_FLIP: {
        FlipArgs a{x->d_chains, x->d_state, op.t, measure, x->opt.rng, x->opt.trace, x->opt.delay,
                   x->flip_groups, x->d_bar, x->nchains, x->opt.precision == LQMC_BF16X3 ? 1 : 0,
                   x->flip_yrows};
        launch_flip(fp32, x->opt.delay > 1, a, x->nchains, x->flip_threads, x->flip_shmem, s);
        break;
      }
    }
    if (rec) {
      (void)hipEventRecord(e1, s);
      x->prof.ev.push_back({e0, e1});
      x->prof.work.push_back(op.work);
    }
  }
}

// ------------------------------------------------------------------------------------------
// C ABI
extern "C" {

const char* lqmc_last_error(void) { return g_err.c_str(); }

int lqmc_create(lqmc_ctx** out, int device, int n_chains, const int32_t* n_sites,
                const int32_t* time_steps, const lqmc_options* opt) {
  if (!out || n_chains <= 0 || !n_sites || !time_steps || !opt) return fail(LQMC_EINVAL, "lqmc_create: bad arguments");
  *out = nullptr;
  lqmc_options o = *opt;
  if (o.mode != LQMC_MODE_DQMC && o.mode != LQMC_MODE_REFERENCE && o.mode != LQMC_MODE_DET)
    return fail(LQMC_EINVAL, "unknown mode");
  if (o.mode == LQMC_MODE_DET && (o.precision != LQMC_FP64 || o.delay != 1))
    return fail(LQMC_EINVAL, "determinant mode runs the FP64 rank-1 sweep only");
  if (o.precision != LQMC_FP64 && o.precision != LQMC_FP32 && o.precision != LQMC_BF16X3)
    return fail(LQMC_EINVAL, "unknown precision");
  if (o.precision == LQMC_BF16X3 && (o.mode != LQMC_MODE_DQMC || o.delay < 4))
    return fail(LQMC_EINVAL, "the split-BF16 form needs the Mode D delayed flip (delay >= 4)");
  if (o.rng != LQMC_RNG_REPLAY && o.rng != LQMC_RNG_PHILOX) return fail(LQMC_EINVAL, "unknown rng");
  if (o.delay < 1 || o.delay > 64) return fail(LQMC_EINVAL, "delay must be in 1..64");
  if (o.delay != 1 && o.delay != 4 && o.delay != 8 && o.delay != 16 && o.delay != 32)
    return fail(LQMC_EINVAL, "delay must be 1 (rank-1) or one of 4, 8, 16, 32 (delayed blocks)");
  if (o.n_stab < 1) return fail(LQMC_EINVAL, "n_stab must be >= 1");
  if (o.time_displaced && o.mode != LQMC_MODE_DQMC)
    return fail(LQMC_EINVAL, "time_displaced Green's functions need Mode D (the UDV stabilisation)");
  o.time_displaced = o.time_displaced ? 1 : 0;
  if (o.mode == LQMC_MODE_REFERENCE && (o.precision != LQMC_FP64 || o.delay != 1))
    return fail(LQMC_EINVAL, "reference mode runs the FP64 rank-1 sweep only");
  if (o.obs_capacity <= 0) o.obs_capacity = 4096;
  if (!(o.max_dg == o.max_dg)) return fail(LQMC_EINVAL, "max_dg must be a number");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(LQMC_EINVAL, "device index out of range");
  HIPCHK(hipSetDevice(device));
  auto* x = new lqmc_ctx();
  x->device = device;
  x->nchains = n_chains;
  x->opt = o;
  x->max_dg = o.max_dg > 0.0 ? o.max_dg : (o.max_dg == 0.0 ? 1e-3 : 0.0);
  x->ch.resize(n_chains);
  for (int c = 0; c < n_chains; ++c) {
    if (n_sites[c] < 1 || n_sites[c] > 1024 || time_steps[c] < 1) {
      delete x;
      return fail(LQMC_EINVAL, "chain sizes must satisfy 1 <= N <= 1024, L >= 1");
    }
    if (o.mode == LQMC_MODE_DET && n_sites[c] > 32) {
      delete x;
      return fail(LQMC_EINVAL, "determinant mode supports at most 32 sites per chain");
    }
    x->ch[c].N = n_sites[c];
    x->ch[c].L = time_steps[c];
    x->ch[c].nb = (time_steps[c] + o.n_stab - 1) / o.n_stab;
    x->nmax = std::max(x->nmax, (int)n_sites[c]);
    x->lmax = std::max(x->lmax, (int)time_steps[c]);
  }
  // The flip kernels keep their dynamic LDS within 64 KB per workgroup (the
  // delayed form's rank-k operands live in a per-chain global workspace), which
  // bounds the supported delay for a given thread count.  Only the delayed flip's
  // staged block rows and Z tiles (16-site blocks, N <= 256) go beyond it, where the
  // device's per-workgroup LDS allows (160 KB on gfx950).
  const int maxsh = 65536;
  const bool fp32 = o.precision != LQMC_FP64;
  const int tmax = o.delay > 1 ? flip_delayed_threads_limit(o.delay) : 1024;
  if (o.flip_threads > 0) x->flip_threads = o.flip_threads;
  else x->flip_threads = x->nmax >= 128 ? tmax : 256;
  const int tmin = o.delay > 1 ? 256 : 64;
  if (x->flip_threads % 64 || x->flip_threads > tmax || x->flip_threads < tmin) {
    char buf[128];
    snprintf(buf, sizeof buf, "flip_threads must be a multiple of 64 in [%d, %d] for this delay", tmin, tmax);
    delete x;
    return fail(LQMC_EINVAL, buf);
  }
  if (o.delay > 1) x->flip_shmem = flip_delayed_shmem(x->nmax, o.delay, x->flip_threads, fp32);
  else x->flip_shmem = sizeof(double) * (4 * (size_t)x->nmax + x->flip_threads);
  if ((int)x->flip_shmem > maxsh) {
    char buf[160];
    snprintf(buf, sizeof buf, "flip kernel needs %zu B of LDS (> %d); lower the delay", x->flip_shmem, maxsh);
    delete x;
    return fail(LQMC_EINVAL, buf);
  }
  // Workgroups per chain of the flip (Mode D: the delayed form with blocks of 16 or 32 sites,
  // whole 16-row strips, or the rank-1 form, rows split evenly): all of them must be resident
  // at once, since they wait for each other, so the grid stays within one workgroup per
  // compute unit.
  size_t lds_dev = 0;  // per-workgroup LDS of the device
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
      delete x;
      return fail(LQMC_EHIP, "hipGetDeviceProperties failed");
    }
    const int cus = prop.multiProcessorCount;
    x->cus = cus;
    lds_dev = prop.sharedMemPerBlock;
    const int padded = 8 * ((n_chains + 7) / 8);
    const bool splittable = o.mode == LQMC_MODE_DQMC && (o.delay == 1 || o.delay == 16 || o.delay == 32);
    const int waves = x->flip_threads / 64;
    if (o.flip_groups < 0 || o.flip_groups > 64) {
      delete x;
      return fail(LQMC_EINVAL, "flip_groups must be in 0..64");
    }
    if (o.flip_groups > 1 && !splittable) {
      delete x;
      return fail(LQMC_EINVAL, "flip_groups > 1 needs the Mode D flip with delay 1, 16 or 32");
    }
    if (o.flip_groups > 1 && flip_grid(n_chains, o.flip_groups) > cus) {
      char buf[160];
      snprintf(buf, sizeof buf, "flip_groups %d: %d workgroups exceed the %d compute units", o.flip_groups,
               flip_grid(n_chains, o.flip_groups), cus);
      delete x;
      return fail(LQMC_EINVAL, buf);
    }
    if (o.flip_groups > 0) x->flip_groups = o.flip_groups;
    else if (splittable && o.delay > 1) {
      // at least one 16-row strip of either spin per wavefront
      const int strips = 2 * ((x->nmax + 15) / 16);
      x->flip_groups = std::max(1, std::min(cus / padded, strips / waves));
    } else if (splittable) {
      // rank-1: at least 32 rows of each spin per workgroup
      x->flip_groups = std::max(1, std::min(cus / padded, x->nmax / 32));
    } else x->flip_groups = 1;
    if (o.delay == 1 && x->flip_groups > 1) {
      if (o.flip_threads <= 0) x->flip_threads = 512;
      x->flip_shmem = flip_rank1_split_shmem(x->nmax, x->flip_threads, x->flip_groups);
      if ((int)x->flip_shmem > maxsh) {
        delete x;
        return fail(LQMC_EINVAL, "split rank-1 flip needs more than 64 KB of LDS");
      }
    }
  }
  // block rows and Z tiles of the delayed flip in LDS (k = 16, nmax <= 256), where they fit
  if (o.delay > 1 && flip_delayed_stage_rows(x->nmax, o.delay, x->flip_groups) &&
      x->flip_shmem + flip_delayed_rows_shmem(x->nmax, o.delay, x->flip_groups) <= lds_dev) {
    x->flip_yrows = 1;
    x->flip_shmem += flip_delayed_rows_shmem(x->nmax, o.delay, x->flip_groups);
  }
  if (o.mode == LQMC_MODE_DET) {
    x->det_shmem = det_sweep_shmem(x->nmax, x->lmax);
    if (x->det_shmem > (size_t)maxsh) {
      delete x;
      return fail(LQMC_EINVAL, "determinant sweep needs more LDS than 64 KB; reduce time_steps");
    }
  }
  // The blocked QR keeps one column panel of the largest chain in LDS.
  if (o.mode == LQMC_MODE_DQMC && x->nmax > 256 && qr_blocked_shmem(x->nmax) > (size_t)maxsh) {
    char buf[160];
    snprintf(buf, sizeof buf, "QR kernel needs %zu B of LDS (> %d) for N = %d; reduce the lattice size",
             qr_blocked_shmem(x->nmax), maxsh, x->nmax);
    delete x;
    return fail(LQMC_EINVAL, buf);
  }
  size_t off = 0;
  for (auto& c : x->ch) layout_chain(x, c, nullptr, off);
  x->arena_bytes = align_up(off, 4096);
  hipError_t e = hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&x->done_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&x->aux_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&x->arena, x->arena_bytes);
  if (e == hipSuccess) e = hipMemset(x->arena, 0, x->arena_bytes);
  if (e == hipSuccess) e = hipMalloc(&x->d_chains, sizeof(ChainDev) * n_chains);
  if (e == hipSuccess) e = hipMalloc(&x->d_state, sizeof(ChainState) * n_chains);
  if (e == hipSuccess) e = hipMemset(x->d_state, 0, sizeof(ChainState) * n_chains);
  if (e == hipSuccess) e = hipMalloc(&x->d_eye_ptrs, sizeof(double*) * n_chains);
  if (e == hipSuccess) e = hipMalloc(&x->d_eye_ns, sizeof(int) * n_chains);
  if (e == hipSuccess) e = hipMalloc(&x->d_bar, sizeof(unsigned) * 4 * n_chains);
  if (e == hipSuccess) e = hipMemset(x->d_bar, 0, sizeof(unsigned) * 4 * n_chains);
  if (e == hipSuccess) e = hipMalloc(&x->d_qrpub, sizeof(unsigned) * 4 * n_chains);
  if (e == hipSuccess) e = hipMemset(x->d_qrpub, 0, sizeof(unsigned) * 4 * n_chains);
  if (e != hipSuccess) {
    std::string m = std::string("allocation failed: ") + hipGetErrorString(e);
    lqmc_destroy(x);
    return fail(LQMC_ENOMEM, m);
  }
  off = 0;
  for (auto& c : x->ch) layout_chain(x, c, x->arena, off);
  x->h_chains.resize(n_chains);
  std::vector<double*> eyes(n_chains);
  std::vector<int> ns(n_chains);
  for (int c = 0; c < n_chains; ++c) { eyes[c] = x->ch[c].eye; ns[c] = x->ch[c].N; }
  if (hipMemcpy(x->d_eye_ptrs, eyes.data(), sizeof(double*) * n_chains, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(x->d_eye_ns, ns.data(), sizeof(int) * n_chains, hipMemcpyHostToDevice) != hipSuccess) {
    lqmc_destroy(x);
    return fail(LQMC_EHIP, "descriptor upload failed");
  }
  *out = x;
  return LQMC_OK;
}

static void refresh_chain_desc(lqmc_ctx* x, int ci) {
  Chain& c = x->ch[ci];
  ChainDev d{};
  d.N = c.N; d.L = c.L; d.mode = x->opt.mode;
  d.lam = c.lam;
  d.e1m = std::exp(-c.lam); d.e1p = std::exp(c.lam);
  d.e2p = std::exp(2.0 * c.lam); d.e2m = std::exp(-2.0 * c.lam);
  d.seed = c.seed;
  d.field = c.field; d.ev = c.ev; d.evi = c.evi; d.G = c.G;
  d.Gacc = c.Gacc; d.Glast = c.Glast; d.Sacc = c.Sacc; d.hop = c.hop; d.uni = c.uni;
  d.expk = c.expk;
  d.tr_ratio = c.tr_ratio; d.tr_acc = c.tr_acc; d.obs = c.obs; d.obs_cap = x->opt.obs_capacity;
  d.dws = c.dws;
  x->h_chains[ci] = d;
}

// E (N x N) = Ex (x) Ey, Ex n1 x n1, Ey n2 x n2, N = n1 n2, site i = a n2 + b?  The factors are read
// off the first rows of blocks, Ex[a][a'] = E[a n2][a' n2] and Ey[b][b'] = E[b][b'] / E[0][0] (so
// that Ex[a][a'] Ey[b][b'] = E[a n2][a' n2] E[b][b'] / E[0][0]), and accepted when every entry of
// E matches its Kronecker product to 1e-13 of max|E|: an E_K = expm(-dtau K) of a hopping matrix
// K = Kx (x) I + I (x) Ky does (the two terms commute), any other fails.  f: [Ex | Ey] as 16 x 16
// row-major blocks, zero outside.
static bool kron_factor(const double* E, int N, int n2, double* f) {
  const int n1 = N / n2;
  const double e00 = E[0];
  if (!(e00 > 0.0)) return false;
  double emax = 0.0;
  for (size_t e = 0; e < (size_t)N * N; ++e) emax = std::max(emax, std::fabs(E[e]));
  std::fill(f, f + 512, 0.0);
  for (int a = 0; a < n1; ++a)
    for (int a2 = 0; a2 < n1; ++a2) f[a * 16 + a2] = E[(size_t)a * n2 * N + (size_t)a2 * n2];
  for (int b = 0; b < n2; ++b)
    for (int b2 = 0; b2 < n2; ++b2) f[256 + b * 16 + b2] = E[(size_t)b * N + b2] / e00;
  const double tol = 1e-13 * emax;
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < N; ++k) {
      const double p = f[(i / n2) * 16 + k / n2] * f[256 + (i % n2) * 16 + k % n2];
      if (!(std::fabs(E[(size_t)i * N + k] - p) <= tol)) return false;
    }
  return true;
}

int lqmc_set_chain(lqmc_ctx* x, int ci, const double* expk, const double* expk_inv, const double* kmat,
                   double lambda, const int8_t* field, uint64_t seed) {
  if (!x || ci < 0 || ci >= x->nchains || !expk || !expk_inv || !kmat || !field)
    return fail(LQMC_EINVAL, "lqmc_set_chain: bad arguments");
  HIPCHK(hipSetDevice(x->device));
  Chain& c = x->ch[ci];
  const size_t N = c.N, L = c.L, n2 = N * N;
  c.lam = lambda;
  c.seed = seed;
  std::vector<int8_t> f(N * L);
  for (size_t i = 0; i < N; ++i)
    for (size_t l = 0; l < L; ++l) {
      const int8_t v = field[i * L + l];
      if (v != 1 && v != -1) return fail(LQMC_EINVAL, "field entries must be +1 or -1");
      f[l * N + i] = v;
    }
  std::vector<double> hop(kmat, kmat + n2);
  for (size_t i = 0; i < N; ++i) hop[i * N + i] = 0.0;
  std::vector<double> eye(n2, 0.0), ones(N, 1.0);
  for (size_t i = 0; i < N; ++i) eye[i * N + i] = 1.0;
  HIPCHK(hipMemcpy(c.field, f.data(), N * L, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.expk, expk, n2 * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.expki, expk_inv, n2 * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.hop, hop.data(), n2 * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.eye, eye.data(), n2 * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.ones, ones.data(), N * 8, hipMemcpyHostToDevice));
  // Kronecker factors of E_K and E_K^-1 (the first split N = n1 n2, 2 <= n1, n2 <= 16, largest
  // n2 first, that both admit), unless dense products are asked for
  c.kn1 = c.kn2 = 0;
  if (!x->opt.dense_ek && x->opt.mode != LQMC_MODE_DET) {
    std::vector<double> kf(1024);
    for (int n2c = 16; n2c >= 2; --n2c) {
      if ((int)N % n2c != 0) continue;
      const int n1c = (int)N / n2c;
      if (n1c < 2 || n1c > 16) continue;
      if (kron_factor(expk, (int)N, n2c, kf.data()) && kron_factor(expk_inv, (int)N, n2c, kf.data() + 512)) {
        c.kn1 = n1c;
        c.kn2 = n2c;
        break;
      }
    }
    if (c.kn1) HIPCHK(hipMemcpy(c.kr, kf.data(), 1024 * 8, hipMemcpyHostToDevice));
  }
  if (x->opt.precision != LQMC_FP64) {
    std::vector<float> a(n2), b(n2);
    for (size_t e = 0; e < n2; ++e) { a[e] = (float)expk[e]; b[e] = (float)expk_inv[e]; }
    HIPCHK(hipMemcpy(c.expk32, a.data(), n2 * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.expki32, b.data(), n2 * 4, hipMemcpyHostToDevice));
  }
  c.set = true;
  refresh_chain_desc(x, ci);
  HIPCHK(hipMemcpy(x->d_chains + ci, &x->h_chains[ci], sizeof(ChainDev), hipMemcpyHostToDevice));
  x->inited = false;
  return LQMC_OK;
}

int lqmc_get_ek_split(lqmc_ctx* x, int ci, int32_t* n1, int32_t* n2) {
  if (!x || ci < 0 || ci >= x->nchains || !n1 || !n2) return fail(LQMC_EINVAL, "lqmc_get_ek_split: bad arguments");
  *n1 = x->ch[ci].kn1;
  *n2 = x->ch[ci].kn2;
  return LQMC_OK;
}

int lqmc_set_uniforms(lqmc_ctx* x, int ci, const double* u, int64_t n_sweeps) {
  if (!x || ci < 0 || ci >= x->nchains || !u || n_sweeps < 0) return fail(LQMC_EINVAL, "lqmc_set_uniforms: bad arguments");
  HIPCHK(hipSetDevice(x->device));
  Chain& c = x->ch[ci];
  const int64_t need = n_sweeps * (int64_t)c.N * c.L;
  HIPCHK(hipStreamSynchronize(x->stream));
  if (need > c.uni_cap) {
    if (c.uni) HIPCHK(hipFree(c.uni));
    c.uni = nullptr;
    c.uni_cap = 0;
    HIPCHK(hipMalloc(&c.uni, sizeof(double) * std::max<int64_t>(need, 1)));
    c.uni_cap = need;
    refresh_chain_desc(x, ci);
    HIPCHK(hipMemcpy(x->d_chains + ci, &x->h_chains[ci], sizeof(ChainDev), hipMemcpyHostToDevice));
  }
  if (need) HIPCHK(hipMemcpy(c.uni, u, sizeof(double) * need, hipMemcpyHostToDevice));
  ChainState st;
  HIPCHK(hipMemcpy(&st, x->d_state + ci, sizeof st, hipMemcpyDeviceToHost));
  st.replay_base = st.sweep;
  st.replay_count = (unsigned long long)n_sweeps;
  HIPCHK(hipMemcpy(x->d_state + ci, &st, sizeof st, hipMemcpyHostToDevice));
  return LQMC_OK;
}

static int check_flags(lqmc_ctx* x) {
  std::vector<ChainState> st(x->nchains);
  HIPCHK(hipMemcpy(st.data(), x->d_state, sizeof(ChainState) * x->nchains, hipMemcpyDeviceToHost));
  for (int c = 0; c < x->nchains; ++c)
    if (st[c].flags) {
      char buf[128];
      snprintf(buf, sizeof buf, "device error flags 0x%x on chain %d", st[c].flags, c);
      return fail(LQMC_ENUMERIC, buf);
    }
  return LQMC_OK;
}

static int upload_program(lqmc_ctx* x) {
  if (x->d_desc) { HIPCHK(hipFree(x->d_desc)); x->d_desc = nullptr; }
  x->desc_bytes = std::max<size_t>(x->h_desc.size(), 256);
  HIPCHK(hipMalloc(&x->d_desc, x->desc_bytes));
  if (!x->h_desc.empty()) HIPCHK(hipMemcpy(x->d_desc, x->h_desc.data(), x->h_desc.size(), hipMemcpyHostToDevice));
  return LQMC_OK;
}

static void drop_graphs(lqmc_ctx* x) {
  for (int i = 0; i < 2; ++i) {
    if (x->gexec[i]) (void)hipGraphExecDestroy(x->gexec[i]);
    if (x->graph[i]) (void)hipGraphDestroy(x->graph[i]);
    x->gexec[i] = nullptr;
    x->graph[i] = nullptr;
  }
}

int lqmc_init(lqmc_ctx* x) {
  if (!x) return fail(LQMC_EINVAL, "null context");
  for (auto& c : x->ch) if (!c.set) return fail(LQMC_ESTATE, "lqmc_init before every chain was set");
  HIPCHK(hipSetDevice(x->device));
  drop_graphs(x);
  build_program(x);
  int rc = upload_program(x);
  if (rc) return rc;
  HIPCHK(hipMemcpy(x->d_chains, x->h_chains.data(), sizeof(ChainDev) * x->nchains, hipMemcpyHostToDevice));
  run_ops(x, x->init_ops, 0, x->stream, false);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(x->stream));
  x->inited = true;
  return check_flags(x);
}

int lqmc_use_graph(lqmc_ctx* x, int on) {
  if (!x) return fail(LQMC_EINVAL, "null context");
  x->use_graph = on != 0;
  if (!x->use_graph) drop_graphs(x);
  return LQMC_OK;
}

int lqmc_sweep(lqmc_ctx* x, int n_sweeps, int measure, void* stream) {
  if (!x || n_sweeps < 0) return fail(LQMC_EINVAL, "lqmc_sweep: bad arguments");
  if (!x->inited) return fail(LQMC_ESTATE, "lqmc_sweep before lqmc_init");
  HIPCHK(hipSetDevice(x->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : x->stream;
  const int mi = measure ? 1 : 0;
  const bool profile = x->prof.kind >= 0;
  if (x->use_graph && !profile) {
    if (!x->gexec[mi]) {
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      run_ops(x, x->sweep_ops, mi, s, false);
      HIPCHK(hipStreamEndCapture(s, &x->graph[mi]));
      HIPCHK(hipGraphInstantiate(&x->gexec[mi], x->graph[mi], nullptr, nullptr, 0));
    }
    for (int k = 0; k < n_sweeps; ++k) HIPCHK(hipGraphLaunch(x->gexec[mi], s));
  } else {
    for (int k = 0; k < n_sweeps; ++k) run_ops(x, x->sweep_ops, mi, s, profile);
  }
  HIPCHK(hipGetLastError());
  if (s != x->stream) {
    // later work on the context stream (the getters, set_uniforms, sync) follows this sweep
    HIPCHK(hipEventRecord(x->done_ev, s));
    HIPCHK(hipStreamWaitEvent(x->stream, x->done_ev, 0));
  }
  return LQMC_OK;
}

int lqmc_sync(lqmc_ctx* x) {
  if (!x) return fail(LQMC_EINVAL, "null context");
  HIPCHK(hipSetDevice(x->device));
  HIPCHK(hipStreamSynchronize(x->stream));
  return check_flags(x);
}

static int green_src(lqmc_ctx* x, int ci, int which, std::vector<double>& out) {
  Chain& c = x->ch[ci];
  const size_t n2 = (size_t)c.N * c.N;
  out.resize(2 * n2);
  HIPCHK(hipStreamSynchronize(x->stream));
  if (which == 0) {
    if (x->opt.precision != LQMC_FP64) {
      std::vector<float> f(2 * n2);
      HIPCHK(hipMemcpy(f.data(), c.G, 4 * 2 * n2, hipMemcpyDeviceToHost));
      for (size_t e = 0; e < 2 * n2; ++e) out[e] = f[e];
    } else {
      HIPCHK(hipMemcpy(out.data(), c.G, 8 * 2 * n2, hipMemcpyDeviceToHost));
    }
  } else if (which == 1) {
    ChainState st;
    HIPCHK(hipMemcpy(&st, x->d_state + ci, sizeof st, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(out.data(), c.Gacc, 8 * 2 * n2, hipMemcpyDeviceToHost));
    const double m = st.measured ? (double)st.measured : 1.0;
    for (auto& v : out) v /= m;
  } else if (which == 2) {
    HIPCHK(hipMemcpy(out.data(), c.Glast, 8 * 2 * n2, hipMemcpyDeviceToHost));
  } else {
    return fail(LQMC_EINVAL, "which must be 0, 1 or 2");
  }
  return LQMC_OK;
}

int lqmc_get_green(lqmc_ctx* x, int ci, double* out, int which) {
  if (!x || ci < 0 || ci >= x->nchains || !out) return fail(LQMC_EINVAL, "lqmc_get_green: bad arguments");
  HIPCHK(hipSetDevice(x->device));
  std::vector<double> v;
  int rc = green_src(x, ci, which, v);
  if (rc) return rc;
  std::memcpy(out, v.data(), sizeof(double) * v.size());
  return LQMC_OK;
}

int lqmc_get_green_dev(lqmc_ctx* x, int ci, double* dev_out, int which, void* stream) {
  if (!x || ci < 0 || ci >= x->nchains || !dev_out) return fail(LQMC_EINVAL, "lqmc_get_green_dev: bad arguments");
  if (which < 0 || which > 2) return fail(LQMC_EINVAL, "which must be 0, 1 or 2");
  HIPCHK(hipSetDevice(x->device));
  Chain& c = x->ch[ci];
  const size_t n = 2 * (size_t)c.N * c.N;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : x->stream;
  if (s != x->stream) {  // after the work already queued on the context stream
    HIPCHK(hipEventRecord(x->aux_ev, x->stream));
    HIPCHK(hipStreamWaitEvent(s, x->aux_ev, 0));
  }
  const bool fp32 = which == 0 && x->opt.precision != LQMC_FP64;
  const void* src = which == 0 ? c.G : (which == 1 ? (const void*)c.Gacc : (const void*)c.Glast);
  launch_green_out(dev_out, src, fp32 ? 1 : 0, n, x->d_state + ci, which == 1 ? 1 : 0, s);
  HIPCHK(hipGetLastError());
  return LQMC_OK;
}

int lqmc_get_recompute_err(lqmc_ctx* x, int ci, double* out, int32_t* n_points) {
  if (!x || ci < 0 || ci >= x->nchains) return fail(LQMC_EINVAL, "lqmc_get_recompute_err: bad arguments");
  if (x->opt.mode != LQMC_MODE_DQMC) return fail(LQMC_ESTATE, "recompute deviations exist in Mode D only");
  HIPCHK(hipSetDevice(x->device));
  Chain& c = x->ch[ci];
  if (n_points) *n_points = c.nb;
  if (out) {
    std::vector<unsigned long long> b(c.nb);
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipMemcpy(b.data(), c.rec, sizeof(unsigned long long) * c.nb, hipMemcpyDeviceToHost));
    for (int k = 0; k < c.nb; ++k) std::memcpy(out + k, &b[k], sizeof(double));
  }
  return LQMC_OK;
}

int lqmc_get_spin_corr(lqmc_ctx* x, int ci, double* out) {
  if (!x || ci < 0 || ci >= x->nchains || !out) return fail(LQMC_EINVAL, "lqmc_get_spin_corr: bad arguments");
  HIPCHK(hipSetDevice(x->device));
  Chain& c = x->ch[ci];
  const size_t n2 = (size_t)c.N * c.N;
  ChainState st;
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(&st, x->d_state + ci, sizeof st, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(out, c.Sacc, 8 * n2, hipMemcpyDeviceToHost));
  const double m = st.measured ? (double)st.measured : 1.0;
  for (size_t e = 0; e < n2; ++e) out[e] /= m;
  return LQMC_OK;
}

int lqmc_get_green_tau(lqmc_ctx* x, int ci, double* out, int32_t* l_out, int32_t* ntau) {
  if (!x || ci < 0 || ci >= x->nchains) return fail(LQMC_EINVAL, "lqmc_get_green_tau: bad arguments");
  if (x->opt.mode != LQMC_MODE_DQMC || !x->opt.time_displaced)
    return fail(LQMC_ESTATE, "context created without the time_displaced option (Mode D only)");
  HIPCHK(hipSetDevice(x->device));
  Chain& c = x->ch[ci];
  const int ns = x->opt.n_stab;
  if (ntau) *ntau = c.nb;
  if (l_out)
    for (int k = 0; k < c.nb; ++k) l_out[k] = c.L - 1 - std::min((k + 1) * ns - 1, c.L - 1);
  if (out) {
    const size_t cnt = (size_t)c.nb * 2 * c.N * c.N;
    ChainState st;
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipMemcpy(&st, x->d_state + ci, sizeof st, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(out, c.Gtd, 8 * cnt, hipMemcpyDeviceToHost));
    const double m = st.measured ? (double)st.measured : 1.0;
    for (size_t e = 0; e < cnt; ++e) out[e] /= m;
  }
  return LQMC_OK;
}

int lqmc_get_field(lqmc_ctx* x, int ci, int8_t* out) {
  if (!x || ci < 0 || ci >= x->nchains || !out) return fail(LQMC_EINVAL, "lqmc_get_field: bad arguments");
  HIPCHK(hipSetDevice(x->device));
  Chain& c = x->ch[ci];
  std::vector<int8_t> f((size_t)c.N * c.L);
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(f.data(), c.field, f.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < c.N; ++i)
    for (int l = 0; l < c.L; ++l) out[(size_t)i * c.L + l] = f[(size_t)l * c.N + i];
  return LQMC_OK;
}

int lqmc_get_stats(lqmc_ctx* x, int ci, lqmc_stats* out) {
  if (!x || ci < -1 || ci >= x->nchains || !out) return fail(LQMC_EINVAL, "lqmc_get_stats: bad arguments");
  HIPCHK(hipSetDevice(x->device));
  std::vector<ChainState> st(x->nchains);
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(st.data(), x->d_state, sizeof(ChainState) * x->nchains, hipMemcpyDeviceToHost));
  lqmc_stats r{};
  for (int c = 0; c < x->nchains; ++c) {
    if (ci >= 0 && c != ci) continue;
    r.attempted += st[c].attempted;
    r.accepted += st[c].accepted;
    r.sweeps = std::max<uint64_t>(r.sweeps, st[c].sweep);
    r.measured = std::max<uint64_t>(r.measured, st[c].measured);
    r.recomputes += st[c].recomputes;
    double e;
    std::memcpy(&e, &st[c].max_err_bits, sizeof e);
    r.max_err = std::max(r.max_err, e);
    if (std::isnan(e)) r.max_err = e;
    r.flags |= st[c].flags;
  }
  *out = r;
  return LQMC_OK;
}

int lqmc_reset_stats(lqmc_ctx* x) {
  if (!x) return fail(LQMC_EINVAL, "null context");
  HIPCHK(hipSetDevice(x->device));
  std::vector<ChainState> st(x->nchains);
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(st.data(), x->d_state, sizeof(ChainState) * x->nchains, hipMemcpyDeviceToHost));
  for (auto& s : st) {
    s.attempted = s.accepted = s.recomputes = 0;
    s.max_err_bits = 0;
    s.flags = 0;
  }
  HIPCHK(hipMemcpy(x->d_state, st.data(), sizeof(ChainState) * x->nchains, hipMemcpyHostToDevice));
  for (auto& c : x->ch)
    if (c.rec) HIPCHK(hipMemset(c.rec, 0, sizeof(unsigned long long) * c.nb));
  return LQMC_OK;
}

int lqmc_get_trace(lqmc_ctx* x, int ci, double* ratio, uint8_t* accept) {
  if (!x || ci < 0 || ci >= x->nchains) return fail(LQMC_EINVAL, "lqmc_get_trace: bad arguments");
  if (!x->opt.trace) return fail(LQMC_ESTATE, "context created without trace");
  HIPCHK(hipSetDevice(x->device));
  Chain& c = x->ch[ci];
  const size_t nl = (size_t)c.N * c.L;
  HIPCHK(hipStreamSynchronize(x->stream));
  if (ratio) HIPCHK(hipMemcpy(ratio, c.tr_ratio, nl * 8, hipMemcpyDeviceToHost));
  if (accept) HIPCHK(hipMemcpy(accept, c.tr_acc, nl, hipMemcpyDeviceToHost));
  return LQMC_OK;
}

int lqmc_get_observables(lqmc_ctx* x, int ci, double* out, int max_rows, int* rows) {
  if (!x || ci < 0 || ci >= x->nchains || !out || !rows || max_rows < 0)
    return fail(LQMC_EINVAL, "lqmc_get_observables: bad arguments");
  HIPCHK(hipSetDevice(x->device));
  Chain& c = x->ch[ci];
  ChainState st;
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(&st, x->d_state + ci, sizeof st, hipMemcpyDeviceToHost));
  const uint64_t cap = (uint64_t)x->opt.obs_capacity;
  const uint64_t have = std::min<uint64_t>(st.measured, cap);
  const uint64_t n = std::min<uint64_t>(have, (uint64_t)max_rows);
  std::vector<double> ring(3 * cap);
  HIPCHK(hipMemcpy(ring.data(), c.obs, 8 * 3 * cap, hipMemcpyDeviceToHost));
  const uint64_t first = st.measured - n;
  for (uint64_t r = 0; r < n; ++r) {
    const uint64_t k = (first + r) % cap;
    for (int q = 0; q < 3; ++q) out[r * 3 + q] = ring[k * 3 + q];
  }
  *rows = (int)n;
  return LQMC_OK;
}

int lqmc_profile(lqmc_ctx* x, int kind) {
  if (!x || kind < -1 || kind > 4) return fail(LQMC_EINVAL, "lqmc_profile: bad kind");
  HIPCHK(hipStreamSynchronize(x->stream));
  for (auto& p : x->prof.ev) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
  x->prof.ev.clear();
  x->prof.work.clear();
  x->prof.kind = kind;
  return LQMC_OK;
}

int lqmc_profile_result(lqmc_ctx* x, double* avg_ms, int64_t* launches, double* work_units) {
  if (!x || !avg_ms || !launches || !work_units) return fail(LQMC_EINVAL, "lqmc_profile_result: bad arguments");
  HIPCHK(hipStreamSynchronize(x->stream));
  double tot = 0, work = 0;
  for (size_t k = 0; k < x->prof.ev.size(); ++k) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, x->prof.ev[k].first, x->prof.ev[k].second));
    tot += ms;
    work += x->prof.work[k];
  }
  *launches = (int64_t)x->prof.ev.size();
  *avg_ms = x->prof.ev.empty() ? 0.0 : tot / x->prof.ev.size();
  *work_units = work;
  return LQMC_OK;
}

int lqmc_destroy(lqmc_ctx* x) {
  if (!x) return LQMC_OK;
  (void)hipSetDevice(x->device);
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  drop_graphs(x);
  for (auto& p : x->prof.ev) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
  for (auto& c : x->ch) if (c.uni) (void)hipFree(c.uni);
  if (x->d_desc) (void)hipFree(x->d_desc);
  if (x->arena) (void)hipFree(x->arena);
  if (x->d_chains) (void)hipFree(x->d_chains);
  if (x->d_state) (void)hipFree(x->d_state);
  if (x->d_eye_ptrs) (void)hipFree(x->d_eye_ptrs);
  if (x->d_eye_ns) (void)hipFree(x->d_eye_ns);
  if (x->d_bar) (void)hipFree(x->d_bar);
  if (x->d_qrpub) (void)hipFree(x->d_qrpub);
  if (x->done_ev) (void)hipEventDestroy(x->done_ev);
  if (x->aux_ev) (void)hipEventDestroy(x->aux_ev);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x;
  return LQMC_OK;
}

}  // extern "C"

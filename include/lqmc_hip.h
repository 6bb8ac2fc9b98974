/*
 * lqmc_hip.h -- C ABI of the MI355X (gfx950) determinant-QMC sweep engine.
 *
 * The reference (KieDani/LatticeQMC) has no native API: its boundary is the
 * Python class LatticeQMC (lqmc/lqmc.py:14-408).  Each entry point below
 * replaces one piece of that class, cited per declaration.  The Python drop-in
 * latticeqmc_amd.LatticeQMC binds these through ctypes.
 *
 * Conventions
 *  - Every entry returns int: 0 on success, a negative LQMC_E* code on error.
 *    Nothing throws or aborts across the ABI; lqmc_last_error() returns a
 *    thread-local message for the last failing call on the calling thread.
 *  - Host pointers are borrowed for the duration of the call and copied.
 *  - A context owns all device workspaces (allocated by lqmc_create /
 *    lqmc_set_chain); sweeps allocate nothing.  A context is bound to one
 *    device and one HIP stream and is not thread-safe; separate contexts may
 *    be driven from separate threads.
 *  - Fields use the reference layout: int8 +-1, shape (N, L) row-major,
 *    element [i*L + l] = config[i, l] (lqmc/configuration.py:89,123-124).
 *  - Spin 0 = up (V = exp(-lambda*s)), spin 1 = down (lqmc/lqmc.py:149).
 */
#ifndef LQMC_HIP_H
#define LQMC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lqmc_ctx lqmc_ctx;

enum {
  LQMC_OK = 0,
  LQMC_EINVAL = -1,    /* bad argument */
  LQMC_EHIP = -2,      /* HIP runtime error */
  LQMC_ENOMEM = -3,    /* device allocation failed */
  LQMC_ESTATE = -4,    /* call out of order (e.g. sweep before set_chain) */
  LQMC_ENUMERIC = -5,  /* device flagged NaN/Inf, zero pivot or replay underrun */
};

enum { LQMC_MODE_DQMC = 0, LQMC_MODE_REFERENCE = 1, LQMC_MODE_DET = 2 };
enum { LQMC_FP64 = 0, LQMC_FP32 = 1, LQMC_BF16X3 = 2 };
enum { LQMC_RNG_REPLAY = 0, LQMC_RNG_PHILOX = 1 };

/* Device-side error flag bits reported in lqmc_stats.flags. */
enum {
  LQMC_FLAG_NAN_RATIO = 1,      /* a Metropolis ratio was NaN/Inf */
  LQMC_FLAG_ZERO_PIVOT = 2,     /* QR or LU met a zero pivot */
  LQMC_FLAG_REPLAY_UNDERRUN = 4,/* a replay sweep ran past the uploaded uniforms */
  LQMC_FLAG_NAN_G = 8,          /* a recompute found NaN/Inf in G */
  LQMC_FLAG_FLIP_BARRIER = 16,  /* a wait between the workgroups of one chain's flip timed out */
  LQMC_FLAG_PRECISION = 32,     /* a recompute found max|G_propagated - G_recomputed| > max_dg */
};

typedef struct {
  int32_t mode;        /* LQMC_MODE_DQMC (textbook, stabilised), LQMC_MODE_REFERENCE, or
                          LQMC_MODE_DET: the reference's brute-force determinant sweep
                          (lqmc/lqmc.py:236-299), FP64 rank-1 only, N <= 32 */
  int32_t precision;   /* LQMC_FP64 or LQMC_FP32: storage of G, wraps and flip updates;
                          LQMC_BF16X3: the FP32 form with the delayed rank-k products G += P Z
                          as three BF16 matrix-core products with FP32 accumulation
                          (P_hi Z_hi + P_hi Z_lo + P_lo Z_hi, x = x_hi + x_lo split in BF16),
                          Mode D, delay >= 4 */
  int32_t delay;       /* 1: rank-1 update per accepted flip; k>1: submatrix blocks of k sites */
  int32_t n_stab;      /* slices between FP64 QR/UDV recomputes (Mode D) */
  int32_t rng;         /* LQMC_RNG_REPLAY or LQMC_RNG_PHILOX */
  int32_t trace;       /* 1: keep per-attempt ratio/accept of the last sweep */
  int32_t obs_capacity;/* per-chain ring of per-sweep observables (0 = 4096) */
  int32_t flip_threads;/* threads per flip workgroup (0 = automatic) */
  int32_t flip_groups; /* workgroups per chain in the delayed flip, delay 16 or 32 only
                          (0 = automatic: fill the compute units; 1 = one per chain) */
  int32_t time_displaced; /* Mode D: 1 accumulates G(tau, 0) at every stabilisation point of
                             the measured sweeps (lqmc_get_green_tau) */
  double max_dg;       /* Mode D: a recompute whose max|G_propagated - G_recomputed| exceeds this
                          raises LQMC_FLAG_PRECISION, so lqmc_sync fails (0 = 1e-3; < 0 = off).
                          The Metropolis ratios between recomputes read the propagated G, so a
                          large deviation (an FP32 form at low temperature) biases the chain. */
  int32_t dense_ek;    /* 1: every product with E_K or E_K^-1 as a dense GEMM.  0 (default): where
                          E_K = Ex (x) Ey and E_K^-1 = Exi (x) Eyi exactly (a hopping matrix that is
                          a Kronecker sum, e.g. the square and cubic lattices; checked entry by
                          entry to 1e-13 of max|E_K| in lqmc_set_chain, n1, n2 <= 16), the FP64
                          products apply the factors: 2 N^2 (n1 + n2) flops in place of 2 N^3,
                          equal to the dense product up to rounding. */
} lqmc_options;

typedef struct {
  uint64_t attempted;   /* attempted flips since create */
  uint64_t accepted;    /* accepted flips since create */
  uint64_t sweeps;      /* completed sweeps */
  uint64_t measured;    /* measured sweeps accumulated in the mean G */
  uint64_t recomputes;  /* FP64 recomputes performed */
  double max_err;       /* max |G_propagated - G_recomputed| since create/reset */
  uint32_t flags;       /* LQMC_FLAG_* bits */
  uint32_t pad;
} lqmc_stats;

/* Replaces LatticeQMC.__init__ (lqmc/lqmc.py:16-79) for n_chains independent
 * chains of sizes N[c] sites x L[c] slices (the batched launch also replaces the
 * chain pool of lqmc/multiprocessing.py:252-289). */
int lqmc_create(lqmc_ctx** ctx, int device, int n_chains, const int32_t* n_sites,
                const int32_t* time_steps, const lqmc_options* opt);

/* Replaces set_beta (lqmc/lqmc.py:93-117) and the Configuration passed in at
 * construction (lqmc/configuration.py:91-124) for chain c: E_K = expm(-dtau K),
 * E_K^-1 = expm(+dtau K), lambda = arccosh(exp(U dtau/2)), the hopping matrix K
 * (N x N, used by the kinetic-energy estimator), the initial field and the
 * Philox seed of the chain. */
int lqmc_set_chain(lqmc_ctx* ctx, int c, const double* expk, const double* expk_inv,
                   const double* kmat, double lambda, const int8_t* field, uint64_t seed);

/* Computes the initial Green's functions: Mode D the stabilised G(L-1)
 * (what _update_step's first inversion provides, lqmc/lqmc.py:303-307). */
int lqmc_init(lqmc_ctx* ctx);

/* The Kronecker split N = n1 x n2 used for chain c's E_K products (after lqmc_set_chain), or
 * 0 x 0 where they run as dense GEMMs.  No counterpart in the reference, which multiplies
 * by the dense expm(-dtau K) (lqmc/lqmc.py:93-117); it reports which path the products take. */
int lqmc_get_ek_split(lqmc_ctx* ctx, int c, int32_t* n1, int32_t* n2);

/* Uniforms for replay mode (the np.random.rand() draws of lqmc/lqmc.py:317):
 * u[s*L*N + t*N + i] for sweeps s = 0..n_sweeps-1 starting at the chain's next
 * sweep, slice l = L-1-t, site i. */
int lqmc_set_uniforms(lqmc_ctx* ctx, int c, const double* u, int64_t n_sweeps);

/* Runs n_sweeps sweeps of every chain (_update_step, lqmc/lqmc.py:301-347, and
 * the loops warmup_loop/measure_loop 349-375). measure != 0 accumulates the
 * end-of-sweep G (lqmc/lqmc.py:371-372) and per-sweep observables.  Asynchronous
 * on the context stream; stream may be NULL (context stream) or a hipStream_t.  On a
 * caller stream the context stream is made to wait for the sweep's end, so every later
 * call on the context (getters, lqmc_set_uniforms, lqmc_sync) sees its results. */
int lqmc_sweep(lqmc_ctx* ctx, int n_sweeps, int measure, void* stream);

/* Waits for the context stream (only: other work on the device is not waited for) and
 * checks the device error flags. */
int lqmc_sync(lqmc_ctx* ctx);

/* Enables (1) or disables (0) replaying whole sweeps from captured hipGraphs. */
int lqmc_use_graph(lqmc_ctx* ctx, int on);

/* Green's functions of chain c into out[2*N*N] (float64, [spin][i][j]).
 * which = 0: current G; 1: mean of measured end-of-sweep G (the return value of
 * measure_loop, lqmc/lqmc.py:375); 2: G of the last measured sweep. */
int lqmc_get_green(lqmc_ctx* ctx, int c, double* out, int which);

/* Same, into device memory (e.g. a torch tensor's data_ptr()): a device-to-device copy
 * (converted to float64, divided by the measured-sweep count for which = 1) issued on
 * stream (NULL: the context stream) after the work queued on the context stream.
 * Asynchronous: dev_out is complete once that stream is synchronised. */
int lqmc_get_green_dev(lqmc_ctx* ctx, int c, double* dev_out, int which, void* stream);

/* Mode D: max |G_propagated - G_recomputed| at each recompute point of the sweep since
 * create / lqmc_reset_stats, out[k] for the k-th point (after slices L-1 .. L-(k+1)n_s, the
 * last after slice 0); *n_points (may be NULL) receives ceil(L/n_s), out may be NULL.  The
 * per-point form of lqmc_stats.max_err: how far the propagation drifts from the FP64
 * stabilised G (lqmc/lqmc.py:303-307 recomputes G once per sweep without this check). */
int lqmc_get_recompute_err(lqmc_ctx* ctx, int c, double* out, int32_t* n_points);

/* Mean over the measured sweeps of the equal-time spin correlation <S^z_i S^z_j>
 * (S^z = n_up - n_dn, out N x N row-major), estimated from the end-of-sweep G of each
 * measured sweep by Wick's theorem.  Extends the measurement of
 * lqmc/lqmc.py:371-375 (which accumulates G only); the reference derives its local
 * moment from the averaged G instead (lqmc/tools.py:87-109). */
int lqmc_get_spin_corr(lqmc_ctx* ctx, int c, double* out);

/* Unequal-time Green's functions G(tau, 0) = (B(tau,0)^-1 + B(beta,tau))^-1 of chain c, the
 * mean over the measured sweeps (time_displaced option, Mode D).  They are taken at the
 * stabilisation points of the sweep, tau = l dtau for l = L-n_s, L-2n_s, ..., 0 (the
 * slices below the point form B(tau,0)), from the same UDV factors and the same QR as the
 * stable equal-time G: G(tau,0) = U_R Db_R^-1 X^-1 Ds_L V_L.  ntau (may be NULL) receives
 * ceil(L/n_s), l_out[ntau] (may be NULL) the slice counts l, out[ntau*2*N*N] (may be NULL;
 * float64, [tau][spin][i][j]) the means.  No counterpart in lqmc/lqmc.py, whose only
 * imaginary-time function is the free-fermion one of lqmc/tools.py:182-200; at U = 0 this
 * is -compute_pole_gf_tau at those tau. */
int lqmc_get_green_tau(lqmc_ctx* ctx, int c, double* out, int32_t* l_out, int32_t* ntau);

/* Current field of chain c into out[N*L] in the reference layout. */
int lqmc_get_field(lqmc_ctx* ctx, int c, int8_t* out);

/* Counters and error flags of chain c (c = -1: summed/maxed over chains). */
int lqmc_get_stats(lqmc_ctx* ctx, int c, lqmc_stats* out);
int lqmc_reset_stats(lqmc_ctx* ctx);

/* Per-attempt trace of the last sweep (trace option): ratio[L*N] and accept[L*N]
 * in attempt order t*N + i (the values _debug logs, lqmc/lqmc.py:217-232). */
int lqmc_get_trace(lqmc_ctx* ctx, int c, double* ratio, uint8_t* accept);

/* Per-sweep observables of the measured sweeps of chain c, oldest first, up to
 * max_rows rows of (density, double occupancy, kinetic energy per site);
 * *rows receives the count. */
int lqmc_get_observables(lqmc_ctx* ctx, int c, double* out, int max_rows, int* rows);

/* Average device duration (ms) of the kernels of one kind over the sweeps run
 * since lqmc_profile(ctx, kind) was called, measured with HIP events on the
 * context stream (graphs are bypassed while profiling).  kind: 0 = flip,
 * 1 = GEMM, 2 = QR, 3 = solve, 4 = Kronecker-factored E_K products; kind = -1
 * stops profiling. */
int lqmc_profile(lqmc_ctx* ctx, int kind);
int lqmc_profile_result(lqmc_ctx* ctx, double* avg_ms, int64_t* launches, double* work_units);

/* Slice wrap of Green's functions held by the caller, outside a context: for m = 0 ..
 * count-1, G_m <- B_m^-1 G_m B_m with B_m = diag(v_m) E_K, the wrap of _update_step
 * (lqmc/lqmc.py:338-345) in the consistent form the Mode D sweep applies after each slice
 * (the reference's own W = V(l-1) E_K wraps with the wrong slice and direction).  All
 * pointers are device memory: expk = E_K and expk_inv = E_K^-1 (n x n, row-major), v
 * (count x n: the diagonal of V for each matrix, e.g. exp(-sigma lambda s[:, l]) of both
 * spins), g (count x n x n, in/out), and workspace of at least
 * lqmc_wrap_green_workspace(n, count) bytes (descriptors and one n x n temporary per
 * matrix).  Two batched FP64 matrix-core products; asynchronous on stream (NULL: the
 * null stream), nothing is copied from the host. */
size_t lqmc_wrap_green_workspace(int n, int count);
int lqmc_wrap_green(int device, int n, int count, const double* expk, const double* expk_inv, const double* v,
                    double* g, void* workspace, size_t ws_bytes, void* stream);

/* lqmc_wrap_green without the workspace argument: the same kernels on a static device
 * workspace of 16 MiB per GPU (allocated with the library's code object, held until the
 * process ends), so the call allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Limit: lqmc_wrap_green_workspace(n, count) <=
 * 16 MiB (e.g. count <= 31 at n = 256); larger calls return LQMC_EINVAL.  All calls on one
 * GPU share its workspace: calls on one stream are ordered, calls that overlap on one GPU
 * (other streams or threads) race. */
int lqmc_wrap_green_static(int device, int n, int count, const double* expk, const double* expk_inv,
                           const double* v, double* g, void* stream);

const char* lqmc_last_error(void);
int lqmc_destroy(lqmc_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* LQMC_HIP_H */

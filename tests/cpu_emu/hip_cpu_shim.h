// CPU emulation of the small HIP subset the kernels use, for bounds checking
// under AddressSanitizer: one OS thread per GPU thread, __syncthreads() as a
// block barrier, the 16x16x4 matrix-core ops emulated across the 64 threads of
// a wavefront with the lane layout measured on gfx950 (inputs: lane l holds
// A[l%16][l/16] and B[l/16][l%16]; outputs: column l%16, row l/16 + 4r for
// FP64 and 4*(l/16) + r for FP32).  emu_launch runs blocks one after another;
// emu_launch_concurrent runs every block at once (kernels whose workgroups wait
// for each other).
#pragma once
#include <barrier>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#define LQMC_CPU_EMU 1
#define __global__
#define __device__
#define __host__
#define __launch_bounds__(...)
#define __shared__ static
#define __restrict__

struct emu_dim3 { unsigned x = 1, y = 1, z = 1; };
inline thread_local emu_dim3 threadIdx, blockIdx;
inline emu_dim3 blockDim, gridDim;
inline thread_local char* emu_lds_buf = nullptr;
inline thread_local std::barrier<>* emu_block_barrier = nullptr;
#define LQMC_DYNAMIC_LDS(type, name) type* name = reinterpret_cast<type*>(emu_lds_buf)
#define LQMC_GLOBAL
#define LQMC_OPAQUE(x, reg) ((void)(x))
inline void __syncthreads() { emu_block_barrier->arrive_and_wait(); }

using std::fabs; using std::fmax; using std::fmin; using std::isfinite; using std::max; using std::min; using std::sqrt;

typedef int hipError_t;
typedef void* hipStream_t;
constexpr int hipSuccess = 0;
constexpr int hipFuncAttributeMaxDynamicSharedMemorySize = 8;
inline int hipFuncSetAttribute(const void*, int, int) { return 0; }
#define hipLaunchKernelGGL(...) ((void)0)

inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicOr(unsigned* p, unsigned v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicExch(unsigned* p, unsigned v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
#define __HIP_MEMORY_SCOPE_AGENT 0
// polled loads acquire, so the host thread sanitizer sees the barrier's ordering
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), __ATOMIC_ACQUIRE)
// a poll waits 100 us, so the kernels' bounded waits allow minutes of lag between emulated blocks
#define __builtin_amdgcn_s_sleep(n) std::this_thread::sleep_for(std::chrono::microseconds(100))
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
  unsigned long long cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
  return cur;
}
inline uint32_t __umulhi(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
inline long long __double_as_longlong(double d) { long long r; std::memcpy(&r, &d, 8); return r; }
inline double __longlong_as_double(long long v) { double r; std::memcpy(&r, &v, 8); return r; }

struct EmuWave { double a[64], b[64]; short a4[64][4], b4[64][4]; std::barrier<>* bar; };
inline thread_local EmuWave* emu_waves = nullptr;

template <typename T, typename V>
inline V emu_mfma(T a, T b, V c, bool fp64) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  EmuWave& s = emu_waves[w];
  s.a[lane] = a;
  s.b[lane] = b;
  s.bar->arrive_and_wait();
  V d;
  for (int r = 0; r < 4; ++r) {
    const int row = fp64 ? (lane >> 4) + 4 * r : ((lane >> 4) << 2) + r;
    const int col = lane & 15;
    double acc = 0.0;
    for (int k = 0; k < 4; ++k) acc += s.a[row + 16 * k] * s.b[col + 16 * k];
    d[r] = (T)((double)c[r] + acc);
  }
  s.bar->arrive_and_wait();
  return d;
}
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) emu_mfma<double>((a), (b), (c), true)
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, x, y, z) emu_mfma<float>((a), (b), (c), false)

// 16x16x16 BF16 product with FP32 accumulation: lane l holds A[l%16][4(l/16) + j] and
// B[4(l/16) + j][l%16] (j = 0..3, BF16 bit patterns); D as the FP32 16x16x4 product.
template <typename S4, typename V>
inline V emu_mfma_bf16(S4 a, S4 b, V c) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  EmuWave& s = emu_waves[w];
  for (int j = 0; j < 4; ++j) { s.a4[lane][j] = a[j]; s.b4[lane][j] = b[j]; }
  s.bar->arrive_and_wait();
  auto f = [](short h) { unsigned u = (unsigned)(unsigned short)h << 16; float x; std::memcpy(&x, &u, 4); return x; };
  V d;
  for (int r = 0; r < 4; ++r) {
    const int row = ((lane >> 4) << 2) + r, col = lane & 15;
    float acc = 0.0f;
    for (int k = 0; k < 16; ++k) acc += f(s.a4[row + 16 * (k >> 2)][k & 3]) * f(s.b4[col + 16 * (k >> 2)][k & 3]);
    d[r] = c[r] + acc;
  }
  s.bar->arrive_and_wait();
  return d;
}
#define __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, x, y, z) emu_mfma_bf16((a), (b), (c))
inline unsigned __float_as_uint(float x) { unsigned u; std::memcpy(&u, &x, 4); return u; }
inline float __uint_as_float(unsigned u) { float x; std::memcpy(&x, &u, 4); return x; }

// Wavefront-level ordering: all 64 lanes of the calling wave meet.
inline void emu_wave_sync() { emu_waves[threadIdx.x >> 6].bar->arrive_and_wait(); }
#define __builtin_amdgcn_wave_barrier() emu_wave_sync()
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define __builtin_amdgcn_sched_barrier(mask) ((void)0)
#define __builtin_amdgcn_s_waitcnt(mask) ((void)0)
// Hardware register reads: XCC_ID (id 20) as the round-robin placement of block b on XCC b mod 8.
inline int emu_getreg(int enc) { return (enc & 63) == 20 ? (int)(blockIdx.x % 8) : 0; }
#define __builtin_amdgcn_s_getreg(enc) emu_getreg(enc)

// Cross-lane exchange within a wavefront.
inline double __shfl_xor(double v, int mask, int width = 64) {
  (void)width;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  EmuWave& s = emu_waves[w];
  s.a[lane] = v;
  s.bar->arrive_and_wait();
  const double r = s.a[lane ^ mask];
  s.bar->arrive_and_wait();
  return r;
}

// Lane broadcast/gather within a wavefront: returns var of lane src.
inline double __shfl(double v, int src, int width = 64) {
  (void)width;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  EmuWave& s = emu_waves[w];
  s.a[lane] = v;
  s.bar->arrive_and_wait();
  const double r = s.a[src & 63];
  s.bar->arrive_and_wait();
  return r;
}

// gfx950 lane-half swaps, returning {x', y'}: permlane32 swaps lanes 32..63 of x with lanes 0..31
// of y; permlane16 swaps the odd 16-lane rows of x with the even rows of y.
struct EmuU2 {
  unsigned v[2];
  unsigned operator[](int i) const { return v[i]; }
};
inline EmuU2 emu_permlane_swap(unsigned x, unsigned y, int half) {
  const int lane = threadIdx.x & 63;
  const unsigned px = (unsigned)__shfl((double)x, lane ^ half);
  const unsigned py = (unsigned)__shfl((double)y, lane ^ half);
  const bool upper = (lane & half) != 0;
  return EmuU2{{upper ? py : x, upper ? y : px}};
}
#define __builtin_amdgcn_permlane32_swap(x, y, fi, bc) emu_permlane_swap((x), (y), 32)
#define __builtin_amdgcn_permlane16_swap(x, y, fi, bc) emu_permlane_swap((x), (y), 16)

// DPP lane exchange within 16-lane rows for the controls the kernels use: quad_perm
// (ctrl < 0x100) and row_ror:n (0x121..0x12F).
inline int emu_mov_dpp(int v, int ctrl) {
  const int lane = threadIdx.x & 63;
  int src = lane;
  if (ctrl < 0x100) src = (lane & ~3) | ((ctrl >> (2 * (lane & 3))) & 3);
  else if (ctrl >= 0x121 && ctrl <= 0x12F) src = (lane & ~15) | ((lane - (ctrl & 15)) & 15);
  else std::abort();
  return (int)(long long)__shfl((double)v, src);
}
#define __builtin_amdgcn_mov_dpp(v, ctrl, rm, bm, bc) emu_mov_dpp((v), (ctrl))

// Scalar broadcast of lane l's value (l uniform across the wavefront).
inline int emu_readlane_i32(int v, int l) { return (int)(long long)__shfl((double)v, l); }
#define __builtin_amdgcn_readlane(v, l) emu_readlane_i32((v), (l))

// Per-block emulation state: LDS, the block barrier and the wavefront exchange areas.
struct EmuBlock {
  char* lds;
  std::barrier<> bar;
  std::vector<EmuWave> waves;
  EmuBlock(int threads, size_t lds_bytes) : bar(threads), waves((threads + 63) / 64) {
    lds = static_cast<char*>(std::malloc(lds_bytes ? lds_bytes : 1));
    std::memset(lds, 0xFF, lds_bytes);  // NaN as doubles: a read of unwritten LDS poisons the result
    for (auto& w : waves) w.bar = new std::barrier<>(64);
  }
  ~EmuBlock() {
    for (auto& w : waves) delete w.bar;
    std::free(lds);
  }
};

template <typename F>
inline void emu_run_blocks(int b_begin, int b_end, int threads, size_t lds_bytes, F& fn) {
  std::vector<EmuBlock*> blk;
  for (int b = b_begin; b < b_end; ++b) blk.push_back(new EmuBlock(threads, lds_bytes));
  std::vector<std::thread> ts;
  for (int b = b_begin; b < b_end; ++b)
    for (int t = 0; t < threads; ++t)
      ts.emplace_back([&, t, b] {
        EmuBlock* e = blk[b - b_begin];
        threadIdx.x = t;
        blockIdx.x = b;
        emu_lds_buf = e->lds;
        emu_block_barrier = &e->bar;
        emu_waves = e->waves.data();
        fn();
      });
  for (auto& t : ts) t.join();
  for (auto* e : blk) delete e;
}

// Run fn() as a block of `threads` threads for every block index in [0, blocks), one block
// after another.
template <typename F>
inline void emu_launch(int blocks, int threads, size_t lds_bytes, F fn) {
  blockDim.x = threads;
  gridDim.x = blocks;
  for (int b = 0; b < blocks; ++b) emu_run_blocks(b, b + 1, threads, lds_bytes, fn);
}

// The blocks with the given indices, all at the same time (a grid whose other blocks return
// at once).
template <typename F>
inline void emu_launch_ids(const std::vector<int>& ids, int grid, int threads, size_t lds_bytes, F fn) {
  blockDim.x = threads;
  gridDim.x = grid;
  std::vector<EmuBlock*> blk;
  for (size_t k = 0; k < ids.size(); ++k) blk.push_back(new EmuBlock(threads, lds_bytes));
  std::vector<std::thread> ts;
  for (size_t k = 0; k < ids.size(); ++k)
    for (int t = 0; t < threads; ++t)
      ts.emplace_back([&, t, k] {
        EmuBlock* e = blk[k];
        threadIdx.x = t;
        blockIdx.x = ids[k];
        emu_lds_buf = e->lds;
        emu_block_barrier = &e->bar;
        emu_waves = e->waves.data();
        fn();
      });
  for (auto& t : ts) t.join();
  for (auto* e : blk) delete e;
}

// As emu_launch, with every block running at the same time.
template <typename F>
inline void emu_launch_concurrent(int blocks, int threads, size_t lds_bytes, F fn) {
  blockDim.x = threads;
  gridDim.x = blocks;
  emu_run_blocks(0, blocks, threads, lds_bytes, fn);
}

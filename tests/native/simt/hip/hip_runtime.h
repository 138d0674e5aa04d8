// CPU SIMT emulation of the HIP subset used by radar_sounder_simulator_amd/
// csrc/ss_kernels.hip -- test infrastructure only.
//
// g++ compiles the kernel source unchanged against this header (it sits on
// the include path as <hip/hip_runtime.h>).  A launch runs its workgroups
// one after another; each workgroup runs as blockDim host threads, one per
// lane, so __syncthreads, wavefront shuffles/ballots and atomics have their
// device meaning:
//   * __syncthreads: a barrier over the lanes that have not returned (an
//     exited lane no longer takes part, as an ended wave on the device);
//   * __shfl_xor / __ballot: exchanged through a per-wave slot array between
//     two wave barriers (all live lanes of the wave must call them);
//   * atomics: compare-and-swap loops on the host words.
// A barrier or wave operation that waits longer than SIM_TIMEOUT_S seconds
// (default 30), or a launch that outlives it, prints where it is stuck and
// aborts the process: a divergent barrier or a loop that never ends on the
// device shows up here as a report instead of a hang.  Device pointers are
// host pointers, so AddressSanitizer sees every out-of-bounds access.
#pragma once

#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>
#include <algorithm>
#include <atomic>

#define __host__
#define __device__
#define __global__
#define __shared__ static
#define __launch_bounds__(...)
#define HIP_SYMBOL(x) (&(x))
#define __align__(n) alignas(n)

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};
struct double2 {
  double x, y;
};
inline double2 make_double2(double x, double y) { return double2{x, y}; }

typedef int hipError_t;
enum { hipSuccess = 0 };
typedef void* hipStream_t;
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1,
                     hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1 };

namespace sim {

inline double timeout_s() {
  const char* e = getenv("SIM_TIMEOUT_S");
  return e ? atof(e) : 30.0;
}

// Barrier over n participants; a participant that leaves lowers n.
struct Barrier {
  std::mutex m;
  std::condition_variable cv;
  int n = 0, arrived = 0;
  long gen = 0;
  void init(int count) { n = count; arrived = 0; gen = 0; }
  void release_locked() { arrived = 0; ++gen; cv.notify_all(); }
  void wait(const char* what) {
    std::unique_lock<std::mutex> lk(m);
    long g = gen;
    if (++arrived >= n) { release_locked(); return; }
    auto dl = std::chrono::steady_clock::now() +
              std::chrono::duration<double>(timeout_s());
    while (gen == g) {
      if (cv.wait_until(lk, dl) == std::cv_status::timeout && gen == g) {
        fprintf(stderr, "sim: %s timed out (%d of %d arrived)\n", what, arrived, n);
        fflush(stderr);
        abort();
      }
    }
  }
  void leave() {
    std::lock_guard<std::mutex> lk(m);
    --n;
    if (n > 0 && arrived >= n) release_locked();
  }
};

struct Block {
  unsigned bx, by;
  int nthreads;
  Barrier bar;
  std::vector<Barrier> wave_bar;
  std::vector<uint64_t> slots;
  explicit Block(int nt) : nthreads(nt), wave_bar((nt + 63) / 64), slots(nt) { reset(); }
  void reset() {
    bar.init(nthreads);
    for (size_t w = 0; w < wave_bar.size(); ++w) {
      int first = (int)w * 64;
      wave_bar[w].init(std::min(64, nthreads - first));
    }
  }
};

struct Ctx {
  dim3 tid, bid;
  Block* blk;
};
inline thread_local Ctx ctx;
inline dim3 g_grid, g_block;

inline uint64_t exchange_raw(uint64_t mine, int src) {
  Block* b = ctx.blk;
  int t = (int)ctx.tid.x, w = t / 64, base = w * 64;
  b->slots[t] = mine;
  b->wave_bar[w].wait("wave exchange (write)");
  uint64_t r = (base + src < b->nthreads) ? b->slots[base + src] : 0;
  b->wave_bar[w].wait("wave exchange (read)");
  return r;
}

// One mutex for every kernel type (a static inside the launch template
// would be one per instantiation).
inline std::mutex& launch_mutex() {
  static std::mutex m;
  return m;
}

// Dynamic LDS of the running launch (HIP_DYNAMIC_SHARED): one buffer, the
// workgroups run one after another.
inline std::vector<double> g_dyn_shared;

template <class F, class... A>
void launch(size_t shm, F f, dim3 grid, dim3 block, A... args) {
  // one launch at a time: launches from several host threads (a serving
  // loop's upload worker and its step loop) share the grid globals below,
  // where the device runs them on separate streams
  std::lock_guard<std::mutex> launch_lock(launch_mutex());
  g_dyn_shared.assign((shm + sizeof(double) - 1) / sizeof(double), 0.0);
  const int nt = (int)(block.x * block.y * block.z);
  const long nblocks = (long)grid.x * grid.y;
  g_grid = grid;
  g_block = block;
  std::atomic<bool> done{false};
  std::thread watchdog([&] {
    auto dl = std::chrono::steady_clock::now() +
              std::chrono::duration<double>(timeout_s() * 4);
    while (!done.load()) {
      if (std::chrono::steady_clock::now() > dl) {
        fprintf(stderr, "sim: launch exceeded %.0f s (a loop that never ends)\n",
                timeout_s() * 4);
        fflush(stderr);
        abort();
      }
      std::this_thread::sleep_for(std::chrono::milliseconds(50));
    }
  });
  // One host thread per lane for the whole launch; the workgroups run one
  // after another, separated by a launch-wide barrier (lane 0 re-arms the
  // workgroup's barriers in between).
  Block blk(nt);
  Barrier step;
  step.init(nt);
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (long b = 0; b < nblocks; ++b) {
        ctx.tid = dim3((unsigned)t, 0, 0);
        ctx.bid = dim3((unsigned)(b % grid.x), (unsigned)(b / grid.x), 0);
        ctx.blk = &blk;
        f(args...);
        blk.bar.leave();
        blk.wave_bar[t / 64].leave();
        step.wait("end of workgroup");
        if (t == 0) blk.reset();
        step.wait("next workgroup");
      }
    });
  for (auto& x : th) x.join();
  done = true;
  watchdog.join();
}

inline int cus() {
  const char* e = getenv("SIM_CUS");
  return e ? atoi(e) : 2;
}

template <typename T>
T atomic_add_fp(T* p, T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "fp atomic");
  using U = typename std::conditional<sizeof(T) == 8, uint64_t, uint32_t>::type;
  U* q = reinterpret_cast<U*>(p);
  U old = __atomic_load_n(q, __ATOMIC_RELAXED);
  for (;;) {
    T o;
    memcpy(&o, &old, sizeof(T));
    T n = o + v;
    U nu;
    memcpy(&nu, &n, sizeof(T));
    if (__atomic_compare_exchange_n(q, &old, nu, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
      return o;
  }
}

}  // namespace sim

#define threadIdx (sim::ctx.tid)
#define blockIdx (sim::ctx.bid)
#define blockDim (sim::g_block)
#define gridDim (sim::g_grid)

inline void __syncthreads() { sim::ctx.blk->bar.wait("__syncthreads"); }

inline int __shfl_xor(int v, int mask, int width = 64) {
  (void)width;
  int lane = (int)(sim::ctx.tid.x & 63);
  uint64_t r = sim::exchange_raw((uint64_t)(uint32_t)v, lane ^ mask);
  return (int)(uint32_t)r;
}

inline unsigned long long __ballot(int pred) {
  sim::Block* b = sim::ctx.blk;
  int t = (int)sim::ctx.tid.x, w = t / 64, base = w * 64;
  b->slots[t] = pred ? 1 : 0;
  b->wave_bar[w].wait("__ballot (write)");
  unsigned long long m = 0;
  for (int l = 0; l < 64 && base + l < b->nthreads; ++l)
    if (b->slots[base + l]) m |= 1ULL << l;
  b->wave_bar[w].wait("__ballot (read)");
  return m;
}

inline int __builtin_amdgcn_readfirstlane(int v) { return v; }

inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }

// sin(pi x), cos(pi x) with the exact values at multiples of 1/2 that the
// device routine returns.
inline void sincospi(double x, double* s, double* c) {
  double r = std::fmod(x, 2.0);
  if (r == 0.0) { *s = 0.0 * (x < 0 ? -1.0 : 1.0); *c = 1.0; return; }
  if (std::fabs(r) == 1.0) { *s = 0.0; *c = -1.0; return; }
  if (r == 0.5 || r == -1.5) { *s = 1.0; *c = 0.0; return; }
  if (r == -0.5 || r == 1.5) { *s = -1.0; *c = 0.0; return; }
  *s = std::sin(M_PI * x);
  *c = std::cos(M_PI * x);
}

inline float atomicAdd(float* p, float v) { return sim::atomic_add_fp(p, v); }
inline double atomicAdd(double* p, double v) { return sim::atomic_add_fp(p, v); }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
}
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }

using std::max;
using std::min;

inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "simulated"; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
// a __device__ variable is a host variable here: its address is the symbol's
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
  *p = const_cast<void*>(sym);
  return hipSuccess;
}
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
  *v = sim::cus();
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  memcpy(d, s, n);
  return hipSuccess;
}
// launches run to completion on the calling thread: a stream is always idle
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }

#define hipLaunchKernelGGL(k, g, b, shm, st, ...) \
  sim::launch((size_t)(shm), k, dim3(g), dim3(b), ##__VA_ARGS__)
#define HIP_DYNAMIC_SHARED(type, var) type* var = reinterpret_cast<type*>(sim::g_dyn_shared.data());

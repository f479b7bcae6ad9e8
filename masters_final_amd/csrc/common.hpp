// Shared device/host helpers for the EAC hot-path library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>

#include "../../include/eac_hip.h"

namespace eac {

constexpr int kWave = 64;

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define EAC_HIP_CHECK(expr)                                                     \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess)                                                       \
      return ::eac::fail(EAC_ERR_HIP, "%s failed: %s (%s:%d)", #expr,           \
                         hipGetErrorString(_e), __FILE__, __LINE__);            \
  } while (0)

#define EAC_LAUNCH_CHECK()                                                      \
  do {                                                                          \
    hipError_t _e = hipGetLastError();                                          \
    if (_e != hipSuccess)                                                       \
      return ::eac::fail(EAC_ERR_HIP, "kernel launch failed: %s (%s:%d)",       \
                         hipGetErrorString(_e), __FILE__, __LINE__);            \
  } while (0)

#define EAC_TRY(expr)                                                           \
  do {                                                                          \
    int _rc = (expr);                                                           \
    if (_rc != EAC_OK) return _rc;                                              \
  } while (0)

// ------------------------------------------------------------ workspace
// Bump allocator over a caller-provided device workspace. Every carve is
// 256-byte aligned so vector loads and 64-bit atomics stay aligned.
struct Arena {
  char* base;
  size_t cap;
  size_t used;
  bool dry;  // size-only pass: base may be null
  Arena(void* b, size_t c, bool dry_run = false)
      : base(static_cast<char*>(b)), cap(c), used(0), dry(dry_run) {}
  template <typename T>
  T* take(size_t count) {
    size_t off = (used + 255) & ~size_t(255);
    used = off + ((count * sizeof(T) + 255) & ~size_t(255));
    if (dry) return nullptr;
    return reinterpret_cast<T*>(base + off);
  }
  bool overflow() const { return !dry && used > cap; }
};

// Small pinned host staging area for the few device->host scalars the
// reference API returns (counts, totals). One per host thread.
void* pinned_scratch(size_t bytes);

// ------------------------------------------------------- kernel profiler
// When enabled (eac_profile_enable), instrumented launches are bracketed by
// HIP events on their own stream; records are resolved at the next
// prof_flush() (called by entry points after they synchronise).
struct ProfRec {
  int slot;
  hipEvent_t a, b;
  int64_t bytes;
};
int tune_value(int knob);  // EAC_TUNE_* setting, 0 = built-in choice
bool prof_on();
void prof_suspend(bool on);  // no records while on (hipGraph capture)
int prof_begin(int slot, int64_t bytes, hipStream_t s);  // returns record id or -1
void prof_end(int rec, hipStream_t s);
void prof_set_bytes(int rec, int64_t bytes);
void prof_flush();

inline int grid_for(int64_t work, int per_block, int cap_blocks = 1 << 20) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap_blocks) g = cap_blocks;
  return static_cast<int>(g);
}

inline int bits_for(uint64_t x) {  // number of bits to hold values 0..x
  int b = 0;
  while (b < 64 && (x >> b) != 0) ++b;
  return b;
}

// --------------------------------------------------------- device helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  int l = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    T o = __shfl_up(v, off, kWave);
    if (l >= off) v += o;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    uint64_t o = __shfl_xor(v, off, kWave);
    v = o < v ? o : v;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread (blockDim.x <= 1024,
// multiple of 64). Returns the exclusive prefix; *total gets the block sum.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* lds_waves, T* total) {
  const int l = lane_id();
  const int w = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  T inc = wave_incl_scan(v);
  if (l == kWave - 1) lds_waves[w] = inc;
  __syncthreads();
  if (w == 0) {
    T x = (l < nw) ? lds_waves[l] : T(0);
    T xi = wave_incl_scan(x);
    if (l < nw) lds_waves[l] = xi - x;
    if (l == nw - 1) lds_waves[nw] = xi;
  }
  __syncthreads();
  T res = lds_waves[w] + inc - v;
  *total = lds_waves[nw];
  __syncthreads();
  return res;
}

}  // namespace eac

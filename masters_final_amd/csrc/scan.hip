// Exclusive prefix sum (reduce-then-scan, 4096-element tiles).
//
// Replaces the reference's three-kernel Blelloch scan
// (MyML/MyML/utils/scan.py:144-249 scan_gpu, prescan :281-343, last_scan
// :347-432, scan_sum :253-277). Differences by design: every integer width
// is scanned in its own type (the reference's shared memory is int32 and
// truncates int64, scan.py:288), sizes 1 and 2*T work (scan.py:174-180
// computes log2(0) there), and an int32->int64 widening variant produces
// row offsets for CSRs with more than 2^31 entries.
//
// Layout: a tile is read coalesced (striped, 256 threads x 16 items), turned
// blocked through padded LDS, scanned sequentially per thread, combined with
// a wave64 shuffle scan across threads, and written back striped. Tile sums
// are scanned recursively with the same kernels.
#include "prims.hpp"

namespace eac {

namespace {

template <typename T>
__device__ __forceinline__ int pad_idx(int i) {
  // 4-byte elements: one pad word per 32; 8-byte: one pad element per 16.
  return sizeof(T) == 4 ? i + (i >> 5) : i + (i >> 4);
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const Tin* __restrict__ in,
                                                            int64_t n,
                                                            Tout* __restrict__ sums) {
  __shared__ Tout lds[kScanBlock / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  Tout s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t i = base + j * kScanBlock + threadIdx.x;
    if (i < n) s += static_cast<Tout>(in[i]);
  }
  s = wave_sum(s);
  if (lane_id() == 0) lds[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    Tout t = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / kWave; ++w) t += lds[w];
    sums[blockIdx.x] = t;
  }
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kScanBlock) void k_scan_down(const Tin* in, Tout* out,
                                                          int64_t n,
                                                          const Tout* __restrict__ offs,
                                                          Tout init,
                                                          Tout* __restrict__ d_total) {
  constexpr int kPad = sizeof(Tout) == 4 ? kScanTile / 32 : kScanTile / 16;
  __shared__ Tout tile[kScanTile + kPad];
  __shared__ Tout wsum[kScanBlock / kWave + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  const Tout off = offs ? offs[blockIdx.x] : init;
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t i = base + j * kScanBlock + t;
    Tout v = (i < n) ? static_cast<Tout>(in[i]) : Tout(0);
    tile[pad_idx<Tout>(j * kScanBlock + t)] = v;
  }
  __syncthreads();
  Tout vals[kScanItems];
  Tout acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    Tout v = tile[pad_idx<Tout>(t * kScanItems + k)];
    vals[k] = acc;
    acc += v;
  }
  Tout total;
  Tout pre = block_excl_scan(acc, wsum, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    tile[pad_idx<Tout>(t * kScanItems + k)] = vals[k] + pre + off;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t i = base + j * kScanBlock + t;
    if (i < n) out[i] = tile[pad_idx<Tout>(j * kScanBlock + t)];
  }
  if (d_total && blockIdx.x == gridDim.x - 1 && t == 0) *d_total = off + total;
}

template <typename T>
__global__ void k_fill(T* p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

size_t scan_ws_bytes(int64_t n, int elem_bytes) {
  size_t total = 0;
  int64_t t = (n + kScanTile - 1) / kScanTile;
  while (t > 1) {
    total += ((size_t)t * elem_bytes + 255) & ~size_t(255);
    t = (t + kScanTile - 1) / kScanTile;
  }
  return total + 256;
}

template <typename T>
int fill(T* p, int64_t n, T value, hipStream_t s) {
  if (n <= 0) return EAC_OK;
  int g = grid_for(n, 256, 4096);
  k_fill<T><<<g, 256, 0, s>>>(p, n, value);
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}

template <typename Tin, typename Tout>
int scan_excl(const Tin* in, Tout* out, int64_t n, Tout init, Tout* d_total,
              Arena& ar, hipStream_t s) {
  if (n < 0) return fail(EAC_ERR_ARG, "scan: negative length %lld", (long long)n);
  if (n == 0) {
    if (d_total) return fill<Tout>(d_total, 1, init, s);
    return EAC_OK;
  }
  int64_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles > (int64_t)INT32_MAX) return fail(EAC_ERR_ARG, "scan: too many tiles");
  if (tiles == 1) {
    k_scan_down<Tin, Tout><<<1, kScanBlock, 0, s>>>(in, out, n, nullptr, init, d_total);
    EAC_LAUNCH_CHECK();
    return EAC_OK;
  }
  Tout* sums = ar.take<Tout>(tiles);
  if (ar.overflow()) return fail(EAC_ERR_WORKSPACE, "scan: workspace too small");
  k_scan_reduce<Tin, Tout><<<(int)tiles, kScanBlock, 0, s>>>(in, n, sums);
  EAC_LAUNCH_CHECK();
  EAC_TRY((scan_excl<Tout, Tout>(sums, sums, tiles, init, nullptr, ar, s)));
  k_scan_down<Tin, Tout><<<(int)tiles, kScanBlock, 0, s>>>(in, out, n, sums, init, d_total);
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}

template int scan_excl<int32_t, int32_t>(const int32_t*, int32_t*, int64_t, int32_t,
                                         int32_t*, Arena&, hipStream_t);
template int scan_excl<int64_t, int64_t>(const int64_t*, int64_t*, int64_t, int64_t,
                                         int64_t*, Arena&, hipStream_t);
template int scan_excl<int32_t, int64_t>(const int32_t*, int64_t*, int64_t, int64_t,
                                         int64_t*, Arena&, hipStream_t);
template int scan_excl<uint8_t, int32_t>(const uint8_t*, int32_t*, int64_t, int32_t,
                                         int32_t*, Arena&, hipStream_t);
template int scan_excl<uint8_t, int64_t>(const uint8_t*, int64_t*, int64_t, int64_t,
                                         int64_t*, Arena&, hipStream_t);
template int fill<int32_t>(int32_t*, int64_t, int32_t, hipStream_t);
template int fill<int64_t>(int64_t*, int64_t, int64_t, hipStream_t);
template int fill<uint64_t>(uint64_t*, int64_t, uint64_t, hipStream_t);
template int fill<uint8_t>(uint8_t*, int64_t, uint8_t, hipStream_t);

}  // namespace eac

// ------------------------------------------------------------------ C ABI
using eac::Arena;

extern "C" size_t eac_scan_workspace_bytes(int64_t n, int32_t elem_bytes) {
  return eac::scan_ws_bytes(n, elem_bytes < 4 ? 4 : elem_bytes);
}

template <typename Tin, typename Tout>
static int scan_entry(const Tin* in, Tout* out, int64_t n, Tout init, Tout* d_total,
                      void* ws, size_t ws_bytes, void* stream) {
  if (n < 0) return eac::fail(EAC_ERR_ARG, "scan: negative length");
  if (n > 0 && (!in || !out)) return eac::fail(EAC_ERR_ARG, "scan: null array");
  size_t need = eac::scan_ws_bytes(n, sizeof(Tout));
  if (ws_bytes < need || (need > 256 && !ws))
    return eac::fail(EAC_ERR_WORKSPACE, "scan: workspace %zu < %zu bytes", ws_bytes, need);
  Arena ar(ws, ws_bytes);
  return eac::scan_excl<Tin, Tout>(in, out, n, init, d_total, ar, (hipStream_t)stream);
}

extern "C" int eac_scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n,
                                      int32_t init, int32_t* d_total, void* ws,
                                      size_t ws_bytes, void* stream) {
  return scan_entry<int32_t, int32_t>(in, out, n, init, d_total, ws, ws_bytes, stream);
}

extern "C" int eac_scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t n,
                                      int64_t init, int64_t* d_total, void* ws,
                                      size_t ws_bytes, void* stream) {
  return scan_entry<int64_t, int64_t>(in, out, n, init, d_total, ws, ws_bytes, stream);
}

extern "C" int eac_scan_exclusive_i32_i64(const int32_t* in, int64_t* out, int64_t n,
                                          int64_t init, int64_t* d_total, void* ws,
                                          size_t ws_bytes, void* stream) {
  if ((const void*)in == (const void*)out && n > 0)
    return eac::fail(EAC_ERR_ARG, "scan i32->i64: in and out must not alias");
  return scan_entry<int32_t, int64_t>(in, out, n, init, d_total, ws, ws_bytes, stream);
}

// Stable LSD radix sort of 32/64-bit unsigned keys with optional 32/64-bit
// values, 8 bits per pass.
//
// Used for the canonical (weight, min, max) order of forest edges in the
// lifetime cuts (the reference sorts MST weights with numpy's unstable
// argsort, MyML/MyML/EAC/eac.py:232 and MyML/MyML/graph/singlelink.py:27),
// for the row order of graphs rebuilt from edge lists
// (MyML/MyML/graph/build.py:49-93) and for per-graph grouping in the batched
// single-link path.
//
// Per pass: (1) per-tile digit histograms (digit-major table), (2) one
// exclusive scan of the table gives every (digit, tile) its output offset,
// (3) each tile ranks its keys stably and scatters them. The stable rank
// inside a tile uses wave64 match masks built from 8 ballots (peers with the
// same digit), so no per-thread digit counters are needed: 4096 keys per
// tile are processed in 16 steps of 256 in index order, each step adding
// the per-wave digit counts of earlier waves.
#include "prims.hpp"
#include "sort.hpp"

namespace eac {
namespace {

constexpr int kRB = 256;               // threads per block
constexpr int kRI = 16;                // keys per thread
constexpr int kRTile = kRB * kRI;      // keys per tile
constexpr int kRadix = 256;

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K k, int shift) {
  return (uint32_t)((k >> shift) & (K)0xff);
}

template <typename K>
__global__ __launch_bounds__(kRB) void k_hist(const K* __restrict__ keys, int64_t n, int shift,
                                              int32_t ntiles, int32_t* __restrict__ table) {
  __shared__ int32_t h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kRTile;
#pragma unroll
  for (int j = 0; j < kRI; ++j) {
    int64_t i = base + j * kRB + threadIdx.x;
    if (i < n) atomicAdd(&h[digit_of(keys[i], shift)], 1);
  }
  __syncthreads();
  table[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

template <typename K, typename V, bool HAS_V>
__global__ __launch_bounds__(kRB) void k_scatter(const K* __restrict__ kin,
                                                 const V* __restrict__ vin, K* __restrict__ kout,
                                                 V* __restrict__ vout, int64_t n, int shift,
                                                 int32_t ntiles,
                                                 const int32_t* __restrict__ offs) {
  __shared__ int32_t s_run[kRadix];
  __shared__ int32_t s_off[kRadix];
  __shared__ int32_t s_wcnt[kRB / kWave][kRadix];
  const int t = threadIdx.x;
  const int w = t / kWave;
  s_run[t] = 0;
  s_off[t] = offs[(int64_t)t * ntiles + blockIdx.x];
#pragma unroll
  for (int ww = 0; ww < kRB / kWave; ++ww) s_wcnt[ww][t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRTile;
  const uint64_t lt = lanemask_lt();
  for (int j = 0; j < kRI; ++j) {
    int64_t i = base + j * kRB + t;
    bool valid = i < n;
    K key = valid ? kin[i] : (K)0;
    V val{};
    if (HAS_V && valid) val = vin[i];
    uint32_t d = digit_of(key, shift);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t bal = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    int rank = __popcll(peers & lt);
    if (valid && rank == 0) s_wcnt[w][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int pos = s_run[d] + rank;
      for (int ww = 0; ww < w; ++ww) pos += s_wcnt[ww][d];
      int64_t o = (int64_t)s_off[d] + pos;
      kout[o] = key;
      if (HAS_V) vout[o] = val;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int ww = 0; ww < kRB / kWave; ++ww) {
      add += s_wcnt[ww][t];
      s_wcnt[ww][t] = 0;
    }
    s_run[t] += add;
    __syncthreads();
  }
}

template <typename T>
__global__ void k_copy(const T* __restrict__ a, T* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

}  // namespace

size_t radix_sort_ws_bytes(int64_t n, int key_bytes, int val_bytes) {
  int64_t ntiles = (n + kRTile - 1) / kRTile;
  size_t table = (size_t)ntiles * kRadix;
  Arena ar(nullptr, 0, true);
  ar.take<char>((size_t)n * key_bytes);
  if (val_bytes) ar.take<char>((size_t)n * val_bytes);
  ar.take<int32_t>(table);
  return ar.used + scan_ws_bytes((int64_t)table, 4) + 1024;
}

template <typename K, typename V>
int radix_sort(const K* kin, K* kout, const V* vin, V* vout, int64_t n, int begin_bit,
               int end_bit, Arena& ar, hipStream_t s) {
  constexpr bool HAS_V = !std::is_same<V, NoValue>::value;
  if (n <= 0) return EAC_OK;
  if (begin_bit < 0 || end_bit > (int)(8 * sizeof(K)) || begin_bit > end_bit)
    return fail(EAC_ERR_ARG, "radix_sort: bad bit range");
  const int64_t ntiles64 = (n + kRTile - 1) / kRTile;
  if (ntiles64 * kRadix > INT32_MAX) return fail(EAC_ERR_ARG, "radix_sort: too many keys");
  const int32_t ntiles = (int32_t)ntiles64;
  K* kbuf = ar.take<K>(n);
  V* vbuf = HAS_V ? ar.take<V>(n) : nullptr;
  int32_t* table = ar.take<int32_t>((size_t)ntiles * kRadix);
  if (ar.overflow()) return fail(EAC_ERR_WORKSPACE, "radix_sort: workspace too small");
  int passes = (end_bit - begin_bit + 7) / 8;
  if (passes == 0) {
    if (kout != kin) {
      k_copy<K><<<grid_for(n, 256, 4096), 256, 0, s>>>(kin, kout, n);
      EAC_LAUNCH_CHECK();
    }
    if (HAS_V && vout != vin) {
      k_copy<V><<<grid_for(n, 256, 4096), 256, 0, s>>>(vin, vout, n);
      EAC_LAUNCH_CHECK();
    }
    return EAC_OK;
  }
  // ping-pong so that the last pass writes kout/vout
  const K* src_k = kin;
  const V* src_v = vin;
  const size_t mark = ar.used;  // each pass's scan reuses the same scratch
  for (int p = 0; p < passes; ++p) {
    int shift = begin_bit + 8 * p;
    bool to_out = ((passes - 1 - p) % 2) == 0;
    K* dk = to_out ? kout : kbuf;
    V* dv = to_out ? vout : vbuf;
    k_hist<K><<<ntiles, kRB, 0, s>>>(src_k, n, shift, ntiles, table);
    EAC_LAUNCH_CHECK();
    EAC_TRY((scan_excl<int32_t, int32_t>(table, table, (int64_t)ntiles * kRadix, 0, nullptr, ar,
                                         s)));
    ar.used = mark;
    k_scatter<K, V, HAS_V><<<ntiles, kRB, 0, s>>>(src_k, src_v, dk, dv, n, shift, ntiles, table);
    EAC_LAUNCH_CHECK();
    src_k = dk;
    src_v = dv;
  }
  return EAC_OK;
}

template int radix_sort<uint64_t, NoValue>(const uint64_t*, uint64_t*, const NoValue*, NoValue*,
                                           int64_t, int, int, Arena&, hipStream_t);
template int radix_sort<uint32_t, NoValue>(const uint32_t*, uint32_t*, const NoValue*, NoValue*,
                                           int64_t, int, int, Arena&, hipStream_t);
template int radix_sort<uint32_t, uint32_t>(const uint32_t*, uint32_t*, const uint32_t*,
                                            uint32_t*, int64_t, int, int, Arena&, hipStream_t);
template int radix_sort<uint32_t, uint64_t>(const uint32_t*, uint32_t*, const uint64_t*,
                                            uint64_t*, int64_t, int, int, Arena&, hipStream_t);
template int radix_sort<uint64_t, uint32_t>(const uint64_t*, uint64_t*, const uint32_t*,
                                            uint32_t*, int64_t, int, int, Arena&, hipStream_t);

}  // namespace eac

// ------------------------------------------------------------------ C ABI
extern "C" size_t eac_sort_workspace_bytes(int64_t n, int32_t key_bytes, int32_t val_bytes) {
  return eac::radix_sort_ws_bytes(n < 1 ? 1 : n, key_bytes, val_bytes);
}

extern "C" int eac_radix_sort_u64(const uint64_t* keys_in, uint64_t* keys_out,
                                  const uint32_t* vals_in, uint32_t* vals_out, int64_t n,
                                  int32_t begin_bit, int32_t end_bit, void* ws, size_t ws_bytes,
                                  void* stream) {
  if (n < 0) return eac::fail(EAC_ERR_ARG, "sort: negative length");
  if (n == 0) return EAC_OK;
  if (!keys_in || !keys_out || (!vals_in) != (!vals_out))
    return eac::fail(EAC_ERR_ARG, "sort: null array");
  size_t need = eac::radix_sort_ws_bytes(n, 8, vals_in ? 4 : 0);
  if (!ws || ws_bytes < need)
    return eac::fail(EAC_ERR_WORKSPACE, "sort: workspace %zu < %zu bytes", ws_bytes, need);
  eac::Arena ar(ws, ws_bytes);
  hipStream_t s = (hipStream_t)stream;
  if (vals_in)
    return eac::radix_sort<uint64_t, uint32_t>(keys_in, keys_out, vals_in, vals_out, n,
                                               begin_bit, end_bit, ar, s);
  return eac::radix_sort<uint64_t, eac::NoValue>(keys_in, keys_out, nullptr, nullptr, n,
                                                 begin_bit, end_bit, ar, s);
}

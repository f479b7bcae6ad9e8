// Boruvka minimum spanning forest, connected components and first-occurrence
// relabelling for gfx950.
//
// Reference algorithm: MyML/MyML/graph/mst.py:143-407 (boruvka_minho_seq and
// its Numba helpers) and mst.py:417-809 (boruvka_minho_gpu and its kernels),
// after Sousa et al. Each round every supervertex picks its lightest edge,
// mirrored picks are resolved by keeping the smaller id as representative
// (mst.py:282-305), the picks are added to the forest (mst.py:308-317), and
// supervertices are merged by pointer jumping (mst.py:333-351).
//
// MI355X design (not a translation of the reference kernels):
//  * The reference rebuilds a contracted CSR every round with atomic
//    counters (countNewEdges/assignInsert, mst.py:755-809), which costs a
//    count pass, a scan and an atomic scatter per round and leaves the edge
//    order nondeterministic. Here the graph is kept as a flat list of
//    undirected edges, each one 64-bit key
//        key = wkey << 2b | min(u,v) << b | max(u,v)      (b = bits(n-1))
//    where wkey is the weight mapped to an order-preserving unsigned integer
//    minus the smallest such value. The key is unique per undirected edge,
//    so (weight, min-id, max-id) is a strict total order: the forest is
//    unique and no cycle longer than 2 can form. One coalesced 8-byte read
//    per surviving edge per round serves both min-edge selection and
//    filtering; endpoints are decoded from the key.
//  * Round 0 streams the CSR itself in balanced 2048-entry chunks (row
//    offsets of the chunk staged in LDS), so vertex degree skew does not
//    unbalance waves. For a symmetric CSR the per-vertex minimum is an LDS
//    segmented min with no global atomics for rows inside the chunk.
//  * Edge lists are chunk-local (chunk c of the CSR owns slots
//    [2048 c, 2048 c + count[c]) of the light and heavy lists), so no pass
//    needs a shared output counter; list order is irrelevant because
//    atomicMin over unique keys is order independent, so every output is
//    deterministic.
//  * A round is two kernels. The hook: every root with a candidate walks
//    its pick tree (root -> other side of its lightest edge -> ...) to the
//    tree's mirror pair and points its parent at the tree's root, recording
//    its edge in its forest slot. The filter: per listed edge, both ends'
//    roots by two parent jumps, the kept edges' ends written back, the
//    cross edges' keys merged per root in an LDS table and flushed with one
//    atomicMin per root and workgroup. Weight levels (a light list filtered
//    round after round, the heavy list routed in when a level is done) keep
//    the early rounds small.
//  * The call's tail gives forest keys in vertex order, their CSR positions
//    and first-occurrence component labels (each root's smallest vertex by
//    a per-tile LDS table and one atomicMin per root and tile). The
//    connected-components entry points use a lock-free union-find instead
//    (larger root under smaller root with atomicCAS, path halving), whose
//    roots are component minima.
#include <climits>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <type_traits>
#include <utility>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "prims.hpp"
#include "sort.hpp"

namespace eac {
namespace {

constexpr uint64_t kInf = ~0ull;
constexpr int kGB = 256;             // threads per block
constexpr int kGI = 8;               // edges per thread per chunk
constexpr int kGTile = kGB * kGI;    // entries per chunk
constexpr int kMaxRows = 1024;       // rows staged in LDS per chunk

// ------------------------------------------------------------ weight order
// Map a weight to an unsigned integer whose order matches the numeric order
// of the weight (the reference compares weights with < and ==).
template <typename W>
__device__ __forceinline__ uint64_t word_of(W w);
template <>
__device__ __forceinline__ uint64_t word_of<int32_t>(int32_t w) {
  return (uint64_t)(int64_t)w ^ (1ull << 63);
}
template <>
__device__ __forceinline__ uint64_t word_of<int64_t>(int64_t w) {
  return (uint64_t)w ^ (1ull << 63);
}
template <>
__device__ __forceinline__ uint64_t word_of<int16_t>(int16_t w) {
  return (uint64_t)(int64_t)w ^ (1ull << 63);
}
template <>
__device__ __forceinline__ uint64_t word_of<uint8_t>(uint8_t w) { return w; }
template <>
__device__ __forceinline__ uint64_t word_of<uint16_t>(uint16_t w) { return w; }
template <>
__device__ __forceinline__ uint64_t word_of<uint32_t>(uint32_t w) { return w; }
template <>
__device__ __forceinline__ uint64_t word_of<float>(float w) {
  uint32_t u = __float_as_uint(w == 0.0f ? 0.0f : w);  // -0 == +0
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return u;
}

__device__ __forceinline__ uint64_t pack_key(uint64_t wk, uint32_t u, uint32_t v, int vb) {
  uint32_t lo = u < v ? u : v, hi = u < v ? v : u;
  return (wk << (2 * vb)) | ((uint64_t)lo << vb) | (uint64_t)hi;
}
__device__ __forceinline__ void unpack_ends(uint64_t key, int vb, int32_t* lo, int32_t* hi) {
  uint64_t mask = (1ull << vb) - 1;
  *hi = (int32_t)(key & mask);
  *lo = (int32_t)((key >> vb) & mask);
}

// ------------------------------------------------------- chunk row mapping
// chunk_row[c] = row holding entry c*kGTile (the last row whose first entry
// is <= it), for c in [0, nchunks]; entries of chunk c belong to rows
// chunk_row[c] .. chunk_row[c+1].
__device__ __forceinline__ int32_t chunk_row_of(const int64_t* __restrict__ fe, int32_t n,
                                                int64_t m, int64_t c) {
  int64_t e = c * kGTile;
  if (e > m - 1) e = m - 1;
  int32_t lo = 0, hi = n;  // upper_bound over fe[0:n]
  while (lo < hi) {
    int32_t mid = lo + ((hi - lo) >> 1);
    if (fe[mid] <= e) lo = mid + 1; else hi = mid;
  }
  const int32_t r = lo - 1;
  return r < 0 ? 0 : r;
}
__global__ void k_chunk_rows(const int64_t* __restrict__ fe, int32_t n, int64_t m,
                             int64_t nchunks, int32_t* __restrict__ chunk_row) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > nchunks) return;
  chunk_row[c] = chunk_row_of(fe, n, m, c);
}

struct ChunkCtx {
  int64_t e0, e1;
  int32_t r_lo, r_cnt;
  bool staged;
};

// Stage row bounds of one chunk in LDS (offsets relative to e0).
template <bool HAS_OD>
__device__ __forceinline__ ChunkCtx stage_chunk(int64_t c, int64_t m, const int64_t* fe,
                                                const int32_t* od, const int32_t* chunk_row,
                                                int32_t* s_beg, int32_t* s_end) {
  ChunkCtx cx;
  cx.e0 = c * kGTile;
  cx.e1 = cx.e0 + kGTile < m ? cx.e0 + kGTile : m;
  cx.r_lo = chunk_row[c];
  int32_t r_hi = chunk_row[c + 1];
  cx.r_cnt = r_hi - cx.r_lo + 1;
  cx.staged = cx.r_cnt <= kMaxRows;
  if (cx.staged) {
    for (int i = threadIdx.x; i < cx.r_cnt; i += blockDim.x) {
      int32_t r = cx.r_lo + i;
      int64_t b = fe[r];
      int64_t e = HAS_OD ? b + od[r] : fe[r + 1];
      s_beg[i] = (int32_t)(b - cx.e0);
      s_end[i] = (int32_t)(e - cx.e0);
    }
  }
  return cx;
}

// Row (index into the staged rows) of entry e, or -1 if e lies in a gap.
template <bool HAS_OD>
__device__ __forceinline__ int32_t row_of(const ChunkCtx& cx, int64_t e, const int64_t* fe,
                                          const int32_t* od, const int32_t* s_beg,
                                          const int32_t* s_end) {
  if (cx.staged) {
    int32_t rel = (int32_t)(e - cx.e0);
    int32_t lo = 0, hi = cx.r_cnt;
    while (lo < hi) {
      int32_t mid = (lo + hi) >> 1;
      if (s_beg[mid] <= rel) lo = mid + 1; else hi = mid;
    }
    int32_t i = lo - 1;
    if (i < 0 || rel >= s_end[i]) return -1;
    return i;
  }
  int32_t lo = cx.r_lo, hi = cx.r_lo + cx.r_cnt;
  while (lo < hi) {
    int32_t mid = lo + ((hi - lo) >> 1);
    if (fe[mid] <= e) lo = mid + 1; else hi = mid;
  }
  int32_t r = lo - 1;
  if (r < 0) return -1;
  int64_t end = HAS_OD ? fe[r] + od[r] : fe[r + 1];
  if (e >= end) return -1;
  return r - cx.r_lo;
}

// ----------------------------------------------------- weight range pass
// Device-scope atomics on one address serialise (tens of ns each), so every
// block of the range pass stores its own (lo, hi) pair and the plan folds
// them (at most kRangeBlocks pairs).
constexpr int kRangeBlocks = 2048;

// Part `bid` of `nblk` of the weight range: rng[2 * bid] = lowest weight
// word, rng[2 * bid + 1] = highest (kInf, 0 when the part holds no entry).
template <typename W, bool HAS_OD>
__device__ void range_part(const W* __restrict__ weight, const int64_t* __restrict__ fe,
                           const int32_t* __restrict__ od, int32_t n, int64_t m,
                           const int32_t* __restrict__ chunk_row, int64_t nchunks, int bid,
                           int nblk, unsigned long long* __restrict__ rng) {
  __shared__ int32_t s_beg[kMaxRows], s_end[kMaxRows];
  __shared__ uint64_t s_mm[2 * (kGB / kWave)];
  uint64_t lo = kInf, hi = 0;
  if (!HAS_OD) {
    // rows tile [fe[0], fe[n]): one flat pass over that range (within [0, m))
    const int64_t b = fe[0] > 0 ? fe[0] : 0, e = fe[n] < m ? fe[n] : m;
    const int64_t tid = (int64_t)bid * kGB + threadIdx.x;
    const int64_t stride = (int64_t)nblk * kGB;
    int64_t s0 = b;
    if (sizeof(W) == 4 && (reinterpret_cast<uintptr_t>(weight + b) & 15) == 0 && e > b) {
      // 16-byte loads over the aligned body
      const int64_t n4 = (e - b) / 4;
      const uint4* w4 = reinterpret_cast<const uint4*>(weight + b);
      for (int64_t q = tid; q < n4; q += stride) {
        const uint4 v = w4[q];
        const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          W x;
          memcpy(&x, &u[t], sizeof(W));
          const uint64_t w = word_of<W>(x);
          lo = w < lo ? w : lo;
          hi = w > hi ? w : hi;
        }
      }
      s0 = b + 4 * n4;
    }
    for (int64_t q = s0 + tid; q < e; q += stride) {
      uint64_t w = word_of<W>(weight[q]);
      lo = w < lo ? w : lo;
      hi = w > hi ? w : hi;
    }
  } else {
    for (int64_t c = bid; c < nchunks; c += nblk) {
      ChunkCtx cx = stage_chunk<HAS_OD>(c, m, fe, od, chunk_row, s_beg, s_end);
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kGI; ++j) {
        int64_t e = cx.e0 + j * kGB + threadIdx.x;
        if (e < cx.e1 && row_of<HAS_OD>(cx, e, fe, od, s_beg, s_end) >= 0) {
          uint64_t w = word_of<W>(weight[e]);
          lo = w < lo ? w : lo;
          hi = w > hi ? w : hi;
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    uint64_t a = __shfl_xor(lo, off, kWave), b = __shfl_xor(hi, off, kWave);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  int w = threadIdx.x / kWave;
  if (lane_id() == 0) { s_mm[2 * w] = lo; s_mm[2 * w + 1] = hi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kGB / kWave; ++i) {
      lo = s_mm[2 * i] < lo ? s_mm[2 * i] : lo;
      hi = s_mm[2 * i + 1] > hi ? s_mm[2 * i + 1] : hi;
    }
    rng[2 * bid] = lo;
    rng[2 * bid + 1] = hi;
  }
}

template <typename W, bool HAS_OD>
__global__ __launch_bounds__(kGB) void k_weight_range(const W* __restrict__ weight,
                                                      const int64_t* __restrict__ fe,
                                                      const int32_t* __restrict__ od,
                                                      int32_t n, int64_t m,
                                                      const int32_t* __restrict__ chunk_row,
                                                      int64_t nchunks,
                                                      unsigned long long* __restrict__ rng) {
  range_part<W, HAS_OD>(weight, fe, od, n, m, chunk_row, nchunks, blockIdx.x, gridDim.x, rng);
}

constexpr int kHistBins = 4096;

// Weight offset, key width and light/heavy threshold of one call, computed
// on the device so that no host round trip sits between the passes.
constexpr int kLevels = 4;  // weight levels of the split (the last is unbounded)
struct MstPlan {
  unsigned long long wlo;        // weight word of the lightest entry
  unsigned long long whi;        // weight word of the heaviest entry
  unsigned long long light_lim;  // keys with (word - wlo) >= light_lim are heavy
  // level thresholds: level L takes the keys with lim[L-1] <= word - wlo <
  // lim[L] (lim[0] = light_lim, lim[kLevels-1] = ~0; ~0 ends the levels)
  unsigned long long lim[kLevels];
  int wbits;                     // bits of whi - wlo
  int err;                       // 1: wbits + 2 * vb > 64
};
// One block: folds the exact weight range, then (levels > 1) estimates the
// weight quantiles frac, frac * step, frac * step^2 ... (levels - 1 of them)
// from kPlanSamples evenly spaced entries with a 4096-bin histogram of the
// range: the level thresholds of the split. They only steer the split; any
// values give the same forest. The samples are gathered by the set-up pass
// (spread over many workgroups: one workgroup alone fetching 8192 scattered
// lines takes several microseconds).
constexpr int kPlanThreads = 256;
constexpr int kPlanSamples = 8192;

// Samples bid, bid + nblk, ... (in units of kGB) of the plan's evenly spaced
// weight entries, as weight words.
template <typename W>
__device__ void sample_part(const W* __restrict__ weight, int64_t m, int bid, int nblk,
                            unsigned long long* __restrict__ samples) {
  if (m <= 0) return;
  for (int i = bid * kGB + threadIdx.x; i < kPlanSamples; i += nblk * kGB) {
    const int64_t e = (int64_t)((double)i * (double)m / kPlanSamples);
    samples[i] = word_of<W>(weight[e < m ? e : m - 1]);
  }
}

template <typename W>
__global__ __launch_bounds__(kPlanThreads) void k_plan(
    int64_t m, const unsigned long long* __restrict__ rng, int nrng,
    const unsigned long long* __restrict__ samples, int vb, int levels, double frac,
    double step, MstPlan* __restrict__ plan) {
  __shared__ unsigned int h[kHistBins];
  __shared__ unsigned int s_tmp[kPlanThreads / kWave + 1];
  __shared__ uint64_t s_mm[2 * (kPlanThreads / kWave)];
  __shared__ int s_found[kLevels];
  constexpr int kPer = kPlanSamples / kPlanThreads;
  const bool sample = levels > 1 && m > 0;  // uniform over the block
  uint64_t sw[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) sw[q] = sample ? samples[threadIdx.x + q * kPlanThreads] : 0ull;
  {
    uint64_t l = kInf, hh = 0;
    for (int i = threadIdx.x; i < nrng; i += kPlanThreads) {
      const uint64_t a = rng[2 * i], b = rng[2 * i + 1];
      l = a < l ? a : l;
      hh = b > hh ? b : hh;
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      const uint64_t a = __shfl_xor(l, off, kWave), b = __shfl_xor(hh, off, kWave);
      l = a < l ? a : l;
      hh = b > hh ? b : hh;
    }
    if (lane_id() == 0) {
      s_mm[2 * (threadIdx.x / kWave)] = l;
      s_mm[2 * (threadIdx.x / kWave) + 1] = hh;
    }
  }
  if (threadIdx.x < kLevels) s_found[threadIdx.x] = kHistBins;
  for (int i = threadIdx.x; i < kHistBins; i += kPlanThreads) h[i] = 0;
  __syncthreads();
  uint64_t wlo = kInf, whi = 0;
  for (int i = 0; i < kPlanThreads / kWave; ++i) {
    wlo = s_mm[2 * i] < wlo ? s_mm[2 * i] : wlo;
    whi = s_mm[2 * i + 1] > whi ? s_mm[2 * i + 1] : whi;
  }
  if (wlo == kInf) { wlo = 0; whi = 0; }
  int wbits = 0;
  while (wbits < 64 && ((whi - wlo) >> wbits) != 0) ++wbits;
  uint64_t lim[kLevels];
#pragma unroll
  for (int L = 0; L < kLevels; ++L) lim[L] = kInf;
  if (sample) {
    const int shift = wbits > 12 ? wbits - 12 : 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const uint64_t w = sw[q];
      if (w >= wlo && w <= whi) atomicAdd(&h[(w - wlo) >> shift], 1u);
    }
    __syncthreads();
    constexpr int per = kHistBins / kPlanThreads;
    unsigned int part = 0;
    for (int q = 0; q < per; ++q) part += h[threadIdx.x * per + q];
    unsigned int total;
    const unsigned int excl = block_excl_scan<unsigned int>(part, s_tmp, &total);
    double f = frac;
    for (int L = 0; L + 1 < levels; ++L, f *= step) {
      const double target = f * kPlanSamples;
      if (f < 1.0 && (double)excl < target && (double)(excl + part) >= target) {
        unsigned int cum = excl;
        for (int q = 0; q < per; ++q) {
          cum += h[threadIdx.x * per + q];
          if ((double)cum >= target) {
            atomicMin(&s_found[L], threadIdx.x * per + q);
            break;
          }
        }
      }
    }
    __syncthreads();
    uint64_t prev = 0;
    for (int L = 0; L + 1 < levels; ++L) {
      uint64_t v = s_found[L] < kHistBins ? (uint64_t)(s_found[L] + 1) << shift : kInf;
      if (v != kInf && v > whi - wlo) v = kInf;
      if (v != kInf && v <= prev) v = prev + 1;  // strictly increasing levels
      if (prev == kInf) v = kInf;
      lim[L] = v;
      prev = v;
    }
  }
  if (threadIdx.x == 0) {
    plan->wlo = wlo;
    plan->whi = whi;
    plan->wbits = wbits;
    plan->err = wbits + 2 * vb > 64;
    plan->light_lim = lim[0];
#pragma unroll
    for (int L = 0; L < kLevels; ++L) plan->lim[L] = lim[L];
  }
}

// -------------------------------------------------------- Boruvka round 0
// Edge lists are chunk-local: chunk c of the CSR (kGTile entries) owns slots
// [c * kGTile, c * kGTile + count[c]) of each list, so no pass ever needs a
// shared output counter; later passes filter each chunk in place.
//
// Round 0 streams the CSR once: every entry becomes a key; light keys are
// round-0 candidates (segmented min per row in LDS) and go to the light
// list, heavy keys go to the heavy list. A symmetric CSR lists each edge
// once (dest > row) and takes each row's minimum from the row itself.
constexpr int kLogSlots = 16;
constexpr int kScatterMax = 64;  // longer rows are located by binary search
template <typename W, bool HAS_OD>
__global__ __launch_bounds__(kGB) void k_round0(
    const int32_t* __restrict__ dest, const W* __restrict__ weight,
    const int64_t* __restrict__ fe, const int32_t* __restrict__ od, int32_t n, int64_t m,
    const int32_t* __restrict__ chunk_row, int64_t nchunks, int vb,
    const MstPlan* __restrict__ plan, int symmetric, unsigned long long* __restrict__ best,
    uint64_t* __restrict__ light, uint64_t* __restrict__ heavy,
    int32_t* __restrict__ cnt_light, int32_t* __restrict__ cnt_heavy,
    unsigned long long* __restrict__ tot, unsigned int* __restrict__ overflow) {
  __shared__ int32_t s_beg[kMaxRows], s_end[kMaxRows];
  __shared__ unsigned long long s_best[kMaxRows];
  __shared__ int16_t s_row[kGTile];  // staged row of each entry, -1: gap or long row
  __shared__ uint32_t s_tmp[kGB / kWave + 1];
  const bool err = plan->err != 0;
  const uint64_t word_lo = plan->wlo, light_lim = plan->light_lim;
  unsigned long long sum_l = 0, sum_h = 0;
  // a chunk's entries are loaded one iteration ahead (software pipelining:
  // their latency overlaps the previous chunk's scan, writes and row staging)
  int32_t dv[kGI];
  W wv[kGI];
  auto load_entries = [&](int64_t cc) {
    const int64_t f0 = cc * kGTile;
    const int64_t f1 = f0 + kGTile < m ? f0 + kGTile : m;
#pragma unroll
    for (int j = 0; j < kGI; ++j) {
      const int64_t e = f0 + j * kGB + threadIdx.x;
      dv[j] = e < f1 ? __builtin_nontemporal_load(dest + e) : -1;  // streamed once
      wv[j] = e < f1 ? __builtin_nontemporal_load(weight + e) : W(0);
    }
  };
  if (!err && blockIdx.x < nchunks) load_entries(blockIdx.x);
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    if (err) {  // key too wide: nothing is listed, the host reports the error
      if (threadIdx.x == 0) { cnt_light[c] = 0; cnt_heavy[c] = 0; }
      continue;
    }
    const int64_t e0 = c * kGTile;
    const int64_t e1 = e0 + kGTile < m ? e0 + kGTile : m;
    ChunkCtx cx = stage_chunk<HAS_OD>(c, m, fe, od, chunk_row, s_beg, s_end);
    if (cx.staged) {
      for (int i = threadIdx.x; i < kGTile; i += kGB) s_row[i] = -1;
      for (int i = threadIdx.x; i < cx.r_cnt; i += kGB) s_best[i] = kInf;
    }
    __syncthreads();
    if (cx.staged) {
      // each short row marks its entries (stage_chunk gave row i to this thread)
      const int32_t clen = (int32_t)(e1 - e0);
      for (int i = threadIdx.x; i < cx.r_cnt; i += kGB) {
        const int32_t rb = s_beg[i] > 0 ? s_beg[i] : 0;
        const int32_t re = s_end[i] < clen ? s_end[i] : clen;
        if (re - rb <= kScatterMax)
          for (int32_t q = rb; q < re; ++q) s_row[q] = (int16_t)i;
      }
      __syncthreads();
    }
    uint64_t keys[kGI];
    uint32_t keep = 0, light_m = 0;
#pragma unroll
    for (int j = 0; j < kGI; ++j) {
      keys[j] = kInf;
      const int64_t e = e0 + j * kGB + threadIdx.x;
      if (e >= e1) continue;
      int32_t ri = cx.staged ? (int32_t)s_row[e - e0] : -1;
      if (ri < 0) ri = row_of<HAS_OD>(cx, e, fe, od, s_beg, s_end);
      if (ri < 0) continue;
      const int32_t u = cx.r_lo + ri;
      const int32_t v = dv[j];
      if (v < 0 || v >= n) {  // invalid id: reported, never dereferenced
        atomicOr(overflow, 2u);
        continue;
      }
      if (v == u) continue;  // self loops never join a forest
      const uint64_t wk = word_of<W>(wv[j]) - word_lo;
      const uint64_t key = pack_key(wk, (uint32_t)u, (uint32_t)v, vb);
      keys[j] = key;
      if (!symmetric || v > u) keep |= 1u << j;
      if (wk >= light_lim) continue;  // heavy: listed, not a round-0 candidate
      light_m |= 1u << j;
      if (cx.staged) {
        atomicMin(&s_best[ri], (unsigned long long)key);
      } else if (key < best[u]) {
        atomicMin(&best[u], (unsigned long long)key);
      }
      if (!symmetric) {
        if (key < best[v]) atomicMin(&best[v], (unsigned long long)key);
      }
    }
    if (c + gridDim.x < nchunks) load_entries(c + gridDim.x);
    // light keys to the light chunk, heavy keys to the heavy chunk, in entry
    // order per thread: one scan of the two counts packed in 16-bit halves
    const uint32_t nl = __popc(keep & light_m), nh = __popc(keep & ~light_m);
    uint32_t total;
    const uint32_t pre = block_excl_scan<uint32_t>(nl | (nh << 16), s_tmp, &total);
    uint32_t pl = pre & 0xffffu, ph = pre >> 16;
    uint64_t* lc = light + c * kGTile;
    uint64_t* hc = heavy + c * kGTile;
#pragma unroll
    for (int j = 0; j < kGI; ++j) {
      if (!(keep & (1u << j))) continue;
      if (light_m & (1u << j)) lc[pl++] = keys[j];
      else hc[ph++] = keys[j];
    }
    if (threadIdx.x == 0) {
      const uint32_t TL = total & 0xffffu, TH = total >> 16;
      cnt_light[c] = (int32_t)TL;
      cnt_heavy[c] = (int32_t)TH;
      sum_l += TL;
      sum_h += TH;
    }
    if (cx.staged) {
      for (int i = threadIdx.x; i < cx.r_cnt; i += blockDim.x) {
        unsigned long long b = s_best[i];
        if (b == kInf) continue;
        int32_t r = cx.r_lo + i;
        bool inside = s_beg[i] >= 0 && s_end[i] <= (int32_t)(e1 - e0);
        if (inside && symmetric) best[r] = b;
        else atomicMin(&best[r], b);
      }
    }
    __syncthreads();
  }
  if (tot && threadIdx.x == 0 && (sum_l | sum_h)) {  // profiling only
    const int slot = blockIdx.x & (kLogSlots - 1);
    atomicAdd(&tot[slot], sum_l);
    atomicAdd(&tot[kLogSlots + slot], sum_h);
  }
}

// ------------------------------------------------------------- union-find
// Parent array with parent[x] <= x (larger roots hook under smaller ones),
// so every walk ends at the smallest vertex of the component.
__device__ __forceinline__ int32_t uf_ld(int32_t* p, int32_t i) {
  return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void uf_st(int32_t* p, int32_t i, int32_t v) {
  __hip_atomic_store(p + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ int32_t uf_find(int32_t* P, int32_t x) {
  for (;;) {
    int32_t p = uf_ld(P, x);
    if (p == x) return x;
    int32_t gp = uf_ld(P, p);
    if (gp == p) return p;
    uf_st(P, x, gp);  // path halving
    x = gp;
  }
}
__device__ void uf_unite(int32_t* P, int32_t a, int32_t b) {
  a = uf_find(P, a);
  b = uf_find(P, b);
  while (a != b) {
    int32_t lo = a < b ? a : b, hi = a < b ? b : a;
    int32_t old = atomicCAS(P + hi, hi, lo);
    if (old == hi) return;
    a = uf_find(P, old);
    b = uf_find(P, lo);
  }
}

// ------------------------------------------------------ device round control
// A round is hook(r) -> filter(r). The control state of round
// r lives in ctl[r & 1]; every block of filter(r) derives round r+1's state
// from it and from had_edge[r] (some representative had a candidate edge in
// hook(r)), and block 0 stores it in ctl[(r + 1) & 1]. Rounds filter the
// light list; when a hook finds no candidate, the next weight level starts:
// one routing pass moves the heavy keys of that level that still cross
// components into the light list (dropping internal ones); after the last
// level the forest is complete.
constexpr int kMaxRounds = kLevels * 34;  // kLevels * (31 + 2 + 1)
struct RoundCtl {
  int phase;            // weight level whose keys the light list holds
  int active;
  int rounds;           // rounds started
  int overflow_rounds;  // stopped at the round limit
};

// A level ends when a hook finds no candidate: the next level (if its lower
// threshold is finite) starts with a routing pass over the heavy list.
__device__ __forceinline__ RoundCtl next_ctl(RoundCtl c, int r, unsigned int had,
                                             int max_rounds, const MstPlan* plan) {
  if (!c.active) return c;
  if (!had) {
    if (c.phase + 1 < kLevels && plan->lim[c.phase] != kInf) c.phase += 1;
    else c.active = 0;
  }
  c.rounds = r + 1;
  if (c.active && r + 1 >= max_rounds) {
    c.active = 0;
    c.overflow_rounds = 1;
  }
  return c;
}

// Filter one list in place against the current components and atomicMin
// every cross edge into both roots' slots of the next round (`best` is that
// round's slot array). A wavefront filters one chunk at a time (4 x 64 keys
// in flight per step); survivors are written back in order at the front of
// the chunk (a write never passes the read position).
// Per-block table of (root, lightest key) pairs: cross edges of one block
// that share a root meet in LDS, and each distinct root costs one global
// atomicMin per block (a component touching most edges -- the giant one --
// would otherwise take one serialised atomic per edge). A root whose slot
// is taken goes straight to the global atomic: while components are many
// (the first rounds, the routing passes) the table is crowded, and probing
// further slots costs more LDS atomics than the global ones it saves
// (measured on the 1M-vertex k-NN graph: 8 probes 0.83 ms per call, 1 probe
// 0.79; 512 to 1024 slots are equal, 256 or 2048 and more are slower).
// The table has four slots per thread (1024 for the 256-thread filter
// block, 4096 for the 1024-thread one).
constexpr int kMinProbes = 1;  // table slots tried before the direct atomic
template <int LOGS>
__device__ __forceinline__ void min_insert(int32_t* s_root, unsigned long long* s_min,
                                           int32_t root, uint64_t key,
                                           unsigned long long* best) {
  uint32_t h = ((uint32_t)root * 2654435761u) >> (32 - LOGS);
  for (int p = 0; p < kMinProbes; ++p, h = (h + 1) & ((1u << LOGS) - 1)) {
    const int32_t k = atomicCAS(&s_root[h], -1, root);
    if (k == -1 || k == root) {
      atomicMin(&s_min[h], (unsigned long long)key);
      return;
    }
  }
  atomicMin(&best[root], (unsigned long long)key);  // table crowded: direct
}

// min_insert from a wavefront: lanes holding the same root (the giant
// component's, typically) are first reduced to one lane, for up to two
// distinct roots per call, so a hot root costs one LDS atomic per wave
// instead of one per lane; a root held by one lane only skips the reduction.
template <int LOGS>
__device__ __forceinline__ void wave_min_insert(int32_t* s_root, unsigned long long* s_min,
                                                int32_t root, uint64_t key, bool active,
                                                unsigned long long* best) {
  uint64_t pend = __ballot(active);
  for (int it = 0; it < 2 && pend; ++it) {
    const int leader = __ffsll((unsigned long long)pend) - 1;
    const int32_t R = __builtin_amdgcn_readlane(root, leader);
    const bool mine = active && root == R;
    // a root no other lane holds (most of them while components are
    // small): every lane inserts its own, no reduction
    if (__popcll(__ballot(mine)) < 2) break;
    uint64_t v = mine ? key : kInf;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      const uint64_t o = __shfl_xor(v, off, kWave);
      v = o < v ? o : v;
    }
    if (lane_id() == leader) min_insert<LOGS>(s_root, s_min, R, v, best);
    pend &= ~__ballot(mine);
    if (mine) active = false;
  }
  if (active) min_insert<LOGS>(s_root, s_min, root, key, best);
}

// Keys per lane per step of the filter (their root lookups in flight together).
#ifndef EAC_FILTER_KEYS
#define EAC_FILTER_KEYS 4
#endif
constexpr int kFilterKeys = EAC_FILTER_KEYS;

// Filters one chunk `cc` (n_in keys) of `list` with one wavefront. ROUTE
// (the first round of a weight level): `list` is the heavy list, whose cross
// keys below `lim` (the level's upper threshold, on the key's weight bits)
// move to the same chunk of the (empty) light list and become candidates,
// while heavier cross keys stay in the heavy chunk; otherwise `list` is the
// light list, filtered in place.
template <bool ROUTE, int LOGS>
__device__ __forceinline__ void filter_chunk(uint64_t* __restrict__ list,
                                             int32_t* __restrict__ cnt,
                                             uint64_t* __restrict__ to,
                                             int32_t* __restrict__ cnt_to, int64_t cc,
                                             int32_t n_in, uint64_t lim, int vb, int32_t* P,
                                             int32_t* s_root, unsigned long long* s_min,
                                             unsigned long long* __restrict__ best,
                                             unsigned long long* sum_in,
                                             unsigned long long* sum_out, bool* any_cand) {
  const int lane = lane_id();
  const uint64_t below = (1ull << lane) - 1ull;
  uint64_t* lc = list + cc * kGTile;
  uint64_t* tc = ROUTE ? to + cc * kGTile : nullptr;
  int32_t out = 0, out_to = 0;
  constexpr int Q = kFilterKeys;
  // the next step's keys are loaded while this step is filtered (they lie
  // past every write of this step: survivors go to [out, base + Q x 64))
  uint64_t kn[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = q * kWave + lane;
    kn[q] = e < n_in ? __builtin_nontemporal_load(lc + e) : kInf;
  }
  for (int base = 0; base < n_in; base += Q * kWave) {
    uint64_t k[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) k[q] = kn[q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = base + Q * kWave + q * kWave + lane;
      kn[q] = e < n_in ? __builtin_nontemporal_load(lc + e) : kInf;
    }
    int32_t x[2 * Q];  // roots of the endpoints of the Q keys
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      int32_t a = 0, b = 0;
      if (k[q] != kInf) unpack_ends(k[q], vb, &a, &b);
      x[2 * q] = a;
      x[2 * q + 1] = b;
    }
    // Light keys: P[x] is x's root when the round started (their ends are
    // written back below) and that root's parent is its root after this
    // round's hooks. A routing pass follows a round without hooks, whose
    // hook pass left P flat.
    bool keep[Q], cand[Q];
    if (ROUTE) {
#pragma unroll
      for (int i = 0; i < 2 * Q; ++i) x[i] = P[x[i]];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        keep[q] = k[q] != kInf && x[2 * q] != x[2 * q + 1];
        cand[q] = keep[q] && (k[q] >> (2 * vb)) < lim;
      }
    } else {
      int32_t p[2 * Q];
#pragma unroll
      for (int i = 0; i < 2 * Q; ++i) p[i] = P[x[i]];
      int32_t rt[2 * Q];
#pragma unroll
      for (int i = 0; i < 2 * Q; ++i) rt[i] = P[p[i]];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        keep[q] = k[q] != kInf && rt[2 * q] != rt[2 * q + 1];
        cand[q] = keep[q];
      }
      // a kept key's ends point at their roots: the next round's hook
      // reads them with one load
#pragma unroll
      for (int i = 0; i < 2 * Q; ++i) {
        if (keep[i / 2] && rt[i] != p[i]) P[x[i]] = rt[i];
        x[i] = rt[i];
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) *any_cand |= cand[q];
#pragma unroll
    for (int i = 0; i < 2 * Q; ++i)
      wave_min_insert<LOGS>(s_root, s_min, x[i], k[i / 2], cand[i / 2], best);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (ROUTE) {
        const uint64_t mt = __ballot(cand[q]);
        if (cand[q]) tc[out_to + __popcll(mt & below)] = k[q];
        out_to += __popcll(mt);
        const bool stay = keep[q] && !cand[q];
        const uint64_t ms = __ballot(stay);
        if (stay) lc[out + __popcll(ms & below)] = k[q];
        out += __popcll(ms);
      } else {
        const uint64_t m = __ballot(keep[q]);
        if (keep[q]) lc[out + __popcll(m & below)] = k[q];
        out += __popcll(m);
      }
    }
  }
  if (lane == 0) {
    cnt[cc] = out;
    if (ROUTE) cnt_to[cc] = out_to;
    *sum_in += (unsigned long long)n_in;
    *sum_out += (unsigned long long)(out + out_to);
  }
}

// Chunks a workgroup takes per batch from its share (b, b + G, b + 2G, ...).
constexpr int kOwnBatch = kWave;

template <int RB>
__global__ __launch_bounds__(RB) void k_round(int r, uint64_t* __restrict__ light,
                                               uint64_t* __restrict__ heavy,
                                               int32_t* __restrict__ cnt_light,
                                               int32_t* __restrict__ cnt_heavy,
                                               int64_t nchunks, RoundCtl* ctl,
                                               const unsigned int* __restrict__ had_edge,
                                               const MstPlan* __restrict__ plan,
                                               int32_t* P, unsigned long long* __restrict__ best,
                                               unsigned int* __restrict__ cand_next, int vb,
                                               int max_rounds,
                                               unsigned long long* __restrict__ rlog) {
  constexpr int LOGS = RB == 1024 ? 12 : 10;
  constexpr int kSlots = 1 << LOGS;
  static_assert(kSlots == 4 * RB, "four table slots per thread");
  __shared__ int32_t s_root[kSlots];
  __shared__ unsigned long long s_min[kSlots];
  __shared__ unsigned long long s_io[2 * (RB / kWave)];
  __shared__ int32_t s_work[kOwnBatch], s_work_n[kOwnBatch];
  __shared__ int s_take, s_nwork;
  const RoundCtl prev = ctl[r & 1];
  const unsigned int had = had_edge[r];
  // The first batch of chunk counts of both lists is read with the control
  // word (one round trip instead of two).
  const int lane = lane_id();
  const int64_t G = gridDim.x;
  const int64_t owned = nchunks > (int64_t)blockIdx.x ? (nchunks - 1 - blockIdx.x) / G + 1 : 0;
  int32_t pre_l = 0, pre_h = 0;
  if (threadIdx.x < kWave && lane < owned) {
    const int64_t cc = blockIdx.x + (int64_t)lane * G;
    pre_l = cnt_light[cc];
    pre_h = cnt_heavy[cc];
  }
  const RoundCtl c = next_ctl(prev, r, had, max_rounds, plan);
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl[(r + 1) & 1] = c;
  if (!c.active) return;
  const bool route = c.phase != prev.phase;
  for (int i = threadIdx.x; i < kSlots; i += RB) {
    s_root[i] = -1;
    s_min[i] = kInf;
  }
  __syncthreads();
  // Workgroup b owns chunks b, b + G, b + 2G, ... (G = grid): the key
  // density of the list varies along it, and interleaving gives every
  // workgroup a sample of the whole list. Per batch of up to 64 owned
  // chunks, the first wavefront lists the non-empty ones in LDS and the
  // wavefronts take them one at a time from an LDS counter, so a wavefront
  // that drew light chunks takes more and empty chunks cost nothing.
  uint64_t* const list = route ? heavy : light;
  int32_t* const cnt = route ? cnt_heavy : cnt_light;
  const uint64_t lim = route ? plan->lim[c.phase] : kInf;
  unsigned long long sum_in = 0, sum_out = 0;
  bool any_cand = false;
  for (int64_t b0 = 0; b0 < owned; b0 += kOwnBatch) {
    if (threadIdx.x < kWave) {
      const int64_t j = b0 + lane;
      const int64_t cc = blockIdx.x + j * G;
      int32_t nk = b0 == 0 ? (route ? pre_h : pre_l) : (j < owned ? cnt[cc] : 0);
      nk = nk < 0 ? 0 : (nk > kGTile ? kGTile : nk);
      const uint64_t m = __ballot(nk > 0);
      if (nk > 0) {
        const int at = __popcll(m & ((1ull << lane) - 1ull));
        s_work[at] = (int32_t)j;
        s_work_n[at] = nk;
      }
      if (lane == 0) {
        s_take = 0;
        s_nwork = __popcll(m);
      }
    }
    __syncthreads();
    for (;;) {
      int w = 0;
      if (lane == 0) w = atomicAdd(&s_take, 1);
      w = __builtin_amdgcn_readfirstlane(w);
      if (w >= s_nwork) break;
      const int64_t cc = blockIdx.x + (int64_t)s_work[w] * G;
      const int32_t n_in = s_work_n[w];
      if (route)
        filter_chunk<true, LOGS>(list, cnt, light, cnt_light, cc, n_in, lim, vb, P, s_root,
                                 s_min, best, &sum_in, &sum_out, &any_cand);
      else
        filter_chunk<false, LOGS>(list, cnt, nullptr, nullptr, cc, n_in, lim, vb, P, s_root,
                                  s_min, best, &sum_in, &sum_out, &any_cand);
    }
    __syncthreads();
  }
  if (__ballot(any_cand) && lane == 0) cand_next[r + 1] = 1u;
  // Flush: one global atomicMin per distinct root of the workgroup. In a
  // routing pass it is skipped when a cached value (an upper bound of the
  // current one) already wins, every thread reading its slots' cached values
  // together (one round trip): there the tables hold thousands of roots that
  // other workgroups have already lowered, and the read saves about three
  // atomics in four. In the other rounds it saves few and costs the round
  // trip (measured on the 1M-vertex k-NN graph: middle light rounds 20-25%
  // shorter without it, the routing pass 7% longer).
  {
    constexpr int J = kSlots / RB;
    int32_t rt[J];
    unsigned long long kk[J], cur[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      rt[j] = s_root[j * RB + threadIdx.x];
      kk[j] = s_min[j * RB + threadIdx.x];
    }
#pragma unroll
    for (int j = 0; j < J; ++j) cur[j] = rt[j] < 0 ? 0ull : route ? best[rt[j]] : kInf;
#pragma unroll
    for (int j = 0; j < J; ++j)
      if (rt[j] >= 0 && kk[j] < cur[j]) atomicMin(&best[rt[j]], kk[j]);
  }
  if (rlog && r < kMaxRounds) {  // profiling only: per-round traffic
    const int w = threadIdx.x / kWave;
    if (lane == 0) { s_io[2 * w] = sum_in; s_io[2 * w + 1] = sum_out; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 1; i < RB / kWave; ++i) { sum_in += s_io[2 * i]; sum_out += s_io[2 * i + 1]; }
      const int slot = blockIdx.x & (kLogSlots - 1);
      if (sum_in) {
        atomicAdd(&rlog[(2 * r) * kLogSlots + slot], sum_in);
        atomicAdd(&rlog[(2 * r + 1) * kLogSlots + slot], sum_out);
      }
    }
  }
}

// Every root v whose slot holds a candidate k (its component's lightest edge
// to another component) joins the component of the other side s. The
// candidates of one round form a forest of pick trees whose only cycle is a
// mirror pair (two roots that picked the same key, mst.py:282-305: the
// smaller id keeps no edge and stays the root). v walks its tree along the
// picks (v -> s -> s's pick -> ...) to that pair, points its parent straight
// at the tree's root and records k in its forest slot (a root hooks at most
// once in its lifetime, so forest slots are written once and their order
// does not depend on timing). Slots of round r are read-only here (the next
// round's slots are cleared instead).
// The ends of every candidate key point straight at their roots when the
// round starts (the filter that listed the key wrote them back), and only
// roots' parents change here, each written once, by its own thread, with
// its tree's root; so P[x] of a key end x reads either x's root or that
// root's tree root, which lies in the same tree: a walk that reads the
// latter jumps ahead and still ends at the pair, and v's own side always
// reads v. Stale reads only make a walk longer. A walk longer than n can
// only come from a graph passed as symmetric that is not (flag 4).
// A round whose previous filter listed no candidate hooks nothing and
// flattens instead: every vertex walks its parent chain (vertices that sat
// in no light key since an earlier round lag behind) and points at its
// root, so the routing pass that follows finds every root with one load.
__global__ __launch_bounds__(kGB) void k_hook(int r, int32_t n, const RoundCtl* __restrict__ ctl,
                                              const unsigned long long* __restrict__ best,
                                              unsigned long long* __restrict__ best_next,
                                              int32_t* P, uint64_t* __restrict__ mst_key,
                                              unsigned int* __restrict__ had_edge,
                                              const unsigned int* __restrict__ cand, int vb,
                                              unsigned int* __restrict__ overflow) {
  const int64_t v0 = (int64_t)blockIdx.x * kGB + threadIdx.x;
  const uint64_t k0 = v0 < n ? best[v0] : kInf;  // read with the control word
  const unsigned int has_cand = r == 0 ? 1u : cand[r];
  if (!ctl[r & 1].active) return;
  if (!has_cand) {
    for (int64_t v = v0; v < n; v += (int64_t)gridDim.x * kGB) {
      best_next[v] = kInf;
      int32_t x = P[v];
      const int32_t p0 = x;
      for (int32_t steps = 0;; ++steps) {
        const int32_t p = P[x];
        if (p == x) break;
        x = p;
        if (steps > n) {
          atomicOr(overflow, 4u);
          break;
        }
      }
      if (x != p0) P[v] = x;
    }
    return;
  }
  bool any = false;
  for (int64_t v = v0; v < n; v += (int64_t)gridDim.x * kGB) {
    best_next[v] = kInf;
    const uint64_t k = v == v0 ? k0 : best[v];
    if (k == kInf) continue;
    any = true;
    int32_t a, b;
    unpack_ends(k, vb, &a, &b);
    // the other side (round 0: every vertex is a root, the ends are roots)
    int32_t t;
    if (r == 0) {
      t = a == (int32_t)v ? b : a;
    } else {
      const int32_t pa = P[a], pb = P[b];
      t = pa == (int32_t)v ? pb : pa;
    }
    if (t == (int32_t)v) continue;
    int32_t y = (int32_t)v, f;
    uint64_t ky = k;
    for (int32_t steps = 0;; ++steps) {
      const uint64_t kt = best[t];
      const int32_t pt = P[t];
      if (kt == ky) {  // mirror pair: the smaller id is the root
        f = y < t ? y : t;
        break;
      }
      if (pt != t || kt == kInf) {  // t's walk already ended (or t does not hook)
        f = pt;
        break;
      }
      int32_t a2, b2;
      unpack_ends(kt, vb, &a2, &b2);
      y = t;
      ky = kt;
      if (r == 0) {
        t = a2 == y ? b2 : a2;
      } else {
        const int32_t qa = P[a2], qb = P[b2];
        if (qa == qb) {  // both ends read the tree's root
          f = qa;
          break;
        }
        t = qa == y ? qb : qa;
      }
      if (steps > n) {
        atomicOr(overflow, 4u);
        f = y;
        break;
      }
    }
    if (f == (int32_t)v) continue;  // v is the tree's root
    P[v] = f;
    mst_key[v] = k;
  }
  if (__ballot(any) && lane_id() == 0) had_edge[r] = 1u;
}

// --------------------------------------------------------- forest output
// Weight whose word (word_of) is `word` (exact except that word_of maps
// -0.0f and +0.0f to one word; callers read a zero float from the graph).
template <typename W>
__device__ __forceinline__ W weight_of_word(uint64_t word) {
  if (std::is_same<W, float>::value) {
    uint32_t u = (uint32_t)word;
    u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    const float f = __uint_as_float(u);
    W w;
    memcpy(&w, &f, sizeof(W));
    return w;
  } else if (std::is_signed<W>::value) {
    return (W)(int64_t)(word ^ (1ull << 63));
  } else {
    return (W)word;
  }
}

// First entry of row [beg, end) with column `col` (and, with CHECK_W, weight
// word `wk`) scanned by the kG lanes of a group, 32 entries per step (each
// load instruction reads 8 consecutive entries of 8 rows); *matches gets
// the number of entries with that column seen (every one when the row is
// scanned to its end, i.e. without CHECK_W).
template <typename W, bool CHECK_W, int kG>
__device__ __forceinline__ int64_t row_find(const int32_t* __restrict__ dest,
                                            const W* __restrict__ weight, int64_t beg,
                                            int64_t end, int32_t col, uint64_t wk,
                                            uint64_t word_lo, int sub, int* matches) {
  constexpr int64_t kNone = INT64_MAX;
  int64_t pos = kNone;
  int cnt = 0;
  for (int64_t p0 = beg; p0 < end && (!CHECK_W || pos == kNone); p0 += 4 * kG) {
    int32_t d[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t e = p0 + q * kG + sub;
      d[q] = e < end ? dest[e] : -1;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t e = p0 + q * kG + sub;
      if (d[q] == col) {
        ++cnt;
        if (e < pos && (!CHECK_W || word_of<W>(weight[e]) - word_lo == wk)) pos = e;
      }
    }
#pragma unroll
    for (int off = kG / 2; off > 0; off >>= 1) {
      const int64_t o = __shfl_xor(pos, off, kG);
      pos = o < pos ? o : pos;
    }
  }
#pragma unroll
  for (int off = kG / 2; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, kG);
  *matches = cnt;
  return pos;
}

// CSR position (in the smaller endpoint's row) and weight of each forest
// edge: 8 lanes per edge scan the row, keeping the first entry with the
// edge's column and weight. A symmetric graph's keys come from the smaller
// endpoint's row (round 0 lists the entries with dest > row), so there a
// column that occurs once in the row is the edge: the row's columns are
// scanned alone and the weights are read only for a repeated column; the
// output weight is the key's (a zero float is read from the graph, -0.0f
// and +0.0f sharing a key word). n_f is bounded by *d_nf when given (the
// count left on the device).
template <typename W, bool HAS_OD>
__device__ void positions_part(
    const uint64_t* __restrict__ fkeys, int64_t n_f, const int32_t* __restrict__ d_nf,
    const int32_t* __restrict__ dest, const W* __restrict__ weight,
    const int64_t* __restrict__ fe, const int32_t* __restrict__ od, int vb, int symmetric,
    const MstPlan* __restrict__ plan, int64_t* __restrict__ ids, W* __restrict__ wout,
    unsigned int* __restrict__ missing, int bid, int nblk) {
  const uint64_t word_lo = plan->wlo;
  constexpr int kG = 8;  // lanes per forest edge
  constexpr int64_t kNone = INT64_MAX;
  if (d_nf) {
    const int64_t c = *d_nf;
    n_f = c < n_f ? c : n_f;
  }
  const int sub = threadIdx.x & (kG - 1);
  const int64_t groups = (int64_t)nblk * (kGB / kG);
  for (int64_t i = ((int64_t)bid * kGB + threadIdx.x) / kG; i < n_f; i += groups) {
    const uint64_t k = fkeys[i];
    int32_t a, b;
    unpack_ends(k, vb, &a, &b);
    const uint64_t wk = k >> (2 * vb);
    const int64_t beg = fe[a];
    const int64_t end = HAS_OD ? beg + od[a] : fe[a + 1];
    int matches = 0;
    int64_t p = kNone;
    bool from_key = false;
    if (symmetric) {
      p = row_find<W, false, kG>(dest, weight, beg, end, b, wk, word_lo, sub, &matches);
      from_key = matches == 1;
    }
    if (!from_key)
      p = row_find<W, true, kG>(dest, weight, beg, end, b, wk, word_lo, sub, &matches);
    if (sub == 0) {
      if (p == kNone) {
        atomicOr(missing, 1u);
        p = 0;
      }
      if (ids) ids[i] = p;
      if (wout) {
        W w = weight_of_word<W>(wk + word_lo);
        if (!from_key || (std::is_same<W, float>::value && w == W(0))) w = weight[p];
        wout[i] = w;
      }
    }
  }
}

// --------------------------------------------------------- call tail
// The tail of a Borůvka call in five launches, over tiles of kGTile
// vertices (thread t of a tile's block holds its vertices t * kGI ... in
// the blocked layout, so index order is kept by one block scan):
//   count   forest keys per tile; with labels, every vertex's root (parent
//           walk) and first[] = INT32_MAX
//   first   first occurrence of every root (smallest vertex of each
//           component); the last block scans the tile counts into offsets
//           and the forest count
//   write   forest keys in vertex order (slots past the count get ~0); with
//           labels, fr[v] = first[root(v)], the rank of each first vertex
//           inside its tile and the tile's number of first vertices
//   pos     forest positions and weights; the last block scans the tiles'
//           first-vertex counts
//   label   labels[v] = rank of fr[v] among first vertices; block 0 packs
//           the call's read-back record

constexpr int kTailSlots = 256;  // per-block (root, first vertex) table

// In-place exclusive scan of a[0 : len) by one block; *total = the sum.
__device__ void block_scan_inplace(int32_t* __restrict__ a, int64_t len, int32_t* total) {
  __shared__ int32_t s_tmp[kGB / kWave + 1];
  int32_t carry = 0;
  for (int64_t b0 = 0; b0 < len; b0 += kGB) {
    const int64_t i = b0 + threadIdx.x;
    const int32_t v = i < len ? a[i] : 0;
    int32_t tot;
    const int32_t pre = block_excl_scan<int32_t>(v, s_tmp, &tot);
    if (i < len) a[i] = carry + pre;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <bool LABELS>
__global__ __launch_bounds__(kGB) void k_tail_count(const uint64_t* __restrict__ mst_key,
                                                    int32_t n, const int32_t* __restrict__ P,
                                                    int32_t* __restrict__ rootv,
                                                    int32_t* __restrict__ first,
                                                    int32_t* __restrict__ fcnt,
                                                    unsigned int* __restrict__ overflow) {
  __shared__ uint32_t s_tmp[kGB / kWave + 1];
  const int64_t t0 = (int64_t)blockIdx.x * kGTile;
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < kGI; ++j) {
    const int64_t i = t0 + j * kGB + threadIdx.x;
    if (i < n) {
      c += mst_key[i] != kInf;
      if (LABELS) {
        int32_t x = P[i];
        for (int32_t steps = 0;; ++steps) {
          const int32_t p = P[x];
          if (p == x) break;
          x = p;
          if (steps > n) {
            atomicOr(overflow, 4u);
            break;
          }
        }
        rootv[i] = x;
        first[i] = INT32_MAX;
      }
    }
  }
  uint32_t total;
  (void)block_excl_scan<uint32_t>(c, s_tmp, &total);
  if (threadIdx.x == 0) fcnt[blockIdx.x] = (int32_t)total;
}

// Grid: the tiles (first occurrences, labels only) + one scan block.
template <bool LABELS>
__global__ __launch_bounds__(kGB) void k_tail_first(const int32_t* __restrict__ rootv,
                                                    int32_t n, int32_t* first,
                                                    int32_t* __restrict__ fcnt, int64_t tiles,
                                                    int32_t* __restrict__ d_nf) {
  if ((int64_t)blockIdx.x == (int64_t)gridDim.x - 1) {
    block_scan_inplace(fcnt, tiles, d_nf);
    return;
  }
  if (!LABELS) return;
  __shared__ int32_t s_key[kTailSlots], s_min[kTailSlots];
  const int64_t t0 = (int64_t)blockIdx.x * kGTile;
  for (int i = threadIdx.x; i < kTailSlots; i += kGB) {
    s_key[i] = -1;
    s_min[i] = INT32_MAX;
  }
  __syncthreads();
  for (int j = 0; j < kGTile; j += kGB) {
    const int64_t i = t0 + j + threadIdx.x;
    const bool ok = i < n;
    const int32_t l = ok ? rootv[i] : -1;
    // lanes with the same root elect their lowest (smallest index) lane
    uint64_t pending = __ballot(ok);
    while (pending) {
      const int leader = __ffsll((unsigned long long)pending) - 1;
      const int32_t ll = __shfl(l, leader, kWave);
      pending &= ~__ballot(ok && l == ll);
      if (lane_id() == leader) {
        uint32_t h = ((uint32_t)ll * 2654435761u) >> 24;  // 8 bits: kTailSlots
        bool done = false;
        for (int p = 0; p < 8 && !done; ++p, h = (h + 1) & (kTailSlots - 1)) {
          const int32_t k = atomicCAS(&s_key[h], -1, ll);
          if (k == -1 || k == ll) {
            atomicMin(&s_min[h], (int32_t)i);
            done = true;
          }
        }
        if (!done) atomicMin(first + ll, (int32_t)i);
      }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < kTailSlots; q += kGB) {
    const int32_t k = s_key[q];
    if (k >= 0) atomicMin(first + k, s_min[q]);
  }
}

template <bool LABELS>
__global__ __launch_bounds__(kGB) void k_tail_write(
    const uint64_t* __restrict__ mst_key, int32_t n, const int32_t* __restrict__ foff,
    const int32_t* __restrict__ d_nf, uint64_t* __restrict__ fkeys,
    const int32_t* __restrict__ rootv, const int32_t* __restrict__ first,
    int32_t* __restrict__ fr, int32_t* __restrict__ rloc, int32_t* __restrict__ lcnt) {
  __shared__ uint32_t s_tmp[kGB / kWave + 1];
  const int64_t t0 = (int64_t)blockIdx.x * kGTile;
  const int64_t base = t0 + (int64_t)threadIdx.x * kGI;
  uint64_t v[kGI];
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < kGI; ++j) {
    const int64_t i = base + j;
    v[j] = i < n ? mst_key[i] : kInf;
    c += v[j] != kInf;
  }
  int32_t f[kGI];
  uint32_t cl = 0;
  if (LABELS) {
#pragma unroll
    for (int j = 0; j < kGI; ++j) {
      const int64_t i = base + j;
      f[j] = i < n ? first[rootv[i]] : -1;
      cl += f[j] == (int32_t)i;
    }
  }
  const int32_t nf = *d_nf;
  uint32_t total;
  const uint32_t pre = block_excl_scan<uint32_t>(c, s_tmp, &total);
  int64_t o = (int64_t)foff[blockIdx.x] + pre;
#pragma unroll
  for (int j = 0; j < kGI; ++j)
    if (v[j] != kInf) fkeys[o++] = v[j];
  // slots [t0, t0 + kGTile) past the forest count hold no key
#pragma unroll
  for (int j = 0; j < kGI; ++j) {
    const int64_t q = t0 + j * kGB + threadIdx.x;
    if (q >= nf && q < n) fkeys[q] = kInf;
  }
  if (LABELS) {
    uint32_t tl;
    uint32_t r = block_excl_scan<uint32_t>(cl, s_tmp, &tl);
#pragma unroll
    for (int j = 0; j < kGI; ++j) {
      const int64_t i = base + j;
      if (i >= n) continue;
      fr[i] = f[j];
      if (f[j] == (int32_t)i) rloc[i] = (int32_t)r++;
    }
    if (threadIdx.x == 0) lcnt[blockIdx.x] = (int32_t)tl;
  }
}

// Grid: position blocks (none when no positions are wanted) + one block
// that scans the tiles' first-vertex counts (labels only).
template <typename W, bool HAS_OD>
__global__ __launch_bounds__(kGB) void k_tail_pos(
    const uint64_t* __restrict__ fkeys, int64_t n_f, const int32_t* __restrict__ d_nf,
    const int32_t* __restrict__ dest, const W* __restrict__ weight,
    const int64_t* __restrict__ fe, const int32_t* __restrict__ od, int vb, int symmetric,
    const MstPlan* __restrict__ plan, int64_t* __restrict__ ids, W* __restrict__ wout,
    unsigned int* __restrict__ missing, int32_t* __restrict__ lcnt, int64_t tiles,
    int32_t* __restrict__ d_ncomp) {
  if (lcnt && blockIdx.x == gridDim.x - 1) {
    block_scan_inplace(lcnt, tiles, d_ncomp);
    return;
  }
  positions_part<W, HAS_OD>(fkeys, n_f, d_nf, dest, weight, fe, od, vb, symmetric, plan, ids,
                            wout, missing, blockIdx.x, gridDim.x - (lcnt ? 1 : 0));
}

// ----------------------------------------------------- union-find kernels
__global__ void k_iota(int32_t* p, int32_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (int32_t)i;
}

// Unites the endpoints of keys[0 : nk], or of keys[0 : min(nk, *d_cnt)]
// when the count lives on the device.
__global__ __launch_bounds__(kGB) void k_unite_keys(const uint64_t* __restrict__ keys,
                                                    int64_t nk, const int32_t* __restrict__ d_cnt,
                                                    int vb, int32_t* P) {
  if (d_cnt) {
    const int64_t c = *d_cnt;
    nk = c < nk ? c : nk;
  }
  for (int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x; i < nk;
       i += (int64_t)gridDim.x * kGB) {
    int32_t a, b;
    unpack_ends(keys[i], vb, &a, &b);
    uf_unite(P, a, b);
  }
}

template <bool HAS_OD>
__global__ __launch_bounds__(kGB) void k_unite_csr(const int32_t* __restrict__ dest,
                                                   const int64_t* __restrict__ fe,
                                                   const int32_t* __restrict__ od,
                                                   int32_t n, int64_t m,
                                                   const int32_t* __restrict__ chunk_row,
                                                   int64_t nchunks, int32_t* P,
                                                   unsigned int* bad) {
  __shared__ int32_t s_beg[kMaxRows], s_end[kMaxRows];
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    ChunkCtx cx = stage_chunk<HAS_OD>(c, m, fe, od, chunk_row, s_beg, s_end);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kGI; ++j) {
      int64_t e = cx.e0 + j * kGB + threadIdx.x;
      if (e >= cx.e1) continue;
      int32_t ri = row_of<HAS_OD>(cx, e, fe, od, s_beg, s_end);
      if (ri < 0) continue;
      int32_t u = cx.r_lo + ri;
      int32_t v = dest[e];
      if (v < 0 || v >= n) {
        atomicOr(bad, 1u);
        continue;
      }
      if (v != u) uf_unite(P, u, v);
    }
    __syncthreads();
  }
}

// Root of every vertex into a separate array (storing it back into P would
// race with other threads' path-halving stores, which may overwrite it with
// a non-root ancestor), plus root flags as int32 0/1 for the scan.
__global__ void k_uf_flatten(int32_t* P, int32_t n, int32_t* __restrict__ root,
                             int32_t* __restrict__ flag) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int32_t r = uf_find(P, (int32_t)i);
    root[i] = r;
    flag[i] = (r == (int32_t)i);
  }
}

// Roots of the Borůvka parents (which need not satisfy parent <= child):
// bounded walks, flag 4 on a cycle.
__global__ void k_gather_rank(const int32_t* __restrict__ root,
                              const int32_t* __restrict__ rank, int32_t n,
                              int32_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = rank[root[i]];
}

// First occurrence of each label value. Lanes of a wave hold consecutive
// indices, so per distinct label in the wave only its lowest lane issues the
// atomicMin: a giant cluster costs one atomic per wave instead of one per
// vertex (same-address atomics from all XCDs serialise).
// first[l] = smallest i with lab[i] == l. Lanes of a wave with the same
// label elect their lowest lane (the smallest index); leaders merge into a
// per-block LDS table of (label, min index), flushed with one global
// atomicMin per distinct label of the tile, so a label held by most
// vertices costs one global atomic per tile instead of one per wave.
constexpr int kFoTile = 2048, kFoSlots = 256;
__global__ __launch_bounds__(256) void k_first_occ(const int32_t* __restrict__ lab, int32_t n,
                                                   int32_t nv, int32_t* first,
                                                   unsigned int* bad) {
  __shared__ int32_t s_key[kFoSlots], s_min[kFoSlots];
  for (int64_t t0 = (int64_t)blockIdx.x * kFoTile; t0 < n; t0 += (int64_t)gridDim.x * kFoTile) {
    for (int i = threadIdx.x; i < kFoSlots; i += blockDim.x) {
      s_key[i] = -1;
      s_min[i] = INT32_MAX;
    }
    __syncthreads();
    for (int j = 0; j < kFoTile; j += blockDim.x) {
      const int64_t i = t0 + j + threadIdx.x;
      const int32_t l = i < n ? lab[i] : -1;
      const bool ok = i < n && l >= 0 && l < nv;
      if (i < n && !ok) atomicOr(bad, 1u);
      uint64_t pending = __ballot(ok);
      while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int32_t ll = __shfl(l, leader, kWave);
        pending &= ~__ballot(ok && l == ll);
        if (lane_id() == leader) {
          uint32_t h = ((uint32_t)ll * 2654435761u) >> 24;  // 8 bits: kFoSlots
          bool done = false;
          for (int p = 0; p < 8 && !done; ++p, h = (h + 1) & (kFoSlots - 1)) {
            const int32_t k = atomicCAS(&s_key[h], -1, ll);
            if (k == -1 || k == ll) {
              atomicMin(&s_min[h], (int32_t)i);
              done = true;
            }
          }
          if (!done) atomicMin(first + ll, (int32_t)i);
        }
      }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kFoSlots; q += blockDim.x) {
      const int32_t k = s_key[q];
      if (k >= 0) atomicMin(first + k, s_min[q]);
    }
    __syncthreads();
  }
}
__global__ void k_first_flag(const int32_t* __restrict__ lab, const int32_t* __restrict__ first,
                             int32_t n, int32_t nv, int32_t* __restrict__ root,
                             int32_t* __restrict__ flag) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int32_t l = lab[i];
    int32_t f = (l >= 0 && l < nv) ? first[l] : (int32_t)i;
    root[i] = f;
    flag[i] = (f == (int32_t)i);
  }
}

// ---------------------------------------------------------------- helpers
inline int persistent_blocks() { return 256 * 8; }

// Grid of a persistent (grid-stride) kernel: every block resident at once,
// i.e. compute units x the blocks per unit that the kernel's registers and
// LDS allow (the runtime's occupancy query), cached per kernel. A larger
// grid runs in several block waves, the last one mostly idle.
template <typename K>
int resident_grid(K kernel, int block) {
  static std::mutex mu;
  static std::unordered_map<const void*, int> cache;
  const void* key = reinterpret_cast<const void*>(kernel);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  int dev = 0, cus = 256, per = 1;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0) != hipSuccess ||
      per < 1)
    per = 1;
  const int g = cus * per;
  std::lock_guard<std::mutex> lk(mu);
  cache[key] = g;
  return g;
}

int d2h_sync(void* host, const void* dev, size_t bytes, hipStream_t s) {
  void* pin = pinned_scratch(bytes);
  if (!pin) return fail(EAC_ERR_HIP, "pinned host allocation failed");
  EAC_HIP_CHECK(hipMemcpyAsync(pin, dev, bytes, hipMemcpyDeviceToHost, s));
  EAC_HIP_CHECK(hipStreamSynchronize(s));
  memcpy(host, pin, bytes);
  return EAC_OK;
}

int weight_bytes(int32_t dt) {
  switch (dt) {
    case EAC_W_I32: case EAC_W_F32: case EAC_W_U32: return 4;
    case EAC_W_U16: case EAC_W_I16: return 2;
    case EAC_W_U8: return 1;
    case EAC_W_I64: return 8;
    default: return 0;
  }
}

// Device state counters of one MST call.
struct MstState {
  unsigned int overflow;
  unsigned int missing;
  MstPlan plan;
  RoundCtl ctl[2];
  unsigned int had_edge[kMaxRounds + 1];
  unsigned long long tot0[2][kLogSlots];                 // round-0 light / heavy keys
  unsigned long long rlog[kMaxRounds][2][kLogSlots];     // per round: keys in / kept
  unsigned int cand[kMaxRounds + 1];  // cand[r]: filter(r - 1) listed a candidate
};
constexpr size_t kStateHead = sizeof(MstState);
constexpr size_t kCtlHead = offsetof(MstState, tot0);  // what the round loop reads

inline unsigned long long fold_slots(const unsigned long long* v) {
  unsigned long long t = 0;
  for (int i = 0; i < kLogSlots; ++i) t += v[i];
  return t;
}

struct MstBuffers {
  int32_t* chunk_row;
  unsigned long long* best[2];  // candidate slots, by round parity
  uint64_t* mst_key;
  int32_t* comp;
  int32_t* hook;
  int32_t* roots[2];     // scratch int32[n] pairs used after the rounds
  int64_t nchunks;       // CSR chunks = list chunks
  uint64_t* light;       // chunk-local lists: chunk c owns [c * kGTile, (c+1) * kGTile)
  uint64_t* heavy;
  int32_t* cnt_light;    // keys held by each chunk
  int32_t* cnt_heavy;
  uint64_t* fkeys;
  int32_t* tile_cnt;
  int32_t* flag;
  int32_t* rank;
  unsigned long long* rng;      // per-block weight ranges (2 x kRangeBlocks)
  unsigned long long* samples;  // the plan's weight samples (kPlanSamples)
  MstState* st;
};

MstBuffers carve_mst(Arena& ar, int32_t n, int64_t m, int symmetric) {
  (void)symmetric;
  MstBuffers b;
  const int64_t nchunks = (m + kGTile - 1) / kGTile;
  b.chunk_row = ar.take<int32_t>(nchunks + 1);
  b.best[0] = ar.take<unsigned long long>(n);
  b.best[1] = ar.take<unsigned long long>(n);
  b.mst_key = ar.take<uint64_t>(n);
  b.comp = ar.take<int32_t>(n);
  b.hook = ar.take<int32_t>(n);
  b.roots[0] = ar.take<int32_t>(n);
  b.roots[1] = ar.take<int32_t>(n);
  b.nchunks = nchunks;
  b.light = ar.take<uint64_t>(nchunks * kGTile);
  b.heavy = ar.take<uint64_t>(nchunks * kGTile);
  b.cnt_light = ar.take<int32_t>(nchunks);
  b.cnt_heavy = ar.take<int32_t>(nchunks);
  b.fkeys = ar.take<uint64_t>(n);
  b.tile_cnt = ar.take<int32_t>((n + kGTile - 1) / kGTile + 2);  // + forest count, + labels count
  b.flag = ar.take<int32_t>(n);
  b.rank = ar.take<int32_t>(n);
  b.rng = ar.take<unsigned long long>(2 * kRangeBlocks);
  b.samples = ar.take<unsigned long long>(kPlanSamples);
  b.st = ar.take<MstState>(1);
  return b;
}

// union-find labels over a key list, numbered by first occurrence
int labels_from_keys(const uint64_t* keys, int64_t nk, int vb, int32_t n, int32_t* parent,
                     int32_t* root, int32_t* flag, int32_t* rank, int32_t* labels,
                     int32_t* d_ncomp, Arena& ar, hipStream_t s,
                     const int32_t* d_nk = nullptr) {
  k_iota<<<grid_for(n, 256, 4096), 256, 0, s>>>(parent, n);
  EAC_LAUNCH_CHECK();
  if (nk > 0) {
    k_unite_keys<<<grid_for(nk, kGB, persistent_blocks()), kGB, 0, s>>>(keys, nk, d_nk, vb,
                                                                        parent);
    EAC_LAUNCH_CHECK();
  }
  k_uf_flatten<<<grid_for(n, 256, 4096), 256, 0, s>>>(parent, n, root, flag);
  EAC_LAUNCH_CHECK();
  EAC_TRY((scan_excl<int32_t, int32_t>(flag, rank, n, 0, d_ncomp, ar, s)));
  if (labels) {
    k_gather_rank<<<grid_for(n, 256, 4096), 256, 0, s>>>(root, rank, n, labels);
    EAC_LAUNCH_CHECK();
  }
  return EAC_OK;
}

struct ForestInfo {
  int64_t n_f;      // forest edges (keys in B.fkeys, vertex order); -1 in graph mode
  int vb;           // vertex-id bits of the key
  int wbits;        // weight-key bits
  uint64_t wlo;     // weight word of the lightest entry
  uint64_t whi;     // weight word of the heaviest entry
  int rounds;
  int enq;               // rounds enqueued: the final control state is ctl[enq & 1]
  const int32_t* d_nf;   // device count of forest edges
  const MstPlan* plan;   // device plan (wlo, whi, wbits)
};

// One block: every field starts at zero except the first control word.
__device__ void state_init_block(MstState* st) {
  uint32_t* w = reinterpret_cast<uint32_t*>(st);
  const size_t words = sizeof(MstState) / 4;
  for (size_t i = threadIdx.x; i < words; i += blockDim.x) w[i] = 0u;
  __syncthreads();
  if (threadIdx.x == 0) st->ctl[0].active = 1;
}

inline int vbits_for(int32_t n) {
  int b = bits_for((uint64_t)(n - 1));
  return b < 1 ? 1 : b;
}

// Weight levels of the split (Filter-Kruskal applied to Borůvka): rounds run
// first on the lightest keys, whose forest is exactly the forest's part below
// the first threshold; each further level starts with one routing pass over
// the remaining (heavier) keys against the components so far, which drops
// the internal ones. Only worth it on graphs with many edges per vertex. The
// first threshold keeps about 1.5 light undirected edges per vertex
// (EAC_TUNE_LIGHT_X10): a fraction 3n/m of the entries of a symmetric CSR,
// 1.5n/m of a one-entry-per-edge CSR; each further level takes step (4,
// EAC_TUNE_LEVEL_STEP) times as many, up to kLevels levels.
struct SplitPlan {
  int levels;   // 1: no split
  double frac;  // sampled fraction below the first threshold
  double step;  // growth of that fraction per level
};
inline SplitPlan split_plan(int32_t n, int64_t m, int symmetric) {
  SplitPlan sp{1, 0.0, 1.0};
  const int light_x10 = tune_value(EAC_TUNE_LIGHT_X10) ? tune_value(EAC_TUNE_LIGHT_X10) : 15;
  if (light_x10 <= 0 || m < 6 * (int64_t)n || m < (1 << 16)) return sp;
  sp.frac = (symmetric ? 2.0 : 1.0) * (light_x10 / 10.0) * n / (double)m;
  const int step_knob = tune_value(EAC_TUNE_LEVEL_STEP);
  const int step = step_knob == 0 ? 4 : step_knob;
  sp.step = step > 1 ? (double)step : 1.0;
  const int max_levels = step > 1 ? kLevels : 2;
  double f = sp.frac;
  sp.levels = 1;
  while (sp.levels < max_levels && f < 1.0) {
    ++sp.levels;
    f *= sp.step;
  }
  return sp;
}

// Rounds that always suffice: every round at least halves the components
// still active in a level, and each level ends with one round whose hook
// finds no candidate, so each level (its routing round included) ends
// within bits(n) + 3 rounds.
inline int mst_round_bound(int32_t n, int levels) {
  const int r = levels * (vbits_for(n) + 3);
  return r < kMaxRounds ? r : kMaxRounds;
}

// The independent set-up work of one call in one launch, by block range:
// [0, nr) the weight range (plain CSR only: with out-degrees it needs the
// chunk rows, and runs as its own pass) and the plan's samples; [nr, nr +
// nv) best[0] = ~0, mst_key = ~0, comp = identity (union-find parents);
// [nr + nv, nr + nv + nc) the chunk rows; the last block clears the state.
template <typename W, bool HAS_OD>
__global__ __launch_bounds__(kGB) void k_mst_setup(
    const W* __restrict__ weight, const int64_t* __restrict__ fe, const int32_t* __restrict__ od,
    int32_t n, int64_t m, int64_t nchunks, int nr, int nv, int nc,
    int32_t* __restrict__ chunk_row, unsigned long long* __restrict__ rng,
    unsigned long long* __restrict__ samples, unsigned long long* __restrict__ best,
    uint64_t* __restrict__ mst_key, int32_t* __restrict__ comp, MstState* st) {
  const int b = blockIdx.x;
  if (b < nr) {
    if (!HAS_OD) range_part<W, false>(weight, fe, od, n, m, chunk_row, nchunks, b, nr, rng);
    sample_part<W>(weight, m, b, nr, samples);
  } else if (b < nr + nv) {
    for (int64_t v = (int64_t)(b - nr) * kGB + threadIdx.x; v < n; v += (int64_t)nv * kGB) {
      best[v] = kInf;
      mst_key[v] = kInf;
      comp[v] = (int32_t)v;
    }
  } else if (b < nr + nv + nc) {
    const int64_t c = (int64_t)(b - nr - nv) * kGB + threadIdx.x;
    if (c <= nchunks) chunk_row[c] = chunk_row_of(fe, n, m, c);
  } else {
    state_init_block(st);
  }
}

// Host-side records of one Borůvka call for the library profiler.
struct MstRecs {
  int pr0 = -1;                // round-0 record
  int64_t round0_bytes = 0;
  std::vector<int> hook, round;
};

// Set-up and round 0: state, chunk rows, weight range and plan, round-0
// candidates and the chunk-local light/heavy lists. Nothing is read back.
template <typename W, bool HAS_OD>
int mst_begin(const int32_t* dest, const W* weight, const int64_t* fe, const int32_t* od,
              int32_t n, int64_t m, int symmetric, MstBuffers& B, MstRecs& rec, hipStream_t s) {
  const int64_t nchunks = B.nchunks;
  const int vb = vbits_for(n);
  const int gchunks = grid_for(nchunks, 1, resident_grid(k_round0<W, HAS_OD>, kGB));
  int cap_r = resident_grid(k_weight_range<W, HAS_OD>, kGB);
  cap_r = cap_r < kRangeBlocks ? cap_r : kRangeBlocks;
  const int grange = grid_for(nchunks, 1, cap_r);
  const int nv = grid_for(n, kGB, 4096), nc = grid_for(nchunks + 1, kGB);
  k_mst_setup<W, HAS_OD><<<grange + nv + nc + 1, kGB, 0, s>>>(
      weight, fe, od, n, m, nchunks, grange, nv, nc, B.chunk_row, B.rng, B.samples, B.best[0],
      B.mst_key, B.comp, B.st);
  EAC_LAUNCH_CHECK();
  if (HAS_OD) {
    k_weight_range<W, HAS_OD><<<grange, kGB, 0, s>>>(weight, fe, od, n, m, B.chunk_row, nchunks,
                                                      B.rng);
    EAC_LAUNCH_CHECK();
  }
  const SplitPlan sp = split_plan(n, m, symmetric);
  k_plan<W><<<1, kPlanThreads, 0, s>>>(m, B.rng, grange, B.samples, vb, sp.levels, sp.frac,
                                       sp.step, &B.st->plan);
  EAC_LAUNCH_CHECK();
  rec.round0_bytes = (int64_t)(4 + sizeof(W)) * m + 8 * ((int64_t)n + 1) +
                     (HAS_OD ? 4 * (int64_t)n : 0) + 8 * (int64_t)n;
  rec.pr0 = prof_begin(EAC_PROF_MST_ROUND0, rec.round0_bytes, s);
  k_round0<W, HAS_OD><<<gchunks, kGB, 0, s>>>(dest, weight, fe, od, n, m, B.chunk_row, nchunks,
                                               vb, &B.st->plan, symmetric, B.best[0], B.light,
                                               B.heavy, B.cnt_light, B.cnt_heavy,
                                               prof_on() ? &B.st->tot0[0][0] : nullptr,
                                               &B.st->overflow);
  EAC_LAUNCH_CHECK();
  prof_end(rec.pr0, s);
  return EAC_OK;
}

// Enqueues rounds [r0, r0 + count) (capped at kMaxRounds); returns the new
// number of enqueued rounds. Rounds after convergence exit at once.
int mst_rounds(MstBuffers& B, int32_t n, int r0, int count, MstRecs& rec, hipStream_t s) {
  const int vb = vbits_for(n);
  // one vertex per thread: a thread's walks run one after another, and a
  // deep walk then holds up no other vertex
  const int gh = grid_for(n, kGB);
  // filter: 256-thread workgroups (as many per compute unit as fit, or the
  // knob's count), or with the knob at -1 one 1024-thread workgroup per
  // unit, whose root table is four times larger and whose flush reaches
  // each root from four times fewer workgroups
  const int bpcu = tune_value(EAC_TUNE_FILTER_BPCU);
  const bool big = bpcu < 0;
  int gr_cap = big ? resident_grid(k_round<1024>, 1024) : resident_grid(k_round<kGB>, kGB);
  if (bpcu != 0) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int want = cus * (big ? 1 : bpcu);
    gr_cap = want < gr_cap ? want : gr_cap;
  }
  const int gr = grid_for(B.nchunks, 1, gr_cap);
  int r = r0;
  for (; r < r0 + count && r < kMaxRounds; ++r) {
    rec.hook.push_back(prof_begin(EAC_PROF_MST_HOOK, 0, s));
    k_hook<<<gh, kGB, 0, s>>>(r, n, B.st->ctl, B.best[r & 1], B.best[(r + 1) & 1], B.comp,
                              B.mst_key, B.st->had_edge, B.st->cand, vb, &B.st->overflow);
    EAC_LAUNCH_CHECK();
    prof_end(rec.hook.back(), s);
    rec.round.push_back(prof_begin(EAC_PROF_MST_ROUND, 0, s));
    if (big)
      k_round<1024><<<gr, 1024, 0, s>>>(r, B.light, B.heavy, B.cnt_light, B.cnt_heavy,
                                        B.nchunks, B.st->ctl, B.st->had_edge, &B.st->plan,
                                        B.comp, B.best[(r + 1) & 1], B.st->cand, vb, kMaxRounds,
                                        prof_on() ? &B.st->rlog[0][0][0] : nullptr);
    else
      k_round<kGB><<<gr, kGB, 0, s>>>(r, B.light, B.heavy, B.cnt_light, B.cnt_heavy,
                                      B.nchunks, B.st->ctl, B.st->had_edge, &B.st->plan,
                                      B.comp, B.best[(r + 1) & 1], B.st->cand, vb, kMaxRounds,
                                      prof_on() ? &B.st->rlog[0][0][0] : nullptr);
    EAC_LAUNCH_CHECK();
    prof_end(rec.round.back(), s);
  }
  return r;
}

// Errors recorded on the device, and whether rounds are still running.
int mst_check(const MstState& hs, int enq, int vb, bool* active) {
  if (hs.plan.err)
    return fail(EAC_ERR_KEY_WIDTH,
                "boruvka: weight range needs %d bits and vertex ids 2x%d bits (> 64)",
                hs.plan.wbits, vb);
  if (hs.overflow & 2u) return fail(EAC_ERR_ARG, "boruvka: dest holds ids outside [0, n)");
  if (hs.overflow & 4u)
    return fail(EAC_ERR_ARG, "boruvka: graph passed as symmetric is not symmetric "
                             "(pass symmetric=0)");
  const RoundCtl& c = hs.ctl[enq & 1];
  if (c.overflow_rounds || (c.active && enq >= kMaxRounds))
    return fail(EAC_ERR_HIP, "boruvka: no convergence after %d rounds", kMaxRounds);
  *active = c.active != 0;
  return EAC_OK;
}

// Algorithmic bytes of the profiled launches, from the device counters.
void mst_prof_bytes(const MstState& hs, int enq, int32_t n, int64_t nchunks, MstRecs& rec) {
  if (!prof_on()) return;
  const unsigned long long l0 = fold_slots(hs.tot0[0]), h0 = fold_slots(hs.tot0[1]);
  if (std::getenv("EAC_ROUND_LOG")) {  // diagnostics: keys per round of profiled calls
    std::fprintf(stderr, "[boruvka] n=%d round0: light %llu heavy %llu\n", n, l0, h0);
    for (int r = 0; r < hs.ctl[enq & 1].rounds && r < kMaxRounds; ++r)
      std::fprintf(stderr, "[boruvka] round %d: candidates %u, keys in %llu, kept %llu\n", r,
                   hs.had_edge[r], fold_slots(hs.rlog[r][0]), fold_slots(hs.rlog[r][1]));
  }
  prof_set_bytes(rec.pr0, rec.round0_bytes + 8 * (int64_t)(l0 + h0));
  for (int r = 0; r < (int)rec.hook.size(); ++r) {
    const bool ran = r < hs.ctl[enq & 1].rounds;
    const int64_t in = ran ? (int64_t)fold_slots(hs.rlog[r][0]) : 0;
    const int64_t out = ran ? (int64_t)fold_slots(hs.rlog[r][1]) : 0;
    prof_set_bytes(rec.hook[r], ran ? 8 * (int64_t)n : 0);
    // key 8 B + two 4 B root lookups per key read, 8 B per key kept, counts
    prof_set_bytes(rec.round[r], ran ? 16 * in + 8 * out + 4 * nchunks : 0);
  }
}

// Forest keys in vertex order (slots past the count stay ~0); the count
// goes to *d_nf on the device.
int mst_forest(MstBuffers& B, int32_t n, int32_t* d_nf, hipStream_t s) {
  const int64_t ftiles = (n + kGTile - 1) / kGTile;
  k_tail_count<false><<<(int)ftiles, kGB, 0, s>>>(B.mst_key, n, B.comp, nullptr, nullptr,
                                                  B.tile_cnt, &B.st->overflow);
  EAC_LAUNCH_CHECK();
  k_tail_first<false><<<1, kGB, 0, s>>>(nullptr, n, nullptr, B.tile_cnt, ftiles, d_nf);
  EAC_LAUNCH_CHECK();
  k_tail_write<false><<<(int)ftiles, kGB, 0, s>>>(B.mst_key, n, B.tile_cnt, d_nf, B.fkeys,
                                                  nullptr, nullptr, nullptr, nullptr, nullptr);
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}

inline int first_batch(int32_t n) {
  // about the rounds a graph of this size needs (idle rounds cost a few
  // microseconds, a host read tens)
  if (tune_value(EAC_TUNE_FIRST_BATCH) > 0) return tune_value(EAC_TUNE_FIRST_BATCH);
  const int vb = vbits_for(n);
  return vb * 3 / 4 > 6 ? vb * 3 / 4 : 6;
}

// Boruvka rounds up to the forest key list (shared by the lifetime cut and
// batched entry points). graph_rounds > 0 selects graph mode: no host
// synchronisation at all (that many device-controlled rounds are enqueued,
// rounds after convergence exit at once, counts stay on the device), so the
// sequence can be captured in a hipGraph.
template <typename W, bool HAS_OD>
int mst_core(const int32_t* dest, const W* weight, const int64_t* fe, const int32_t* od,
             int32_t n, int64_t m, int symmetric, MstBuffers& B, Arena& ar, ForestInfo* fi,
             hipStream_t s, int graph_rounds = 0) {
  const bool graph_mode = graph_rounds > 0;
  const int vb = vbits_for(n);
  MstRecs rec;
  EAC_TRY((mst_begin<W, HAS_OD>(dest, weight, fe, od, n, m, symmetric, B, rec, s)));
  const int fixed_rounds = graph_rounds < kMaxRounds ? graph_rounds : kMaxRounds;
  int enq = 0, batch = graph_mode ? fixed_rounds : first_batch(n);
  MstState hs;
  for (;;) {
    enq = mst_rounds(B, n, enq, batch, rec, s);
    if (graph_mode) break;
    EAC_TRY(d2h_sync(&hs, B.st, prof_on() ? kStateHead : kCtlHead, s));
    bool active = false;
    EAC_TRY(mst_check(hs, enq, vb, &active));
    if (!active) break;
    batch = 4;
  }
  if (!graph_mode) mst_prof_bytes(hs, enq, n, B.nchunks, rec);
  const int64_t ftiles = (n + kGTile - 1) / kGTile;
  int32_t* d_nf = B.tile_cnt + ftiles;
  EAC_TRY(mst_forest(B, n, d_nf, s));
  fi->vb = vb;
  fi->d_nf = d_nf;
  fi->plan = &B.st->plan;
  fi->enq = enq;
  if (graph_mode) {
    fi->n_f = -1;
    fi->wbits = 64 - 2 * vb;  // upper bound: sort passes must be host-known
    fi->wlo = 0;
    fi->whi = 0;
    fi->rounds = fixed_rounds;
    return EAC_OK;
  }
  int32_t n_f = 0;
  EAC_TRY(d2h_sync(&n_f, d_nf, sizeof(int32_t), s));
  fi->n_f = n_f;
  fi->wbits = hs.plan.wbits;
  fi->wlo = hs.plan.wlo;
  fi->whi = hs.plan.whi;
  fi->rounds = hs.ctl[enq & 1].rounds;
  return EAC_OK;
}

// Everything a speculative Borůvka call needs to read back, packed by one
// kernel: forest count, lookup flag, error flags, weight bits, control.
struct MstTail {
  int32_t n_f, missing, overflow, plan_err, wbits, active, overflow_rounds, rounds;
};
__device__ void pack_tail(const MstState* __restrict__ st, int enq,
                          const int32_t* __restrict__ d_nf, MstTail* __restrict__ out) {
  const RoundCtl c = st->ctl[enq & 1];
  out->n_f = *d_nf;
  out->missing = (int32_t)st->missing;
  out->overflow = (int32_t)st->overflow;
  out->plan_err = st->plan.err;
  out->wbits = st->plan.wbits;
  out->active = c.active;
  out->overflow_rounds = c.overflow_rounds;
  out->rounds = c.rounds;
}

__global__ void k_pack_tail(const MstState* __restrict__ st, int enq,
                            const int32_t* __restrict__ d_nf, MstTail* __restrict__ out) {
  pack_tail(st, enq, d_nf, out);
}

__global__ __launch_bounds__(kGB) void k_tail_label(const int32_t* __restrict__ fr,
                                                    const int32_t* __restrict__ rloc,
                                                    const int32_t* __restrict__ lpre, int32_t n,
                                                    int32_t* __restrict__ labels,
                                                    const MstState* __restrict__ st, int enq,
                                                    const int32_t* __restrict__ d_nf,
                                                    MstTail* __restrict__ tail) {
  for (int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kGB) {
    const int32_t u = fr[i];
    labels[i] = lpre[u / kGTile] + rloc[u];
  }
  if (tail && blockIdx.x == 0 && threadIdx.x == 0) pack_tail(st, enq, d_nf, tail);
}

// The forest keys, their positions and weights, first-occurrence component
// labels and the read-back record, enqueued after the rounds (nothing is
// read back here).
template <typename W, bool HAS_OD>
int mst_tail(const int32_t* dest, const W* weight, const int64_t* fe, const int32_t* od,
             int32_t n, int symmetric, MstBuffers& B, int32_t* d_nf, int64_t* ids, W* wout,
             int32_t* labels, int enq, MstTail* tail, hipStream_t s) {
  const int vb = vbits_for(n);
  const int64_t ftiles = (n + kGTile - 1) / kGTile;
  int32_t* rootv = B.roots[1];
  int32_t* first = B.roots[0];
  int32_t* fr = B.hook;
  int32_t* lcnt = B.flag;
  int32_t* d_ncomp = B.tile_cnt + ftiles + 1;  // component count after the forest count
  const bool pos = n > 1 && (ids || wout);
  // the forest count stays on the device: positions run over the bound n-1
  const int gpos = pos ? grid_for(8 * (int64_t)(n - 1), kGB,
                                  resident_grid(k_tail_pos<W, HAS_OD>, kGB))
                       : 0;
  if (labels) {
    k_tail_count<true><<<(int)ftiles, kGB, 0, s>>>(B.mst_key, n, B.comp, rootv, first,
                                                   B.tile_cnt, &B.st->overflow);
    EAC_LAUNCH_CHECK();
    k_tail_first<true><<<(int)ftiles + 1, kGB, 0, s>>>(rootv, n, first, B.tile_cnt, ftiles,
                                                       d_nf);
    EAC_LAUNCH_CHECK();
    k_tail_write<true><<<(int)ftiles, kGB, 0, s>>>(B.mst_key, n, B.tile_cnt, d_nf, B.fkeys,
                                                   rootv, first, fr, B.rank, lcnt);
    EAC_LAUNCH_CHECK();
    k_tail_pos<W, HAS_OD><<<gpos + 1, kGB, 0, s>>>(B.fkeys, n - 1, d_nf, dest, weight, fe, od,
                                                   vb, symmetric, &B.st->plan, ids, wout,
                                                   &B.st->missing, lcnt, ftiles, d_ncomp);
    EAC_LAUNCH_CHECK();
    k_tail_label<<<grid_for(n, kGB, 4096), kGB, 0, s>>>(fr, B.rank, lcnt, n, labels, B.st, enq,
                                                        d_nf, tail);
    EAC_LAUNCH_CHECK();
  } else {
    EAC_TRY(mst_forest(B, n, d_nf, s));
    if (pos) {
      k_tail_pos<W, HAS_OD><<<gpos, kGB, 0, s>>>(B.fkeys, n - 1, d_nf, dest, weight, fe, od, vb,
                                                 symmetric, &B.st->plan, ids, wout,
                                                 &B.st->missing, nullptr, 0, nullptr);
      EAC_LAUNCH_CHECK();
    }
    k_pack_tail<<<1, 1, 0, s>>>(B.st, enq, d_nf, tail);
    EAC_LAUNCH_CHECK();
  }
  return EAC_OK;
}

// eac_boruvka_mst. Speculative: the first batch of rounds (usually all a
// graph of this size needs), the forest keys, positions and labels are all
// enqueued before the one read-back; if the rounds had not converged, the
// remaining rounds run with a read per batch and the outputs are redone.
template <typename W, bool HAS_OD>
int mst_impl(const int32_t* dest, const W* weight, const int64_t* fe, const int32_t* od,
             int32_t n, int64_t m, int symmetric, int64_t* ids, W* wout, int32_t* labels,
             int64_t* h_out, void* ws, size_t ws_bytes, hipStream_t s) {
  Arena ar(ws, ws_bytes);
  MstBuffers B = carve_mst(ar, n, m, symmetric);
  MstTail* tail = ar.take<MstTail>(1);
  if (ar.overflow()) return fail(EAC_ERR_WORKSPACE, "boruvka: workspace too small");
  const size_t mark = ar.used;
  const int vb = vbits_for(n);
  const int64_t ftiles = (n + kGTile - 1) / kGTile;
  int32_t* d_nf = B.tile_cnt + ftiles;
  MstRecs rec;
  EAC_TRY((mst_begin<W, HAS_OD>(dest, weight, fe, od, n, m, symmetric, B, rec, s)));
  int enq = mst_rounds(B, n, 0, first_batch(n), rec, s);
  MstTail ht;
  for (int attempt = 0;; ++attempt) {
    ar.used = mark;
    int pr = prof_begin(EAC_PROF_MST_FINAL, 16 * (int64_t)n, s);
    EAC_TRY((mst_tail<W, HAS_OD>(dest, weight, fe, od, n, symmetric, B, d_nf, ids, wout, labels,
                                 enq, tail, s)));
    prof_end(pr, s);
    EAC_TRY(d2h_sync(&ht, tail, sizeof(ht), s));
    if (ht.plan_err)
      return fail(EAC_ERR_KEY_WIDTH,
                  "boruvka: weight range needs %d bits and vertex ids 2x%d bits (> 64)",
                  ht.wbits, vb);
    if (ht.overflow & 2) return fail(EAC_ERR_ARG, "boruvka: dest holds ids outside [0, n)");
    if (ht.overflow & 4)
      return fail(EAC_ERR_ARG, "boruvka: graph passed as symmetric is not symmetric "
                               "(pass symmetric=0)");
    if (ht.overflow_rounds)
      return fail(EAC_ERR_HIP, "boruvka: no convergence after %d rounds", kMaxRounds);
    if (!ht.active) break;
    if (attempt > 0 || enq >= kMaxRounds)
      return fail(EAC_ERR_HIP, "boruvka: no convergence after %d rounds", enq);
    // not converged yet: finish the rounds with a read per batch, then redo
    MstState hs;
    for (;;) {
      enq = mst_rounds(B, n, enq, 4, rec, s);
      EAC_TRY(d2h_sync(&hs, B.st, kCtlHead, s));
      bool active = false;
      EAC_TRY(mst_check(hs, enq, vb, &active));
      if (!active) break;
    }
  }
  if (prof_on()) {
    MstState hs;
    EAC_TRY(d2h_sync(&hs, B.st, kStateHead, s));
    mst_prof_bytes(hs, enq, n, B.nchunks, rec);
  }
  prof_flush();
  if (ht.missing) return fail(EAC_ERR_HIP, "boruvka: forest edge not found in its row");
  const int64_t n_f = ht.n_f;
  if (h_out) {
    h_out[0] = n_f;
    h_out[1] = n - n_f;
    h_out[2] = ht.rounds;
    h_out[3] = ht.wbits;
  }
  return EAC_OK;
}

// ------------------------------------------------------------ lifetime cuts
// Numeric weight of a weight word (inverse of word_of).
template <typename W>
__device__ __forceinline__ double weight_value(uint64_t word) {
  if (std::is_same<W, float>::value) {
    uint32_t u = (uint32_t)word;
    u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    return (double)__uint_as_float(u);
  } else if (std::is_signed<W>::value) {
    return (double)(int64_t)(word ^ (1ull << 63));
  } else {
    return (double)word;
  }
}

struct CutState {
  unsigned long long max_lt_bits;   // bits of the largest lifetime (>= 0)
  unsigned long long min_lt_bits;   // bits of the smallest lifetime (>= 0)
  unsigned long long argmax;        // first index of the largest lifetime
  unsigned long long zero_w;        // forest edges of weight exactly 0
  long long keep;                   // forest edges kept (prefix of the sorted order)
  long long nc_stable;
  long long err;                    // 1: the reference raises ValueError
  long long labels_empty;
  double th;
};

// Sorted forest weights w[i]; lifetimes w[i+1]-w[i] are >= 0, so their
// IEEE bit patterns order like the values.
template <typename W>
__global__ __launch_bounds__(kGB) void k_lifetime_stats(const uint64_t* __restrict__ sk,
                                                        int64_t nf, int vb, uint64_t wlo,
                                                        CutState* cs) {
  __shared__ unsigned long long s_mx[kGB / kWave], s_mn[kGB / kWave], s_z[kGB / kWave];
  unsigned long long mx = 0, mn = ~0ull, z = 0;
  for (int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x; i < nf;
       i += (int64_t)gridDim.x * kGB) {
    double wi = weight_value<W>((sk[i] >> (2 * vb)) + wlo);
    z += wi == 0.0;
    if (i + 1 < nf) {
      double wj = weight_value<W>((sk[i + 1] >> (2 * vb)) + wlo);
      unsigned long long b = (unsigned long long)__double_as_longlong(wj - wi);
      mx = b > mx ? b : mx;
      mn = b < mn ? b : mn;
    }
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    unsigned long long a = __shfl_xor(mx, off, kWave), b = __shfl_xor(mn, off, kWave);
    mx = a > mx ? a : mx;
    mn = b < mn ? b : mn;
    z += __shfl_xor(z, off, kWave);
  }
  if (lane_id() == 0) {
    s_mx[threadIdx.x / kWave] = mx;
    s_mn[threadIdx.x / kWave] = mn;
    s_z[threadIdx.x / kWave] = z;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kGB / kWave; ++i) {
      mx = s_mx[i] > mx ? s_mx[i] : mx;
      mn = s_mn[i] < mn ? s_mn[i] : mn;
      z += s_z[i];
    }
    atomicMax(&cs->max_lt_bits, mx);
    atomicMin(&cs->min_lt_bits, mn);
    if (z) atomicAdd(&cs->zero_w, z);
  }
}

template <typename W>
__global__ __launch_bounds__(kGB) void k_lifetime_argmax(const uint64_t* __restrict__ sk,
                                                         int64_t nf, int vb, uint64_t wlo,
                                                         CutState* cs) {
  const unsigned long long target = cs->max_lt_bits;
  for (int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x; i + 1 < nf;
       i += (int64_t)gridDim.x * kGB) {
    double wi = weight_value<W>((sk[i] >> (2 * vb)) + wlo);
    double wj = weight_value<W>((sk[i + 1] >> (2 * vb)) + wlo);
    if ((unsigned long long)__double_as_longlong(wj - wi) == target)
      atomicMin(&cs->argmax, (unsigned long long)i);
  }
}

// Drop forest edges of weight exactly 0 (Kruskal's eliminate_zeros,
// scipy_numba/.../_min_spanning_tree.py:163-175), keeping the order.
template <typename W>
__global__ __launch_bounds__(kGB) void k_mark_nonzero(const uint64_t* __restrict__ sk,
                                                      int64_t nf, int vb, uint64_t wlo,
                                                      int32_t* __restrict__ flag) {
  for (int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x; i < nf;
       i += (int64_t)gridDim.x * kGB)
    flag[i] = weight_value<W>((sk[i] >> (2 * vb)) + wlo) != 0.0;
}
// Batched form of the same rule: zero-weight forest keys become kInf
// (absent) in place, so no count leaves the device.
template <typename W>
__global__ __launch_bounds__(kGB) void k_drop_zero(uint64_t* __restrict__ keys, int64_t n, int vb,
                                                   const MstPlan* __restrict__ plan) {
  const uint64_t wlo = plan->wlo;
  for (int64_t i = (int64_t)blockIdx.x * kGB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kGB) {
    const uint64_t k = keys[i];
    if (k != kInf && weight_value<W>((k >> (2 * vb)) + wlo) == 0.0) keys[i] = kInf;
  }
}
__global__ void k_compact_flagged(const uint64_t* __restrict__ in, const int32_t* __restrict__ flag,
                                  const int32_t* __restrict__ pos, int64_t n,
                                  uint64_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (flag[i]) out[pos[i]] = in[i];
}

// Decision of eac.py:220-265 (rule 0) or singlelink.py:16-48 (rule 1) for
// one graph from the statistics of its sorted forest weights sk[0:nf].
struct CutOut {
  int64_t keep, nc_stable;
  int err, labels_empty;
  double th;
};
template <typename W>
__device__ CutOut cut_decide(const uint64_t* __restrict__ sk, int64_t nf, int64_t n, int vb,
                             uint64_t wlo, uint64_t gmax_word, int rule, double max_dist,
                             int32_t n_clusters, unsigned long long max_lt_bits,
                             unsigned long long min_lt_bits, unsigned long long argmax) {
  CutOut o{nf, 0, 0, 0, 0.0};
  const double gmax = weight_value<W>(gmax_word);  // largest graph weight
  if (nf < 2 && (rule == 1 || n_clusters == 0)) { o.err = 1; return o; }
  const int64_t a = (int64_t)argmax;
  const double th = nf >= 2 ? __longlong_as_double((long long)max_lt_bits) : 0.0;
  const double wmaxf = nf >= 1 ? weight_value<W>((sk[nf - 1] >> (2 * vb)) + wlo) : 0.0;
  o.th = th;
  if (rule == 1) {
    // singlelink.py:14-48
    double disconnect = max_dist < 0 ? gmax : max_dist;
    double lt_threshold = disconnect - wmaxf;
    o.keep = th >= lt_threshold ? a : nf;
    return o;
  }
  // eac.py:199-267
  double md = max_dist < 0 ? gmax + 1 : max_dist + 1;
  int64_t n_disc = n - nf;
  int64_t nc_stable;
  if (n_clusters == 0) {
    // lifetime_n_clusters, eac.py:366-391
    double wa = weight_value<W>((sk[a] >> (2 * vb)) + wlo);
    int64_t lo = a, hi = nf;  // first index with w > w[a]
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if (weight_value<W>((sk[mid] >> (2 * vb)) + wlo) > wa) hi = mid; else lo = mid + 1;
    }
    int64_t above = nf - lo;
    int64_t cont = above == 0 ? 1 : above + 1;
    double mn = __longlong_as_double((long long)min_lt_bits);
    if (!(mn <= 1e-8)) { o.err = 1; return o; }  // np.min over an empty selection
    if (th < 2 * mn) cont = 1;
    if (n_disc > 1) {
      double disconnect_lifetime = md + 1 - wmaxf;
      if (th > disconnect_lifetime) cont += n_disc - 1;
      else cont = n_disc;
    }
    nc_stable = cont;
  } else {
    nc_stable = n_clusters;
  }
  int64_t keep = nf;
  if (nc_stable > n_disc) {
    keep = nf - (nc_stable - n_disc);
    if (keep < 0) keep = 0;
  }
  o.keep = keep;
  o.nc_stable = nc_stable;
  o.labels_empty = nc_stable > 1 ? 0 : 1;
  return o;
}

template <typename W>
__global__ void k_cut_decide(const uint64_t* __restrict__ sk, int64_t nf, int32_t n, int vb,
                             uint64_t wlo, uint64_t whi, int rule, double max_dist,
                             int32_t n_clusters, CutState* cs) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  CutOut o = cut_decide<W>(sk, nf, n, vb, wlo, whi, rule, max_dist, n_clusters, cs->max_lt_bits,
                           cs->min_lt_bits, cs->argmax);
  cs->err = o.err;
  cs->labels_empty = o.labels_empty;
  cs->keep = o.keep;
  cs->nc_stable = o.nc_stable;
  cs->th = o.th;
}

// ------------------------------------------------- batched single-link cut
// Graph of a forest key (by its smaller endpoint) for disjoint-union batches;
// absent keys (kInf) get the sentinel id G.
__global__ void k_graph_of_key(const uint64_t* __restrict__ sk, int64_t nf, int vb,
                               const int64_t* __restrict__ goff, int G,
                               uint32_t* __restrict__ gid) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t key = sk[i];
    if (key == kInf) {  // dropped key (zero weight) or padding: sentinel graph G
      gid[i] = (uint32_t)G;
      continue;
    }
    int32_t a, b;
    unpack_ends(key, vb, &a, &b);
    int lo = 0, hi = G;
    while (lo + 1 < hi) {
      int mid = (lo + hi) >> 1;
      if (goff[mid] <= a) lo = mid; else hi = mid;
    }
    gid[i] = (uint32_t)lo;
  }
}

// Segment starts of the graph-sorted keys: seg[g] = first i with gid[i] >= g
// (g = 0..G; seg[G] = keys of real graphs). No counters: binary searches.
__global__ void k_seg_bounds(const uint32_t* __restrict__ gid_s, int64_t nf, int G,
                             int64_t* __restrict__ seg) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = nf;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (gid_s[mid] < (uint32_t)g) lo = mid + 1; else hi = mid;
    }
    seg[g] = lo;
  }
}

// Per-graph largest weight word over the graph's CSR entries. Rows of a
// graph are contiguous, so (offsets without gaps) graph g owns entries
// [fe[goff[g]], fe[goff[g+1]]): every block takes an equal entry range and
// folds the part of each graph it overlaps, one atomicMax per (block, graph).
// With explicit out-degrees (rows may have gaps) one block walks each graph.
template <typename W, bool HAS_OD>
__global__ __launch_bounds__(kGB) void k_graph_wmax(const W* __restrict__ weight,
                                                    const int64_t* __restrict__ fe,
                                                    const int32_t* __restrict__ od,
                                                    const int64_t* __restrict__ goff, int G,
                                                    int64_t m,
                                                    unsigned long long* __restrict__ gmax) {
  __shared__ unsigned long long s_m[kGB / kWave];
  if (!HAS_OD) {
    const int64_t per = (m + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = (int64_t)blockIdx.x * per;
    const int64_t b1 = b0 + per < m ? b0 + per : m;
    if (b0 >= b1) return;
    // graph holding entry b0: last g with fe[goff[g]] <= b0
    int lo = 0, hi = G;
    while (lo + 1 < hi) {
      const int mid = (lo + hi) >> 1;
      if (fe[goff[mid]] <= b0) lo = mid; else hi = mid;
    }
    for (int g = lo; g < G; ++g) {  // uniform over the block
      const int64_t gs = fe[goff[g]], ge = fe[goff[g + 1]];
      if (gs >= b1) break;
      const int64_t e0 = gs > b0 ? gs : b0, e1 = ge < b1 ? ge : b1;
      unsigned long long mx = 0;
      for (int64_t q = e0 + threadIdx.x; q < e1; q += kGB) {
        const unsigned long long x = word_of<W>(weight[q]);
        mx = x > mx ? x : mx;
      }
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(mx, off, kWave);
        mx = o > mx ? o : mx;
      }
      if (lane_id() == 0) s_m[threadIdx.x / kWave] = mx;
      __syncthreads();
      if (threadIdx.x == 0) {
        for (int q = 1; q < kGB / kWave; ++q) mx = s_m[q] > mx ? s_m[q] : mx;
        if (e1 > e0) atomicMax(&gmax[g], mx);
      }
      __syncthreads();
    }
    return;
  }
  for (int g = blockIdx.x; g < G; g += gridDim.x) {
    unsigned long long mx = 0;
    const int64_t v0 = goff[g], v1 = goff[g + 1];
    for (int64_t v = v0; v < v1; ++v) {
      int64_t b = fe[v], e = b + od[v];
      for (int64_t q = b + threadIdx.x; q < e; q += kGB) {
        unsigned long long x = word_of<W>(weight[q]);
        mx = x > mx ? x : mx;
      }
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      unsigned long long o = __shfl_xor(mx, off, kWave);
      mx = o > mx ? o : mx;
    }
    if (lane_id() == 0) s_m[threadIdx.x / kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 1; q < kGB / kWave; ++q) mx = s_m[q] > mx ? s_m[q] : mx;
      atomicMax(&gmax[g], mx);
    }
    __syncthreads();
  }
}

// Lifetime statistics and the cut decision of every graph: one block per
// graph over its segment sk[seg[g] : seg[g+1]] (sorted by key).
struct GraphCut {
  long long keep, nc_stable, err, labels_empty;
};
template <typename W>
__global__ __launch_bounds__(kGB) void k_seg_cut(const uint64_t* __restrict__ sk,
                                                 const int64_t* __restrict__ seg,
                                                 const int64_t* __restrict__ goff, int G, int vb,
                                                 const MstPlan* __restrict__ plan,
                                                 const unsigned long long* __restrict__ gmax,
                                                 int rule, double max_dist, int32_t n_clusters,
                                                 GraphCut* __restrict__ out) {
  const uint64_t wlo = plan->wlo;
  __shared__ unsigned long long s_a[kGB / kWave], s_b[kGB / kWave];
  __shared__ unsigned long long s_mx, s_mn, s_arg;
  for (int g = blockIdx.x; g < G; g += gridDim.x) {
    const uint64_t* k = sk + seg[g];
    const int64_t nf = seg[g + 1] - seg[g];
    unsigned long long mx = 0, mn = ~0ull;
    for (int64_t i = threadIdx.x; i + 1 < nf; i += kGB) {
      double wi = weight_value<W>((k[i] >> (2 * vb)) + wlo);
      double wj = weight_value<W>((k[i + 1] >> (2 * vb)) + wlo);
      unsigned long long b = (unsigned long long)__double_as_longlong(wj - wi);
      mx = b > mx ? b : mx;
      mn = b < mn ? b : mn;
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      unsigned long long a = __shfl_xor(mx, off, kWave), b = __shfl_xor(mn, off, kWave);
      mx = a > mx ? a : mx;
      mn = b < mn ? b : mn;
    }
    if (lane_id() == 0) { s_a[threadIdx.x / kWave] = mx; s_b[threadIdx.x / kWave] = mn; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 1; q < kGB / kWave; ++q) {
        mx = s_a[q] > mx ? s_a[q] : mx;
        mn = s_b[q] < mn ? s_b[q] : mn;
      }
      s_mx = mx;
      s_mn = mn;
      s_arg = ~0ull;
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i + 1 < nf; i += kGB) {
      double wi = weight_value<W>((k[i] >> (2 * vb)) + wlo);
      double wj = weight_value<W>((k[i + 1] >> (2 * vb)) + wlo);
      if ((unsigned long long)__double_as_longlong(wj - wi) == s_mx)
        atomicMin(&s_arg, (unsigned long long)i);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      CutOut o = cut_decide<W>(k, nf, goff[g + 1] - goff[g], vb, wlo, gmax[g], rule, max_dist,
                               n_clusters, s_mx, s_mn, s_arg);
      out[g].keep = o.keep;
      out[g].nc_stable = o.nc_stable;
      out[g].err = o.err;
      out[g].labels_empty = o.labels_empty;
    }
    __syncthreads();
  }
}

__global__ void k_seg_keep_flags(const uint32_t* __restrict__ gid, int64_t nf, int G,
                                 const int64_t* __restrict__ seg,
                                 const GraphCut* __restrict__ cut, int32_t* __restrict__ flag) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t g = gid[i];
    flag[i] = g < (uint32_t)G && (i - seg[g]) < cut[g].keep;
  }
}

// per-graph labels from union labels numbered by first occurrence (graphs
// occupy consecutive vertex ranges, so graph g's labels start at the label of
// its first vertex): base[g] = label of the first vertex, base[G] = total.
__global__ void k_graph_base(const int32_t* __restrict__ lab, const int64_t* __restrict__ goff,
                             int G, const int32_t* __restrict__ total,
                             int32_t* __restrict__ base) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += gridDim.x * blockDim.x) {
    if (g == G) { base[g] = *total; continue; }
    base[g] = goff[g + 1] > goff[g] ? lab[goff[g]] : -1;
  }
}
__global__ void k_graph_sub(int32_t* __restrict__ lab, const int64_t* __restrict__ goff, int G,
                            const int32_t* __restrict__ base, const GraphCut* __restrict__ cut,
                            int64_t* __restrict__ gout) {
  for (int g = blockIdx.x; g < G; g += gridDim.x) {
    const int64_t v0 = goff[g], v1 = goff[g + 1];
    const int32_t b = base[g];
    for (int64_t v = v0 + threadIdx.x; v < v1; v += blockDim.x) lab[v] -= b;
    if (threadIdx.x == 0) {
      int32_t nb = -1;
      for (int q = g + 1; q <= G; ++q)
        if (base[q] >= 0) { nb = base[q]; break; }
      gout[4 * g + 0] = v1 > v0 ? (int64_t)(nb - b) : 0;
      gout[4 * g + 1] = cut[g].labels_empty;
      gout[4 * g + 2] = cut[g].nc_stable;
      gout[4 * g + 3] = cut[g].err ? -1 : cut[g].keep;
    }
  }
}

template <typename W, bool HAS_OD>
int sl_impl(const int32_t* dest, const W* weight, const int64_t* fe, const int32_t* od,
            int32_t n, int64_t m, int symmetric, int rule, double max_dist, int32_t n_clusters,
            int32_t* labels, int64_t* h_out, void* ws, size_t ws_bytes, hipStream_t s) {
  Arena ar(ws, ws_bytes);
  MstBuffers B = carve_mst(ar, n, m, symmetric);
  uint64_t* sk = ar.take<uint64_t>(n);
  CutState* cs = ar.take<CutState>(1);
  if (ar.overflow()) return fail(EAC_ERR_WORKSPACE, "sl_lifetime: workspace too small");
  ForestInfo fi;
  EAC_TRY((mst_core<W, HAS_OD>(dest, weight, fe, od, n, m, symmetric, B, ar, &fi, s)));
  int64_t nf = fi.n_f;
  const int vb = fi.vb;
  // canonical (w, min, max) order = ascending key order
  EAC_TRY((radix_sort<uint64_t, NoValue>(B.fkeys, sk, nullptr, nullptr, nf, 0,
                                         fi.wbits + 2 * vb, ar, s)));
  CutState init;
  memset(&init, 0, sizeof(init));
  init.min_lt_bits = ~0ull;
  init.argmax = ~0ull;
  {
    void* pin = pinned_scratch(sizeof(CutState));
    if (!pin) return fail(EAC_ERR_HIP, "pinned host allocation failed");
    EAC_HIP_CHECK(hipStreamSynchronize(s));
    memcpy(pin, &init, sizeof(init));
    EAC_HIP_CHECK(hipMemcpyAsync(cs, pin, sizeof(CutState), hipMemcpyHostToDevice, s));
  }
  const int g = grid_for(nf, kGB, persistent_blocks());
  if (nf > 0) {
    k_lifetime_stats<W><<<g, kGB, 0, s>>>(sk, nf, vb, fi.wlo, cs);
    EAC_LAUNCH_CHECK();
  }
  CutState hc;
  if (rule == 0 && nf > 0) {
    EAC_TRY(d2h_sync(&hc, cs, sizeof(CutState), s));
    if (hc.zero_w > 0) {  // rare: zero weights leave the forest, stats again
      k_mark_nonzero<W><<<g, kGB, 0, s>>>(sk, nf, vb, fi.wlo, B.flag);
      EAC_LAUNCH_CHECK();
      int32_t* d_cnt = B.tile_cnt;
      EAC_TRY((scan_excl<int32_t, int32_t>(B.flag, B.rank, nf, 0, d_cnt, ar, s)));
      k_compact_flagged<<<g, kGB, 0, s>>>(sk, B.flag, B.rank, nf, B.fkeys);
      EAC_LAUNCH_CHECK();
      nf -= (int64_t)hc.zero_w;
      EAC_HIP_CHECK(hipMemcpyAsync(sk, B.fkeys, sizeof(uint64_t) * nf, hipMemcpyDeviceToDevice, s));
      {
        void* pin = pinned_scratch(sizeof(CutState));
        EAC_HIP_CHECK(hipStreamSynchronize(s));
        memcpy(pin, &init, sizeof(init));
        EAC_HIP_CHECK(hipMemcpyAsync(cs, pin, sizeof(CutState), hipMemcpyHostToDevice, s));
      }
      if (nf > 0) {
        k_lifetime_stats<W><<<g, kGB, 0, s>>>(sk, nf, vb, fi.wlo, cs);
        EAC_LAUNCH_CHECK();
      }
    }
  }
  if (nf > 1) {
    k_lifetime_argmax<W><<<g, kGB, 0, s>>>(sk, nf, vb, fi.wlo, cs);
    EAC_LAUNCH_CHECK();
  }
  k_cut_decide<W><<<1, 64, 0, s>>>(sk, nf, n, vb, fi.wlo, fi.whi, rule, max_dist, n_clusters,
                                   cs);
  EAC_LAUNCH_CHECK();
  EAC_TRY(d2h_sync(&hc, cs, sizeof(CutState), s));
  if (hc.err)
    return fail(EAC_ERR_VALUE, rule == 1 ? "sl_mst_lifetime: fewer than 2 forest edges "
                                           "(numpy argmax of an empty array)"
                                         : "sp_sl_lifetime: lifetime rule undefined (fewer "
                                           "than 2 forest edges or no zero lifetime)");
  int64_t n_comp = 1;
  const int64_t ftiles = (n + kGTile - 1) / kGTile;
  if (!hc.labels_empty) {
    EAC_TRY(labels_from_keys(sk, hc.keep, vb, n, B.hook, B.comp, B.flag, B.rank, labels,
                             B.tile_cnt + ftiles, ar, s));
    int32_t nc = 0;
    EAC_TRY(d2h_sync(&nc, B.tile_cnt + ftiles, sizeof(int32_t), s));
    n_comp = nc;
  }
  prof_flush();
  if (h_out) {
    h_out[0] = n_comp;
    h_out[1] = hc.labels_empty;
    h_out[2] = hc.nc_stable;
    h_out[3] = hc.keep;
    h_out[4] = nf;
    h_out[5] = fi.rounds;
  }
  return EAC_OK;
}

// Many independent graphs given as one disjoint union (graph g owns the
// vertex range [goff[g], goff[g+1]) and its rows): one Boruvka run over the
// union (keys on global ids keep each graph's own (w, min, max) order, so
// each graph's forest is its own canonical forest), then per-graph sorted
// forests (global key sort + stable sort by graph), per-graph lifetime cut,
// and one union-find labelling numbered by first occurrence inside each
// graph. gout[g] = {components, labels_empty, nc_stable, kept edges or -1
// where the reference raises ValueError}.
// Graph offsets as the batched kernels read them: each clamped to [0, N]
// (so no kernel indexes outside the union); bad[0] = 1 when the input was
// not 0 = goff[0] <= ... <= goff[G] = N.
__global__ void k_goff_prep(const int64_t* __restrict__ in, int G, int32_t N,
                            int64_t* __restrict__ out, int32_t* __restrict__ bad) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += gridDim.x * blockDim.x) {
    const int64_t v = in[g];
    out[g] = v < 0 ? 0 : (v > N ? N : v);
    const bool ok = (g > 0 || v == 0) && (g < G || v == N) && (g == 0 || in[g - 1] <= v);
    if (!ok) *bad = 1;
  }
}

// Run summary of a batched cut, written on the device: {components,
// forest edges (nonzero weight under rule 0), Boruvka rounds, kept edges,
// error flags}. Flags: 1 edge list overflow, 2 dest ids outside [0, N),
// 4 graph passed as symmetric is not, 8 key wider than 64 bits,
// 16 rounds not converged, 32 bad graph offsets.
__global__ void k_batch_summary(const MstState* __restrict__ st, int enq,
                                const int32_t* __restrict__ d_ncomp,
                                const int64_t* __restrict__ seg_total,
                                const int32_t* __restrict__ misc, int64_t* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  out[0] = *d_ncomp;
  out[1] = *seg_total;
  const RoundCtl c = st->ctl[enq & 1];
  out[2] = c.rounds;
  out[3] = misc[0];
  int64_t flags = st->overflow & 7u;
  if (misc[1]) flags |= 32;
  if (st->plan.err) flags |= 8;
  if (c.active || c.overflow_rounds) flags |= 16;
  out[4] = flags;
}

// Error message of summary flags (0 when none).
int summary_error(int64_t flags) {
  if (flags & 32) return fail(EAC_ERR_ARG, "sl_batched: graph offsets not 0 <= ... <= N");
  if (flags & 8) return fail(EAC_ERR_KEY_WIDTH, "boruvka: weight range and vertex ids need > 64 key bits");
  if (flags & 2) return fail(EAC_ERR_ARG, "boruvka: dest holds ids outside [0, n)");
  if (flags & 4)
    return fail(EAC_ERR_ARG, "boruvka: graph passed as symmetric is not symmetric (pass symmetric=0)");
  if (flags & 1)
    return fail(EAC_ERR_ARG, "boruvka: more edges with dest > row than a symmetric CSR has; "
                             "pass symmetric=0");
  if (flags & 16) return fail(EAC_ERR_HIP, "boruvka: rounds did not converge");
  return EAC_OK;
}

// Batched single-link over a disjoint union of G graphs (graph g owns the
// vertex range [goff[g], goff[g+1]) and its rows): one Boruvka run over the
// union (keys on global ids keep each graph's own (w, min, max) order, so
// each graph's forest is its own canonical forest), then per-graph sorted
// forests (global key sort + stable sort by graph), per-graph lifetime cut,
// and one union-find labelling numbered by first occurrence inside each
// graph. gout[g] = {components, labels_empty, nc_stable, kept edges or -1
// where the reference raises ValueError}; summary (device int64[5]) as
// k_batch_summary writes it.
//
// graph_mode: the sequence never synchronises with the host (fixed number
// of Boruvka rounds, every count kept on the device, absent keys as kInf in
// fixed-size buffers), so it can be captured into a hipGraph; d_goff must
// then be a device array and h_out is not written.
template <typename W, bool HAS_OD>
int sl_batched_impl(const int32_t* dest, const W* weight, const int64_t* fe, const int32_t* od,
                    int32_t N, int64_t M, int symmetric, const int64_t* h_goff,
                    const int64_t* d_goff, int G, int rule, double max_dist, int32_t n_clusters,
                    int32_t* labels, int64_t* gout, int64_t* summary, int64_t* h_out, void* ws,
                    size_t ws_bytes, hipStream_t s, int graph_rounds = 0) {
  const bool graph_mode = graph_rounds > 0;
  Arena ar(ws, ws_bytes);
  MstBuffers B = carve_mst(ar, N, M, symmetric);
  uint64_t* sk = ar.take<uint64_t>(N);
  uint64_t* sk2 = ar.take<uint64_t>(N);
  uint32_t* gid = ar.take<uint32_t>(N);
  uint32_t* gid_s = ar.take<uint32_t>(N);
  int64_t* goff_ws = ar.take<int64_t>(G + 1);
  int64_t* seg = ar.take<int64_t>(G + 2);
  unsigned long long* gmax = ar.take<unsigned long long>(G);
  GraphCut* cut = ar.take<GraphCut>(G);
  int32_t* base = ar.take<int32_t>(G + 1);
  int64_t* sum_ws = ar.take<int64_t>(8);
  int32_t* misc = ar.take<int32_t>(8);
  if (ar.overflow()) return fail(EAC_ERR_WORKSPACE, "sl_batched: workspace too small");
  if (!summary) summary = sum_ws;
  const size_t mark0 = ar.used;
  EAC_TRY(fill<int32_t>(misc, 2, 0, s));
  if (d_goff) {
    k_goff_prep<<<grid_for(G + 1, 256), 256, 0, s>>>(d_goff, G, N, goff_ws, misc + 1);
    EAC_LAUNCH_CHECK();
  } else {
    if (graph_mode) return fail(EAC_ERR_ARG, "sl_batched: graph mode needs device goff");
    EAC_HIP_CHECK(hipMemcpyAsync(goff_ws, h_goff, sizeof(int64_t) * (G + 1),
                                 hipMemcpyHostToDevice, s));
    EAC_HIP_CHECK(hipStreamSynchronize(s));  // h_goff may be pageable
  }
  const int64_t* goff = goff_ws;
  ForestInfo fi;
  EAC_TRY((mst_core<W, HAS_OD>(dest, weight, fe, od, N, M, symmetric, B, ar, &fi, s,
                               graph_rounds)));
  ar.used = mark0;
  const int vb = fi.vb;
  // graph mode sorts all N slots of the forest buffer (kInf past the count)
  // over the widest key the weight type allows
  int64_t nf = graph_mode ? N : fi.n_f;
  int key_bits = fi.wbits + 2 * vb;
  if (graph_mode) {
    key_bits = 8 * (int)sizeof(W) + 2 * vb;
    if (key_bits > 64) key_bits = 64;
  }
  const int g1 = grid_for(nf, kGB, persistent_blocks());
  // zero weights leave the forest under rule 0 (Kruskal's eliminate_zeros);
  // the absent keys sort after every present one (a present key never has
  // all key bits set since its endpoints differ)
  if (rule == 0 && nf > 0) {
    k_drop_zero<W><<<g1, kGB, 0, s>>>(B.fkeys, nf, vb, fi.plan);
    EAC_LAUNCH_CHECK();
  }
  EAC_TRY((radix_sort<uint64_t, NoValue>(B.fkeys, sk, nullptr, nullptr, nf, 0, key_bits, ar, s)));
  ar.used = mark0;
  if (nf > 0) {
    k_graph_of_key<<<g1, kGB, 0, s>>>(sk, nf, vb, goff, G, gid);
    EAC_LAUNCH_CHECK();
    const int gb = bits_for((uint64_t)G);
    EAC_TRY((radix_sort<uint32_t, uint64_t>(gid, gid_s, sk, sk2, nf, 0, gb < 1 ? 1 : gb, ar, s)));
    ar.used = mark0;
  }
  // seg[G] = present forest keys; they lead sk2, grouped by graph
  k_seg_bounds<<<grid_for(G + 1, 256), 256, 0, s>>>(gid_s, nf, G, seg);
  EAC_LAUNCH_CHECK();
  EAC_TRY(fill<uint64_t>((uint64_t*)gmax, G, 0ull, s));
  k_graph_wmax<W, HAS_OD><<<HAS_OD ? (G < 4096 ? G : 4096)
                                    : grid_for(M, 1, resident_grid(k_graph_wmax<W, HAS_OD>, kGB)),
                            kGB, 0, s>>>(weight, fe, od, goff, G, M, gmax);
  EAC_LAUNCH_CHECK();
  k_seg_cut<W><<<G < 4096 ? G : 4096, kGB, 0, s>>>(sk2, seg, goff, G, vb, fi.plan, gmax, rule,
                                                  max_dist, n_clusters, cut);
  EAC_LAUNCH_CHECK();
  EAC_TRY(fill<int32_t>(misc, 1, 0, s));
  if (nf > 0) {
    k_seg_keep_flags<<<g1, kGB, 0, s>>>(gid_s, nf, G, seg, cut, B.flag);
    EAC_LAUNCH_CHECK();
    EAC_TRY((scan_excl<int32_t, int32_t>(B.flag, B.rank, nf, 0, misc, ar, s)));
    ar.used = mark0;
    k_compact_flagged<<<g1, kGB, 0, s>>>(sk2, B.flag, B.rank, nf, B.fkeys);
    EAC_LAUNCH_CHECK();
  }
  const int64_t ftiles = (N + kGTile - 1) / kGTile;
  int32_t* d_ncomp = B.tile_cnt + ftiles;
  EAC_TRY(labels_from_keys(B.fkeys, nf, vb, N, B.hook, B.comp, B.flag, B.rank, labels, d_ncomp,
                           ar, s, misc));
  k_graph_base<<<grid_for(G + 1, 256), 256, 0, s>>>(labels, goff, G, d_ncomp, base);
  EAC_LAUNCH_CHECK();
  k_graph_sub<<<G < 4096 ? G : 4096, 256, 0, s>>>(labels, goff, G, base, cut, gout);
  EAC_LAUNCH_CHECK();
  k_batch_summary<<<1, 64, 0, s>>>(B.st, fi.enq, d_ncomp, seg + G, misc, summary);
  EAC_LAUNCH_CHECK();
  if (graph_mode) return EAC_OK;
  int64_t hs[5];
  EAC_TRY(d2h_sync(hs, summary, sizeof(hs), s));
  prof_flush();
  EAC_TRY(summary_error(hs[4]));
  if (h_out) {
    h_out[0] = hs[0];
    h_out[1] = hs[1];
    h_out[2] = fi.rounds;
    h_out[3] = hs[3];
  }
  return EAC_OK;
}

// Calls F::template run<W, HAS_OD>(...) for the runtime weight dtype.
template <typename F, typename... A>
int dispatch_weight(int32_t dt, bool has_od, A&&... a) {
#define EAC_DISPATCH_OD(T)                                                   \
  return has_od ? F::template run<T, true>(std::forward<A>(a)...)          \
                : F::template run<T, false>(std::forward<A>(a)...)
  switch (dt) {
    case EAC_W_I32: EAC_DISPATCH_OD(int32_t);
    case EAC_W_F32: EAC_DISPATCH_OD(float);
    case EAC_W_U16: EAC_DISPATCH_OD(uint16_t);
    case EAC_W_U8: EAC_DISPATCH_OD(uint8_t);
    case EAC_W_I64: EAC_DISPATCH_OD(int64_t);
    case EAC_W_I16: EAC_DISPATCH_OD(int16_t);
    case EAC_W_U32: EAC_DISPATCH_OD(uint32_t);
  }
#undef EAC_DISPATCH_OD
  return fail(EAC_ERR_ARG, "unsupported weight dtype %d", dt);
}

struct MstRun {
  template <typename W, bool HAS_OD>
  static int run(const int32_t* dest, const void* weight, const int64_t* fe, const int32_t* od,
                 int32_t n, int64_t m, int symmetric, int64_t* ids, void* wout, int32_t* labels,
                 int64_t* h_out, void* ws, size_t ws_bytes, hipStream_t s) {
    return mst_impl<W, HAS_OD>(dest, (const W*)weight, fe, od, n, m, symmetric, ids, (W*)wout,
                               labels, h_out, ws, ws_bytes, s);
  }
};

struct SlBatchRun {
  template <typename W, bool HAS_OD>
  static int run(const int32_t* dest, const void* weight, const int64_t* fe, const int32_t* od,
                 int32_t N, int64_t M, int symmetric, const int64_t* h_goff, int G, int rule,
                 double max_dist, int32_t n_clusters, int32_t* labels, int64_t* gout,
                 int64_t* h_out, void* ws, size_t ws_bytes, hipStream_t s) {
    return sl_batched_impl<W, HAS_OD>(dest, (const W*)weight, fe, od, N, M, symmetric, h_goff,
                                      nullptr, G, rule, max_dist, n_clusters, labels, gout,
                                      nullptr, h_out, ws, ws_bytes, s);
  }
};

struct SlBatchGraphRun {
  template <typename W, bool HAS_OD>
  static int run(const int32_t* dest, const void* weight, const int64_t* fe, const int32_t* od,
                 int32_t N, int64_t M, int symmetric, const int64_t* d_goff, int G, int rule,
                 double max_dist, int32_t n_clusters, int32_t* labels, int64_t* gout,
                 int64_t* summary, int rounds, void* ws, size_t ws_bytes, hipStream_t s) {
    return sl_batched_impl<W, HAS_OD>(dest, (const W*)weight, fe, od, N, M, symmetric, nullptr,
                                      d_goff, G, rule, max_dist, n_clusters, labels, gout,
                                      summary, nullptr, ws, ws_bytes, s, rounds);
  }
};

struct SlRun {
  template <typename W, bool HAS_OD>
  static int run(const int32_t* dest, const void* weight, const int64_t* fe, const int32_t* od,
                 int32_t n, int64_t m, int symmetric, int rule, double max_dist,
                 int32_t n_clusters, int32_t* labels, int64_t* h_out, void* ws, size_t ws_bytes,
                 hipStream_t s) {
    return sl_impl<W, HAS_OD>(dest, (const W*)weight, fe, od, n, m, symmetric, rule, max_dist,
                              n_clusters, labels, h_out, ws, ws_bytes, s);
  }
};

}  // namespace

size_t mst_ws_bytes(int32_t n, int64_t m) {
  Arena ar(nullptr, 0, true);
  (void)carve_mst(ar, n, m, 0);
  return ar.used + scan_ws_bytes(n, 4) + 4096;
}

}  // namespace eac

// ------------------------------------------------------------------ C ABI
using namespace eac;

extern "C" size_t eac_boruvka_workspace_bytes(int32_t n, int64_t m) {
  return eac::mst_ws_bytes(n < 1 ? 1 : n, m < 1 ? 1 : m);
}

extern "C" int eac_boruvka_mst(const int32_t* dest, const void* weight, int32_t weight_dtype,
                               const int64_t* first_edge, const int32_t* out_degree,
                               int32_t n, int64_t m, int32_t symmetric,
                               int64_t* mst_edge_ids, void* mst_weight, int32_t* labels,
                               int64_t* h_out, void* ws, size_t ws_bytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n < 0 || m < 0) return fail(EAC_ERR_ARG, "boruvka: negative size");
  if (weight_bytes(weight_dtype) == 0)
    return fail(EAC_ERR_ARG, "boruvka: unsupported weight dtype %d", weight_dtype);
  if (n == 0) {
    if (h_out) { h_out[0] = 0; h_out[1] = 0; h_out[2] = 0; h_out[3] = 0; }
    return EAC_OK;
  }
  if (!first_edge || (m > 0 && (!dest || !weight)))
    return fail(EAC_ERR_ARG, "boruvka: null graph array");
  if (m == 0) {
    if (labels) {
      k_iota<<<grid_for(n, 256, 4096), 256, 0, s>>>(labels, n);
      EAC_LAUNCH_CHECK();
    }
    if (h_out) { h_out[0] = 0; h_out[1] = n; h_out[2] = 0; h_out[3] = 0; }
    return EAC_OK;
  }
  size_t need = eac_boruvka_workspace_bytes(n, m);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "boruvka: workspace %zu < %zu bytes", ws_bytes, need);
  symmetric = symmetric ? 1 : 0;
  return dispatch_weight<MstRun>(weight_dtype, out_degree != nullptr, dest, weight, first_edge,
                                 out_degree, n, m, symmetric, mst_edge_ids, mst_weight, labels,
                                 h_out, ws, ws_bytes, s);
}

extern "C" size_t eac_sl_lifetime_workspace_bytes(int32_t n, int64_t m) {
  size_t b = eac_boruvka_workspace_bytes(n, m);
  b += ((size_t)(n < 1 ? 1 : n) * 8 + 255) & ~size_t(255);
  b += 512 + radix_sort_ws_bytes(n < 1 ? 1 : n, 8, 0);
  return b;
}

extern "C" int eac_sl_lifetime(const int32_t* dest, const void* weight, int32_t weight_dtype,
                               const int64_t* first_edge, const int32_t* out_degree, int32_t n,
                               int64_t m, int32_t symmetric, int32_t rule, double max_dist,
                               int32_t n_clusters, int32_t* labels, int64_t* h_out, void* ws,
                               size_t ws_bytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n < 1 || m < 0) return fail(EAC_ERR_ARG, "sl_lifetime: need n >= 1 and m >= 0");
  if (rule != EAC_CUT_EAC && rule != EAC_CUT_SINGLELINK)
    return fail(EAC_ERR_ARG, "sl_lifetime: unknown rule %d", rule);
  if (n_clusters < 0) return fail(EAC_ERR_ARG, "sl_lifetime: negative n_clusters");
  if (weight_bytes(weight_dtype) == 0)
    return fail(EAC_ERR_ARG, "sl_lifetime: unsupported weight dtype %d", weight_dtype);
  if (m == 0)
    return fail(EAC_ERR_VALUE, "sl_lifetime: graph without edges (the reference raises)");
  if (!first_edge || !dest || !weight || !labels)
    return fail(EAC_ERR_ARG, "sl_lifetime: null array");
  size_t need = eac_sl_lifetime_workspace_bytes(n, m);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "sl_lifetime: workspace %zu < %zu bytes", ws_bytes, need);
  return dispatch_weight<SlRun>(weight_dtype, out_degree != nullptr, dest, weight, first_edge,
                                out_degree, n, m, symmetric ? 1 : 0, (int)rule, max_dist,
                                n_clusters, labels, h_out, ws, ws_bytes, s);
}

// ------------------------------------------------- connected components
extern "C" size_t eac_cc_workspace_bytes(int32_t n, int64_t m) {
  int64_t nchunks = (m + kGTile - 1) / kGTile;
  Arena ar(nullptr, 0, true);
  ar.take<int32_t>(nchunks + 2);
  ar.take<int32_t>(n > 0 ? n : 1);  // parent
  ar.take<int32_t>(n > 0 ? n : 1);  // root
  ar.take<int32_t>(n > 0 ? n : 1);  // flag
  ar.take<int32_t>(n > 0 ? n : 1);  // rank
  ar.take<int32_t>(4);
  return ar.used + eac::scan_ws_bytes(n, 4) + 4096;
}

namespace eac {
namespace {
// The components launch sequence over scratch carved from `ar`; d_nc[0]
// receives the number of components, d_nc[1] is set when dest holds ids
// outside [0, n). Enqueue only: no synchronisation, no host memory.
int cc_enqueue(const int32_t* dest, const int64_t* first_edge, const int32_t* out_degree,
               int32_t n, int64_t m, int32_t* labels, int32_t* d_nc, Arena& ar, bool profile,
               hipStream_t s) {
  int64_t nchunks = (m + kGTile - 1) / kGTile;
  int32_t* chunk_row = ar.take<int32_t>(nchunks + 2);
  int32_t* parent = ar.take<int32_t>(n);
  int32_t* root = ar.take<int32_t>(n);
  int32_t* flag = ar.take<int32_t>(n);
  int32_t* rank = ar.take<int32_t>(n);
  if (ar.overflow()) return fail(EAC_ERR_WORKSPACE, "cc: workspace too small");
  EAC_TRY(fill<int32_t>(d_nc, 2, 0, s));
  k_iota<<<grid_for(n, 256, 4096), 256, 0, s>>>(parent, n);
  EAC_LAUNCH_CHECK();
  if (m > 0) {
    k_chunk_rows<<<grid_for(nchunks + 1, 256), 256, 0, s>>>(first_edge, n, m, nchunks, chunk_row);
    EAC_LAUNCH_CHECK();
    int g = grid_for(nchunks, 1, persistent_blocks());
    int pr = profile ? prof_begin(EAC_PROF_CC, 4 * m + 8 * (int64_t)n, s) : -1;
    if (out_degree)
      k_unite_csr<true><<<g, kGB, 0, s>>>(dest, first_edge, out_degree, n, m, chunk_row, nchunks,
                                          parent, (unsigned int*)(d_nc + 1));
    else
      k_unite_csr<false><<<g, kGB, 0, s>>>(dest, first_edge, out_degree, n, m, chunk_row, nchunks,
                                           parent, (unsigned int*)(d_nc + 1));
    EAC_LAUNCH_CHECK();
    if (profile) prof_end(pr, s);
  }
  k_uf_flatten<<<grid_for(n, 256, 4096), 256, 0, s>>>(parent, n, root, flag);
  EAC_LAUNCH_CHECK();
  EAC_TRY((scan_excl<int32_t, int32_t>(flag, rank, n, 0, d_nc, ar, s)));
  k_gather_rank<<<grid_for(n, 256, 4096), 256, 0, s>>>(root, rank, n, labels);
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}

// Scratch of eac_connected_components_noalloc: a fixed device array that
// every GPU the library runs on holds until the process ends. Host code
// never uses this name as an address; it asks the runtime for each
// device's own copy (cc_scratch_ptr).
constexpr size_t kCcScratchBytes = 64u << 20;
static __device__ __align__(256) unsigned char scratch[kCcScratchBytes];

constexpr int kMaxDevices = 64;
void* g_cc_scratch[kMaxDevices] = {nullptr};
std::mutex g_cc_scratch_mu;

// The address of the current device's copy of `scratch`, looked up once per
// device (hipGetSymbolAddress with that device current).
int cc_scratch_ptr(void** out) {
  int dev = 0;
  EAC_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(EAC_ERR_UNSUPPORTED, "cc noalloc: device %d", dev);
  std::lock_guard<std::mutex> lk(g_cc_scratch_mu);
  if (!g_cc_scratch[dev]) {
    void* p = nullptr;
    EAC_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(EAC_ERR_HIP, "cc noalloc: no address for the scratch array");
    g_cc_scratch[dev] = p;
  }
  *out = g_cc_scratch[dev];
  return EAC_OK;
}
}  // namespace
}  // namespace eac

extern "C" int eac_connected_components(const int32_t* dest, const int64_t* first_edge,
                                        const int32_t* out_degree, int32_t n, int64_t m,
                                        int32_t* labels, int64_t* h_ncomp, void* ws,
                                        size_t ws_bytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n < 0 || m < 0) return fail(EAC_ERR_ARG, "cc: negative size");
  if (n == 0) {
    if (h_ncomp) *h_ncomp = 0;
    return EAC_OK;
  }
  if (!first_edge || !labels || (m > 0 && !dest)) return fail(EAC_ERR_ARG, "cc: null array");
  size_t need = eac_cc_workspace_bytes(n, m);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "cc: workspace %zu < %zu bytes", ws_bytes, need);
  Arena ar(ws, ws_bytes);
  int32_t* d_nc = ar.take<int32_t>(4);
  EAC_TRY(cc_enqueue(dest, first_edge, out_degree, n, m, labels, d_nc, ar, true, s));
  int32_t hc[2] = {0, 0};
  EAC_TRY(d2h_sync(hc, d_nc, sizeof(hc), s));
  prof_flush();
  if (hc[1]) return fail(EAC_ERR_ARG, "cc: dest holds ids outside [0, n)");
  if (h_ncomp) *h_ncomp = hc[0];
  return EAC_OK;
}

extern "C" size_t eac_connected_components_noalloc_max_bytes(void) { return kCcScratchBytes; }

extern "C" int eac_connected_components_noalloc(const int32_t* dest, const int64_t* first_edge,
                                                const int32_t* out_degree, int32_t n, int64_t m,
                                                int32_t* labels, int32_t* d_out, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n < 1 || m < 0) return fail(EAC_ERR_ARG, "cc noalloc: need n >= 1, m >= 0");
  if (!first_edge || !labels || !d_out || (m > 0 && !dest))
    return fail(EAC_ERR_ARG, "cc noalloc: null array");
  const size_t need = eac_cc_workspace_bytes(n, m);
  if (need > kCcScratchBytes)
    return fail(EAC_ERR_UNSUPPORTED,
                "cc noalloc: this graph needs %zu bytes of scratch, over the fixed %zu", need,
                kCcScratchBytes);
  void* ws = nullptr;
  EAC_TRY(cc_scratch_ptr(&ws));
  Arena ar(ws, kCcScratchBytes);
  return cc_enqueue(dest, first_edge, out_degree, n, m, labels, d_out, ar, false, s);
}

// ------------------------------------------------ first-occurrence relabel
extern "C" size_t eac_relabel_workspace_bytes(int32_t n, int32_t n_values) {
  Arena ar(nullptr, 0, true);
  ar.take<int32_t>(n_values > 0 ? n_values : 1);
  ar.take<int32_t>(n > 0 ? n : 1);
  ar.take<int32_t>(n > 0 ? n : 1);
  ar.take<int32_t>(n > 0 ? n : 1);
  ar.take<uint32_t>(4);
  return ar.used + eac::scan_ws_bytes(n, 4) + 4096;
}

extern "C" int eac_canonical_relabel(const int32_t* labels_in, int32_t n, int32_t n_values,
                                     int32_t* out, int64_t* h_count, void* ws,
                                     size_t ws_bytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n < 0 || n_values < 0) return fail(EAC_ERR_ARG, "relabel: negative size");
  if (n == 0) {
    if (h_count) *h_count = 0;
    return EAC_OK;
  }
  if (!labels_in || !out) return fail(EAC_ERR_ARG, "relabel: null array");
  size_t need = eac_relabel_workspace_bytes(n, n_values);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "relabel: workspace %zu < %zu bytes", ws_bytes, need);
  Arena ar(ws, ws_bytes);
  int32_t* first = ar.take<int32_t>(n_values > 0 ? n_values : 1);
  int32_t* root = ar.take<int32_t>(n);
  int32_t* flag = ar.take<int32_t>(n);
  int32_t* rank = ar.take<int32_t>(n);
  uint32_t* misc = ar.take<uint32_t>(4);
  EAC_TRY(fill<int32_t>(first, n_values, INT32_MAX, s));
  EAC_TRY(fill<int32_t>((int32_t*)misc, 4, 0, s));
  int g = grid_for(n, 256, 4096);
  k_first_occ<<<grid_for(n, kFoTile, 4096), 256, 0, s>>>(labels_in, n, n_values, first, misc);
  EAC_LAUNCH_CHECK();
  k_first_flag<<<g, 256, 0, s>>>(labels_in, first, n, n_values, root, flag);
  EAC_LAUNCH_CHECK();
  EAC_TRY((scan_excl<int32_t, int32_t>(flag, rank, n, 0, (int32_t*)misc + 1, ar, s)));
  k_gather_rank<<<g, 256, 0, s>>>(root, rank, n, out);
  EAC_LAUNCH_CHECK();
  uint32_t hm[4];
  EAC_TRY(d2h_sync(hm, misc, sizeof(hm), s));
  if (hm[0]) return fail(EAC_ERR_ARG, "relabel: label outside [0, n_values)");
  if (h_count) *h_count = (int32_t)hm[1];
  return EAC_OK;
}

extern "C" size_t eac_sl_lifetime_batched_workspace_bytes(int32_t N, int64_t M, int32_t G) {
  size_t b = eac_boruvka_workspace_bytes(N, M);
  size_t n = (size_t)(N < 1 ? 1 : N);
  b += 4 * ((n * 8 + 255) & ~size_t(255)) + 64 * (size_t)(G + 2) + 8192;
  size_t r1 = radix_sort_ws_bytes(n, 8, 0), r2 = radix_sort_ws_bytes(n, 4, 8);
  return b + (r1 > r2 ? r1 : r2);
}

extern "C" int eac_sl_lifetime_batched(const int32_t* dest, const void* weight,
                                       int32_t weight_dtype, const int64_t* first_edge,
                                       const int32_t* out_degree, int32_t N, int64_t M,
                                       int32_t symmetric, const int64_t* h_goff, int32_t G,
                                       int32_t rule, double max_dist, int32_t n_clusters,
                                       int32_t* labels, int64_t* gout, int64_t* h_out, void* ws,
                                       size_t ws_bytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (N < 1 || M < 1 || G < 1) return fail(EAC_ERR_ARG, "sl_batched: need N, M, G >= 1");
  if (rule != EAC_CUT_EAC && rule != EAC_CUT_SINGLELINK)
    return fail(EAC_ERR_ARG, "sl_batched: unknown rule %d", rule);
  if (!dest || !weight || !first_edge || !h_goff || !labels || !gout)
    return fail(EAC_ERR_ARG, "sl_batched: null array");
  if (h_goff[0] != 0 || h_goff[G] != N) return fail(EAC_ERR_ARG, "sl_batched: bad graph offsets");
  for (int g = 0; g < G; ++g)
    if (h_goff[g + 1] < h_goff[g]) return fail(EAC_ERR_ARG, "sl_batched: offsets not ascending");
  if (weight_bytes(weight_dtype) == 0) return fail(EAC_ERR_ARG, "sl_batched: bad weight dtype");
  size_t need = eac_sl_lifetime_batched_workspace_bytes(N, M, G);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "sl_batched: workspace %zu < %zu bytes", ws_bytes, need);
  return dispatch_weight<SlBatchRun>(weight_dtype, out_degree != nullptr, dest, weight, first_edge,
                                     out_degree, N, M, symmetric ? 1 : 0, h_goff, (int)G,
                                     (int)rule, max_dist, n_clusters, labels, gout, h_out, ws,
                                     ws_bytes, s);
}

extern "C" int eac_mst_round_bound(int32_t n, int64_t m, int32_t symmetric) {
  if (n < 1 || m < 0) return 0;
  return eac::mst_round_bound(n, eac::split_plan(n, m, symmetric).levels);
}

// ------------------------------------------- batched cut as one hipGraph
struct eac_sl_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int device = 0;
  int rounds = 0;
};

extern "C" int eac_sl_lifetime_batched_graph_create(
    const int32_t* dest, const void* weight, int32_t weight_dtype, const int64_t* first_edge,
    const int32_t* out_degree, int32_t N, int64_t M, int32_t symmetric, const int64_t* goff,
    int32_t G, int32_t rule, double max_dist, int32_t n_clusters, int32_t max_rounds,
    int32_t* labels, int64_t* gout, int64_t* summary, void* ws, size_t ws_bytes,
    eac_sl_graph** out) {
  using namespace eac;
  if (!out) return fail(EAC_ERR_ARG, "sl_graph: null output handle");
  *out = nullptr;
  if (N < 1 || M < 1 || G < 1) return fail(EAC_ERR_ARG, "sl_graph: need N, M, G >= 1");
  if (rule != EAC_CUT_EAC && rule != EAC_CUT_SINGLELINK)
    return fail(EAC_ERR_ARG, "sl_graph: unknown rule %d", rule);
  if (!dest || !weight || !first_edge || !goff || !labels || !gout || !summary)
    return fail(EAC_ERR_ARG, "sl_graph: null array");
  if (weight_bytes(weight_dtype) == 0) return fail(EAC_ERR_ARG, "sl_graph: bad weight dtype");
  size_t need = eac_sl_lifetime_batched_workspace_bytes(N, M, G);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "sl_graph: workspace %zu < %zu bytes", ws_bytes, need);
  const int bound = mst_round_bound(N, split_plan(N, M, symmetric).levels);
  const int rounds = max_rounds <= 0 || max_rounds > bound ? bound : max_rounds;
  eac_sl_graph* g = new eac_sl_graph();
  g->rounds = rounds;
  EAC_HIP_CHECK(hipGetDevice(&g->device));
  // captured on a stream of the library's own: nothing runs now
  hipStream_t cs = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
    delete g;
    return fail(EAC_ERR_HIP, "sl_graph: stream creation failed");
  }
  prof_suspend(true);
  int rc = EAC_OK;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = fail(EAC_ERR_HIP, "sl_graph: begin capture failed");
  } else {
    rc = dispatch_weight<SlBatchGraphRun>(weight_dtype, out_degree != nullptr, dest, weight,
                                          first_edge, out_degree, N, M, symmetric ? 1 : 0, goff,
                                          (int)G, (int)rule, max_dist, n_clusters, labels, gout,
                                          summary, rounds, ws, ws_bytes, cs);
    hipError_t e = hipStreamEndCapture(cs, &g->graph);
    if (rc == EAC_OK && e != hipSuccess)
      rc = fail(EAC_ERR_HIP, "sl_graph: end capture: %s", hipGetErrorString(e));
  }
  prof_suspend(false);
  if (rc == EAC_OK && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
    rc = fail(EAC_ERR_HIP, "sl_graph: instantiation failed");
  (void)hipStreamDestroy(cs);
  if (rc != EAC_OK) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return rc;
  }
  *out = g;
  return EAC_OK;
}

extern "C" int eac_sl_lifetime_batched_graph_launch(eac_sl_graph* g, void* stream) {
  using namespace eac;
  if (!g || !g->exec) return fail(EAC_ERR_ARG, "sl_graph: null graph");
  int dev = -1;
  EAC_HIP_CHECK(hipGetDevice(&dev));
  if (dev != g->device)
    return fail(EAC_ERR_ARG, "sl_graph: captured on device %d, current is %d", g->device, dev);
  EAC_HIP_CHECK(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return EAC_OK;
}

extern "C" int eac_sl_lifetime_batched_graph_rounds(const eac_sl_graph* g) {
  return g ? g->rounds : 0;
}

extern "C" int eac_sl_lifetime_batched_graph_destroy(eac_sl_graph* g) {
  if (!g) return EAC_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return EAC_OK;
}

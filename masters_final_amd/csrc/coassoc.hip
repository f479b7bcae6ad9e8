// Sparse co-association matrix of an ensemble of partitions (EAC) on gfx950.
//
// Reference: MyML/MyML/EAC/sparse.py EAC_CSR (:12-304): for every partition
// and cluster, every pair of members gets +1 (update_cluster_fp_condensed
// :445-465 for the first partition, update_cluster_sorted_surgical_condensed
// :956-1040 with binary_search_interval :1069-1099 for the others); the
// condensed form keeps the upper triangle (col > row), rows sorted by column;
// _condense (:212-225, condense_eac_csr :336-349) packs the rows with a
// sequential offset sum; make_diassoc (:295-311) turns counts into
// (max + 1) - count. The reference caps each row at max_assocs and silently
// drops associations beyond it (sparse.py:501-503, :995-997); the default
// build here is exact (no cap, int64 row offsets), and a capacity-capped
// selection (eac_coassoc_cap, at the end of this file) reproduces the
// reference's lossy matrix from the exact one plus per-entry partition ranks.
//
// MI355X design:
//  * Plan: a stable radix sort of (global cluster id, point) turns labels
//    [P][n] into member lists (ascending point id per cluster) plus each
//    point's position in its list. A point's later co-members in partition p
//    are then the suffix of its cluster's list after its own position: no
//    search and no pair materialisation.
//  * Count/fill: one workgroup per row (rows handed out by atomic counters,
//    one per XCD over its own stretch of the rows). Columns are covered in windows of 32768; each window has a
//    dense array of u16 counters (two per u32 word, 64 KiB of LDS). Per window every partition's
//    suffix cursor advances over the members inside the window (flattened
//    over partitions so that lanes stream consecutive members), each member
//    an LDS atomic increment. The window is then read in column order: pass 1
//    counts non-zeros (row nnz, max count), pass 2 writes (col, value) after
//    a block scan, so rows come out sorted without a per-row sort.
//  * Row offsets: exclusive scan of the per-row nnz into int64 (nnz > 2^31
//    for the large ensembles).
#include <cstring>
#include <vector>

#include "prims.hpp"
#include "sort.hpp"

namespace eac {
namespace {

constexpr int kCB = 256;       // threads per row workgroup
constexpr int kWin = 32768;    // columns per LDS window (two u16 counters per u32 word)
constexpr int kMaxP = 1024;    // partitions staged in LDS
constexpr int kMaxRanges = 8;  // row ranges (one per XCD)
constexpr int kCtrs = kMaxRanges + 2;  // row counters, then the sampled nnz sum

// keys: global cluster id (unassigned points -> ktot, sorted last);
// values: p * n + i
__global__ void k_member_keys(const int32_t* __restrict__ labels, int64_t n, int P,
                              const int64_t* __restrict__ k_off, int64_t ktot,
                              uint32_t* __restrict__ key, uint32_t* __restrict__ val,
                              int32_t* __restrict__ csize, unsigned int* __restrict__ bad) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)P * n;
       t += (int64_t)gridDim.x * blockDim.x) {
    int p = (int)(t / n);
    int32_t l = labels[t];
    int64_t kp = k_off[p + 1] - k_off[p];
    uint32_t g;
    if (l < 0) {
      g = (uint32_t)ktot;
    } else if (l >= kp) {
      atomicOr(bad, 1u);
      g = (uint32_t)ktot;
    } else {
      g = (uint32_t)(k_off[p] + l);
    }
    key[t] = g;
    val[t] = (uint32_t)t;
    atomicAdd(&csize[g], 1);
  }
}

// members[q] = point id; pos[p*n + i] = q (or -1 when unassigned)
__global__ void k_member_pos(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sval,
                             int64_t total, int64_t n, int64_t ktot,
                             int32_t* __restrict__ members, int32_t* __restrict__ pos) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    uint32_t v = sval[q];
    members[q] = (int32_t)(v % (uint32_t)n);
    pos[v] = skey[q] == (uint32_t)ktot ? -1 : (int32_t)q;
  }
}

// Window table: wtab[g * (nw + 1) + w] = first position in cluster g's
// member list whose member is >= w * win (windows aligned to multiples of
// win), so a window's end in every partition is one load instead of a
// search. Built from the sorted (cluster, point) keys: each position writes
// the windows that start after its predecessor's member and up to its own,
// the last member of a cluster also the windows after it.
__global__ void k_win_table(const uint32_t* __restrict__ skey, const int32_t* __restrict__ members,
                            const int32_t* __restrict__ cstart, int64_t total, int64_t ktot,
                            int win, int nw, int32_t* __restrict__ wtab) {
  const int64_t ts = nw + 1;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t g = skey[q];
    if (g >= (uint32_t)ktot) continue;  // unassigned points sort last
    const int wq = members[q] / win;
    const int wp = q == cstart[g] ? -1 : members[q - 1] / win;
    int32_t* t = wtab + (int64_t)g * ts;
    for (int w = wp + 1; w <= wq; ++w) t[w] = (int32_t)q;
    if (q + 1 == cstart[g + 1])
      for (int w = wq + 1; w <= nw; ++w) t[w] = (int32_t)(q + 1);
  }
}

// 64-bit words of a window's run-start bitmap: the short-run path covers
// fewer than 32 flattened members per partition (rounded up to an even
// count: the counters after it stay 16-byte aligned).
__host__ __device__ inline int co_bm_words(int P) { return ((P / 2 + 1) + 1) & ~1; }

struct CoPlan {
  const int32_t* labels;
  int64_t n;
  int P;
  const int64_t* k_off;   // [P+1] device
  const int32_t* members;
  const int32_t* pos;     // [P*n]
  const int32_t* cstart;  // [ktot+1] start of each global cluster in members
  const int32_t* wtab;    // window table (nullptr: search the window ends)
  int nw;                 // windows of the table: ceil(n / win)
  const int32_t* first;   // ranked fill: the partition processed first (device)
  int ranges;             // row ranges with a counter each (1..kMaxRanges)
  int64_t row_stride;     // > 1: only rows 0, stride, 2*stride, ... (a sample)
};

// Outputs of the row kernel (pointers a pass does not use are null).
struct CoOut {
  unsigned long long* row_ctr;  // [kMaxRanges] row counters, zeroed
  int32_t* row_nnz;             // PASS 0 / 2: nnz of each row
  unsigned int* max_count;      // PASS 0 / 2: largest count
  unsigned long long* nnz_sum;  // PASS 0: sum of the visited rows' nnz
  const int64_t* indptr;        // PASS 1: row offsets
  int32_t w_offset;             // PASS 1: data = w_offset - count when non-zero
  int32_t* col;                 // PASS 1: columns
  int32_t* val;                 // PASS 1: counts (or w_offset - count)
  int32_t* rank;                // PASS 1 with RANK: first partition rank
  uint32_t* pool;               // PASS 2: staged entries (count << 16 | column - c0)
  uint64_t* seg;                // PASS 2: per (row, window) offset << 16 | entries
  unsigned long long* pool_ctr; // PASS 2: next free pool chunk
  int64_t pool_cap;             // PASS 2: pool entries
  int64_t pool_chunk;           // PASS 2: entries per chunk (>= win)
  unsigned int* overflow;       // PASS 2: set when the pool ran out
};

// Staged entries are taken from the pool in chunks (one global atomic per
// chunk) of up to this many, or 16 windows when that is more.
constexpr int64_t kPoolChunk = 65536;

// Segment-table slot of (row i, first window of the row): rows keep the
// windows from w0(i) = (i + 1) / win on (every window with full rows), so
// the table packs them with sum_{r < i} (nw - w0(r)) slots before row i;
// sum_{m=1..i} floor(m / win) = win * q (q - 1) / 2 + q (i - q win + 1)
// with q = floor(i / win).
__host__ __device__ inline int64_t seg_base(int64_t i, int64_t win, int64_t nw, int full_rows) {
  if (full_rows) return i * nw;
  const int64_t q = i / win;
  return i * nw - (win * q * (q - 1) / 2 + q * (i - q * win + 1));
}

// The partition the reference processes first: the one with the fewest
// non-empty clusters, the first such (sparse.py:158: np.argmin over the
// cluster-list lengths; clusters are the labels present, partition.py:73-86).
// One block; cstart gives every global cluster's member count.
__global__ __launch_bounds__(256) void k_first_partition(const int32_t* __restrict__ cstart,
                                                         const int64_t* __restrict__ k_off,
                                                         int P, int32_t* __restrict__ first) {
  __shared__ unsigned long long s_best;
  if (threadIdx.x == 0) s_best = ~0ull;
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    uint32_t nonempty = 0;
    for (int64_t g = k_off[p]; g < k_off[p + 1]; ++g) nonempty += cstart[g + 1] > cstart[g];
    atomicMin(&s_best, ((unsigned long long)nonempty << 32) | (unsigned)p);
  }
  __syncthreads();
  if (threadIdx.x == 0) first[0] = (int32_t)(s_best & 0xffffffffu);
}

// One row: walk every partition's co-members > i (or all co-members != i for
// full rows) window by window. PASS 0: per-row nnz and max count. PASS 1:
// write (col, value) at indptr[i]. RANK (PASS 1 only): also write, per entry,
// the processing rank of the first partition in which the pair co-occurs
// (0 = the partition pl.first[0], then the others in index order), kept in a
// second LDS window with one atomicMin per pair visit. PASS 2 (one-pass
// build): each window's entries go to the workgroup's current pool chunk as
// (count << 16 | column - c0), the window's (pool offset, entries) to the
// segment table, and the row's nnz and the max count as in PASS 0; a
// workgroup that finds the pool used up sets the overflow flag and stops.
// U: member loads in flight per lane while visiting (4 or 8, a tuning knob).
template <int PASS, bool RANK = false, int U = 4>
__global__ __launch_bounds__(kCB) void k_coassoc_rows(CoPlan pl, int full_rows, int win,
                                                      CoOut out) {
  unsigned long long* __restrict__ row_ctr = out.row_ctr;
  const int64_t* __restrict__ indptr = out.indptr;
  const int32_t w_offset = out.w_offset;
  int32_t* __restrict__ out_col = out.col;
  int32_t* __restrict__ out_val = out.val;
  int32_t* __restrict__ out_rank = out.rank;
  // launch-time LDS: this window's run-start bitmap (co_bm_words(P)
  // 64-bit words), the window counters (win / 2 words), with RANK the rank
  // window (win words), then per partition the next member position, the
  // cluster end, this window's member end and the flattened-range prefix
  // (P + 1), the row's global cluster, this window's member position minus
  // flattened index, each partition's rank; per run (non-empty partition
  // range of this window) the position delta and the partition's rank; per
  // bitmap word the run starts before it
  extern __shared__ __align__(16) uint32_t s_dyn[];
  uint64_t* s_bm = reinterpret_cast<uint64_t*>(s_dyn);
  const int bmw = co_bm_words(pl.P);
  uint32_t* s_cnt = s_dyn + 2 * bmw;
  uint32_t* s_rank = s_cnt + win / 2;
  int32_t* s_cur = reinterpret_cast<int32_t*>(s_cnt + win / 2 + (RANK ? win : 0));
  int32_t* s_end = s_cur + pl.P;
  int32_t* s_hi = s_end + pl.P;
  int32_t* s_pre = s_hi + pl.P;
  int32_t* s_g = s_pre + pl.P + 1;  // global cluster of the row in each partition
  int32_t* s_d = s_g + pl.P;        // this window: member position - flattened index
  int32_t* s_prank = s_d + pl.P;
  int32_t* s_dr = s_prank + pl.P;   // per run: member position - flattened index
  int32_t* s_rp = s_dr + pl.P;      // per run: rank of its partition
  int32_t* s_wpre = s_rp + pl.P;    // per bitmap word: run starts in the words before
  __shared__ uint32_t s_tmp[kCB / kWave + 1];
  __shared__ int64_t s_row;
  __shared__ unsigned long long s_chunk;
  __shared__ unsigned int s_max;
  const int t = threadIdx.x;
  const int64_t n = pl.n;
  const int P = pl.P;
  // the window is cleared once here; afterwards every window clears the
  // words it used while reading them, so a clean window costs no pass
  for (int v = t; v < win / 2; v += kCB) s_cnt[v] = 0;
  if (RANK) {
    for (int v = t; v < win; v += kCB) s_rank[v] = ~0u;
    const int f = pl.first[0];
    for (int p = t; p < P; p += kCB) s_prank[p] = p == f ? 0 : (p < f ? p + 1 : p);
  }
  // Rows are taken in the order of partition 0's member list (points
  // grouped by their partition-0 cluster) when every point is assigned
  // there: rows in flight at the same time then share most of their
  // clusters, so their member runs are read from cache. The output does not
  // depend on the order (each row writes its own segment).
  // The row order is cut into pl.ranges contiguous ranges with a counter
  // each; workgroup b starts on range b % ranges. Workgroups are dispatched
  // to the XCDs round-robin, so with 8 ranges each XCD (own L2) works
  // through its own stretch of the order and its resident rows share their
  // clusters' member lists in that L2. A workgroup whose range is used up
  // moves on to the next ranges (thread 0 keeps the cursor).
  const int32_t* order = nullptr;
  {
    const int32_t a = pl.cstart[pl.k_off[0]], b = pl.cstart[pl.k_off[1]];
    if ((int64_t)(b - a) == n) order = pl.members + a;
  }
  // a sample (row_stride > 1) takes rows 0, stride, 2 * stride, ...
  const int64_t stride = pl.row_stride > 1 ? pl.row_stride : 1;
  const int64_t nrows = (n + stride - 1) / stride;
  if (stride > 1) order = nullptr;
  int xr = (int)(blockIdx.x % (unsigned)pl.ranges), tries = 0;
  // PASS 2: the workgroup's pool chunk [chunk_pos, chunk_end) (uniform)
  int64_t chunk_pos = 0, chunk_end = 0;
  uint32_t mx_row = 0;  // PASS 2: largest count this lane emitted
  if (t == 0) s_max = 0;  // a workgroup may get no row
  for (;;) {
    if (t == 0) {
      int64_t r = nrows;
      while (tries < pl.ranges) {
        const int64_t lo = nrows * xr / pl.ranges, hi = nrows * (xr + 1) / pl.ranges;
        const int64_t q = lo + (int64_t)atomicAdd(&row_ctr[xr], 1ull);
        if (q < hi) {
          r = q;
          break;
        }
        xr = xr + 1 == pl.ranges ? 0 : xr + 1;
        ++tries;
      }
      s_row = r;
    }
    __syncthreads();
    const int64_t r = s_row;
    if (r >= nrows) break;
    const int64_t i = order ? (int64_t)order[r] : r * stride;
    if (t == 0) s_max = 0;
    for (int p = t; p < P; p += kCB) {
      int32_t q = pl.pos[(int64_t)p * n + i];
      if (q < 0) {  // unassigned in partition p
        s_cur[p] = 0;
        s_end[p] = 0;
      } else {
        int64_t g = pl.k_off[p] + pl.labels[(int64_t)p * n + i];
        s_cur[p] = full_rows ? pl.cstart[g] : q + 1;
        s_end[p] = pl.cstart[g + 1];
        s_g[p] = (int32_t)g;
      }
    }
    int64_t row_base = PASS == 1 ? indptr[i] : 0;
    int64_t written = 0;
    // PASS 2: the row's first segment-table slot and its first window
    const int64_t w_first = full_rows ? 0 : (i + 1) / win;
    const int64_t seg0 = PASS == 2 ? seg_base(i, win, pl.nw, full_rows) : 0;
    uint32_t nnz_local = 0;
    // windows aligned to multiples of win (columns <= i get no counts)
    for (int64_t c0 = full_rows ? 0 : ((i + 1) / win) * win; c0 < n; c0 += win) {
      const int64_t c1 = c0 + win < n ? c0 + win : n;
      for (int v = t; v < bmw; v += kCB) s_bm[v] = 0ull;  // read back after the next barrier
      // members of this window: [cur, hi) with member < c1 (lists ascending);
      // the last window takes every remaining member, no search. Otherwise
      // an exponential probe from cur (a window holds few members of each
      // partition), then a binary search inside the bracket.
      if (c1 == n) {
        for (int p = t; p < P; p += kCB) s_hi[p] = s_end[p];
      } else if (pl.wtab) {
        const int64_t w1 = c1 / win;
        for (int p = t; p < P; p += kCB)
          s_hi[p] = s_cur[p] >= s_end[p] ? s_end[p]
                                         : pl.wtab[(int64_t)s_g[p] * (pl.nw + 1) + w1];
      } else {
        for (int p = t; p < P; p += kCB) {
          int32_t lo = s_cur[p], hi = s_end[p];
          for (int32_t step = 16; lo < hi; step <<= 1) {
            const int32_t b = hi - lo > step ? lo + step : hi;
            if (pl.members[b - 1] < c1) {
              lo = b;  // all of [lo, b) lie in the window
            } else {
              hi = b;
              break;
            }
          }
          while (lo < hi) {
            int32_t mid = (lo + hi) >> 1;
            if (pl.members[mid] < c1) lo = mid + 1; else hi = mid;
          }
          s_hi[p] = lo;
        }
      }
      __syncthreads();
      // flattened ranges: prefix of (hi - cur) over partitions; the runs
      // (non-empty ranges) in partition order, each start marked in the
      // bitmap over the flattened indices (short windows), and per bitmap
      // word the run starts before it. One wavefront: its LDS operations
      // complete in order, so its bitmap reads see its own marks.
      if (t < kWave) {
        int32_t carry = 0, rcarry = 0;
        const uint64_t below = (1ull << t) - 1ull;
        for (int p0 = 0; p0 < P; p0 += kWave) {
          int p = p0 + t;
          int32_t len = p < P ? s_hi[p] - s_cur[p] : 0;
          int32_t inc = wave_incl_scan(len);
          const uint64_t ne = __ballot(len > 0);
          if (p < P) {
            const int32_t pre = carry + inc - len;
            s_pre[p] = pre;
            s_d[p] = s_cur[p] - pre;
            if (len > 0) {
              const int r = rcarry + __popcll(ne & below);
              s_dr[r] = s_cur[p] - pre;
              if (RANK) s_rp[r] = s_prank[p];
              if (pre < 64 * bmw) atomicOr(&s_bm[pre >> 6], 1ull << (pre & 63));
            }
          }
          carry += __shfl(inc, kWave - 1, kWave);
          rcarry += __popcll(ne);
        }
        if (t == 0) s_pre[P] = carry;
        if (carry < 32 * P) {  // the bitmap covers the window: word prefix counts
          const int W = (carry + 63) >> 6;
          int32_t wc = 0;
          for (int w0 = 0; w0 < W; w0 += kWave) {
            const int w = w0 + t;
            const int32_t c = w < W ? (int32_t)__popcll(s_bm[w]) : 0;
            const int32_t ci = wave_incl_scan(c);
            if (w < W) s_wpre[w] = wc + ci - c;
            wc += __shfl(ci, kWave - 1, kWave);
          }
        }
      }
      __syncthreads();
      const int32_t tot = s_pre[P];
      // Flattened member index f -> partition p with s_pre[p] <= f <
      // s_pre[p+1], member position f + s_d[p]; U member loads in flight
      // per lane, consecutive lanes on consecutive f (coalesced). Long runs
      // (32+ members per partition on average): lanes stride the block
      // (f = t, t + kCB, ...) and walk the few boundaries each step crosses.
      // Short runs: each wavefront takes U * 64 consecutive f per step; a
      // wave-uniform cursor finds the partition of the step's first f by a
      // ballot over 64 boundaries at a time; the next 64 boundaries are held
      // one per lane, and for each 64-wide sub-block one ballot counts
      // those before it while the few falling inside it are compared one by
      // one (a wave-uniform readlane each), so the search costs a few
      // instructions per boundary crossed and no LDS traffic (steps
      // spanning more than 64 partitions walk instead).
      if (tot >= 32 * P) {
        int p = 0;
        for (int32_t f0 = t; f0 < tot; f0 += U * kCB) {
          int32_t j[U];
          int pu[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int32_t f = f0 + u * kCB;
            j[u] = -1;
            if (f < tot) {
              while (s_pre[p + 1] <= f) ++p;
              j[u] = pl.members[f + s_d[p]];
            }
            pu[u] = p;
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            if (j[u] < 0 || (full_rows && j[u] == (int32_t)i)) continue;
            const int jj = (int)(j[u] - c0);
            atomicAdd(&s_cnt[jj >> 1], 1u << (16 * (jj & 1)));
            if (RANK) atomicMin(&s_rank[jj], (uint32_t)s_prank[pu[u]]);
          }
        }
      } else {
        // short runs: each wavefront takes U * 64 consecutive f per step;
        // f's run = (run starts at or before f) - 1: the starts before its
        // 64-bit bitmap word (prefix) plus a popcount of the word up to f
        const int wv = t / kWave, ln = lane_id();
        const uint64_t upto = ln == kWave - 1 ? ~0ull : ((2ull << ln) - 1ull);
        for (int32_t base = wv * U * kWave; base < tot; base += U * kCB) {
          int32_t j[U];
          int32_t rk[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int32_t f0 = base + u * kWave, f = f0 + ln;
            j[u] = -1;
            rk[u] = 0;
            if (f < tot) {
              const int w = f0 >> 6;
              const int r = s_wpre[w] + (int)__popcll(s_bm[w] & upto) - 1;
              j[u] = pl.members[f + s_dr[r]];
              if (RANK) rk[u] = s_rp[r];
            }
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            if (j[u] < 0 || (full_rows && j[u] == (int32_t)i)) continue;
            const int jj = (int)(j[u] - c0);
            atomicAdd(&s_cnt[jj >> 1], 1u << (16 * (jj & 1)));
            if (RANK) atomicMin(&s_rank[jj], (uint32_t)rk[u]);
          }
        }
      }
      __syncthreads();
      for (int p2 = t; p2 < P; p2 += kCB) s_cur[p2] = s_hi[p2];
      // read the window four words (eight columns) per lane and load, clearing
      // the words it used; PASS 1 emits them in column order
      const int width = (int)(c1 - c0);
      const int words = (width + 1) >> 1;
      const int groups = (words + 3) >> 2;  // 4-word groups, the last may be partial
      uint4* s4 = reinterpret_cast<uint4*>(s_cnt);
      // words 4g..4g+3 of group g (zeros past the window), cleared when used
      auto read_group = [&](int g, uint32_t* xw) {
        if (4 * g + 4 <= words) {
          const uint4 x = s4[g];
          xw[0] = x.x; xw[1] = x.y; xw[2] = x.z; xw[3] = x.w;
          if (x.x | x.y | x.z | x.w) s4[g] = make_uint4(0u, 0u, 0u, 0u);
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int v = 4 * g + q;
            xw[q] = v < words ? s_cnt[v] : 0u;
            if (xw[q]) s_cnt[v] = 0u;
          }
        }
      };
      if (PASS == 0) {
        uint32_t c = 0, mx = 0;
        for (int g = t; g < groups; g += kCB) {
          uint32_t xw[4];
          read_group(g, xw);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint32_t lo = xw[q] & 0xffffu, hi = xw[q] >> 16;
            c += (lo != 0) + (hi != 0);
            mx = lo > mx ? lo : mx;
            mx = hi > mx ? hi : mx;
          }
        }
        nnz_local += c;
        if (mx) atomicMax(&s_max, mx);
      } else {  // PASS 1 and 2
        // wave w emits groups [w*RG, (w+1)*RG) in column order
        const int w = t / kWave, l = lane_id();
        const int RG = ((groups + 4 * kWave - 1) / (4 * kWave)) * kWave;
        const int ga = w * RG, gb = (w + 1) * RG < groups ? (w + 1) * RG : groups;
        // non-zero 16-bit halves of a group: the high bit of each half of
        // ((x & 0x7fff7fff) + 0x7fff7fff) | x is set iff the half is non-zero
        auto nz_count = [](const uint32_t* xw) {
          uint32_t c = 0;
#pragma unroll
          for (int q = 0; q < 4; ++q)
            c += __popc((((xw[q] & 0x7fff7fffu) + 0x7fff7fffu) | xw[q]) & 0x80008000u);
          return c;
        };
        uint32_t c = 0;
        for (int g = ga + l; g < gb; g += kWave) {
          uint32_t xw[4];
          if (4 * g + 4 <= words) {
            const uint4 x = s4[g];
            xw[0] = x.x; xw[1] = x.y; xw[2] = x.z; xw[3] = x.w;
          } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) xw[q] = 4 * g + q < words ? s_cnt[4 * g + q] : 0u;
          }
          c += nz_count(xw);
        }
        c = wave_sum(c);
        if (l == 0) s_tmp[w] = c;
        __syncthreads();
        int64_t o = row_base + written;
        uint32_t wtot = 0;
        for (int q = 0; q < kCB / kWave; ++q) {
          if (q < w) o += s_tmp[q];
          wtot += s_tmp[q];
        }
        if (PASS == 2) {  // block-uniform: wtot, chunk_pos, chunk_end
          if (chunk_pos + (int64_t)wtot > chunk_end) {
            if (t == 0) s_chunk = atomicAdd(out.pool_ctr, (unsigned long long)out.pool_chunk);
            __syncthreads();
            chunk_pos = (int64_t)s_chunk;
            chunk_end = chunk_pos + out.pool_chunk;
            if (chunk_end > out.pool_cap) {
              if (t == 0) atomicOr(out.overflow, 1u);
              return;  // the whole workgroup: the caller rebuilds in two passes
            }
          }
          o += chunk_pos - row_base - written;
          if (t == 0)
            out.seg[seg0 + c0 / win - w_first] =
                wtot ? ((uint64_t)chunk_pos << 16) | (uint64_t)wtot : 0ull;
          chunk_pos += wtot;
        }
        for (int g0 = ga; g0 < gb; g0 += kWave) {
          const int g = g0 + l;
          uint32_t xw[4] = {0u, 0u, 0u, 0u};
          if (g < gb) read_group(g, xw);
          const uint32_t cnt = nz_count(xw);
          if (__ballot(cnt != 0u) == 0ull) continue;  // wave-uniform: no entry in these groups
          const uint32_t incl = wave_incl_scan(cnt);
          int64_t at = o + (incl - cnt);
          if (PASS == 2) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const uint32_t lo = xw[q] & 0xffffu, hi = xw[q] >> 16;
              const uint32_t cw = 2u * (uint32_t)(4 * g + q);  // column - c0
              if (lo) out.pool[at++] = (lo << 16) | cw;
              if (hi) out.pool[at++] = (hi << 16) | (cw + 1u);
              mx_row = lo > mx_row ? lo : mx_row;
              mx_row = hi > mx_row ? hi : mx_row;
            }
            o += __shfl(incl, kWave - 1, kWave);
            continue;
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint32_t lo = xw[q] & 0xffffu, hi = xw[q] >> 16;
            const int32_t col = (int32_t)(c0 + 2 * (4 * g + q));
            if (lo) {
              out_col[at] = col;
              out_val[at] = w_offset ? w_offset - (int32_t)lo : (int32_t)lo;
              if (RANK) {
                out_rank[at] = (int32_t)s_rank[2 * (4 * g + q)];
                s_rank[2 * (4 * g + q)] = ~0u;
              }
              ++at;
            }
            if (hi) {
              out_col[at] = col + 1;
              out_val[at] = w_offset ? w_offset - (int32_t)hi : (int32_t)hi;
              if (RANK) {
                out_rank[at] = (int32_t)s_rank[2 * (4 * g + q) + 1];
                s_rank[2 * (4 * g + q) + 1] = ~0u;
              }
              ++at;
            }
          }
          o += __shfl(incl, kWave - 1, kWave);
        }
        written += wtot;
      }
      __syncthreads();
    }
    if (PASS == 0) {
      uint32_t rt;
      (void)block_excl_scan<uint32_t>(nnz_local, s_tmp, &rt);
      if (t == 0) {
        out.row_nnz[i] = (int32_t)rt;
        if (out.nnz_sum) atomicAdd(out.nnz_sum, (unsigned long long)rt);
        if (s_max) atomicMax(out.max_count, s_max);
      }
    }
    if (PASS == 2 && t == 0) out.row_nnz[i] = (int32_t)written;
    __syncthreads();
  }
  if (PASS == 2) {
    if (mx_row) atomicMax(&s_max, mx_row);
    __syncthreads();
    if (t == 0 && s_max) atomicMax(out.max_count, s_max);
  }
}

// One-pass build, last step: copies each row's staged segments, in window
// order, to indices/data at indptr[i] (one wavefront per row; the segment
// headers are wave-uniform loads).
// With ctl (the capturable build's device results: nnz, max count, error
// bits) nothing is written when an error bit is set, and w_offset = -1
// takes max count + 1 from ctl.
__global__ __launch_bounds__(256) void k_coassoc_compact(
    const unsigned char* __restrict__ stage, const int64_t* __restrict__ indptr, int64_t n,
    int32_t w_offset, int32_t* __restrict__ out_col, int32_t* __restrict__ out_val,
    const int64_t* __restrict__ ctl = nullptr) {
  if (ctl) {
    if (ctl[2] != 0) return;
    if (w_offset == -1) w_offset = (int32_t)ctl[1] + 1;
  }
  const int64_t* hdr = reinterpret_cast<const int64_t*>(stage);
  const int win = (int)hdr[0], nw = (int)hdr[1], full_rows = (int)hdr[5];
  const uint64_t* __restrict__ seg = reinterpret_cast<const uint64_t*>(stage + 256);
  const uint32_t* __restrict__ pool = reinterpret_cast<const uint32_t*>(stage + hdr[4]);
  const int l = lane_id();
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  for (int64_t i = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; i < n;
       i += waves) {
    const int64_t w_first = full_rows ? 0 : (i + 1) / win;
    const uint64_t* sg = seg + seg_base(i, win, nw, full_rows);
    int64_t dst = indptr[i];
    for (int64_t w = w_first; w < nw; ++w) {
      const uint64_t h = sg[w - w_first];
      const int cnt = (int)(h & 0xffffu);
      if (!cnt) continue;
      const uint32_t* src = pool + (h >> 16);
      const int32_t c0 = (int32_t)(w * win);
      for (int e = l; e < cnt; e += kWave) {
        const uint32_t x = src[e];
        const int32_t c = (int32_t)(x >> 16);
        out_col[dst + e] = c0 + (int32_t)(x & 0xffffu);
        out_val[dst + e] = w_offset ? w_offset - c : c;
      }
      dst += cnt;
    }
  }
}

struct CoBufs {
  int64_t* k_off;
  uint32_t* key;
  uint32_t* val;
  uint32_t* skey;
  uint32_t* sval;
  int32_t* members;
  int32_t* pos;
  int32_t* csize;
  int32_t* cstart;
  int32_t* row_nnz;
  unsigned long long* ctr;  // [kCtrs]: row counters, [kMaxRanges] sampled nnz sum
  unsigned int* misc;  // [0] bad label, [1] max count
};

CoBufs carve_co(Arena& ar, int64_t n, int P, int64_t ktot) {
  CoBufs b;
  int64_t items = (int64_t)P * n;
  b.k_off = ar.take<int64_t>(P + 1);
  b.members = ar.take<int32_t>(items);
  b.pos = ar.take<int32_t>(items);
  b.csize = ar.take<int32_t>(ktot + 2);
  b.cstart = ar.take<int32_t>(ktot + 2);
  b.row_nnz = ar.take<int32_t>(n + 1);
  b.ctr = ar.take<unsigned long long>(kCtrs);
  b.misc = ar.take<unsigned int>(4);
  // sort staging last: reused by later passes
  b.key = ar.take<uint32_t>(items);
  b.val = ar.take<uint32_t>(items);
  b.skey = ar.take<uint32_t>(items);
  b.sval = ar.take<uint32_t>(items);
  return b;
}

int d2h(void* host, const void* dev, size_t bytes, hipStream_t s) {
  void* pin = pinned_scratch(bytes);
  if (!pin) return fail(EAC_ERR_HIP, "pinned host allocation failed");
  EAC_HIP_CHECK(hipMemcpyAsync(pin, dev, bytes, hipMemcpyDeviceToHost, s));
  EAC_HIP_CHECK(hipStreamSynchronize(s));
  memcpy(host, pin, bytes);
  return EAC_OK;
}

int co_check(const int32_t* labels, int64_t n, const int64_t* h_koff, int32_t P) {
  if (n < 1 || P < 1) return fail(EAC_ERR_ARG, "coassoc: need n, P >= 1");
  if (P > kMaxP) return fail(EAC_ERR_UNSUPPORTED, "coassoc: more than %d partitions", kMaxP);
  if (!labels || !h_koff) return fail(EAC_ERR_ARG, "coassoc: null array");
  if (h_koff[0] != 0) return fail(EAC_ERR_ARG, "coassoc: k_off[0] must be 0");
  for (int p = 0; p < P; ++p)
    if (h_koff[p + 1] < h_koff[p]) return fail(EAC_ERR_ARG, "coassoc: k_off not ascending");
  if ((int64_t)P * n >= (int64_t)INT32_MAX)
    return fail(EAC_ERR_UNSUPPORTED, "coassoc: P*n too large");
  return EAC_OK;
}

}  // namespace
// Columns per LDS window (even, <= kWin): the whole row range when n is
// small (fewer LDS bytes, more resident workgroups), else kWin, or the
// EAC_TUNE_CO_WINDOW setting (the autotuner trades windows per row against
// workgroups per compute unit).
int co_window(int64_t n) {
  int w = tune_value(EAC_TUNE_CO_WINDOW);
  if (w <= 0) w = kWin;
  if (w > kWin) w = kWin;
  const int64_t need = ((n + 63) / 64) * 64;
  if (need < w) w = (int)need;
  return w < 64 ? 64 : (w & ~1);
}
// Member loads in flight per lane in the pair visits (EAC_TUNE_CO_UNROLL:
// 4 by default, or 8).
inline int co_unroll() { return tune_value(EAC_TUNE_CO_UNROLL) == 8 ? 8 : 4; }
// Row ranges of the co-association row loop (EAC_TUNE_CO_RANGES: 1..8;
// default one per XCD, or one range below 65536 rows: there every
// workgroup takes few rows and the end of the pass, where each row fetch
// tries every range's counter, dominates; 10,000 rows x 100 partitions:
// the EAC pipeline's graph replay 1.51 ms with one range, 1.74 with eight).
inline int co_ranges(int64_t n) {
  const int r = tune_value(EAC_TUNE_CO_RANGES);
  if (r > 0 && r <= kMaxRanges) return r;
  return n < 65536 ? 1 : kMaxRanges;
}
inline size_t co_lds(int win, int P, bool rank = false) {
  return (size_t)(2 * co_bm_words(P) + win / 2 + (rank ? win : 0) + 9 * P + 1 +
                  (P / 2 + 1)) * 4;
}
// The ranked fill keeps a 4-byte rank per column next to the counters: its
// window is at most half as wide.
constexpr int kWinRank = kWin / 2;

// Builds the window table in the sort staging (free after the plan) when it
// fits; returns nullptr otherwise (the rows then search their window ends).
int32_t* build_win_table(const CoBufs& B, int64_t n, int P, int64_t ktot, int win, int* nw,
                         hipStream_t s) {
  *nw = (int)((n + win - 1) / win);
  const int64_t items = (int64_t)P * n;
  if (ktot * (int64_t)(*nw + 1) > items) return nullptr;
  int32_t* wtab = reinterpret_cast<int32_t*>(B.key);
  k_win_table<<<grid_for(items, 256, 16384), 256, 0, s>>>(B.skey, B.members, B.cstart, items,
                                                          ktot, win, *nw, wtab);
  if (hipGetLastError() != hipSuccess) return nullptr;
  return wtab;
}
// Every workgroup resident at once (rows are handed out dynamically).
template <int PASS, bool RANK = false, int U = 4>
int co_grid(int win, int P) {
  int dev = 0, cus = 256, per = 1;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // allow launch-time LDS beyond 64 KB (the window of large graphs)
  const size_t lds = co_lds(win, P, RANK);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_coassoc_rows<PASS, RANK, U>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_coassoc_rows<PASS, RANK, U>, kCB,
                                                   lds) != hipSuccess || per < 1)
    per = 1;
  return cus * per;
}

}  // namespace eac

// ------------------------------------------------------------------ C ABI
using namespace eac;

extern "C" size_t eac_coassoc_workspace_bytes(int64_t n, int32_t P, int64_t k_total) {
  if (n < 1) n = 1;
  if (P < 1) P = 1;
  Arena ar(nullptr, 0, true);
  (void)carve_co(ar, n, P, k_total);
  size_t s1 = radix_sort_ws_bytes((int64_t)P * n, 4, 4);
  size_t s2 = scan_ws_bytes(k_total + 2, 4);
  size_t s3 = scan_ws_bytes(n + 1, 8);
  size_t m = s1 > s2 ? s1 : s2;
  m = m > s3 ? m : s3;
  return ar.used + m + 4096;
}

namespace eac {
namespace {
// The member plan in ws (shared by every pass over the same labels):
// global cluster keys, the stable sort of (cluster, point), cluster starts,
// each point's position in its cluster's list; zeroes the row counters and
// flags.
int co_plan_build(const int32_t* labels, int64_t n, const int64_t* h_koff, int32_t P,
                  void* ws, size_t ws_bytes, CoBufs* out, size_t* mark_out, Arena& ar,
                  hipStream_t s, const int64_t* d_koff = nullptr) {
  EAC_TRY(co_check(labels, n, h_koff, P));
  const int64_t ktot = h_koff[P];
  size_t need = eac_coassoc_workspace_bytes(n, P, ktot);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "coassoc: workspace %zu < %zu bytes", ws_bytes, need);
  CoBufs B = carve_co(ar, n, P, ktot);
  const size_t mark = ar.used;
  const int64_t items = (int64_t)P * n;
  if (d_koff) {  // already on the device: a device copy (capturable)
    EAC_HIP_CHECK(hipMemcpyAsync(B.k_off, d_koff, sizeof(int64_t) * (P + 1),
                                 hipMemcpyDeviceToDevice, s));
  } else {
    EAC_HIP_CHECK(hipMemcpyAsync(B.k_off, h_koff, sizeof(int64_t) * (P + 1),
                                 hipMemcpyHostToDevice, s));
    EAC_HIP_CHECK(hipStreamSynchronize(s));
  }
  EAC_TRY(fill<int32_t>(B.csize, ktot + 2, 0, s));
  EAC_TRY(fill<int32_t>((int32_t*)B.misc, 4, 0, s));
  EAC_TRY(fill<int64_t>((int64_t*)B.ctr, kCtrs, 0, s));
  int pr = prof_begin(EAC_PROF_CO_PLAN, 16 * items, s);
  k_member_keys<<<grid_for(items, 256, 8192), 256, 0, s>>>(labels, n, P, B.k_off, ktot, B.key,
                                                           B.val, B.csize, B.misc);
  EAC_LAUNCH_CHECK();
  const int bits = bits_for((uint64_t)ktot);
  EAC_TRY((radix_sort<uint32_t, uint32_t>(B.key, B.skey, B.val, B.sval, items, 0,
                                          bits < 1 ? 1 : bits, ar, s)));
  ar.used = mark;
  EAC_TRY((scan_excl<int32_t, int32_t>(B.csize, B.cstart, ktot + 2, 0, nullptr, ar, s)));
  ar.used = mark;
  k_member_pos<<<grid_for(items, 256, 8192), 256, 0, s>>>(B.skey, B.sval, items, n, ktot,
                                                          B.members, B.pos);
  EAC_LAUNCH_CHECK();
  prof_end(pr, s);
  *out = B;
  *mark_out = mark;
  return EAC_OK;
}

// Launches the row kernel for PASS 0 or 2 (unranked) with the unroll knob.
template <int PASS>
int co_launch(const CoPlan& pl, int full_rows, int win, const CoOut& o, hipStream_t s) {
  if (co_unroll() == 8)
    k_coassoc_rows<PASS, false, 8><<<co_grid<PASS, false, 8>(win, pl.P), kCB,
                                     co_lds(win, pl.P), s>>>(pl, full_rows, win, o);
  else
    k_coassoc_rows<PASS><<<co_grid<PASS>(win, pl.P), kCB, co_lds(win, pl.P), s>>>(pl, full_rows,
                                                                                 win, o);
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}
}  // namespace
}  // namespace eac

// Pass 1: builds the member plan in ws (keep ws unchanged for the fill),
// writes indptr[n+1] (int64) and returns nnz and the largest count.
extern "C" int eac_coassoc_count(const int32_t* labels, int64_t n, const int64_t* h_koff,
                                 int32_t P, int32_t full_rows, int64_t* indptr, int64_t* h_nnz,
                                 int32_t* h_max_count, void* ws, size_t ws_bytes, void* stream) {
  EAC_TRY(co_check(labels, n, h_koff, P));
  if (!indptr) return fail(EAC_ERR_ARG, "coassoc: null indptr");
  hipStream_t s = (hipStream_t)stream;
  const int64_t ktot = h_koff[P];
  Arena ar(ws, ws_bytes);
  CoBufs B;
  size_t mark = 0;
  EAC_TRY(co_plan_build(labels, n, h_koff, P, ws, ws_bytes, &B, &mark, ar, s));
  const int win = co_window(n);
  int nw = 0;
  const int32_t* wtab = build_win_table(B, n, P, ktot, win, &nw, s);
  CoPlan pl{labels, n, P, B.k_off, B.members, B.pos, B.cstart, wtab, nw, nullptr, co_ranges(n), 1};
  CoOut o{};
  o.row_ctr = B.ctr;
  o.row_nnz = B.row_nnz;
  o.max_count = B.misc + 1;
  int pr2 = prof_begin(EAC_PROF_CO_COUNT, 0, s);
  EAC_TRY(co_launch<0>(pl, full_rows, win, o, s));
  prof_end(pr2, s);
  EAC_TRY((scan_excl<int32_t, int64_t>(B.row_nnz, indptr, n + 1, 0, nullptr, ar, s)));
  ar.used = mark;
  unsigned int misc[4];
  EAC_TRY(d2h(misc, B.misc, sizeof(misc), s));
  int64_t nnz = 0;
  EAC_TRY(d2h(&nnz, indptr + n, sizeof(int64_t), s));
  prof_flush();
  if (misc[0]) return fail(EAC_ERR_ARG, "coassoc: label >= k of its partition");
  if (h_nnz) *h_nnz = nnz;
  if (h_max_count) *h_max_count = (int32_t)misc[1];
  return EAC_OK;
}

// Pass 2 (same ws as pass 1): indices[nnz] (int32 columns, ascending per
// row) and data[nnz] = count, or w_offset - count when w_offset != 0
// (make_diassoc, sparse.py:295-311, with w_offset = max + 1).
namespace eac {
namespace {
int coassoc_fill(const int32_t* labels, int64_t n, const int64_t* h_koff, int32_t P,
                 int32_t full_rows, const int64_t* indptr, int32_t w_offset, int32_t* indices,
                 int32_t* data, int32_t* rank, void* ws, size_t ws_bytes, hipStream_t s) {
  EAC_TRY(co_check(labels, n, h_koff, P));
  if (!indptr || !indices || !data) return fail(EAC_ERR_ARG, "coassoc: null output");
  const int64_t ktot = h_koff[P];
  size_t need = eac_coassoc_workspace_bytes(n, P, ktot);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "coassoc: workspace %zu < %zu bytes", ws_bytes, need);
  Arena ar(ws, ws_bytes);
  CoBufs B = carve_co(ar, n, P, ktot);
  EAC_TRY(fill<int64_t>((int64_t*)B.ctr, kCtrs, 0, s));
  int64_t nnz = 0;
  EAC_TRY(d2h(&nnz, indptr + n, sizeof(int64_t), s));
  int win = co_window(n);
  if (rank && win > kWinRank) win = kWinRank;
  int nw = 0;
  const int32_t* wtab = build_win_table(B, n, P, ktot, win, &nw, s);
  int32_t* d_first = reinterpret_cast<int32_t*>(B.misc + 2);
  CoPlan pl{labels, n, P, B.k_off, B.members, B.pos, B.cstart, wtab, nw, d_first, co_ranges(n), 1};
  CoOut o{};
  o.row_ctr = B.ctr;
  o.indptr = indptr;
  o.w_offset = w_offset;
  o.col = indices;
  o.val = data;
  o.rank = rank;
  int pr = prof_begin(EAC_PROF_CO_FILL, 8 * nnz + 8 * (n + 1), s);
  if (rank) {
    k_first_partition<<<1, 256, 0, s>>>(B.cstart, B.k_off, P, d_first);
    EAC_LAUNCH_CHECK();
    k_coassoc_rows<1, true><<<co_grid<1, true>(win, P), kCB, co_lds(win, P, true), s>>>(
        pl, full_rows, win, o);
  } else if (co_unroll() == 8) {
    k_coassoc_rows<1, false, 8><<<co_grid<1, false, 8>(win, P), kCB, co_lds(win, P), s>>>(
        pl, full_rows, win, o);
  } else {
    k_coassoc_rows<1><<<co_grid<1>(win, P), kCB, co_lds(win, P), s>>>(pl, full_rows, win, o);
  }
  EAC_LAUNCH_CHECK();
  prof_end(pr, s);
  EAC_HIP_CHECK(hipStreamSynchronize(s));
  prof_flush();
  return EAC_OK;
}
}  // namespace
}  // namespace eac

extern "C" int eac_coassoc_fill(const int32_t* labels, int64_t n, const int64_t* h_koff,
                                int32_t P, int32_t full_rows, const int64_t* indptr,
                                int32_t w_offset, int32_t* indices, int32_t* data, void* ws,
                                size_t ws_bytes, void* stream) {
  return coassoc_fill(labels, n, h_koff, P, full_rows, indptr, w_offset, indices, data, nullptr,
                      ws, ws_bytes, (hipStream_t)stream);
}

// Pass 2 with, per entry, the processing rank of the first partition in
// which the pair co-occurs (the input of the capacity-capped selection).
extern "C" int eac_coassoc_fill_ranked(const int32_t* labels, int64_t n, const int64_t* h_koff,
                                       int32_t P, int32_t full_rows, const int64_t* indptr,
                                       int32_t* indices, int32_t* data, int32_t* rank, void* ws,
                                       size_t ws_bytes, void* stream) {
  if (!rank) return fail(EAC_ERR_ARG, "coassoc: null rank output");
  return coassoc_fill(labels, n, h_koff, P, full_rows, indptr, 0, indices, data, rank, ws,
                      ws_bytes, (hipStream_t)stream);
}

// ------------------------------------------------------- one-pass build
// The count pass visits every pair once only to size the output; the
// one-pass build visits them once: the plan plus a sampled count estimate
// the nnz, the caller allocates a staging buffer from it, the staged pass
// writes every window's entries (4 bytes each) to chunks of that buffer
// and the row sizes, and a compaction copies the segments into the CSR.
// A pool that turns out too small is reported (h_nnz = -1): the caller
// then runs eac_coassoc_count + eac_coassoc_fill.
namespace eac {
namespace {
// staging buffer: header (int64: win, nw, pool counter, overflow flag,
// pool byte offset, full_rows), then the segment table, then the pool
// (4-byte entries). The compaction takes the layout from the header.
constexpr size_t kStageHdr = 256;
__global__ void k_stage_header(int64_t* hdr, int win, int nw, int64_t pool_at, int full_rows) {
  if (threadIdx.x == 0) {
    hdr[0] = win;
    hdr[1] = nw;
    hdr[2] = 0;
    hdr[3] = 0;
    hdr[4] = pool_at;
    hdr[5] = full_rows;
  }
}
struct StageLayout {
  int win, nw;
  int64_t slots;     // segment-table entries
  size_t pool_at;    // byte offset of the pool
  int64_t chunk;     // pool entries per chunk: n / 8 rounded up to a power
                     // of two, at most kPoolChunk, and at least 16 windows
                     // (a chunk loses at most one window's width at its end)
};
StageLayout stage_layout(int64_t n, int full_rows) {
  StageLayout L;
  L.win = co_window(n);
  L.nw = (int)((n + L.win - 1) / L.win);
  L.slots = seg_base(n, L.win, L.nw, full_rows);
  L.pool_at = kStageHdr + (((size_t)L.slots * 8 + 255) & ~(size_t)255);
  int64_t c = 1;
  while (c < n / 8 && c < kPoolChunk) c <<= 1;
  L.chunk = c < 16 * (int64_t)L.win ? 16 * (int64_t)L.win : c;
  return L;
}
}  // namespace
}  // namespace eac

// Plan (as eac_coassoc_count) plus the count pass over rows 0, stride,
// 2 * stride, ...: h_estimate = their nnz sum x (rows / sampled rows),
// rounded up. Keep ws for eac_coassoc_stage.
extern "C" int eac_coassoc_estimate(const int32_t* labels, int64_t n, const int64_t* h_koff,
                                    int32_t P, int32_t full_rows, int64_t sample_stride,
                                    int64_t* h_estimate, void* ws, size_t ws_bytes,
                                    void* stream) {
  EAC_TRY(co_check(labels, n, h_koff, P));
  if (sample_stride < 1) return fail(EAC_ERR_ARG, "coassoc: sample stride must be >= 1");
  if (!h_estimate) return fail(EAC_ERR_ARG, "coassoc: null estimate");
  hipStream_t s = (hipStream_t)stream;
  const int64_t ktot = h_koff[P];
  Arena ar(ws, ws_bytes);
  CoBufs B;
  size_t mark = 0;
  EAC_TRY(co_plan_build(labels, n, h_koff, P, ws, ws_bytes, &B, &mark, ar, s));
  const int win = co_window(n);
  int nw = 0;
  const int32_t* wtab = build_win_table(B, n, P, ktot, win, &nw, s);
  CoPlan pl{labels, n, P,  B.k_off,     B.members,    B.pos, B.cstart,
            wtab,   nw, nullptr, co_ranges(n), sample_stride};
  CoOut o{};
  o.row_ctr = B.ctr;
  o.row_nnz = B.row_nnz;
  o.max_count = B.misc + 1;
  o.nnz_sum = B.ctr + kMaxRanges;
  int pr = prof_begin(EAC_PROF_CO_COUNT, 0, s);
  EAC_TRY(co_launch<0>(pl, full_rows, win, o, s));
  prof_end(pr, s);
  unsigned int misc[4];
  EAC_TRY(d2h(misc, B.misc, sizeof(misc), s));
  unsigned long long sum = 0;
  EAC_TRY(d2h(&sum, B.ctr + kMaxRanges, sizeof(sum), s));
  prof_flush();
  if (misc[0]) return fail(EAC_ERR_ARG, "coassoc: label >= k of its partition");
  const int64_t rows = (n + sample_stride - 1) / sample_stride;
  *h_estimate = (int64_t)(((long double)sum * n + rows - 1) / rows);
  return EAC_OK;
}

// Bytes of the staging buffer with room for pool_entries staged entries
// beyond the partly used chunks of the workgroups (the segment table and
// the chunk size depend on n, full_rows and the window knob).
extern "C" size_t eac_coassoc_stage_bytes(int64_t n, int32_t full_rows, int32_t P,
                                          int64_t pool_entries) {
  if (n < 1) n = 1;
  if (P < 1) P = 1;
  if (pool_entries < 0) pool_entries = 0;
  const StageLayout L = stage_layout(n, full_rows);
  const int64_t grid = co_unroll() == 8 ? co_grid<2, false, 8>(L.win, P) : co_grid<2>(L.win, P);
  return L.pool_at + (size_t)(pool_entries + (grid + 1) * L.chunk) * 4;
}

// Staged pass (ws from eac_coassoc_estimate): writes indptr[n+1] and, in
// stage, every row's window segments; h_nnz = nnz and h_max_count the
// largest count, or h_nnz = -1 when the pool was too small.
extern "C" int eac_coassoc_stage(const int32_t* labels, int64_t n, const int64_t* h_koff,
                                 int32_t P, int32_t full_rows, int64_t* indptr, void* stage,
                                 size_t stage_bytes, int64_t* h_nnz, int32_t* h_max_count,
                                 void* ws, size_t ws_bytes, void* stream) {
  EAC_TRY(co_check(labels, n, h_koff, P));
  if (!indptr || !stage || !h_nnz) return fail(EAC_ERR_ARG, "coassoc: null argument");
  const int64_t ktot = h_koff[P];
  size_t need = eac_coassoc_workspace_bytes(n, P, ktot);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "coassoc: workspace %zu < %zu bytes", ws_bytes, need);
  const StageLayout SL = stage_layout(n, full_rows);
  if (stage_bytes < SL.pool_at + (size_t)SL.chunk * 4)
    return fail(EAC_ERR_WORKSPACE, "coassoc: staging buffer %zu < %zu bytes", stage_bytes,
                SL.pool_at + (size_t)SL.chunk * 4);
  hipStream_t s = (hipStream_t)stream;
  Arena ar(ws, ws_bytes);
  CoBufs B = carve_co(ar, n, P, ktot);
  const size_t mark = ar.used;
  unsigned char* st = static_cast<unsigned char*>(stage);
  int64_t* hdr = reinterpret_cast<int64_t*>(st);
  k_stage_header<<<1, 64, 0, s>>>(hdr, SL.win, SL.nw, (int64_t)SL.pool_at, full_rows);
  EAC_LAUNCH_CHECK();
  EAC_TRY(fill<int64_t>((int64_t*)B.ctr, kCtrs, 0, s));
  EAC_TRY(fill<int32_t>((int32_t*)B.misc, 4, 0, s));
  int nw = 0;
  const int32_t* wtab = build_win_table(B, n, P, ktot, SL.win, &nw, s);
  CoPlan pl{labels, n, P, B.k_off, B.members, B.pos, B.cstart, wtab, nw, nullptr, co_ranges(n), 1};
  CoOut o{};
  o.row_ctr = B.ctr;
  o.row_nnz = B.row_nnz;
  o.max_count = B.misc + 1;
  o.seg = reinterpret_cast<uint64_t*>(st + kStageHdr);
  o.pool = reinterpret_cast<uint32_t*>(st + SL.pool_at);
  o.pool_ctr = reinterpret_cast<unsigned long long*>(hdr + 2);
  o.pool_cap = (int64_t)((stage_bytes - SL.pool_at) / 4);
  o.pool_chunk = SL.chunk;
  o.overflow = reinterpret_cast<unsigned int*>(hdr + 3);
  int pr = prof_begin(EAC_PROF_CO_STAGE, 0, s);
  EAC_TRY(co_launch<2>(pl, full_rows, SL.win, o, s));
  prof_end(pr, s);
  EAC_TRY((scan_excl<int32_t, int64_t>(B.row_nnz, indptr, n + 1, 0, nullptr, ar, s)));
  ar.used = mark;
  unsigned int misc[4];
  EAC_TRY(d2h(misc, B.misc, sizeof(misc), s));
  int64_t tail[2] = {0, 0};  // overflow flag (hdr[3])
  EAC_TRY(d2h(tail, hdr + 2, sizeof(tail), s));
  int64_t nnz = 0;
  EAC_TRY(d2h(&nnz, indptr + n, sizeof(int64_t), s));
  prof_flush();
  *h_nnz = tail[1] ? -1 : nnz;
  if (h_max_count) *h_max_count = (int32_t)misc[1];
  return EAC_OK;
}

// Compaction of a staged build: indices[nnz] / data[nnz] as
// eac_coassoc_fill writes them (data = count, or w_offset - count).
extern "C" int eac_coassoc_compact(int64_t n, const int64_t* indptr, const void* stage,
                                   size_t stage_bytes, int32_t w_offset, int32_t* indices,
                                   int32_t* data, void* stream) {
  if (n < 1) return fail(EAC_ERR_ARG, "coassoc: need n >= 1");
  if (!indptr || !stage || !indices || !data) return fail(EAC_ERR_ARG, "coassoc: null argument");
  if (stage_bytes < kStageHdr)
    return fail(EAC_ERR_WORKSPACE, "coassoc: staging buffer %zu < %zu bytes", stage_bytes,
                kStageHdr);
  hipStream_t s = (hipStream_t)stream;
  int pr = prof_begin(EAC_PROF_CO_COMPACT, 0, s);
  const int64_t blocks = (n + 3) / 4;
  k_coassoc_compact<<<(unsigned)(blocks < 65536 ? blocks : 65536), 256, 0, s>>>(
      static_cast<const unsigned char*>(stage), indptr, n, w_offset, indices, data, nullptr);
  EAC_LAUNCH_CHECK();
  prof_end(pr, s);
  EAC_HIP_CHECK(hipStreamSynchronize(s));
  prof_flush();
  return EAC_OK;
}

// ------------------------------------------- capturable one-pass build
// Results of the capturable build in the caller's device int64[4]: nnz,
// largest count, error bits (1: pool too small, 2: more than cap_nnz
// entries, 4: a label >= k of its partition), 0.
namespace eac {
namespace {
__global__ void k_co_finish(const int64_t* __restrict__ indptr, int64_t n,
                            const unsigned int* __restrict__ misc,
                            const int64_t* __restrict__ hdr, int64_t cap_nnz,
                            int64_t* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t nnz = indptr[n];
  int64_t err = 0;
  if (hdr[3]) err |= 1;
  if (nnz > cap_nnz) err |= 2;
  if (misc[0]) err |= 4;
  out[0] = nnz;
  out[1] = (int64_t)misc[1];
  out[2] = err;
  out[3] = 0;
}
}  // namespace
}  // namespace eac

// The one-pass build with no host read, copy or synchronisation (for
// capture in a HIP graph): the member plan from d_koff (h_koff still gives
// the sizes), the staged pass into `stage` (eac_coassoc_stage_bytes of the
// pool the caller can afford; no estimate), the row offsets, and the
// compaction into indices/data of capacity cap_nnz. w_offset as in
// eac_coassoc_fill, or -1 for (largest count + 1) taken on the device.
// d_out (device int64[4]): nnz, largest count, error bits (1: pool too
// small, 2: nnz > cap_nnz, 4: label >= k; no entry is written then).
extern "C" int eac_coassoc_graph(const int32_t* labels, int64_t n, const int64_t* h_koff,
                                 const int64_t* d_koff, int32_t P, int32_t full_rows,
                                 int32_t w_offset, int64_t* indptr, int32_t* indices,
                                 int32_t* data, int64_t cap_nnz, void* stage,
                                 size_t stage_bytes, int64_t* d_out, void* ws, size_t ws_bytes,
                                 void* stream) {
  EAC_TRY(co_check(labels, n, h_koff, P));
  if (!d_koff || !indptr || !indices || !data || !stage || !d_out)
    return fail(EAC_ERR_ARG, "coassoc graph: null argument");
  if (w_offset < -1) return fail(EAC_ERR_ARG, "coassoc graph: w_offset < -1");
  const int64_t ktot = h_koff[P];
  const StageLayout SL = stage_layout(n, full_rows);
  if (stage_bytes < SL.pool_at + (size_t)SL.chunk * 4)
    return fail(EAC_ERR_WORKSPACE, "coassoc graph: staging buffer %zu < %zu bytes", stage_bytes,
                SL.pool_at + (size_t)SL.chunk * 4);
  hipStream_t s = (hipStream_t)stream;
  Arena ar(ws, ws_bytes);
  CoBufs B;
  size_t mark = 0;
  EAC_TRY(co_plan_build(labels, n, h_koff, P, ws, ws_bytes, &B, &mark, ar, s, d_koff));
  unsigned char* st = static_cast<unsigned char*>(stage);
  int64_t* hdr = reinterpret_cast<int64_t*>(st);
  k_stage_header<<<1, 64, 0, s>>>(hdr, SL.win, SL.nw, (int64_t)SL.pool_at, full_rows);
  EAC_LAUNCH_CHECK();
  int nw = 0;
  const int32_t* wtab = build_win_table(B, n, P, ktot, SL.win, &nw, s);
  CoPlan pl{labels, n, P, B.k_off, B.members, B.pos, B.cstart, wtab, nw, nullptr, co_ranges(n), 1};
  CoOut o{};
  o.row_ctr = B.ctr;
  o.row_nnz = B.row_nnz;
  o.max_count = B.misc + 1;
  o.seg = reinterpret_cast<uint64_t*>(st + kStageHdr);
  o.pool = reinterpret_cast<uint32_t*>(st + SL.pool_at);
  o.pool_ctr = reinterpret_cast<unsigned long long*>(hdr + 2);
  o.pool_cap = (int64_t)((stage_bytes - SL.pool_at) / 4);
  o.pool_chunk = SL.chunk;
  o.overflow = reinterpret_cast<unsigned int*>(hdr + 3);
  int pr = prof_begin(EAC_PROF_CO_STAGE, 0, s);
  EAC_TRY(co_launch<2>(pl, full_rows, SL.win, o, s));
  prof_end(pr, s);
  EAC_TRY((scan_excl<int32_t, int64_t>(B.row_nnz, indptr, n + 1, 0, nullptr, ar, s)));
  ar.used = mark;
  k_co_finish<<<1, 64, 0, s>>>(indptr, n, B.misc, hdr, cap_nnz, d_out);
  EAC_LAUNCH_CHECK();
  int pr2 = prof_begin(EAC_PROF_CO_COMPACT, 0, s);
  const int64_t blocks = (n + 3) / 4;
  k_coassoc_compact<<<(unsigned)(blocks < 65536 ? blocks : 65536), 256, 0, s>>>(
      st, indptr, n, w_offset, indices, data, d_out);
  EAC_LAUNCH_CHECK();
  prof_end(pr2, s);
  return EAC_OK;
}

// ------------------------------------------------ capacity-capped selection
// The reference's EAC_CSR gives row r a fixed capacity (cap_off[r+1] -
// cap_off[r] condensed, max_assocs with full rows) and, processing the
// partitions in their order (the one with the fewest clusters first), adds
// a new column to a row only while the row is below capacity; a column met
// when the row is full is discarded, and discarded again at each later
// co-occurrence (sparse.py:956-1040 / :866-950). Rows never shrink, so a row
// keeps exactly the first `cap` of its columns in the order (rank of the
// first partition with the pair, column), and each kept column ends with
// its full count. From the exact matrix with per-entry ranks, the kept set
// of a row is {rank < R*} plus the first `need` columns of rank R*, where
// R* is the smallest rank whose cumulative entry count reaches cap. Full
// rows above id 405109 keep only the first partition's entries at count 1
// (the update returns before them, sparse.py:877-878). A first partition
// that gives a row more entries than its capacity makes the reference write
// past the row: reported as an error.
namespace eac {
namespace {
constexpr int kCapB = 256;
constexpr int64_t kDebugRow = 405109;

__global__ __launch_bounds__(kCapB) void k_cap_select(const int64_t* __restrict__ indptr,
                                                      const int32_t* __restrict__ rank,
                                                      int64_t n, int P, int full_rows,
                                                      const int64_t* __restrict__ cap_off,
                                                      int64_t max_assocs,
                                                      int32_t* __restrict__ kept,
                                                      int32_t* __restrict__ rstar,
                                                      int32_t* __restrict__ need,
                                                      unsigned long long* __restrict__ bad_row) {
  extern __shared__ int32_t s_hist[];  // P bins
  __shared__ int32_t s_tmp[kCapB / kWave + 1];
  __shared__ int32_t s_sel[2];
  const int t = threadIdx.x;
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const int64_t a = indptr[i], b = indptr[i + 1], total = b - a;
    const int64_t cap = full_rows ? max_assocs : cap_off[i + 1] - cap_off[i];
    const bool debug_row = full_rows && i > kDebugRow;
    if (total <= cap && !debug_row) {  // block-uniform: the row fits
      if (t == 0) {
        kept[i] = (int32_t)total;
        rstar[i] = INT32_MAX;
        need[i] = 0;
      }
      continue;
    }
    for (int v = t; v < P; v += kCapB) s_hist[v] = 0;
    if (t == 0) { s_sel[0] = 0; s_sel[1] = 0; }
    __syncthreads();
    for (int64_t e = a + t; e < b; e += kCapB) {
      const int32_t R = rank[e];
      if ((uint32_t)R < (uint32_t)P) atomicAdd(&s_hist[R], 1);
      else atomicAdd(&bad_row[3], 1ull);  // a rank outside [0, P): rejected by the caller
    }
    __syncthreads();
    const int32_t n0 = s_hist[0];
    if (t == 0 && n0 > cap) atomicMin(bad_row, (unsigned long long)i);
    if (debug_row) {
      if (t == 0) {
        kept[i] = n0;
        rstar[i] = 1;  // keep rank 0 only
        need[i] = 0;
      }
    } else {
      // contiguous bins per thread, block scan of their sums
      const int per = (P + kCapB - 1) / kCapB;
      const int b0 = t * per, b1 = b0 + per < P ? b0 + per : P;
      int32_t mine = 0;
      for (int r = b0; r < b1; ++r) mine += s_hist[r];
      int32_t tot;
      int32_t cum = block_excl_scan<int32_t>(mine, s_tmp, &tot);
      for (int r = b0; r < b1; ++r) {
        if (cum < cap && cum + s_hist[r] >= cap) {
          s_sel[0] = r;
          s_sel[1] = (int32_t)(cap - cum);
        }
        cum += s_hist[r];
      }
      __syncthreads();
      if (t == 0) {
        kept[i] = (int32_t)cap;
        rstar[i] = cap > 0 ? s_sel[0] : 0;
        need[i] = cap > 0 ? s_sel[1] : 0;
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kCapB) void k_cap_compact(
    const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
    const int32_t* __restrict__ count, const int32_t* __restrict__ rank, int64_t n,
    int full_rows, const int32_t* __restrict__ rstar, const int32_t* __restrict__ need,
    const int64_t* __restrict__ out_off, int32_t* __restrict__ out_col,
    int32_t* __restrict__ out_val, unsigned long long* __restrict__ discarded,
    unsigned int* __restrict__ max_kept) {
  __shared__ int32_t s_tmp[kCapB / kWave + 1];
  const int t = threadIdx.x;
  unsigned long long disc = 0;
  unsigned int mx = 0;
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const int64_t a = indptr[i], b = indptr[i + 1], o = out_off[i];
    const int32_t rs = rstar[i], nd = need[i];
    const bool debug_row = full_rows && i > kDebugRow;
    if (rs == INT32_MAX) {  // the whole row is kept
      for (int64_t e = a + t; e < b; e += kCapB) {
        const int32_t v = count[e];
        out_col[o + (e - a)] = indices[e];
        out_val[o + (e - a)] = v;
        mx = (unsigned)v > mx ? (unsigned)v : mx;
      }
      continue;
    }
    int32_t carry_star = 0, carry_keep = 0;
    for (int64_t base = a; base < b; base += kCapB) {  // block-uniform trip count
      const int64_t e = base + t;
      const bool valid = e < b;
      const int32_t R = valid ? rank[e] : INT32_MAX;
      const int32_t is_star = valid && R == rs;
      int32_t star_tot, keep_tot;
      const int32_t idx = carry_star + block_excl_scan<int32_t>(is_star, s_tmp, &star_tot);
      const int32_t keep = valid && (R < rs || (is_star && idx < nd));
      const int32_t pos = carry_keep + block_excl_scan<int32_t>(keep, s_tmp, &keep_tot);
      if (keep) {
        const int32_t v = debug_row ? 1 : count[e];
        out_col[o + pos] = indices[e];
        out_val[o + pos] = v;
        mx = (unsigned)v > mx ? (unsigned)v : mx;
      } else if (valid && !debug_row) {
        disc += (unsigned long long)count[e];
      }
      carry_star += star_tot;
      carry_keep += keep_tot;
    }
  }
  disc = wave_sum(disc);
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const unsigned int o2 = __shfl_xor(mx, off, kWave);
    mx = o2 > mx ? o2 : mx;
  }
  if (lane_id() == 0) {
    if (disc) atomicAdd(discarded, disc);
    if (mx) atomicMax(max_kept, mx);
  }
}

struct CapBufs {
  int32_t* kept;
  int32_t* rstar;
  int32_t* need;
  unsigned long long* misc;  // [0] bad row (min), [1] discarded, [2] max kept, [3] bad ranks
};

CapBufs carve_cap(Arena& ar, int64_t n) {
  CapBufs c;
  c.kept = ar.take<int32_t>(n + 1);
  c.rstar = ar.take<int32_t>(n);
  c.need = ar.take<int32_t>(n);
  c.misc = ar.take<unsigned long long>(4);
  return c;
}
}  // namespace
}  // namespace eac

extern "C" size_t eac_coassoc_cap_workspace_bytes(int64_t n) {
  if (n < 1) n = 1;
  Arena ar(nullptr, 0, true);
  (void)carve_cap(ar, n);
  return ar.used + scan_ws_bytes(n + 1, 8) + 4096;
}

extern "C" int eac_coassoc_cap(const int64_t* indptr, const int32_t* indices,
                               const int32_t* count, const int32_t* rank, int64_t n, int32_t P,
                               int32_t full_rows, const int64_t* cap_off, int64_t max_assocs,
                               int64_t out_capacity, int64_t* out_indptr, int32_t* out_indices,
                               int32_t* out_data, int64_t* h_out, void* ws, size_t ws_bytes,
                               void* stream) {
  if (n < 1 || P < 1 || P > kMaxP) return fail(EAC_ERR_ARG, "coassoc cap: need n >= 1, 1 <= P <= %d", kMaxP);
  if (!indptr || !count || !rank || !out_indptr || (!full_rows && !cap_off))
    return fail(EAC_ERR_ARG, "coassoc cap: null array");
  if (full_rows && max_assocs < 0) return fail(EAC_ERR_ARG, "coassoc cap: max_assocs < 0");
  const size_t need_ws = eac_coassoc_cap_workspace_bytes(n);
  if (!ws || ws_bytes < need_ws)
    return fail(EAC_ERR_WORKSPACE, "coassoc cap: workspace %zu < %zu bytes", ws_bytes, need_ws);
  hipStream_t s = (hipStream_t)stream;
  Arena ar(ws, ws_bytes);
  CapBufs B = carve_cap(ar, n);
  const size_t mark = ar.used;
  uint64_t* misc64 = reinterpret_cast<uint64_t*>(B.misc);
  EAC_TRY(fill<uint64_t>(misc64, 4, 0ull, s));
  EAC_TRY(fill<uint64_t>(misc64, 1, ~0ull, s));
  EAC_TRY(fill<int32_t>(B.kept + n, 1, 0, s));
  const int g = grid_for(n, 1, 4096);
  k_cap_select<<<g, kCapB, sizeof(int32_t) * P, s>>>(indptr, rank, n, P, full_rows, cap_off,
                                                     max_assocs, B.kept, B.rstar, B.need, B.misc);
  EAC_LAUNCH_CHECK();
  EAC_TRY((scan_excl<int32_t, int64_t>(B.kept, out_indptr, n + 1, 0, nullptr, ar, s)));
  ar.used = mark;
  unsigned long long bad[4] = {0, 0, 0, 0};
  int64_t total = 0;
  EAC_TRY(d2h(bad, B.misc, sizeof(bad), s));
  EAC_TRY(d2h(&total, out_indptr + n, sizeof(total), s));
  if (bad[3]) return fail(EAC_ERR_ARG, "coassoc cap: %llu ranks outside [0, P)", bad[3]);
  if (bad[0] != ~0ull)
    return fail(EAC_ERR_VALUE,
                "coassoc cap: the first partition gives row %llu more entries than its "
                "capacity (the reference writes past the row)", bad[0]);
  if (total > out_capacity)
    return fail(EAC_ERR_ARG, "coassoc cap: %lld kept entries exceed the output capacity %lld",
                (long long)total, (long long)out_capacity);
  if (total > 0 && (!out_indices || !out_data))
    return fail(EAC_ERR_ARG, "coassoc cap: null output");
  k_cap_compact<<<g, kCapB, 0, s>>>(indptr, indices, count, rank, n, full_rows, B.rstar, B.need,
                                    out_indptr, out_indices, out_data, B.misc + 1,
                                    reinterpret_cast<unsigned int*>(B.misc + 2));
  EAC_LAUNCH_CHECK();
  unsigned long long misc[3];
  EAC_TRY(d2h(misc, B.misc, sizeof(misc), s));
  if (h_out) {
    h_out[0] = total;
    h_out[1] = (int64_t)misc[1];
    h_out[2] = (int64_t)(unsigned int)misc[2];
    h_out[3] = -1;
  }
  return EAC_OK;
}

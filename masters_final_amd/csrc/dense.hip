// Dense co-association (EAC_FULL) and single linkage over it, for small n.
//
// Reference: MyML/MyML/EAC/full.py:8-181 (EAC_FULL: counts of every pair of
// samples, condensed or n x n, make_diassoc val - count),
// MyML/MyML/EAC/eac.py:337-391 (full_sl_lifetime: SLINK, the lifetime rule
// lifetime_n_clusters on the float32 heights, labels from Z),
// MyML/MyML/cluster/linkage.py:144-149 (SL: the SLINK wrapper returning the
// float32 linkage matrix Z) and :303-357 (labels_from_Z_numba: the first
// n - n_clusters merges, clusters numbered by first occurrence),
// scipy_numba/scipy_numba/cluster/_hierarchy_eac.py:842-927 (SLINK and
// from_pointer_representation).
//
// MI355X design:
//  * Counts: tiles of 64 x 64 sample pairs. The tile's 64 row and 64 column
//    labels of 32 partitions at a time are staged in LDS; every thread keeps
//    4 x 4 pair counts in registers and compares labels (n^2 P / 2
//    comparisons in all, far fewer operations than a one-hot product).
//  * The single-linkage dendrogram is the minimum spanning tree's merges in
//    height order. The dense graph is complete, so Borůvka needs at most
//    bits(n - 1) rounds; each round streams the n x n matrix once (a
//    workgroup per row, the columns' component ids read alongside) and takes
//    one 64-bit atomicMin per row on its component's slot. Keys (weight
//    word, min id, max id) are a strict total order: the tree is unique and
//    a mutual choice is the only cycle (the smaller root stays root).
//  * The lifetime rule runs in one workgroup over the sorted heights; the
//    labels of the cut are a parallel union-find (roots are component
//    minima, so a scan of the root flags numbers components by first
//    occurrence). The linkage matrix Z, when asked for, is built by one
//    thread from the sorted tree edges with a union-find in LDS; rows of
//    equal height come in (height, min id, max id) order.
#include "prims.hpp"
#include "sort.hpp"

namespace eac {
namespace {

constexpr int kDT = 64;   // dense count tile
constexpr int kDP = 32;   // partitions staged per step
constexpr int64_t kDenseMaxN = 65536;  // 16-bit ids in the Borůvka keys

__host__ __device__ __forceinline__ int64_t cidx(int64_t n, int64_t i, int64_t j) {
  return n * i - (i * (i + 1)) / 2 + (j - i - 1);  // i < j
}

template <typename T>
__global__ __launch_bounds__(256) void k_dense_counts(const int32_t* __restrict__ labels,
                                                      int64_t n, int P, int condensed,
                                                      T* __restrict__ out,
                                                      unsigned int* __restrict__ max_count) {
  const int ti = blockIdx.y, tj = blockIdx.x;
  if (ti > tj) return;  // block-uniform
  __shared__ int32_t s_r[kDP][kDT];
  __shared__ int32_t s_c[kDP][kDT];
  __shared__ unsigned int s_mx;
  const int t = threadIdx.x, ty = t / 16, tx = t % 16;
  const int64_t i0 = (int64_t)ti * kDT, j0 = (int64_t)tj * kDT;
  uint32_t cnt[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) cnt[a][b] = 0;
  if (t == 0) s_mx = 0;
  for (int p0 = 0; p0 < P; p0 += kDP) {
    const int pc = P - p0 < kDP ? P - p0 : kDP;
    __syncthreads();
    for (int v = t; v < kDP * kDT; v += 256) {
      const int pp = v / kDT, x = v % kDT;
      int32_t lr = -1, lc = -2;  // not listed / past the end: never equal
      if (pp < pc) {
        if (i0 + x < n) lr = labels[(int64_t)(p0 + pp) * n + i0 + x];
        if (j0 + x < n) lc = labels[(int64_t)(p0 + pp) * n + j0 + x];
      }
      s_r[pp][x] = lr < 0 ? -1 : lr;
      s_c[pp][x] = lc < 0 ? -2 : lc;
    }
    __syncthreads();
    for (int pp = 0; pp < pc; ++pp) {
      int32_t r[4], c[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) r[a] = s_r[pp][ty * 4 + a];
#pragma unroll
      for (int b = 0; b < 4; ++b) c[b] = s_c[pp][tx * 4 + b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) cnt[a][b] += r[a] == c[b];
    }
  }
  unsigned int mx = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t i = i0 + ty * 4 + a;
    if (i >= n) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t j = j0 + tx * 4 + b;
      if (j >= n) continue;
      const T v = (T)cnt[a][b];  // uint8 wraps as numpy's += 1 does
      if (condensed) {
        if (j > i) {
          out[cidx(n, i, j)] = v;
          mx = (unsigned)v > mx ? (unsigned)v : mx;
        }
      } else if (i == j) {
        out[i * n + i] = (T)0;
      } else {
        out[i * n + j] = v;
        out[j * n + i] = v;
        mx = (unsigned)v > mx ? (unsigned)v : mx;
      }
    }
  }
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const unsigned int o = __shfl_xor(mx, off, kWave);
    mx = o > mx ? o : mx;
  }
  if (lane_id() == 0 && mx) atomicMax(&s_mx, mx);
  __syncthreads();
  if (t == 0 && s_mx) atomicMax(max_count, s_mx);
}

// make_diassoc_1d / _2d (full.py:98-112): a[i] = val - a[i], stored in the
// matrix dtype (integer types wrap, as the reference's store does).
template <typename T>
__global__ void k_dense_diassoc(T* __restrict__ a, int64_t len, double val) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (T(0.5) == T(0))  // integer type: int64 arithmetic, modular store
      a[i] = (T)((int64_t)val - (int64_t)a[i]);
    else
      a[i] = (T)(val - (double)a[i]);
  }
}

// Square copy of a condensed matrix (both triangles; the diagonal is not
// read): 64 x 64 tiles through LDS so both the row-major read and the
// transposed write are coalesced.
template <typename T>
__global__ __launch_bounds__(256) void k_square_from_condensed(const T* __restrict__ cond,
                                                               int64_t n, T* __restrict__ sq) {
  const int ti = blockIdx.y, tj = blockIdx.x;
  if (ti > tj) return;
  __shared__ T s[kDT][kDT + 1];
  const int64_t i0 = (int64_t)ti * kDT, j0 = (int64_t)tj * kDT;
  for (int v = threadIdx.x; v < kDT * kDT; v += 256) {
    const int a = v / kDT, b = v % kDT;
    const int64_t i = i0 + a, j = j0 + b;
    T x = (T)0;
    if (i < n && j < n && j > i) x = cond[cidx(n, i, j)];
    else if (i < n && j < n && j < i) x = cond[cidx(n, j, i)];  // diagonal tile only
    s[a][b] = x;
    if (i < n && j < n) sq[i * n + j] = x;
  }
  if (ti == tj) return;
  __syncthreads();
  for (int v = threadIdx.x; v < kDT * kDT; v += 256) {
    const int a = v / kDT, b = v % kDT;  // write row j0 + a, column i0 + b
    const int64_t j = j0 + a, i = i0 + b;
    if (i < n && j < n) sq[j * n + i] = s[b][a];
  }
}

// Order-preserving 32-bit word of a weight.
__device__ __forceinline__ uint32_t wword(uint8_t w) { return w; }
__device__ __forceinline__ uint32_t wword(uint16_t w) { return w; }
__device__ __forceinline__ uint32_t wword(int32_t w) { return (uint32_t)w ^ 0x80000000u; }
__device__ __forceinline__ uint32_t wword(float w) {
  const uint32_t b = __float_as_uint(w);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
// float32 value of a weight word (the reference's float32 Z heights)
template <typename T>
__device__ __forceinline__ float wvalue(uint32_t x);
template <>
__device__ __forceinline__ float wvalue<uint8_t>(uint32_t x) { return (float)x; }
template <>
__device__ __forceinline__ float wvalue<uint16_t>(uint32_t x) { return (float)x; }
template <>
__device__ __forceinline__ float wvalue<int32_t>(uint32_t x) {
  return (float)(int32_t)(x ^ 0x80000000u);
}
template <>
__device__ __forceinline__ float wvalue<float>(uint32_t x) {
  return __uint_as_float((x & 0x80000000u) ? (x & 0x7fffffffu) : ~x);
}

// One Borůvka step's minimum: the lightest key from row u to another
// component, into its component's slot.
template <typename T>
__global__ __launch_bounds__(256) void k_dense_rowmin(const T* __restrict__ sq, int64_t n,
                                                      const int32_t* __restrict__ comp,
                                                      unsigned long long* __restrict__ best) {
  __shared__ unsigned long long s_w[256 / kWave];
  for (int64_t u = blockIdx.x; u < n; u += gridDim.x) {
    const int32_t cu = comp[u];
    const T* row = sq + u * n;
    unsigned long long k = ~0ull;
    for (int64_t v = threadIdx.x; v < n; v += 256) {
      if (comp[v] == cu) continue;
      const uint64_t lo = u < v ? (uint64_t)u : (uint64_t)v, hi = u < v ? (uint64_t)v : (uint64_t)u;
      const unsigned long long key = ((uint64_t)wword(row[v]) << 32) | (lo << 16) | hi;
      k = key < k ? key : k;
    }
    k = wave_min_u64(k);
    if (lane_id() == 0) s_w[threadIdx.x / kWave] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < 256 / kWave; ++w) k = s_w[w] < k ? s_w[w] : k;
      if (k != ~0ull) atomicMin(&best[cu], k);
    }
    __syncthreads();
  }
}

__global__ void k_dense_hook(int64_t n, const int32_t* __restrict__ comp,
                             const unsigned long long* __restrict__ best,
                             int32_t* __restrict__ parent, unsigned long long* __restrict__ tree) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || comp[r] != r) return;
  const unsigned long long b = best[r];
  if (b == ~0ull) return;
  const int32_t a = (int32_t)((b >> 16) & 0xffffu), c = (int32_t)(b & 0xffffu);
  const int32_t other = comp[a] == (int32_t)r ? comp[c] : comp[a];
  if (best[other] == b && r < other) return;  // mutual choice: the smaller root stays
  parent[r] = other;
  tree[r] = b;
}

__global__ void k_dense_flatten(int64_t n, const int32_t* __restrict__ parent,
                                int32_t* __restrict__ comp) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int32_t x = (int32_t)v;
  while (parent[x] != x) x = parent[x];
  comp[v] = x;
}

__global__ void k_iota(int32_t* __restrict__ a, int64_t n) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) a[v] = (int32_t)v;
}

// Lifetime rule of eac.py:366-391 on the float32 heights of the n - 1 sorted
// tree edges (one workgroup): st[0] = clusters, st[1] = merges to apply,
// st[2] = 1 where numpy raises ValueError (fewer than two heights, or no
// lifetime within 1e-8 of zero).
template <typename T>
__global__ __launch_bounds__(1024) void k_dense_lifetime(const unsigned long long* __restrict__ sk,
                                                         int64_t n, int32_t n_clusters,
                                                         int64_t* __restrict__ st) {
  __shared__ unsigned long long s_arg;  // (lifetime word, n - index): max = first max
  __shared__ unsigned int s_min_zero;   // float bits of the smallest close-to-zero lifetime
  __shared__ int s_have_zero;
  __shared__ unsigned long long s_above;
  const int t = threadIdx.x;
  if (n_clusters > 0) {
    if (t == 0) {
      st[0] = n_clusters;
      st[1] = n - n_clusters > 0 ? n - n_clusters : 0;
      st[2] = 0;
    }
    return;
  }
  if (n < 3) {
    if (t == 0) { st[0] = 0; st[1] = 0; st[2] = 1; }
    return;
  }
  if (t == 0) { s_arg = 0; s_min_zero = 0x7f800000u; s_have_zero = 0; s_above = 0; }
  __syncthreads();
  const int64_t nl = n - 2;  // lifetimes
  for (int64_t i = t; i < nl; i += blockDim.x) {
    const float w0 = wvalue<T>((uint32_t)(sk[i] >> 32)), w1 = wvalue<T>((uint32_t)(sk[i + 1] >> 32));
    const float lt = w1 - w0;  // >= 0: heights ascend
    const uint32_t bits = __float_as_uint(lt);  // non-negative floats order as integers
    atomicMax(&s_arg, ((unsigned long long)bits << 32) | (unsigned long long)(nl - i));
    if (fabs((double)lt) <= 1e-8) {
      s_have_zero = 1;
      atomicMin(&s_min_zero, bits);
    }
  }
  __syncthreads();
  const int64_t m = nl - (int64_t)(s_arg & 0xffffffffull);
  const float th = __uint_as_float((uint32_t)(s_arg >> 32));
  const float wm = wvalue<T>((uint32_t)(sk[m] >> 32));
  unsigned long long above = 0;
  for (int64_t i = t; i < n - 1; i += blockDim.x) above += wvalue<T>((uint32_t)(sk[i] >> 32)) > wm;
  atomicAdd(&s_above, above);
  __syncthreads();
  if (t == 0) {
    if (!s_have_zero) {
      st[0] = 0; st[1] = 0; st[2] = 1;
      return;
    }
    int64_t nc = s_above == 0 ? 1 : (int64_t)s_above + 1;
    if (th < 2.0f * __uint_as_float(s_min_zero)) nc = 1;
    st[0] = nc;
    st[1] = n - nc;
    st[2] = 0;
  }
}

__device__ __forceinline__ int32_t uf_find(int32_t* parent, int32_t x) {
  for (;;) {
    const int32_t p = parent[x];
    if (p == x) return x;
    const int32_t g = parent[p];
    if (g != p) parent[x] = g;  // path halving (a benign race: g is an ancestor)
    x = p;
  }
}

// Union of the endpoints of the first st[1] sorted edges (ends decoded from
// the key, or from node pairs in idx when given); the larger root hooks
// under the smaller, so every root is its component's smallest vertex.
__global__ void k_uf_link(const unsigned long long* __restrict__ sk, const int32_t* __restrict__ ends,
                          const int64_t* __restrict__ st, int32_t* __restrict__ parent) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= st[1]) return;
  int32_t a, b;
  if (ends) {
    a = ends[2 * e];
    b = ends[2 * e + 1];
  } else {
    a = (int32_t)((sk[e] >> 16) & 0xffffu);
    b = (int32_t)(sk[e] & 0xffffu);
  }
  for (;;) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    if (a > b) { const int32_t t = a; a = b; b = t; }
    if (atomicCAS(&parent[b], b, a) == b) return;
  }
}

__global__ void k_uf_roots(int64_t n, int32_t* __restrict__ parent, int32_t* __restrict__ flag) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int32_t r = uf_find(parent, (int32_t)v);
  flag[v] = r == (int32_t)v;
}

__global__ void k_uf_labels(int64_t n, int32_t* __restrict__ parent, const int32_t* __restrict__ rank,
                            int32_t* __restrict__ labels) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  labels[v] = rank[uf_find(parent, (int32_t)v)];
}

// Linkage matrix from the sorted tree edges: one thread, union-find with the
// cluster's node id and size at its root (in LDS when 12 n bytes fit).
template <typename T, bool LDS>
__global__ void k_dense_linkage(const unsigned long long* __restrict__ sk, int64_t n,
                                int32_t* __restrict__ g_par, int32_t* __restrict__ g_nid,
                                int32_t* __restrict__ g_size, float* __restrict__ Z) {
  extern __shared__ int32_t s_uf[];
  int32_t* par = LDS ? s_uf : g_par;
  int32_t* nid = LDS ? s_uf + n : g_nid;
  int32_t* size = LDS ? s_uf + 2 * n : g_size;
  for (int64_t v = threadIdx.x; v < n; v += blockDim.x) {
    par[v] = (int32_t)v;
    nid[v] = (int32_t)v;
    size[v] = 1;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int64_t i = 0; i + 1 < n; ++i) {
    const unsigned long long k = sk[i];
    int32_t a = (int32_t)((k >> 16) & 0xffffu), b = (int32_t)(k & 0xffffu);
    while (par[a] != a) { par[a] = par[par[a]]; a = par[a]; }
    while (par[b] != b) { par[b] = par[par[b]]; b = par[b]; }
    const int32_t ia = nid[a], ib = nid[b];
    const int32_t sz = size[a] + size[b];
    Z[4 * i + 0] = (float)(ia < ib ? ia : ib);
    Z[4 * i + 1] = (float)(ia < ib ? ib : ia);
    Z[4 * i + 2] = wvalue<T>((uint32_t)(k >> 32));
    Z[4 * i + 3] = (float)sz;
    if (size[a] < size[b]) { const int32_t t = a; a = b; b = t; }
    par[b] = a;
    size[a] = sz;
    nid[a] = (int32_t)(n + i);
  }
}

// labels_from_Z on a given Z: the leaf reached by following left children
// from every internal node (pointer jumping, one workgroup), then the union
// of the first `merges` rows' two sides.
__global__ __launch_bounds__(1024) void k_z_leftleaf(const float* __restrict__ Z, int64_t n,
                                                     int32_t* __restrict__ rep,
                                                     int32_t* __restrict__ bad) {
  __shared__ int s_changed;
  for (int64_t i = threadIdx.x; i + 1 < n; i += blockDim.x) {
    const float l = Z[4 * i];
    const int64_t li = (int64_t)l;
    if (!(l >= 0.0f) || (float)li != l || li >= n + i) {  // not a node formed before row i
      atomicOr(bad, 1);
      rep[i] = 0;
    } else {
      rep[i] = (int32_t)li;
    }
  }
  __syncthreads();
  for (;;) {
    if (threadIdx.x == 0) s_changed = 0;
    __syncthreads();
    for (int64_t i = threadIdx.x; i + 1 < n; i += blockDim.x) {
      const int32_t r = rep[i];
      if (r >= n) {
        rep[i] = rep[r - n];
        s_changed = 1;
      }
    }
    __syncthreads();
    if (!s_changed) break;
    __syncthreads();
  }
}

__global__ void k_z_ends(const float* __restrict__ Z, int64_t n, const int64_t* __restrict__ st,
                         const int32_t* __restrict__ rep, int32_t* __restrict__ ends,
                         int32_t* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st[1]) return;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const float x = Z[4 * i + s];
    const int64_t xi = (int64_t)x;
    int32_t leaf = 0;
    if (!(x >= 0.0f) || (float)xi != x || xi >= n + i) atomicOr(bad, 1);
    else leaf = xi < n ? (int32_t)xi : rep[xi - n];
    ends[2 * i + s] = leaf;
  }
}

// Non-zero associations per row (full.py:83-95 get_degree with
// full_get_assoc_degree :190-212 / full_condensed_assoc_degree :214-226: a
// condensed row counts only its columns j > i). One workgroup per row.
template <typename T>
__global__ __launch_bounds__(256) void k_dense_degree(const T* __restrict__ mat, int64_t n,
                                                      int condensed, int32_t* __restrict__ degree,
                                                      unsigned long long* __restrict__ nnz) {
  __shared__ unsigned int s_w[256 / kWave];
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const T* row = condensed ? (i + 1 < n ? mat + cidx(n, i, i + 1) : mat) : mat + i * n;
    const int64_t len = condensed ? n - 1 - i : n;
    unsigned int c = 0;
    for (int64_t j = threadIdx.x; j < len; j += 256) c += row[j] != (T)0;
    c = wave_sum(c);
    if (lane_id() == 0) s_w[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < 256 / kWave; ++w) c += s_w[w];
      degree[i] = (int32_t)c;
      if (c) atomicAdd(nnz, (unsigned long long)c);
    }
    __syncthreads();
  }
}

template <typename T>
int launch_rowmin(const void* sq, int64_t n, const int32_t* comp, unsigned long long* best,
                  hipStream_t s) {
  k_dense_rowmin<T><<<grid_for(n, 1, 8192), 256, 0, s>>>(static_cast<const T*>(sq), n, comp,
                                                         best);
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}

int elem_bytes(int32_t dtype) {
  switch (dtype) {
    case EAC_W_U8: return 1;
    case EAC_W_U16: return 2;
    case EAC_W_I32: return 4;
    case EAC_W_F32: return 4;
    default: return 0;
  }
}

struct DenseBufs {
  void* sq;        // square copy (condensed input only)
  int32_t* comp;
  int32_t* parent;
  unsigned long long* best;
  unsigned long long* tree;
  unsigned long long* sorted;
  int64_t* st;     // lifetime state
  int32_t* flag;
  int32_t* rank;
  int32_t* misc;
};

DenseBufs carve_dense(Arena& ar, int64_t n, int eb, bool condensed) {
  DenseBufs b;
  b.sq = condensed ? static_cast<void*>(ar.take<uint8_t>((size_t)n * n * eb)) : nullptr;
  b.comp = ar.take<int32_t>(n);
  b.parent = ar.take<int32_t>(n);
  b.best = ar.take<unsigned long long>(n);
  b.tree = ar.take<unsigned long long>(n);
  b.sorted = ar.take<unsigned long long>(n);
  b.st = ar.take<int64_t>(4);
  b.flag = ar.take<int32_t>(n + 1);
  b.rank = ar.take<int32_t>(n + 1);
  b.misc = ar.take<int32_t>(4);
  return b;
}

int dense_check(const void* d, int32_t dtype, int64_t n) {
  if (!d) return fail(EAC_ERR_ARG, "dense: null matrix");
  if (!elem_bytes(dtype)) return fail(EAC_ERR_ARG, "dense: dtype %d not supported", dtype);
  if (n < 1 || n > kDenseMaxN) return fail(EAC_ERR_UNSUPPORTED, "dense: n = %lld outside [1, %lld]",
                                           (long long)n, (long long)kDenseMaxN);
  return EAC_OK;
}

// Borůvka over the dense matrix, tree edges sorted into B.sorted (n - 1
// keys, then ~0).
int dense_tree(const void* dists, int32_t dtype, int64_t n, int32_t condensed, DenseBufs& B,
               Arena& ar, hipStream_t s) {
  const int eb = elem_bytes(dtype);
  const void* sq = dists;
  if (condensed) {
    const int tiles = (int)((n + kDT - 1) / kDT);
    const dim3 g(tiles, tiles);
    switch (eb) {
      case 1:
        k_square_from_condensed<uint8_t><<<g, 256, 0, s>>>(static_cast<const uint8_t*>(dists), n,
                                                          static_cast<uint8_t*>(B.sq));
        break;
      case 2:
        k_square_from_condensed<uint16_t><<<g, 256, 0, s>>>(static_cast<const uint16_t*>(dists),
                                                           n, static_cast<uint16_t*>(B.sq));
        break;
      default:
        k_square_from_condensed<uint32_t><<<g, 256, 0, s>>>(static_cast<const uint32_t*>(dists),
                                                           n, static_cast<uint32_t*>(B.sq));
        break;
    }
    EAC_LAUNCH_CHECK();
    sq = B.sq;
  }
  const int gv = grid_for(n, 256);
  k_iota<<<gv, 256, 0, s>>>(B.comp, n);
  k_iota<<<gv, 256, 0, s>>>(B.parent, n);
  EAC_LAUNCH_CHECK();
  EAC_TRY(fill<uint64_t>(reinterpret_cast<uint64_t*>(B.tree), n, ~0ull, s));
  const int rounds = bits_for((uint64_t)(n > 1 ? n - 1 : 1)) + 1;
  for (int r = 0; r < rounds && n > 1; ++r) {
    EAC_TRY(fill<uint64_t>(reinterpret_cast<uint64_t*>(B.best), n, ~0ull, s));
    switch (dtype) {
      case EAC_W_U8: EAC_TRY(launch_rowmin<uint8_t>(sq, n, B.comp, B.best, s)); break;
      case EAC_W_U16: EAC_TRY(launch_rowmin<uint16_t>(sq, n, B.comp, B.best, s)); break;
      case EAC_W_I32: EAC_TRY(launch_rowmin<int32_t>(sq, n, B.comp, B.best, s)); break;
      default: EAC_TRY(launch_rowmin<float>(sq, n, B.comp, B.best, s)); break;
    }
    k_dense_hook<<<gv, 256, 0, s>>>(n, B.comp, B.best, B.parent, B.tree);
    k_dense_flatten<<<gv, 256, 0, s>>>(n, B.parent, B.comp);
    EAC_LAUNCH_CHECK();
  }
  const size_t mark = ar.used;
  EAC_TRY((radix_sort<uint64_t, NoValue>(reinterpret_cast<const uint64_t*>(B.tree),
                                          reinterpret_cast<uint64_t*>(B.sorted), nullptr, nullptr,
                                          n, 0, 64, ar, s)));
  ar.used = mark;
  return EAC_OK;
}

template <typename T>
void launch_lifetime(const unsigned long long* sk, int64_t n, int32_t nc, int64_t* st,
                     hipStream_t s) {
  k_dense_lifetime<T><<<1, 1024, 0, s>>>(sk, n, nc, st);
}

template <typename T>
int launch_linkage(const unsigned long long* sk, int64_t n, int32_t* par, int32_t* nid,
                   int32_t* size, float* Z, hipStream_t s) {
  const size_t lds = (size_t)12 * n;
  if (lds <= 150 * 1024) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_dense_linkage<T, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_dense_linkage<T, true><<<1, 256, lds, s>>>(sk, n, par, nid, size, Z);
  } else {
    k_dense_linkage<T, false><<<1, 256, 0, s>>>(sk, n, par, nid, size, Z);
  }
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}

int d2h_small(void* host, const void* dev, size_t bytes, hipStream_t s) {
  EAC_HIP_CHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s));
  EAC_HIP_CHECK(hipStreamSynchronize(s));
  return EAC_OK;
}

// Labels of the first st[1] merges (edges from sk, or node pairs in ends).
int cut_labels(const unsigned long long* sk, const int32_t* ends, int64_t n, DenseBufs& B,
               int32_t* labels, Arena& ar, hipStream_t s) {
  const int gv = grid_for(n, 256);
  k_iota<<<gv, 256, 0, s>>>(B.parent, n);
  k_uf_link<<<gv, 256, 0, s>>>(sk, ends, B.st, B.parent);
  k_uf_roots<<<gv, 256, 0, s>>>(n, B.parent, B.flag);
  EAC_LAUNCH_CHECK();
  const size_t mark = ar.used;
  EAC_TRY((scan_excl<int32_t, int32_t>(B.flag, B.rank, n, 0, nullptr, ar, s)));
  ar.used = mark;
  k_uf_labels<<<gv, 256, 0, s>>>(n, B.parent, B.rank, labels);
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}

size_t dense_ws(int64_t n, int32_t dtype, bool condensed) {
  Arena ar(nullptr, 0, true);
  (void)carve_dense(ar, n, elem_bytes(dtype) ? elem_bytes(dtype) : 4, condensed);
  const size_t a = radix_sort_ws_bytes(n, 8, 0), b = scan_ws_bytes(n + 1, 4);
  return ar.used + (a > b ? a : b) + 3 * sizeof(int32_t) * (size_t)n + 4096;
}

}  // namespace
}  // namespace eac

using namespace eac;

extern "C" int eac_coassoc_dense(const int32_t* labels, int64_t n, int32_t P, int32_t condensed,
                                 int32_t dtype, void* out, int64_t* h_max, void* ws,
                                 size_t ws_bytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!labels || !out) return fail(EAC_ERR_ARG, "coassoc dense: null array");
  if (n < 1 || P < 1) return fail(EAC_ERR_ARG, "coassoc dense: need n, P >= 1");
  if (n > (int64_t)1 << 20) return fail(EAC_ERR_UNSUPPORTED, "coassoc dense: n too large");
  if (!ws || ws_bytes < 256) return fail(EAC_ERR_WORKSPACE, "coassoc dense: workspace < 256 bytes");
  unsigned int* mx = static_cast<unsigned int*>(ws);
  EAC_TRY(fill<int32_t>(reinterpret_cast<int32_t*>(mx), 1, 0, s));
  const int tiles = (int)((n + kDT - 1) / kDT);
  const dim3 g(tiles, tiles);
  switch (dtype) {
    case EAC_W_U8:
      k_dense_counts<uint8_t><<<g, 256, 0, s>>>(labels, n, P, condensed, (uint8_t*)out, mx);
      break;
    case EAC_W_U16:
      k_dense_counts<uint16_t><<<g, 256, 0, s>>>(labels, n, P, condensed, (uint16_t*)out, mx);
      break;
    case EAC_W_I32:
      k_dense_counts<int32_t><<<g, 256, 0, s>>>(labels, n, P, condensed, (int32_t*)out, mx);
      break;
    case EAC_W_F32:
      k_dense_counts<float><<<g, 256, 0, s>>>(labels, n, P, condensed, (float*)out, mx);
      break;
    default:
      return fail(EAC_ERR_ARG, "coassoc dense: dtype %d not supported", dtype);
  }
  EAC_LAUNCH_CHECK();
  if (h_max) {
    unsigned int v = 0;
    EAC_TRY(d2h_small(&v, mx, sizeof(v), s));
    *h_max = v;
  }
  return EAC_OK;
}

extern "C" int eac_dense_diassoc(void* mat, int64_t len, int32_t dtype, double val,
                                 void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (len < 0 || (len > 0 && !mat)) return fail(EAC_ERR_ARG, "dense diassoc: bad array");
  if (len == 0) return EAC_OK;
  const int g = grid_for(len, 256, 8192);
  switch (dtype) {
    case EAC_W_U8: k_dense_diassoc<uint8_t><<<g, 256, 0, s>>>((uint8_t*)mat, len, val); break;
    case EAC_W_U16: k_dense_diassoc<uint16_t><<<g, 256, 0, s>>>((uint16_t*)mat, len, val); break;
    case EAC_W_I32: k_dense_diassoc<int32_t><<<g, 256, 0, s>>>((int32_t*)mat, len, val); break;
    case EAC_W_F32: k_dense_diassoc<float><<<g, 256, 0, s>>>((float*)mat, len, val); break;
    default: return fail(EAC_ERR_ARG, "dense diassoc: dtype %d not supported", dtype);
  }
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}

extern "C" int eac_dense_degree(const void* mat, int32_t dtype, int64_t n, int32_t condensed,
                                int32_t* degree, int64_t* h_nnz, void* ws, size_t ws_bytes,
                                void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!mat || !degree || n < 1) return fail(EAC_ERR_ARG, "dense degree: bad arguments");
  if (!ws || ws_bytes < 256) return fail(EAC_ERR_WORKSPACE, "dense degree: workspace < 256 bytes");
  uint64_t* nnz = static_cast<uint64_t*>(ws);
  EAC_TRY(fill<uint64_t>(nnz, 1, 0ull, s));
  unsigned long long* c = reinterpret_cast<unsigned long long*>(nnz);
  const int g = grid_for(n, 1, 8192);
  switch (dtype) {
    case EAC_W_U8: k_dense_degree<uint8_t><<<g, 256, 0, s>>>((const uint8_t*)mat, n, condensed, degree, c); break;
    case EAC_W_U16: k_dense_degree<uint16_t><<<g, 256, 0, s>>>((const uint16_t*)mat, n, condensed, degree, c); break;
    case EAC_W_I32: k_dense_degree<int32_t><<<g, 256, 0, s>>>((const int32_t*)mat, n, condensed, degree, c); break;
    case EAC_W_F32: k_dense_degree<float><<<g, 256, 0, s>>>((const float*)mat, n, condensed, degree, c); break;
    default: return fail(EAC_ERR_ARG, "dense degree: dtype %d not supported", dtype);
  }
  EAC_LAUNCH_CHECK();
  if (h_nnz) {
    uint64_t v = 0;
    EAC_TRY(d2h_small(&v, nnz, sizeof(v), s));
    *h_nnz = (int64_t)v;
  }
  return EAC_OK;
}

extern "C" size_t eac_dense_sl_workspace_bytes(int64_t n, int32_t dtype, int32_t condensed) {
  if (n < 1) n = 1;
  return dense_ws(n, dtype, condensed != 0);
}

extern "C" int eac_full_sl_lifetime(const void* dists, int32_t dtype, int64_t n,
                                    int32_t condensed, int32_t n_clusters, float* Z,
                                    int32_t* labels, int64_t* h_out, void* ws, size_t ws_bytes,
                                    void* stream) {
  hipStream_t s = (hipStream_t)stream;
  EAC_TRY(dense_check(dists, dtype, n));
  if (!labels) return fail(EAC_ERR_ARG, "full_sl_lifetime: null labels");
  if (n_clusters < 0) return fail(EAC_ERR_ARG, "full_sl_lifetime: n_clusters < 0");
  const size_t need = dense_ws(n, dtype, condensed != 0);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "full_sl_lifetime: workspace %zu < %zu bytes", ws_bytes, need);
  Arena ar(ws, ws_bytes);
  DenseBufs B = carve_dense(ar, n, elem_bytes(dtype), condensed != 0);
  EAC_TRY(dense_tree(dists, dtype, n, condensed, B, ar, s));
  switch (dtype) {
    case EAC_W_U8: launch_lifetime<uint8_t>(B.sorted, n, n_clusters, B.st, s); break;
    case EAC_W_U16: launch_lifetime<uint16_t>(B.sorted, n, n_clusters, B.st, s); break;
    case EAC_W_I32: launch_lifetime<int32_t>(B.sorted, n, n_clusters, B.st, s); break;
    default: launch_lifetime<float>(B.sorted, n, n_clusters, B.st, s); break;
  }
  EAC_LAUNCH_CHECK();
  if (Z && n > 1) {
    int32_t* par = ar.take<int32_t>(n);
    int32_t* nid = ar.take<int32_t>(n);
    int32_t* size = ar.take<int32_t>(n);
    switch (dtype) {
      case EAC_W_U8: EAC_TRY(launch_linkage<uint8_t>(B.sorted, n, par, nid, size, Z, s)); break;
      case EAC_W_U16: EAC_TRY(launch_linkage<uint16_t>(B.sorted, n, par, nid, size, Z, s)); break;
      case EAC_W_I32: EAC_TRY(launch_linkage<int32_t>(B.sorted, n, par, nid, size, Z, s)); break;
      default: EAC_TRY(launch_linkage<float>(B.sorted, n, par, nid, size, Z, s)); break;
    }
  }
  int64_t st[3];
  EAC_TRY(d2h_small(st, B.st, sizeof(st), s));
  if (st[2]) return fail(EAC_ERR_VALUE, "full_sl_lifetime: the lifetime rule raises (no lifetime "
                                       "within 1e-8 of zero, or fewer than two heights)");
  const int64_t nc = st[0];
  if (h_out) {
    h_out[0] = nc;
    h_out[1] = nc > 1 ? 0 : 1;
  }
  if (nc > 1) EAC_TRY(cut_labels(B.sorted, nullptr, n, B, labels, ar, s));
  return EAC_OK;
}

extern "C" int eac_slink(const void* dists, int32_t dtype, int64_t n, int32_t condensed,
                         float* Z, void* ws, size_t ws_bytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  EAC_TRY(dense_check(dists, dtype, n));
  if (!Z) return fail(EAC_ERR_ARG, "slink: null Z");
  const size_t need = dense_ws(n, dtype, condensed != 0);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "slink: workspace %zu < %zu bytes", ws_bytes, need);
  if (n < 2) return EAC_OK;
  Arena ar(ws, ws_bytes);
  DenseBufs B = carve_dense(ar, n, elem_bytes(dtype), condensed != 0);
  EAC_TRY(dense_tree(dists, dtype, n, condensed, B, ar, s));
  int32_t* par = ar.take<int32_t>(n);
  int32_t* nid = ar.take<int32_t>(n);
  int32_t* size = ar.take<int32_t>(n);
  switch (dtype) {
    case EAC_W_U8: EAC_TRY(launch_linkage<uint8_t>(B.sorted, n, par, nid, size, Z, s)); break;
    case EAC_W_U16: EAC_TRY(launch_linkage<uint16_t>(B.sorted, n, par, nid, size, Z, s)); break;
    case EAC_W_I32: EAC_TRY(launch_linkage<int32_t>(B.sorted, n, par, nid, size, Z, s)); break;
    default: EAC_TRY(launch_linkage<float>(B.sorted, n, par, nid, size, Z, s)); break;
  }
  return EAC_OK;
}

extern "C" size_t eac_labels_from_Z_workspace_bytes(int64_t n) {
  if (n < 1) n = 1;
  Arena ar(nullptr, 0, true);
  (void)carve_dense(ar, n, 4, false);
  (void)ar.take<int32_t>(n);      // rep
  (void)ar.take<int32_t>(2 * n);  // ends
  return ar.used + scan_ws_bytes(n + 1, 4) + 4096;
}

extern "C" int eac_labels_from_Z(const float* Z, int64_t n, int32_t n_clusters, int32_t* labels,
                                 void* ws, size_t ws_bytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n < 1 || (n > 1 && !Z) || !labels) return fail(EAC_ERR_ARG, "labels_from_Z: bad arguments");
  if (n_clusters < 1 || n_clusters > n)
    return fail(EAC_ERR_ARG, "labels_from_Z: n_clusters %d outside [1, n]", n_clusters);
  const size_t need = eac_labels_from_Z_workspace_bytes(n);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "labels_from_Z: workspace %zu < %zu bytes", ws_bytes, need);
  Arena ar(ws, ws_bytes);
  DenseBufs B = carve_dense(ar, n, 4, false);
  int32_t* rep = ar.take<int32_t>(n);
  int32_t* ends = ar.take<int32_t>(2 * n);
  EAC_TRY(fill<int32_t>(B.misc, 1, 0, s));
  int64_t st[2] = {n_clusters, n - n_clusters};
  EAC_HIP_CHECK(hipMemcpyAsync(B.st, st, sizeof(st), hipMemcpyHostToDevice, s));
  if (n > 1) {
    k_z_leftleaf<<<1, 1024, 0, s>>>(Z, n, rep, B.misc);
    k_z_ends<<<grid_for(n, 256), 256, 0, s>>>(Z, n, B.st, rep, ends, B.misc);
    EAC_LAUNCH_CHECK();
  }
  EAC_TRY(cut_labels(nullptr, ends, n, B, labels, ar, s));
  int32_t bad = 0;
  EAC_TRY(d2h_small(&bad, B.misc, sizeof(bad), s));
  if (bad) return fail(EAC_ERR_ARG, "labels_from_Z: Z holds a node id that is not a leaf or an "
                                    "earlier row");
  return EAC_OK;
}

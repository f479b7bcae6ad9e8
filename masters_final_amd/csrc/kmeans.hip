// K-Means labeling and centroid update for gfx950.
//
// Reference: MyML/MyML/cluster/KMeans.py:117-189 (K_Means.fit),
// KMeans.py:453-490 (py_label/numba_label: direct squared differences, strict
// < so the lowest k wins exact ties), KMeans.py:586-629 (the CUDA labeling
// kernel) and KMeans.py:709-754 (centroid recomputation: float32 sums in
// point order, division by the count in double, empty clusters reseeded with
// the points of smallest distance).
//
// Labeling modes (all batched over P partitions that share the data, each
// with its own centroid range [k_off[p], k_off[p+1]) of one centroid array):
//  * F32: plain FP32 direct difference, one thread per point, centroid tiles
//    staged in LDS (the "plain per-call FP32" path).
//  * F64: FP64 direct difference in ascending dimension order with strict <,
//    i.e. the oracle's arithmetic; exact by construction.
//  * BF16X3: the distance ||x||^2 - 2 x.c + ||c||^2 with the contraction on
//    MFMA. Each operand is split into two bf16 halves (x = xh + xl + ex,
//    xh = bf16(x), xl = bf16(x - xh)); three 16x16x32 bf16 MFMAs
//    (xh.ch + xh.cl + xl.ch) accumulate in fp32. The epilogue keeps the two
//    smallest distances per point. A rigorous per-point bound E on
//    |approx - FP64 oracle| (split residuals, fp32 accumulation with a
//    truncation-safe 2^-23 unit, epilogue rounding, oracle rounding) certifies
//    the argmin when second - best > 2E; every other point is re-scored in
//    FP64 against all centroids of its partition. Labels therefore equal the
//    FP64 oracle's; distances are then recomputed in FP64 for every point
//    (d flops) so they equal the oracle's too.
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "prims.hpp"
#include "sort.hpp"

namespace eac {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_val(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Partition of a global centroid index (k_off ascending, P small).
__device__ __forceinline__ int part_of(const int64_t* k_off, int P, int64_t c) {
  int lo = 0, hi = P;
  while (lo + 1 < hi) {
    int mid = (lo + hi) >> 1;
    if (k_off[mid] <= c) lo = mid; else hi = mid;
  }
  return lo;
}

// Per-partition maxima of centroid norms, as float bits (non-negative floats
// order like their bit patterns): [0] ||c||, [1] ||ch||, [2] ||cl||, [3] ||ec||.
constexpr int kStat = 4;

// Split every centroid into bf16 halves (row length KD, zero padded) and
// record norms. The per-partition maxima are reduced over the wavefront
// first when its lanes share one partition (one atomic per statistic instead
// of one per centroid: same-address atomics serialise).
__global__ void k_cent_prep(const float* __restrict__ C, int64_t ktot, int d, int KD,
                            const int64_t* __restrict__ k_off, int P,
                            uint16_t* __restrict__ Ch, uint16_t* __restrict__ Cl,
                            float* __restrict__ Ct, float* __restrict__ ncf,
                            unsigned int* __restrict__ pstat) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < ktot; base += stride) {
    const int64_t c = base + threadIdx.x;
    const bool valid = c < ktot;
    double n2 = 0, h2 = 0, l2 = 0, e2 = 0;
    int p = -1;
    if (valid) {
      for (int j = 0; j < KD; ++j) {
        float x = j < d ? C[c * d + j] : 0.f;
        uint16_t h = bf16_rne(x);
        float r = x - bf16_val(h);  // exact
        uint16_t l = bf16_rne(r);
        float e = r - bf16_val(l);  // exact
        Ch[c * KD + j] = h;
        Cl[c * KD + j] = l;
        if (j < d) Ct[(int64_t)j * ktot + c] = x;  // transposed copy for the re-scoring
        n2 += (double)x * x;
        h2 += (double)bf16_val(h) * bf16_val(h);
        l2 += (double)bf16_val(l) * bf16_val(l);
        e2 += (double)e * e;
      }
      ncf[c] = (float)n2;
      p = part_of(k_off, P, c);
    }
    // norms rounded up so the bound stays rigorous; non-negative floats order
    // like their bit patterns, and 0 is the identity of the maximum
    unsigned int v[kStat] = {
        valid ? __float_as_uint((float)sqrt(n2) * (1.f + 1e-6f)) : 0u,
        valid ? __float_as_uint((float)sqrt(h2) * (1.f + 1e-6f)) : 0u,
        valid ? __float_as_uint((float)sqrt(l2) * (1.f + 1e-6f)) : 0u,
        valid ? __float_as_uint((float)sqrt(e2) * (1.f + 1e-6f)) : 0u};
    int pmin = valid ? p : INT_MAX, pmax = p;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      pmin = min(pmin, __shfl_xor(pmin, off, kWave));
      pmax = max(pmax, __shfl_xor(pmax, off, kWave));
    }
    if (pmin == pmax) {  // every valid lane in one partition (uniform)
#pragma unroll
      for (int q = 0; q < kStat; ++q)
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) v[q] = max(v[q], __shfl_xor(v[q], off, kWave));
      if (lane_id() == 0)
#pragma unroll
        for (int q = 0; q < kStat; ++q) atomicMax(&pstat[pmax * kStat + q], v[q]);
    } else if (valid) {
#pragma unroll
      for (int q = 0; q < kStat; ++q) atomicMax(&pstat[p * kStat + q], v[q]);
    }
  }
}

constexpr int kLB = 256;       // threads per labeling block (4 waves)
constexpr int kRowsW = 32;     // points per wave
constexpr int kBM = 128;       // points per block
constexpr int kBN = 64;        // centroids per LDS tile
constexpr int kFlagSlots = 64; // flag list regions (one counter each)

// Rigorous bound on |approx - oracle| for one point and one partition.
__device__ __forceinline__ float point_bound(float nx, float xn, float xhn, float xln, float exn,
                                             const float* st, int KD, int d) {
  const float u = 5.9604645e-08f;            // 2^-24
  const float ut = 1.1920929e-07f;           // 2^-23: fp32 accumulation, any rounding
  const float Cn = st[0], Chn = st[1], Cln = st[2], Cen = st[3];
  // nonzero products per dot: 3 d (columns d..KD-1 are zero in both
  // operands; adding an exact zero neither rounds nor truncates)
  const float nterms = 3.f * (d < KD ? d : KD);
  const float gam = nterms * ut / (1.f - nterms * ut);
  float cross = gam * (xhn * (Chn + Cln) + xln * Chn) + xhn * Cen + xln * (Cln + Cen) + exn * Cn;
  float epi = 6.f * u * (nx + Cn * Cn);
  float orc = 4.f * d * 1.1102230e-16f * (xn + Cn) * (xn + Cn);
  return (2.f * cross + epi + orc) * 1.0001f + 1e-30f;
}

// Waves per SIMD the compiler must fit (WPE): by default 3 for KD <= 64 (a
// few values spill; measured faster than 2 waves without spills),
// otherwise no constraint (1); the EAC_TUNE_KM_WAVES knob picks 2, 3 or 4.
constexpr int label_waves(int KD) { return KD <= 64 ? 3 : 1; }
template <int KD, int WPE = label_waves(KD)>
__global__ __launch_bounds__(kLB) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_label_bf16x3(
    const float* __restrict__ X, int64_t n, int d, const uint16_t* __restrict__ Ch,
    const uint16_t* __restrict__ Cl, const float* __restrict__ ncf,
    const int64_t* __restrict__ k_off, int P, const unsigned int* __restrict__ pstat,
    int32_t* __restrict__ labels, uint64_t* __restrict__ flags,
    unsigned long long* __restrict__ nflag, int64_t region) {
  constexpr int KB = KD / 32;
  constexpr int LROW = KD + 8;  // padded LDS row (bf16 elements)
  __shared__ uint16_t s_h[2][kBN * LROW];
  __shared__ uint16_t s_l[2][kBN * LROW];
  __shared__ float s_nc[2][kBN];
  const int t = threadIdx.x;
  const int w = t / kWave;
  const int l = lane_id();
  const int lr = l & 15;  // row (A layout) / column (output layout) inside a 16-block
  const int lq = l >> 4;  // k-group (A layout) / row quad (output layout)
  const int64_t p0 = (int64_t)blockIdx.x * kBM + w * kRowsW;

  // A fragments: rows p0 + rb*16 + lr, k = kb*32 + lq*8 .. +8 (hi and lo)
  bf16x8 Ah[2][KB], Al[2][KB];
  float nx_a[2], xn_a[2], xhn_a[2], xln_a[2], exn_a[2];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    int64_t row = p0 + rb * 16 + lr;
    float s2 = 0, sh = 0, sl = 0, se = 0;
    double s2d = 0;
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        int col = kb * 32 + lq * 8 + j;
        float x = (row < n && col < d) ? X[row * d + col] : 0.f;
        uint16_t h = bf16_rne(x);
        float r = x - bf16_val(h);
        uint16_t lo = bf16_rne(r);
        float e = r - bf16_val(lo);
        Ah[rb][kb][j] = __builtin_bit_cast(__bf16, h);
        Al[rb][kb][j] = __builtin_bit_cast(__bf16, lo);
        s2d += (double)x * x;
        sh += bf16_val(h) * bf16_val(h);
        sl += bf16_val(lo) * bf16_val(lo);
        se += e * e;
      }
    }
    // the 4 lanes lr, lr+16, lr+32, lr+48 hold one row
#pragma unroll
    for (int off = 16; off < 64; off <<= 1) {
      s2d += __shfl_xor(s2d, off, kWave);
      sh += __shfl_xor(sh, off, kWave);
      sl += __shfl_xor(sl, off, kWave);
      se += __shfl_xor(se, off, kWave);
    }
    nx_a[rb] = (float)s2d;
    xn_a[rb] = (float)sqrt(s2d) * (1.f + 1e-6f);
    xhn_a[rb] = sqrtf(sh) * (1.f + 1e-5f);
    xln_a[rb] = sqrtf(sl) * (1.f + 1e-5f);
    exn_a[rb] = sqrtf(se) * (1.f + 1e-5f);
  }

  // Centroid tiles of every partition in sequence, double buffered in LDS:
  // the next tile is fetched into registers before the current one is
  // multiplied and stored to the other buffer after it, so one barrier per
  // tile and the L2 latency overlaps the MFMAs.
  constexpr int VPR = KD / 8;  // uint4 per row
  static_assert((kBN * VPR) % kLB == 0 && kBN <= kLB, "a tile splits evenly over the block");
  constexpr int NV = kBN * VPR / kLB;
  uint4 ph[NV], pl[NV];
  float pnc;
  // Rows past the partition repeat its last centroid (every partition has
  // one; the host checks k >= 1) and get ||c||^2 = +inf, so their D is +inf
  // and the epilogue never takes them. Fetch and store are branch-free: the
  // loads are unconditional, every thread loads a norm (kLB / kBN threads
  // store the same value to each entry) and the +inf default is chosen when
  // the tile is stored. A register write to a load's destination next to
  // the load, or a load skipped on some lanes, would make the wave wait for
  // the new tile there, before the MFMAs, and undo the double buffering.
  // Row indices are 32-bit offsets from the tile's first centroid.
  int32_t f_rem = 0;
  auto fetch = [&](int64_t c0, int64_t kend) {
    const int32_t rem = (int32_t)(kend - c0);  // >= 1 rows left in the range
    const uint4* bh = reinterpret_cast<const uint4*>(Ch + c0 * KD);
    const uint4* bl = reinterpret_cast<const uint4*>(Cl + c0 * KD);
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int v = t + i * kLB;
      const int32_t r = min(v / VPR, rem - 1);
      ph[i] = bh[r * VPR + v % VPR];
      pl[i] = bl[r * VPR + v % VPR];
    }
    pnc = ncf[c0 + min(t % kBN, rem - 1)];
    f_rem = rem;
  };
  // (the stored values pass through a mask: stored straight from the load
  // registers, the compiler keeps the prefetched tile in scratch memory)
  auto masked = [](uint4 a, uint32_t m) { return make_uint4(a.x & m, a.y & m, a.z & m, a.w & m); };
  auto stash = [&](int b) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int v = t + i * kLB;
      const uint32_t m = v / VPR < f_rem ? ~0u : 0u;
      *reinterpret_cast<uint4*>(&s_h[b][(v / VPR) * LROW + (v % VPR) * 8]) = masked(ph[i], m);
      *reinterpret_cast<uint4*>(&s_l[b][(v / VPR) * LROW + (v % VPR) * 8]) = masked(pl[i], m);
    }
    s_nc[b][t % kBN] = (t % kBN) < f_rem ? pnc : INFINITY;
  };
  int p = 0, buf = 0;
  int64_t c0 = k_off[0];
  fetch(c0, k_off[1]);
  stash(0);
  __syncthreads();
  for (; p < P; ++p) {
    const int64_t kbeg = k_off[p], kend = k_off[p + 1];
    float d1[2][4], d2[2][4];
    int32_t k1[2][4];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int i = 0; i < 4; ++i) { d1[rb][i] = FLT_MAX; d2[rb][i] = FLT_MAX; k1[rb][i] = -1; }
    for (c0 = kbeg; c0 < kend; c0 += kBN) {
      // the next tile: this partition's, else the next partition's first
      int64_t nc0 = c0 + kBN, nend = kend;
      if (nc0 >= kend && p + 1 < P) {
        nc0 = kend;
        nend = k_off[p + 2];
      }
      const bool more = nc0 < nend;
      if (more) fetch(nc0, nend);
      f32x4 acc[2][4];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
      const uint16_t* sh = s_h[buf];
      const uint16_t* sl = s_l[buf];
      // per 32-wide K block: its B fragments (8 LDS reads), then the MFMAs
      // in rounds of 8 independent accumulators, so no MFMA waits on the one
      // before it (fragments of one K block at a time keep the register count
      // within 3 waves per SIMD)
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        bf16x8 Bh[4], Bl[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
          const int r = cb * 16 + lr;
          Bh[cb] = *reinterpret_cast<const bf16x8*>(&sh[r * LROW + kb * 32 + lq * 8]);
          Bl[cb] = *reinterpret_cast<const bf16x8*>(&sl[r * LROW + kb * 32 + lq * 8]);
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
          for (int rb = 0; rb < 2; ++rb)
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[rb][kb], Bh[cb],
                                                                  acc[rb][cb], 0, 0, 0);
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
          for (int rb = 0; rb < 2; ++rb)
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[rb][kb], Bl[cb],
                                                                  acc[rb][cb], 0, 0, 0);
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
          for (int rb = 0; rb < 2; ++rb)
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al[rb][kb], Bh[cb],
                                                                  acc[rb][cb], 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
      // epilogue: lane holds column cb*16 + lr, rows 4*lq + i. The running
      // top two per row compare D' = ||c||^2 - 2 x.c (the row's ||x||^2 is
      // common to all columns, so it changes neither the argmin nor the
      // margin); columns past the partition carry ||c||^2 = +inf. Five VALU
      // operations per element: fma, compare, med3 (new second) and two
      // selects (best value, best column).
      const int32_t colb = (int32_t)(c0 - kbeg) + lr;
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        const float nc = s_nc[buf][cb * 16 + lr];
        const int32_t col = colb + cb * 16;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float D = fmaf(-2.f, acc[rb][cb][i], nc);
            const bool take = D < d1[rb][i];
            d2[rb][i] = __builtin_amdgcn_fmed3f(d1[rb][i], d2[rb][i], D);
            d1[rb][i] = take ? D : d1[rb][i];
            k1[rb][i] = take ? col : k1[rb][i];
          }
      }
      // the epilogue stays ahead of the tile store: its VALU work covers the
      // latency of the next tile's loads (left alone, the compiler sinks it
      // below the store and the barrier, which then wait right after the MFMAs)
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          asm volatile("" : "+v"(d1[rb][i]), "+v"(d2[rb][i]), "+v"(k1[rb][i])::"memory");
      if (more) {
        stash(buf ^ 1);  // last read in the previous tile, before its barrier
        __syncthreads();
        buf ^= 1;
      }
    }
    // merge the 16 lanes that share rows (same lq)
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
          float od1 = __shfl_xor(d1[rb][i], off, kWave);
          float od2 = __shfl_xor(d2[rb][i], off, kWave);
          int32_t ok1 = __shfl_xor(k1[rb][i], off, kWave);
          float hi1 = fmaxf(d1[rb][i], od1);
          float nd2 = fminf(fminf(d2[rb][i], od2), hi1);
          bool take = od1 < d1[rb][i] || (od1 == d1[rb][i] && ok1 >= 0 &&
                                          (k1[rb][i] < 0 || ok1 < k1[rb][i]));
          if (take) { d1[rb][i] = od1; k1[rb][i] = ok1; }
          d2[rb][i] = nd2;
        }
      }
    // Bounds in output layout, then write. After the merge the 16 lanes of a
    // row group hold the same top two for each of the group's 8 rows; lane
    // lr < 8 finishes row slot (rb, i) = (lr / 4, lr % 4), so each lane
    // computes one bound instead of the group's first lane computing eight.
    float st[kStat];
#pragma unroll
    for (int q = 0; q < kStat; ++q) st[q] = __uint_as_float(pstat[p * kStat + q]);
    const int rs = lr & 7, rbs = rs >> 2, is = rs & 3;
    float D1 = d1[0][0], D2 = d2[0][0];
    int32_t K1 = k1[0][0];
#pragma unroll
    for (int q = 1; q < 8; ++q)
      if (rs == q) {
        D1 = d1[q >> 2][q & 3];
        D2 = d2[q >> 2][q & 3];
        K1 = k1[q >> 2][q & 3];
      }
    // row statistics: lanes lr' = lq * 4 + i hold row lq * 4 + i of each block
    const int src = lq * 4 + is;
    float xs[2][5];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      xs[rb][0] = __shfl(nx_a[rb], src, kWave);
      xs[rb][1] = __shfl(xn_a[rb], src, kWave);
      xs[rb][2] = __shfl(xhn_a[rb], src, kWave);
      xs[rb][3] = __shfl(xln_a[rb], src, kWave);
      xs[rb][4] = __shfl(exn_a[rb], src, kWave);
    }
    const int64_t row = p0 + rbs * 16 + lq * 4 + is;
    bool flagged = false;
    if (lr < 8 && row < n && K1 >= 0) {
      const float E = rbs ? point_bound(xs[1][0], xs[1][1], xs[1][2], xs[1][3], xs[1][4], st, KD, d)
                          : point_bound(xs[0][0], xs[0][1], xs[0][2], xs[0][3], xs[0][4], st, KD, d);
      labels[(int64_t)p * n + row] = K1;
      flagged = !(D2 - D1 > 2.f * E);
    }
    // append flagged (p, row) pairs: one atomic per wave
    const uint32_t cnt = flagged ? 1u : 0u;
    const uint32_t incl = wave_incl_scan(cnt);
    const uint32_t tot = __shfl(incl, kWave - 1, kWave);
    unsigned long long base = 0;
    // flags go to one of kFlagSlots regions (by block): counters on one
    // address would serialise one atomic per wave and partition
    const int slot = blockIdx.x % kFlagSlots;
    if (l == kWave - 1 && tot) base = atomicAdd(&nflag[slot], (unsigned long long)tot);
    base = __shfl(base, kWave - 1, kWave);
    if (flagged) {
      const unsigned long long o = base + incl - 1;
      if (o < (unsigned long long)region)
        flags[(int64_t)slot * region + (int64_t)o] = ((uint64_t)p << 40) | (uint64_t)row;
    }
  }
}

// FP64 re-scoring of flagged points: one wave per (p, row), lanes stride the
// partition's centroids; strict < in ascending k, lowest k on ties.
// acc + diff^2 with separately rounded multiply and add in FP64 (the
// oracle's arithmetic: no fused multiply-add), plain expression otherwise.
// HIP's __dmul_rn / __dadd_rn are plain operators that the compiler may
// contract into one fma; the pragma keeps the two roundings.
__device__ __forceinline__ double sq_acc(double acc, double diff) {
#pragma clang fp contract(off)
  return acc + diff * diff;
}
__device__ __forceinline__ float sq_acc(float acc, float diff) { return acc + diff * diff; }

// Squared distance of two float rows in FP64, accumulated in ascending
// dimension order (the oracle's arithmetic); 16-byte loads when both rows
// are 16-byte aligned and d is a multiple of 4.
__device__ __forceinline__ double dist2_f64(const float* __restrict__ x,
                                           const float* __restrict__ c, int d, bool vec4) {
  double acc = 0.0;
  if (vec4) {
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const float4* c4 = reinterpret_cast<const float4*>(c);
    for (int j = 0; j < d / 4; ++j) {
      const float4 xv = x4[j], cv = c4[j];
      acc = sq_acc(acc, (double)xv.x - (double)cv.x);
      acc = sq_acc(acc, (double)xv.y - (double)cv.y);
      acc = sq_acc(acc, (double)xv.z - (double)cv.z);
      acc = sq_acc(acc, (double)xv.w - (double)cv.w);
    }
  } else {
    for (int j = 0; j < d; ++j) {
      acc = sq_acc(acc, (double)x[j] - (double)c[j]);
    }
  }
  return acc;
}

// The lanes stride the partition's centroids and read them from the
// transposed copy Ct [d][ktot] (64 consecutive floats per load, the point's
// coordinates are wave-uniform). A wave takes F flagged points of one
// partition at a time (F > 1 only for single-partition calls, where every
// flag shares the centroids): each centroid value loaded serves F
// independent FP64 chains (on 4M x 64 points, k = 2048, 6263 flags: 225 us
// with 2 per wave against 254 with 1).
template <int F>
__global__ __launch_bounds__(256) void k_rescore_f64(const float* __restrict__ X, int64_t n,
                                                     int d, const float* __restrict__ Ct,
                                                     int64_t ktot,
                                                     const int64_t* __restrict__ k_off,
                                                     const uint64_t* __restrict__ flags,
                                                     int64_t nflag, int32_t* __restrict__ labels,
                                                     const int64_t* __restrict__ d_nflag = nullptr) {
  const int l = lane_id();
  if (d_nflag) nflag = *d_nflag;  // count left on the device (capturable calls)
  const int64_t ngroups = (nflag + F - 1) / F;
  for (int64_t gi = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; gi < ngroups;
       gi += ((int64_t)gridDim.x * blockDim.x) / kWave) {
    const float* x[F];
    int64_t row[F];
    bool has[F];
    int p = 0;
#pragma unroll
    for (int q = 0; q < F; ++q) {
      const int64_t f = gi * F + q;
      has[q] = f < nflag;
      const uint64_t fl = flags[has[q] ? f : gi * F];  // a missing flag repeats the first
      if (q == 0) p = (int)(fl >> 40);
      row[q] = (int64_t)(fl & ((1ull << 40) - 1));
      x[q] = X + row[q] * d;
    }
    const int64_t kbeg = k_off[p], kend = k_off[p + 1];
    double best[F];
    int64_t bk[F];
#pragma unroll
    for (int q = 0; q < F; ++q) { best[q] = DBL_MAX; bk[q] = -1; }
    for (int64_t c = kbeg + l; c < kend; c += kWave) {
      const float* ct = Ct + c;
      double acc[F];
#pragma unroll
      for (int q = 0; q < F; ++q) acc[q] = 0.0;
      int j = 0;
      for (; j + 8 <= d; j += 8) {  // eight loads in flight, then the chains
        float cv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) cv[u] = ct[(int64_t)(j + u) * ktot];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int q = 0; q < F; ++q) acc[q] = sq_acc(acc[q], (double)x[q][j + u] - (double)cv[u]);
      }
      for (; j < d; ++j) {
        const float cj = ct[(int64_t)j * ktot];
#pragma unroll
        for (int q = 0; q < F; ++q) acc[q] = sq_acc(acc[q], (double)x[q][j] - (double)cj);
      }
#pragma unroll
      for (int q = 0; q < F; ++q)
        if (bk[q] < 0 || acc[q] < best[q]) { best[q] = acc[q]; bk[q] = c; }
    }
#pragma unroll
    for (int q = 0; q < F; ++q) {
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) {
        double ob = __shfl_xor(best[q], off, kWave);
        int64_t ok = __shfl_xor(bk[q], off, kWave);
        if (ok >= 0 && (bk[q] < 0 || ob < best[q] || (ob == best[q] && ok < bk[q]))) {
          best[q] = ob;
          bk[q] = ok;
        }
      }
      if (l == 0 && has[q]) labels[(int64_t)p * n + row[q]] = (int32_t)(bk[q] - kbeg);
    }
  }
}

// FP64 re-scoring with the partition's centroids staged in LDS, for flags
// sorted by partition: a block takes 64 flags; four threads per flag split
// the centroids (c = s, s+4, ... ascending, two chains at a time) and read
// them from LDS tiles shared by all 64 flags, so each centroid is read from
// memory once per 64 flagged points instead of once per point.
//
// Screen first: the FP32 direct difference S = sum fl(x_j - c_j)^2 (fused
// multiply-add) is within a factor (1 +- g) of the exact distance D, g =
// gamma_{d+2} (d + 2 roundings, unit 2^-23 for safety), plus an absolute
// underflow slack A; the oracle's FP64 sum is within (1 +- g64) of D. With m
// the smallest screened value, a centroid whose screened value exceeds
// T = (m + A) R + A, R = (1+g)(1+g64) / ((1-g)(1-g64)), has an oracle distance
// strictly above the oracle distance of the true nearest centroid, so it can
// be neither the oracle's minimum nor tied with it. Each thread keeps its two
// smallest screened values; the candidates (screened value <= T, usually one
// per flag) are re-scored with the oracle's FP64 arithmetic, lowest index on
// equal values (the oracle's first minimum). A flag where some thread's
// second-smallest value is also <= T may have more candidates than were kept:
// the block then re-scores its run in FP64 against every centroid, strict < in
// ascending order inside a subset, lowest index among equal minima across the
// four subsets.
constexpr int kRtF = 64;  // flags per block
constexpr int kRtS = 4;   // threads per flag
constexpr int kRtT = 64;  // centroids per LDS tile
// point rows: stride d + 1 (lanes read different rows: conflict-free);
// centroid rows: stride rounded up to 4 floats, zero padded (every lane of a
// wave reads the same row: broadcast 16-byte reads)
inline int rescore_cstride(int d) { return (d + 3) & ~3; }
inline size_t rescore_tiled_lds(int d) {
  return ((size_t)kRtF * (d + 1) + (size_t)kRtT * rescore_cstride(d)) * sizeof(float);
}
// DX > 0: the screen keeps the point's d <= DX coordinates in registers.
template <int DX>
__global__ __launch_bounds__(kRtF * kRtS) void k_rescore_tiled(
    const float* __restrict__ X, int64_t n, int d, const float* __restrict__ C,
    const int64_t* __restrict__ k_off, const uint64_t* __restrict__ flags, int64_t nflag,
    int32_t* __restrict__ labels, const int64_t* __restrict__ d_nflag = nullptr) {
  extern __shared__ __align__(16) float s_dyn[];
  __shared__ double s_best[kRtS][kRtF];
  __shared__ int32_t s_bk[kRtS][kRtF];
  __shared__ float s_v1[kRtS][kRtF];
  constexpr uint64_t kRowMask = (1ull << 40) - 1;
  const int RS = d + 1;
  const int RC = (d + 3) & ~3;
  float* s_x = s_dyn;
  float* s_c = s_dyn + kRtF * RS;  // 16-byte aligned: kRtF * RS * 4 = 256 (d + 1) bytes
  const int t = threadIdx.x, g = t & (kRtF - 1), sub = t / kRtF;
  if (d_nflag) nflag = *d_nflag;  // count left on the device: tiles in a loop
  for (int64_t f0 = (int64_t)blockIdx.x * kRtF; f0 < nflag; f0 += (int64_t)gridDim.x * kRtF) {
  const int cnt = nflag - f0 < kRtF ? (int)(nflag - f0) : kRtF;
  // screen bounds (double, then rounded up into float)
  const double ug = 1.1920928955078125e-07, u64 = 1.1102230246251565e-16;
  const double gs = (d + 2) * ug / (1.0 - (d + 2) * ug);
  const double g64 = (d + 2) * 2 * u64 / (1.0 - (d + 2) * 2 * u64);
  const float R = (float)((1.0 + gs) * (1.0 + g64) / ((1.0 - gs) * (1.0 - g64)) * (1.0 + 1e-6));
  const float A = (float)d * 4.7e-38f;  // >= d * 2^-124
  for (int done = 0; done < cnt;) {  // runs of one partition (block-uniform)
    const int p = (int)(flags[f0 + done] >> 40);
    int end = done + 1;
    while (end < cnt && (int)(flags[f0 + end] >> 40) == p) ++end;
    for (int v = t; v < (end - done) * d; v += kRtF * kRtS) {
      const int r = v / d, j = v - r * d;
      s_x[r * RS + j] = X[(int64_t)(flags[f0 + done + r] & kRowMask) * d + j];
    }
    const bool active = g >= done && g < end;
    const float* xa = s_x + (active ? g - done : 0) * RS;
    const int64_t kbeg = k_off[p], kend = k_off[p + 1];
    // pass 1: FP32 screen, two smallest values per thread
    float v1 = INFINITY, v2 = INFINITY;
    int64_t k1 = -1;
    auto keep = [&](float v, int64_t k) {
      const bool lt1 = v < v1;
      v2 = __builtin_amdgcn_fmed3f(v1, v2, v);
      k1 = lt1 ? k : k1;
      v1 = lt1 ? v : v1;
    };
    // staged centroid rows, zero padded to RC columns
    auto stage = [&](int64_t c0, int rows) {
      for (int v = t; v < rows * RC; v += kRtF * kRtS) {
        const int r = v / RC, j = v - r * RC;
        s_c[v] = j < d ? C[(c0 + r) * d + j] : 0.f;
      }
    };
    constexpr int NX = DX > 0 ? DX : 1;
    float xr[NX];
    __syncthreads();  // the point rows are staged
    if constexpr (DX > 0) {
#pragma unroll
      for (int j = 0; j < NX; ++j) xr[j] = j < d ? xa[j] : 0.f;
    }
    const int nq = RC / 4;
    // screened value of one staged row (padding adds exact zeros)
    auto screen = [&](const float* cr) {
      float a = 0.f;
      if constexpr (DX > 0) {
        const float4* c4 = reinterpret_cast<const float4*>(cr);
#pragma unroll
        for (int q = 0; q < NX / 4; ++q) {
          if (q < nq) {
            const float4 u = c4[q];
            float e = xr[4 * q] - u.x;
            a = fmaf(e, e, a);
            e = xr[4 * q + 1] - u.y;
            a = fmaf(e, e, a);
            e = xr[4 * q + 2] - u.z;
            a = fmaf(e, e, a);
            e = xr[4 * q + 3] - u.w;
            a = fmaf(e, e, a);
          }
        }
      } else {
        for (int j = 0; j < d; ++j) {
          const float e = xa[j] - cr[j];
          a = fmaf(e, e, a);
        }
      }
      return a;
    };
    for (int64_t c0 = kbeg; c0 < kend; c0 += kRtT) {
      const int rows = kend - c0 < kRtT ? (int)(kend - c0) : kRtT;
      __syncthreads();  // the rows are staged, the previous tile is consumed
      stage(c0, rows);
      __syncthreads();
      if (!active) continue;
      int c = sub;
      for (; c + kRtS < rows; c += 2 * kRtS) {
        const float a0 = screen(s_c + c * RC);
        const float a1 = screen(s_c + (c + kRtS) * RC);
        keep(a0, c0 + c);
        keep(a1, c0 + c + kRtS);
      }
      for (; c < rows; c += kRtS) keep(screen(s_c + c * RC), c0 + c);
    }
    s_v1[sub][g] = v1;
    __syncthreads();
    bool full = false;
    {
      float m = s_v1[0][g];
      for (int q = 1; q < kRtS; ++q) m = fminf(m, s_v1[q][g]);
      const float T = (m + A) * R + A;
      double best = DBL_MAX;
      int32_t bk = -1;
      if (active && k1 >= 0 && v1 <= T) {
        best = dist2_f64(xa, C + k1 * d, d, false);
        bk = (int32_t)(k1 - kbeg);
      }
      full = active && (k1 < 0 || !(v2 > T));  // (NaN screens take the full pass)
      s_best[sub][g] = best;
      s_bk[sub][g] = bk;
    }
    // pass 2 (rare): every centroid in FP64 for the whole run
    if (__syncthreads_or(full)) {
      double best = DBL_MAX;
      int64_t bk = -1;
      for (int64_t c0 = kbeg; c0 < kend; c0 += kRtT) {
        const int rows = kend - c0 < kRtT ? (int)(kend - c0) : kRtT;
        __syncthreads();
        stage(c0, rows);
        __syncthreads();
        if (!active) continue;
        for (int c = sub; c < rows; c += kRtS) {
          const float* ca = s_c + c * RC;
          double a0 = 0.0;
          for (int j = 0; j < d; ++j) a0 = sq_acc(a0, (double)xa[j] - (double)ca[j]);
          if (bk < 0 || a0 < best) { best = a0; bk = c0 + c; }
        }
      }
      s_best[sub][g] = best;
      s_bk[sub][g] = bk < 0 ? -1 : (int32_t)(bk - kbeg);
      __syncthreads();
    }
    if (active && sub == 0) {
      double b = s_best[0][g];
      int32_t k = s_bk[0][g];
      for (int q = 1; q < kRtS; ++q) {
        const double ob = s_best[q][g];
        const int32_t ok = s_bk[q][g];
        if (ok >= 0 && (k < 0 || ob < b || (ob == b && ok < k))) { b = ob; k = ok; }
      }
      labels[(int64_t)p * n + (int64_t)(flags[f0 + g] & kRowMask)] = k;
    }
    __syncthreads();
    done = end;
  }
  }  // tiles
}

// Concatenates the kFlagSlots flag regions (counts cnt[slot]) into one list.
__global__ void k_gather_flags(const uint64_t* __restrict__ regions, int64_t region,
                               const unsigned long long* __restrict__ cnt,
                               uint64_t* __restrict__ out, int64_t total,
                               const int64_t* __restrict__ d_total = nullptr) {
  __shared__ unsigned long long s_pre[kFlagSlots + 1];
  if (d_total) total = *d_total;
  if (threadIdx.x == 0) {
    unsigned long long a = 0;
    for (int q = 0; q < kFlagSlots; ++q) {
      s_pre[q] = a;
      a += cnt[q];
    }
    s_pre[kFlagSlots] = a;
  }
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = kFlagSlots;  // slot q with s_pre[q] <= i < s_pre[q+1]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (s_pre[mid] <= (unsigned long long)i) lo = mid; else hi = mid;
    }
    out[i] = regions[(int64_t)lo * region + (i - (int64_t)s_pre[lo])];
  }
}

// FP64 distance of every point to its labelled centroid (oracle arithmetic),
// stored as float32 like the reference's dists array. LPP = 4 lanes share a
// point (the addition chain has d/4 steps whatever LPP is, and fewer lanes
// per point run more points per wave: on 4M x 64 points, 303 us against 342
// with 8 lanes, 500 with 16, 416 with 2 and 785 with 1):
// lane s loads dimensions 4s..4s+3 of each (4 LPP)-dimension chunk of the
// point's row and its centroid's row (coalesced) and squares the
// differences in parallel; the running sum then visits the lanes in order
// (4 adds in lane s, then a shuffle to lane s+1), so the additions happen in
// ascending dimension order exactly as in the oracle. Padding dimensions add
// exact zeros. A block takes kPG partitions (grid.y over partition groups)
// with the partitions innermost, so a point's row is read from memory once
// per group and from cache for the other partitions.
constexpr int kPG = 8;
template <int LPP>
__global__ __launch_bounds__(256) void k_exact_dists(const float* __restrict__ X, int64_t n,
                                                     int d, const float* __restrict__ C,
                                                     const int64_t* __restrict__ k_off, int P,
                                                     const int32_t* __restrict__ labels,
                                                     float* __restrict__ dists) {
#pragma clang fp contract(off)  // the oracle's separately rounded products and sums
  constexpr int PPB = 256 / LPP;  // points per block and step
  const int lane = lane_id(), sub = lane & (LPP - 1), base = lane & ~(LPP - 1);
  const int64_t groups = (int64_t)gridDim.x * PPB;
  const bool v4 = (d & 3) == 0 && ((reinterpret_cast<uintptr_t>(X) |
                                    reinterpret_cast<uintptr_t>(C)) & 15) == 0;
  for (int pg = blockIdx.y * kPG; pg < P; pg += gridDim.y * kPG) {
    const int np = P - pg < kPG ? P - pg : kPG;
    // wave-uniform trip count (the shuffles need every lane)
    const int64_t g0 = (int64_t)blockIdx.x * PPB + threadIdx.x / LPP;
    const int64_t first = g0 - (threadIdx.x & (kWave - 1)) / LPP;  // group of lane 0
    for (int64_t gb = first; gb < n; gb += groups) {
      const int64_t i = gb + base / LPP;
      const bool valid = i < n;
      const float* xr = X + (valid ? i : 0) * d;
#pragma unroll
      for (int q = 0; q < kPG; ++q) {
        if (q >= np) break;  // uniform
        const int p = pg + q;
        const int64_t kbeg = k_off[p];
        const int64_t c = valid ? kbeg + labels[(int64_t)p * n + i] : kbeg;
        const float* cr = C + c * d;
        double acc = 0.0;
        for (int j0 = 0; j0 < d; j0 += 4 * LPP) {
          double sq[4];
          const int jb = j0 + 4 * sub;
          if (v4 && jb + 4 <= d) {  // one 16-byte load per row
            const float4 xv = *reinterpret_cast<const float4*>(xr + jb);
            const float4 cv = *reinterpret_cast<const float4*>(cr + jb);
            const double e0 = (double)xv.x - (double)cv.x, e1 = (double)xv.y - (double)cv.y;
            const double e2 = (double)xv.z - (double)cv.z, e3 = (double)xv.w - (double)cv.w;
            sq[0] = e0 * e0;
            sq[1] = e1 * e1;
            sq[2] = e2 * e2;
            sq[3] = e3 * e3;
          } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int j = jb + u;
              const double diff = j < d ? (double)xr[j] - (double)cr[j] : 0.0;
              sq[u] = diff * diff;
            }
          }
          const int steps = d - j0 >= 4 * LPP ? LPP : (d - j0 + 3) / 4;
          for (int st = 0; st < steps; ++st) {
            if (sub == st) {
#pragma unroll
              for (int u = 0; u < 4; ++u) acc = acc + sq[u];
            }
            acc = __shfl(acc, base + st, kWave);
          }
        }
        if (valid && sub == 0) dists[(int64_t)p * n + i] = (float)acc;
      }
    }
  }
}

// Direct-difference labeling: one thread per (point, partition), blocks of
// 256 points x one partition (grid.y over partitions, so a small data set
// with many partitions still fills the chip). The point sits in registers
// (DX = d exactly for d <= 4, else zero padded to DX: padding adds exact
// zeros) and the partition's centroids are staged in LDS, already converted
// to T, a tile of at most 16 KiB at a time. T = float: plain FP32; T =
// double: the oracle's FP64 arithmetic (ascending dimensions, strict <).
template <typename T, int DX>
__global__ __launch_bounds__(256) void k_label_direct(const float* __restrict__ X, int64_t n,
                                                      int d, const float* __restrict__ C,
                                                      const int64_t* __restrict__ k_off, int P,
                                                      int32_t* __restrict__ labels,
                                                      float* __restrict__ dists) {
  constexpr int kTileRaw = 16384 / (DX * (int)sizeof(T));
  constexpr int kTile = kTileRaw > 256 ? 256 : kTileRaw;
  __shared__ T s_c[kTile * DX];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  T xr[DX];
#pragma unroll
  for (int j = 0; j < DX; ++j) xr[j] = (valid && j < d) ? (T)X[i * d + j] : T(0);
  for (int p = blockIdx.y; p < P; p += gridDim.y) {
    const int64_t kbeg = k_off[p], kend = k_off[p + 1];
    T best = 0;
    int32_t bl = -1;
    for (int64_t c0 = kbeg; c0 < kend; c0 += kTile) {
      const int rows = (int)((kend - c0) < kTile ? (kend - c0) : kTile);
      __syncthreads();
      for (int v = threadIdx.x; v < rows * DX; v += blockDim.x) {
        const int r = v / DX, j = v - (v / DX) * DX;
        s_c[v] = j < d ? (T)C[(c0 + r) * d + j] : T(0);
      }
      __syncthreads();
      if (valid) {
        for (int r = 0; r < rows; ++r) {
          T dist = 0;
#pragma unroll
          for (int j = 0; j < DX; ++j) dist = sq_acc(dist, xr[j] - s_c[r * DX + j]);
          if (bl < 0 || dist < best) { best = dist; bl = (int32_t)(c0 + r - kbeg); }
        }
      }
    }
    if (valid) {
      labels[(int64_t)p * n + i] = bl;
      if (dists) dists[(int64_t)p * n + i] = (float)best;
    }
  }
}

// numpy's float32 sum of one short row v[0..d), d <= 128 (numpy's add
// reduction along the last axis: 0 + the pairwise sum of the row; a row of
// at most PW_BLOCKSIZE = 128 elements is one leaf: fewer than 8 elements
// summed in order from 0, otherwise eight partial sums over v[j], v[j+8],
// ... up to the last multiple of 8, combined as ((r0+r1)+(r2+r3)) +
// ((r4+r5)+(r6+r7)), then the remaining elements added in order). Every
// operation is a separately rounded float32 addition.
// (x - c) ** 2 in float32: a rounded difference, then a rounded square; the
// pragma keeps the square from being fused into the row sum's additions.
__device__ __forceinline__ float np_square_diff(float x, float c) {
#pragma clang fp contract(off)
  const float df = x - c;
  return df * df;
}

template <int DX>
__device__ __forceinline__ float np_row_sum(const float* v, int d) {
#pragma clang fp contract(off)
  if (d < 8) {
    float r = 0.0f;
#pragma unroll
    for (int j = 0; j < (DX < 8 ? DX : 8); ++j)
      if (j < d) r = __fadd_rn(r, v[j]);
    return r;
  }
  float r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = j < DX ? v[j < DX ? j : 0] : 0.0f;
  const int full = d - (d & 7);
#pragma unroll
  for (int i = 8; i + 8 <= DX; i += 8)
    if (i < full) {
#pragma unroll
      for (int j = 0; j < 8; ++j) r[j] = __fadd_rn(r[j], v[i + j]);
    }
  float res = __fadd_rn(__fadd_rn(__fadd_rn(r[0], r[1]), __fadd_rn(r[2], r[3])),
                        __fadd_rn(__fadd_rn(r[4], r[5]), __fadd_rn(r[6], r[7])));
#pragma unroll
  for (int i = 0; i < DX; ++i)
    if (i >= full && i < d) res = __fadd_rn(res, v[i]);
  return __fadd_rn(0.0f, res);
}

// Labeling in numpy's float32 arithmetic (KMeans.py:493-522 np_label):
// dist_c = ((x - c) ** 2).sum(axis=1) with a float32 difference, a float32
// square and numpy's row sum (np_row_sum); a centroid replaces the best one
// only when strictly closer (labels[dist < best_dist] = c), so the lowest k
// wins ties. Same grid and LDS staging as k_label_direct.
template <int DX>
__global__ __launch_bounds__(256) void k_label_np(const float* __restrict__ X, int64_t n, int d,
                                                  const float* __restrict__ C,
                                                  const int64_t* __restrict__ k_off, int P,
                                                  int32_t* __restrict__ labels,
                                                  float* __restrict__ dists) {
  constexpr int kTile = 16384 / (DX * 4) > 256 ? 256 : 16384 / (DX * 4);
  __shared__ float s_c[kTile * DX];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  float xr[DX];
#pragma unroll
  for (int j = 0; j < DX; ++j) xr[j] = (valid && j < d) ? X[i * d + j] : 0.0f;
  for (int p = blockIdx.y; p < P; p += gridDim.y) {
    const int64_t kbeg = k_off[p], kend = k_off[p + 1];
    float best = 0.0f;
    int32_t bl = -1;
    for (int64_t c0 = kbeg; c0 < kend; c0 += kTile) {
      const int rows = (int)((kend - c0) < kTile ? (kend - c0) : kTile);
      __syncthreads();
      for (int v = threadIdx.x; v < rows * DX; v += blockDim.x) {
        const int r = v / DX, j = v - (v / DX) * DX;
        s_c[v] = j < d ? C[(c0 + r) * d + j] : 0.0f;
      }
      __syncthreads();
      if (valid) {
        for (int r = 0; r < rows; ++r) {
          float sq[DX];
#pragma unroll
          for (int j = 0; j < DX; ++j) sq[j] = np_square_diff(xr[j], s_c[r * DX + j]);
          const float dist = np_row_sum<DX>(sq, d);
          if (bl < 0 || dist < best) { best = dist; bl = (int32_t)(c0 + r - kbeg); }
        }
      }
    }
    if (valid) {
      labels[(int64_t)p * n + i] = bl;
      if (dists) dists[(int64_t)p * n + i] = best;
    }
  }
}

int launch_np(const float* X, int64_t n, int d, const float* C, const int64_t* k_off, int P,
              int32_t* labels, float* dists, hipStream_t s) {
  const dim3 g(grid_for(n, 256), P < 65535 ? P : 65535);
  if (d == 1) k_label_np<1><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d == 2) k_label_np<2><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d <= 4) k_label_np<4><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d <= 8) k_label_np<8><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d <= 16) k_label_np<16><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d <= 32) k_label_np<32><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d <= 64) k_label_np<64><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d <= 128) k_label_np<128><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else return fail(EAC_ERR_UNSUPPORTED, "kmeans numpy arithmetic: d=%d > 128", d);
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}

// numpy's float32 sum of each partition's dists: the inertia of
// KMeans.py:164 (new_inertia = self._dists.sum()). numpy's add reduction
// starts from 0 and visits the array in chunks of its buffer size (8192
// elements), adding each chunk's pairwise sum in order
// (numpy/_core/src/umath/loops_utils.h.src @TYPE@_pairwise_sum,
// PW_BLOCKSIZE 128): a range of len < 8 is summed sequentially from 0; len
// <= 128 keeps eight partial sums over a[j], a[j+8], ... combined as
// ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then adds the tail in order; longer
// ranges split at n2 = len/2 rounded down to a multiple of 8. The tree of a
// chunk is a pure function of its length, so one wavefront evaluates it
// bottom-up: heap node h at depth t has its range found by descending from
// the root along the bits of h, each level's nodes are computed in parallel
// from the level below (held in LDS), and leaves (len <= 128) sum their
// range directly. Every addition is a float32 addition, so the result is
// numpy's to the bit.
constexpr int kNpBuf = 8192;
constexpr int kNpDepth = 8;  // chunk trees of len <= 8192 are at most 7 deep

__device__ __forceinline__ bool np_node(int len0, int h, int t, int& off, int& len) {
  off = 0;
  len = len0;
  for (int q = t - 1; q >= 0; --q) {
    if (len <= 128) return false;  // an ancestor is a leaf
    int n2 = len >> 1;
    n2 -= n2 & 7;
    if ((h >> q) & 1) { off += n2; len -= n2; } else { len = n2; }
  }
  return true;
}

__device__ __forceinline__ float np_leaf(const float* __restrict__ a, int len) {
  if (len < 8) {
    float res = 0.0f;
    for (int i = 0; i < len; ++i) res = __fadd_rn(res, a[i]);
    return res;
  }
  float r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = a[j];
  int i = 8;
  for (; i < len - (len & 7); i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = __fadd_rn(r[j], a[i + j]);
  }
  float res = __fadd_rn(__fadd_rn(__fadd_rn(r[0], r[1]), __fadd_rn(r[2], r[3])),
                        __fadd_rn(__fadd_rn(r[4], r[5]), __fadd_rn(r[6], r[7])));
  for (; i < len; ++i) res = __fadd_rn(res, a[i]);
  return res;
}

// One wavefront per (chunk, row): csum[row * nchunks + chunk].
__global__ __launch_bounds__(kWave) void k_np_chunk_sums(const float* __restrict__ x, int64_t n,
                                                         int64_t nchunks,
                                                         float* __restrict__ csum) {
  __shared__ float v[2 << kNpDepth];
  const int64_t c = blockIdx.x;
  const int64_t row = blockIdx.y;
  const int64_t base = c * kNpBuf;
  const int len0 = (int)(n - base < kNpBuf ? n - base : kNpBuf);
  const float* a = x + row * n + base;
  for (int t = kNpDepth; t >= 0; --t) {
    for (int h = (1 << t) + (int)threadIdx.x; h < (2 << t); h += kWave) {
      int off, len;
      if (!np_node(len0, h, t, off, len)) continue;
      v[h] = len <= 128 ? np_leaf(a + off, len) : __fadd_rn(v[2 * h], v[2 * h + 1]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) csum[row * nchunks + c] = v[1];
}

// out[row] = ((0 + csum[row][0]) + csum[row][1]) + ...
__global__ void k_np_finish(const float* __restrict__ csum, int64_t nchunks, int64_t rows,
                            float* __restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  float r = 0.0f;
  for (int64_t c = 0; c < nchunks; ++c) r = __fadd_rn(r, csum[row * nchunks + c]);
  out[row] = r;
}

int64_t np_chunks(int64_t n) { return n > 0 ? (n + kNpBuf - 1) / kNpBuf : 0; }

// numpy float32 sums of `rows` rows of length n; csum holds rows * np_chunks(n)
int np_sum_rows(const float* x, int64_t n, int64_t rows, float* csum, float* out,
                hipStream_t s) {
  const int64_t nc = np_chunks(n);
  if (nc > 0) {
    for (int64_t r0 = 0; r0 < rows; r0 += 65535) {
      const int64_t rr = rows - r0 < 65535 ? rows - r0 : 65535;
      k_np_chunk_sums<<<dim3((unsigned)nc, (unsigned)rr), kWave, 0, s>>>(x + r0 * n, n, nc,
                                                                          csum + r0 * nc);
      EAC_LAUNCH_CHECK();
    }
  }
  k_np_finish<<<grid_for(rows, 256), 256, 0, s>>>(csum, nc, rows, out);
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}

template <typename T>
int launch_direct(const float* X, int64_t n, int d, const float* C, const int64_t* k_off, int P,
                  int32_t* labels, float* dists, hipStream_t s) {
  const dim3 g(grid_for(n, 256), P < 65535 ? P : 65535);
  if (d == 1) k_label_direct<T, 1><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d == 2) k_label_direct<T, 2><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d == 3) k_label_direct<T, 3><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d == 4) k_label_direct<T, 4><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d <= 8) k_label_direct<T, 8><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d <= 16) k_label_direct<T, 16><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d <= 32) k_label_direct<T, 32><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d <= 64) k_label_direct<T, 64><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else if (d <= 128) k_label_direct<T, 128><<<g, 256, 0, s>>>(X, n, d, C, k_off, P, labels, dists);
  else return fail(EAC_ERR_UNSUPPORTED, "kmeans direct: d=%d > 128", d);
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}

// ------------------------------------------------------- centroid update
// Sort keys (global centroid id) and point indices of every (partition,
// point), and the member count of every centroid. Blocks cover a run of one
// partition's points (grid.y over partitions); when that run is long next to
// the partition's k, counts meet in LDS first (one global atomic per
// centroid and block instead of one per point).
constexpr int kKeyLds = 4096;
__global__ __launch_bounds__(256) void k_label_keys(const int32_t* __restrict__ labels,
                                                    int64_t n, int P,
                                                    const int64_t* __restrict__ k_off,
                                                    uint32_t* __restrict__ key,
                                                    uint32_t* __restrict__ idx,
                                                    int32_t* __restrict__ count) {
  __shared__ int32_t s_cnt[kKeyLds];
  const int64_t per_block = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * per_block;
  const int64_t i1 = i0 + per_block < n ? i0 + per_block : n;
  for (int p = blockIdx.y; p < P; p += gridDim.y) {
    const int64_t kb = k_off[p], kp = k_off[p + 1] - kb;
    const bool lds = kp <= kKeyLds && 4 * kp <= i1 - i0;  // uniform over the block
    if (lds) {
      for (int c = threadIdx.x; c < kp; c += blockDim.x) s_cnt[c] = 0;
      __syncthreads();
    }
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
      const int64_t t = (int64_t)p * n + i;
      const int32_t l = labels[t];
      key[t] = (uint32_t)(kb + l);
      idx[t] = (uint32_t)i;
      if (lds) atomicAdd(&s_cnt[l], 1);
      else atomicAdd(&count[kb + l], 1);
    }
    if (lds) {
      __syncthreads();
      for (int c = threadIdx.x; c < kp; c += blockDim.x)
        if (s_cnt[c]) atomicAdd(&count[kb + c], s_cnt[c]);
      __syncthreads();
    }
  }
}

// new[c][j] = fp32 sequential sum over members in point order, / count in
// double (KMeans.py:736-745). One single-wave block per centroid and group
// of 128 dimensions: the lanes gather up to 2048 member coordinates at a
// time into LDS (many loads in flight), then lane j adds dimension j's
// values in member order (lane j + 64 too), so the serial chain runs on LDS
// reads.
constexpr int kSumStage = 2048;
constexpr int kSumDims = 2 * kWave;
__global__ __launch_bounds__(kWave) void k_centroid_sums(const float* __restrict__ X, int d,
                                                         const uint32_t* __restrict__ members,
                                                         const int64_t* __restrict__ mstart,
                                                         const int32_t* __restrict__ count,
                                                         int64_t ktot, float* __restrict__ Cnew) {
  __shared__ float s_x[kSumStage];
  const int lane = threadIdx.x;
  for (int64_t c = blockIdx.x; c < ktot; c += gridDim.x) {
    const int32_t cnt = count[c];
    if (cnt == 0) continue;  // uniform over the block
    const uint32_t* m = members + mstart[c];
    for (int j0 = 0; j0 < d; j0 += kSumDims) {
      const int dd = d - j0 < kSumDims ? d - j0 : kSumDims;
      const int q_max = kSumStage / dd;  // members per stage
      float s0 = 0.f, s1 = 0.f;
      for (int32_t b = 0; b < cnt; b += q_max) {
        const int q = cnt - b < q_max ? cnt - b : q_max;
        __syncthreads();
        for (int v = lane; v < q * dd; v += kWave) {
          const int r = v / dd;
          s_x[v] = X[(int64_t)m[b + r] * d + j0 + (v - r * dd)];
        }
        __syncthreads();
        if (lane < dd)
          for (int r = 0; r < q; ++r) s0 = s0 + s_x[r * dd + lane];
        if (lane + kWave < dd)
          for (int r = 0; r < q; ++r) s1 = s1 + s_x[r * dd + lane + kWave];
      }
      if (lane < dd) Cnew[c * d + j0 + lane] = (float)((double)s0 / (double)cnt);
      if (lane + kWave < dd)
        Cnew[c * d + j0 + lane + kWave] = (float)((double)s1 / (double)cnt);
    }
  }
}

// The same sums for d <= 64 with L lanes per centroid (L = the power of two
// >= d, at least 4; 64 / L centroids per wavefront): lane j of a group adds
// dimension j of its centroid's members in point order straight from X, 16
// member rows loaded ahead of the (in-order) additions. The one-wave-per-
// centroid form above leaves 64 - d lanes idle and waits on LDS for every
// member.
template <int L>
__global__ __launch_bounds__(256) void k_centroid_sums_narrow(
    const float* __restrict__ X, int d, const uint32_t* __restrict__ members,
    const int64_t* __restrict__ mstart, const int32_t* __restrict__ count, int64_t ktot,
    float* __restrict__ Cnew) {
  constexpr int U = 16;
  const int j = threadIdx.x % L;
  const int64_t groups = (int64_t)gridDim.x * (blockDim.x / L);
  for (int64_t c = (int64_t)blockIdx.x * (blockDim.x / L) + threadIdx.x / L; c < ktot;
       c += groups) {
    const int32_t cnt = count[c];
    if (cnt == 0 || j >= d) continue;
    const uint32_t* m = members + mstart[c];
    const float* xj = X + j;
    float s = 0.f;
    int32_t r = 0;
    for (; r + U <= cnt; r += U) {
      uint32_t mi[U];
      float v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) mi[u] = m[r + u];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = xj[(int64_t)mi[u] * d];
#pragma unroll
      for (int u = 0; u < U; ++u) s = s + v[u];
    }
    for (; r < cnt; ++r) s = s + xj[(int64_t)m[r] * d];
    Cnew[c * d + j] = (float)((double)s / (double)cnt);
  }
}

// Copies the new centroids of the partitions still running (upd[p] != 0).
__global__ void k_copy_running(float* __restrict__ C, const float* __restrict__ Cnew, int d,
                               const int64_t* __restrict__ k_off, int P,
                               const int32_t* __restrict__ upd) {
  for (int p = blockIdx.y; p < P; p += gridDim.y) {
    if (!upd[p]) continue;
    const int64_t a = k_off[p] * d, b = k_off[p + 1] * d;
    for (int64_t t = a + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < b;
         t += (int64_t)gridDim.x * blockDim.x)
      C[t] = Cnew[t];
  }
}

// KMeans.py:117-189 stopping rule of iteration `it` for every partition
// still running: new_inertia = numpy float32 sum of the dists (inertia[p]),
// err = |new_inertia - previous| in float32 (previous starts at +inf); stop
// when err <= tol (compared in float32, as numpy compares a float32 scalar
// with a Python float) or it >= max_iter, else mark the partition for a
// centroid update. any[0] = some partition continues.
__global__ void k_converge(const float* __restrict__ inertia, float* __restrict__ prev,
                           int32_t* __restrict__ active, int32_t* __restrict__ iters,
                           int32_t* __restrict__ upd, int P, int it, int max_iter, float tol,
                           int32_t* __restrict__ any) {
  __shared__ int s_any;
  if (threadIdx.x == 0) s_any = 0;
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    upd[p] = 0;
    if (!active[p]) continue;
    iters[p] = it;
    const float err = fabsf(__fsub_rn(inertia[p], prev[p]));
    prev[p] = inertia[p];
    if (err <= tol || it >= max_iter) {
      active[p] = 0;
      continue;
    }
    upd[p] = 1;
    s_any = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) any[0] = s_any;
}

__global__ void k_fit_init(float* __restrict__ prev, int32_t* __restrict__ active,
                           int32_t* __restrict__ iters, int P) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
    prev[p] = INFINITY;
    active[p] = 1;
    iters[p] = 0;
  }
}

__global__ void k_empty_flags(const int32_t* __restrict__ count, int64_t ktot,
                              int32_t* __restrict__ flag) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ktot;
       c += (int64_t)gridDim.x * blockDim.x)
    flag[c] = count[c] == 0;
}

// keys (dist bits << 32 | index) of one partition: ascending = (dist, index)
__global__ void k_dist_keys(const float* __restrict__ dists, int64_t n,
                            uint64_t* __restrict__ keys) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    keys[i] = ((uint64_t)__float_as_uint(dists[i]) << 32) | (uint64_t)i;
}

// empty clusters c of one partition (ascending) take the points of the
// smallest (dist, index), in that order
__global__ void k_reseed(const float* __restrict__ X, int d, const uint64_t* __restrict__ sorted,
                         const int32_t* __restrict__ count, const int32_t* __restrict__ erank,
                         int64_t kbeg, int64_t kend, float* __restrict__ Cnew) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (kend - kbeg) * d;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t c = kbeg + t / d;
    int j = (int)(t % d);
    if (count[c] != 0) continue;
    int32_t r = erank[c] - erank[kbeg];
    uint32_t src = (uint32_t)(sorted[r] & 0xffffffffu);
    Cnew[c * d + j] = X[(int64_t)src * d + j];
  }
}

// ----------------------------------------- capturable fit (no host reads)
// Total of the flag regions (clamped to the list), error bit 1 when a
// region or the list overflowed. One wavefront.
__global__ void k_flags_total(const unsigned long long* __restrict__ cnt, int64_t region,
                              int64_t cap, int64_t* __restrict__ d_nf,
                              unsigned int* __restrict__ err) {
  const int l = lane_id();
  const unsigned long long c = l < kFlagSlots ? cnt[l] : 0ull;
  const bool over = (int64_t)c > region;
  unsigned long long tot = c;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) tot += __shfl_xor(tot, off, kWave);
  const bool any_over = __ballot(over) != 0ull;
  if (l == 0) {
    if (any_over || (int64_t)tot > cap) atomicOr(err, 1u);
    d_nf[0] = (int64_t)tot < cap ? (int64_t)tot : cap;
  }
}

// Counting sort of the flags by partition (flag = partition << 40 | row),
// stable-free but deterministic: kSortBlocks workgroups each take the same
// grid-stride share of the flags in both passes; pass 1 counts its share's
// flags per partition in LDS (bcount[p][b]); one workgroup scans bcount in
// (partition, workgroup) order into offsets; pass 2 places the share's flags
// from LDS cursors started at those offsets. No global atomics (their
// serialisation on a few hot partitions dominated).
constexpr int kSortBlocks = 256;
constexpr int kSortMaxP = 4096;  // partitions counted in LDS
__global__ __launch_bounds__(256) void k_flag_bcount(const uint64_t* __restrict__ flags,
                                                     const int64_t* __restrict__ d_nf, int P,
                                                     int32_t* __restrict__ bcount) {
  __shared__ int32_t h[kSortMaxP];
  for (int p = threadIdx.x; p < P; p += blockDim.x) h[p] = 0;
  __syncthreads();
  const int64_t nf = *d_nf;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < nf;
       f += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)(flags[f] >> 40);
    if (p < P) atomicAdd(&h[p], 1);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += blockDim.x) bcount[(int64_t)p * gridDim.x + blockIdx.x] = h[p];
}
__global__ __launch_bounds__(1024) void k_flag_bscan(int32_t* __restrict__ bcount, int64_t len) {
  __shared__ int32_t s_tmp[1024 / kWave + 1];
  int32_t carry = 0;
  for (int64_t i0 = 0; i0 < len; i0 += 1024) {
    const int64_t i = i0 + threadIdx.x;
    const int32_t v = i < len ? bcount[i] : 0;
    int32_t tot;
    const int32_t pre = block_excl_scan<int32_t>(v, s_tmp, &tot);
    if (i < len) bcount[i] = carry + pre;
    carry += tot;
    __syncthreads();
  }
}
__global__ __launch_bounds__(256) void k_flag_bscatter(const uint64_t* __restrict__ flags,
                                                       const int64_t* __restrict__ d_nf, int P,
                                                       const int32_t* __restrict__ boff,
                                                       uint64_t* __restrict__ out) {
  __shared__ int32_t cur[kSortMaxP];
  for (int p = threadIdx.x; p < P; p += blockDim.x)
    cur[p] = boff[(int64_t)p * gridDim.x + blockIdx.x];
  __syncthreads();
  const int64_t nf = *d_nf;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < nf;
       f += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t fl = flags[f];
    const int p = (int)(fl >> 40);
    if (p < P) out[atomicAdd(&cur[p], 1)] = fl;
  }
}

// Device reseeding of every partition's empty clusters (the same rule as
// k_dist_keys + radix sort + k_reseed): the r-th empty cluster of partition
// p (ascending) takes the point of the r-th smallest key (dist bits << 32 |
// index). One workgroup per partition with empties: the e-th smallest key by
// a radix select (eight 8-bit digit passes over the partition's dists), the
// e keys at or below it collected in LDS and ranked among themselves. More
// than kMaxReseed empties in one partition set error bit 2.
constexpr int kMaxReseed = 2048;
__global__ __launch_bounds__(1024) void k_reseed_dev(
    const float* __restrict__ X, int d, const float* __restrict__ dists, int64_t n,
    const int64_t* __restrict__ k_off, int P, const int32_t* __restrict__ count,
    const int32_t* __restrict__ erank, float* __restrict__ Cnew, unsigned int* __restrict__ err) {
  __shared__ unsigned int hist[256];
  __shared__ unsigned long long s_sel[kMaxReseed];
  __shared__ int32_t s_emp[kMaxReseed];
  __shared__ unsigned long long s_prefix;
  __shared__ int64_t s_want;
  __shared__ int s_nsel;
  const int t = threadIdx.x;
  for (int p = blockIdx.x; p < P; p += gridDim.x) {
    const int64_t kb = k_off[p], ke = k_off[p + 1];
    const int64_t e = (int64_t)erank[ke] - erank[kb];
    if (e == 0) continue;
    if (e > n || e > kMaxReseed) {
      if (t == 0) atomicOr(err, 2u);
      continue;
    }
    const float* dp = dists + (int64_t)p * n;
    if (t == 0) {
      s_prefix = 0ull;
      s_want = e;  // rank (1-based) of the wanted key among those left
    }
    for (int shift = 56; shift >= 0; shift -= 8) {
      for (int b = t; b < 256; b += blockDim.x) hist[b] = 0u;
      __syncthreads();
      const unsigned long long pre = s_prefix;
      for (int64_t i = t; i < n; i += blockDim.x) {
        const unsigned long long key =
            ((unsigned long long)__float_as_uint(dp[i]) << 32) | (unsigned long long)i;
        if (shift == 56 || ((key ^ pre) >> (shift + 8)) == 0ull)
          atomicAdd(&hist[(key >> shift) & 255u], 1u);
      }
      __syncthreads();
      if (t == 0) {
        int64_t want = s_want, cum = 0;
        int b = 0;
        for (; b < 255 && cum + (int64_t)hist[b] < want; ++b) cum += hist[b];
        s_want = want - cum;
        s_prefix = pre | ((unsigned long long)b << shift);
      }
      __syncthreads();
    }
    const unsigned long long T = s_prefix;  // the e-th smallest key (keys are unique)
    if (t == 0) s_nsel = 0;
    for (int64_t c = kb + t; c < ke; c += blockDim.x)
      if (count[c] == 0) s_emp[erank[c] - erank[kb]] = (int32_t)c;
    __syncthreads();
    for (int64_t i = t; i < n; i += blockDim.x) {
      const unsigned long long key =
          ((unsigned long long)__float_as_uint(dp[i]) << 32) | (unsigned long long)i;
      if (key <= T) {
        const int at = atomicAdd(&s_nsel, 1);
        if (at < kMaxReseed) s_sel[at] = key;
      }
    }
    __syncthreads();
    const int ns = s_nsel < (int)e ? s_nsel : (int)e;
    for (int j = t; j < ns; j += blockDim.x) {
      const unsigned long long key = s_sel[j];
      int r = 0;
      for (int q = 0; q < ns; ++q) r += s_sel[q] < key;
      const int64_t c = s_emp[r];
      const int64_t pt = (int64_t)(key & 0xffffffffull);
      for (int jj = 0; jj < d; ++jj) Cnew[c * d + jj] = X[pt * d + jj];
    }
    __syncthreads();
  }
}

// Running totals of a capturable fit: info[0] += flags re-scored,
// info[1] += empty clusters reseeded (device; one thread), counted only
// while some partition still runs (any[0], the last stopping-rule result),
// as the fit that stops when none does counts them.
__global__ void k_fit_tally(int64_t* __restrict__ info, const int64_t* __restrict__ d_nf,
                            const int32_t* __restrict__ erank, int64_t ktot,
                            const int32_t* __restrict__ any) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && any[0]) {
    if (d_nf) info[0] += *d_nf;
    if (erank) info[1] += erank[ktot];
  }
}

int kd_for(int d) {
  if (d <= 32) return 32;
  if (d <= 64) return 64;
  if (d <= 128) return 128;
  return 0;
}

int d2h(void* host, const void* dev, size_t bytes, hipStream_t s) {
  void* pin = pinned_scratch(bytes);
  if (!pin) return fail(EAC_ERR_HIP, "pinned host allocation failed");
  EAC_HIP_CHECK(hipMemcpyAsync(pin, dev, bytes, hipMemcpyDeviceToHost, s));
  EAC_HIP_CHECK(hipStreamSynchronize(s));
  memcpy(host, pin, bytes);
  return EAC_OK;
}

int h2d(void* dev, const void* host, size_t bytes, hipStream_t s) {
  void* pin = pinned_scratch(bytes);
  if (!pin) return fail(EAC_ERR_HIP, "pinned host allocation failed");
  EAC_HIP_CHECK(hipStreamSynchronize(s));
  memcpy(pin, host, bytes);
  EAC_HIP_CHECK(hipMemcpyAsync(dev, pin, bytes, hipMemcpyHostToDevice, s));
  EAC_HIP_CHECK(hipStreamSynchronize(s));  // pinned buffer is reused
  return EAC_OK;
}

struct LabelBufs {
  int64_t* k_off;
  uint16_t* Ch;
  uint16_t* Cl;
  float* Ct;  // centroids transposed: [d][ktot]
  float* ncf;
  unsigned int* pstat;
  uint64_t* flags_r;  // kFlagSlots regions of `region` flags
  uint64_t* flags;    // the flags of all regions, contiguous
  unsigned long long* nflag;  // per-region counts
  float* csum;                // numpy-order chunk sums of the dists
  int64_t flag_cap;
  int64_t region;
};

LabelBufs carve_label(Arena& ar, int64_t n, int d, int64_t ktot, int P) {
  LabelBufs b;
  int KD = kd_for(d) ? kd_for(d) : 32;
  b.k_off = ar.take<int64_t>(P + 1);
  b.Ch = ar.take<uint16_t>((size_t)ktot * KD + 64);
  b.Cl = ar.take<uint16_t>((size_t)ktot * KD + 64);
  b.Ct = ar.take<float>((size_t)ktot * d);
  b.ncf = ar.take<float>(ktot);
  b.pstat = ar.take<unsigned int>((size_t)P * kStat);
  b.flag_cap = (int64_t)P * n;
  // a region holds every (partition, row) its blocks can flag
  const int64_t nblocks = (n + kBM - 1) / kBM;
  b.region = ((nblocks + kFlagSlots - 1) / kFlagSlots) * kBM * (int64_t)P;
  b.flags_r = ar.take<uint64_t>((size_t)kFlagSlots * b.region);
  b.flags = ar.take<uint64_t>(b.flag_cap);
  b.nflag = ar.take<unsigned long long>(kFlagSlots);
  b.csum = ar.take<float>((size_t)P * np_chunks(n) + 1);
  return b;
}

}  // namespace

size_t kmeans_label_ws_bytes(int64_t n, int d, int64_t ktot, int P) {
  Arena ar(nullptr, 0, true);
  (void)carve_label(ar, n, d, ktot, P);
  return ar.used + 4096;
}

// Device-side state of a capturable fit: flag count, per-partition flag
// counts and cursors, flags sorted by partition, running totals (info[0]
// flags re-scored, info[1] clusters reseeded), error bits.
struct KmGraph {
  const int32_t* any;  // some partition still runs (the labeling counts only then)
  int64_t* d_nf;
  int32_t* bcount;     // [P][kSortBlocks] flags per (partition, sorting workgroup)
  uint64_t* sorted;
  int64_t* info;
  unsigned int* err;
};

// Label all partitions; dists (nullable) receive FP64-exact distances for
// the F64 / BF16X3 modes and fp32 distances for F32; inertia[p] (device,
// nullable) their per-partition numpy float32 sums.
// With g (capturable calls) nothing is read back: the flag count stays on
// the device and sizes the re-scoring from there.
int kmeans_label(const float* X, int64_t n, int d, const float* C, const int64_t* h_koff,
                 int P, int mode, int32_t* labels, float* dists, float* inertia_out,
                 int64_t* n_rescored, Arena& ar, hipStream_t s,
                 const int64_t* d_koff = nullptr, const KmGraph* kg = nullptr) {
  const int64_t ktot = h_koff[P];
  LabelBufs B = carve_label(ar, n, d, ktot, P);
  if (ar.overflow()) return fail(EAC_ERR_WORKSPACE, "kmeans: workspace too small");
  if (d_koff) {
    B.k_off = const_cast<int64_t*>(d_koff);  // read only
  } else {
    EAC_HIP_CHECK(hipMemcpyAsync(B.k_off, h_koff, sizeof(int64_t) * (P + 1),
                                 hipMemcpyHostToDevice, s));
    EAC_HIP_CHECK(hipStreamSynchronize(s));  // h_koff may be a pageable stack array
  }
  if (n_rescored) *n_rescored = 0;
  if (mode == EAC_KM_F32 || mode == EAC_KM_F64 || mode == EAC_KM_F32_NUMPY ||
      (mode == EAC_KM_BF16X3 && d < 16)) {
    int pr = prof_begin(EAC_PROF_KM_LABEL, 2 * n * (int64_t)d * ktot, s);
    if (mode == EAC_KM_F32_NUMPY)
      EAC_TRY(launch_np(X, n, d, C, B.k_off, P, labels, dists, s));
    else if (mode == EAC_KM_F32)
      EAC_TRY(launch_direct<float>(X, n, d, C, B.k_off, P, labels, dists, s));
    else
      EAC_TRY(launch_direct<double>(X, n, d, C, B.k_off, P, labels, dists, s));
    prof_end(pr, s);
    if (inertia_out) {
      if (!dists) return fail(EAC_ERR_ARG, "kmeans: inertia needs dists");
      EAC_TRY(np_sum_rows(dists, n, P, B.csum, inertia_out, s));
    }
    return EAC_OK;
  }
  if (mode != EAC_KM_BF16X3) return fail(EAC_ERR_ARG, "kmeans: unknown mode %d", mode);
  const int KD = kd_for(d);
  if (!KD) return fail(EAC_ERR_UNSUPPORTED, "kmeans bf16x3: d=%d > 128", d);
  EAC_TRY(fill<int32_t>((int32_t*)B.pstat, (int64_t)P * kStat, 0, s));
  EAC_TRY(fill<uint64_t>((uint64_t*)B.nflag, kFlagSlots, 0ull, s));
  k_cent_prep<<<grid_for(ktot, 256, 8192), 256, 0, s>>>(C, ktot, d, KD, B.k_off, P, B.Ch, B.Cl,
                                                        B.Ct, B.ncf, B.pstat);
  EAC_LAUNCH_CHECK();
  int g = (int)((n + kBM - 1) / kBM);
  int pr = prof_begin(EAC_PROF_KM_LABEL, 2 * n * (int64_t)d * ktot, s);
  {
    // waves per SIMD the kernel is compiled for: the built-in choice, or
    // EAC_TUNE_KM_WAVES (2, 3 or 4; the same labels, other register
    // budgets and occupancy)
    const int wk = tune_value(EAC_TUNE_KM_WAVES);
    const int w = (wk >= 2 && wk <= 4) ? wk : 0;
    auto go = [&](auto kern) {
      kern<<<g, kLB, 0, s>>>(X, n, d, B.Ch, B.Cl, B.ncf, B.k_off, P, B.pstat, labels, B.flags_r,
                             B.nflag, B.region);
    };
    if (KD == 32) {
      if (w == 2) go(k_label_bf16x3<32, 2>);
      else if (w == 4) go(k_label_bf16x3<32, 4>);
      else go(k_label_bf16x3<32, 3>);
    } else if (KD == 64) {
      if (w == 2) go(k_label_bf16x3<64, 2>);
      else if (w == 4) go(k_label_bf16x3<64, 4>);
      else go(k_label_bf16x3<64, 3>);
    } else {
      if (w == 2) go(k_label_bf16x3<128, 2>);
      else if (w == 3) go(k_label_bf16x3<128, 3>);
      else go(k_label_bf16x3<128, 1>);
    }
  }
  EAC_LAUNCH_CHECK();
  prof_end(pr, s);
  if (kg) {
    // flags gathered, sorted by partition and re-scored with device counts
    // (grids sized for the whole list, blocks past the count exit at once)
    k_flags_total<<<1, kWave, 0, s>>>(B.nflag, B.region, B.flag_cap, kg->d_nf, kg->err);
    EAC_LAUNCH_CHECK();
    const int gp = grid_for(B.flag_cap, 256, 4096);
    k_gather_flags<<<gp, 256, 0, s>>>(B.flags_r, B.region, B.nflag, B.flags, 0, kg->d_nf);
    EAC_LAUNCH_CHECK();
    const uint64_t* fl = B.flags;
    const bool tiled = rescore_tiled_lds(d) <= 64 * 1024;
    if (tiled && P > 1 && P <= kSortMaxP) {
      k_flag_bcount<<<kSortBlocks, 256, 0, s>>>(B.flags, kg->d_nf, P, kg->bcount);
      EAC_LAUNCH_CHECK();
      k_flag_bscan<<<1, 1024, 0, s>>>(kg->bcount, (int64_t)P * kSortBlocks);
      EAC_LAUNCH_CHECK();
      k_flag_bscatter<<<kSortBlocks, 256, 0, s>>>(B.flags, kg->d_nf, P, kg->bcount, kg->sorted);
      EAC_LAUNCH_CHECK();
      fl = kg->sorted;
    }
    int pr2 = prof_begin(EAC_PROF_KM_RESCORE, 0, s);
    if (tiled) {
      const size_t lds = rescore_tiled_lds(d);
      int dev = 0, cus = 256;
      if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      // as many workgroups as the eager launch of a few million flags
      // would have; those past the device count exit at once
      const int64_t tiles = (B.flag_cap + kRtF - 1) / kRtF;
      const unsigned nb = (unsigned)(tiles < 64 * (int64_t)cus ? tiles : 64 * (int64_t)cus);
      auto go = [&](auto kern) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        kern<<<nb, kRtF * kRtS, lds, s>>>(X, n, d, C, B.k_off, fl, 0, labels, kg->d_nf);
      };
      if (d <= 16) go(k_rescore_tiled<16>);
      else if (d <= 32) go(k_rescore_tiled<32>);
      else if (d <= 64) go(k_rescore_tiled<64>);
      else go(k_rescore_tiled<0>);
    } else {
      if (P == 1)
        k_rescore_f64<2><<<grid_for(B.flag_cap * kWave / 2, 256, 16384), 256, 0, s>>>(
            X, n, d, B.Ct, ktot, B.k_off, B.flags, 0, labels, kg->d_nf);
      else
        k_rescore_f64<1><<<grid_for(B.flag_cap * kWave, 256, 16384), 256, 0, s>>>(
            X, n, d, B.Ct, ktot, B.k_off, B.flags, 0, labels, kg->d_nf);
    }
    EAC_LAUNCH_CHECK();
    prof_end(pr2, s);
    k_fit_tally<<<1, 64, 0, s>>>(kg->info, kg->d_nf, nullptr, 0, kg->any);
    EAC_LAUNCH_CHECK();
  } else {
  unsigned long long hcnt[kFlagSlots];
  EAC_TRY(d2h(hcnt, B.nflag, sizeof(hcnt), s));
  unsigned long long nf = 0;
  for (int q = 0; q < kFlagSlots; ++q) {
    if ((int64_t)hcnt[q] > B.region) return fail(EAC_ERR_HIP, "kmeans: flag list overflow");
    nf += hcnt[q];
  }
  if ((int64_t)nf > B.flag_cap) return fail(EAC_ERR_HIP, "kmeans: flag list overflow");
  if (nf > 0) {
    k_gather_flags<<<grid_for((int64_t)nf, 256, 65536), 256, 0, s>>>(B.flags_r, B.region, B.nflag,
                                                                     B.flags, (int64_t)nf);
    EAC_LAUNCH_CHECK();
    int pr2 = prof_begin(EAC_PROF_KM_RESCORE, 3 * (int64_t)nf * d * (ktot / P), s);
    // Many flags (at least 1024 blocks of 64) and d <= 127 (LDS tiles up to
    // 64 KiB): flags sorted by partition (one radix pass for up to 256
    // partitions) feed the LDS-tiled re-scoring. Few flags, or no room for the sort: one wavefront per flag
    // over the transposed centroids (more parallelism per flag).
    const uint64_t* fl = B.flags;
    bool tiled = nf >= (unsigned long long)kRtF * 1024 && rescore_tiled_lds(d) <= 64 * 1024;
    if (tiled && P > 1) {
      const size_t mark2 = ar.used;
      uint64_t* sorted = ar.take<uint64_t>(nf);
      if (ar.overflow() || ar.cap - ar.used < radix_sort_ws_bytes((int64_t)nf, 8, 0)) {
        tiled = false;
        ar.used = mark2;
      } else {
        const int pb = bits_for((uint64_t)(P - 1));
        EAC_TRY((radix_sort<uint64_t, NoValue>(B.flags, sorted, nullptr, nullptr, (int64_t)nf, 40,
                                               40 + (pb < 1 ? 1 : pb), ar, s)));
        fl = sorted;
      }
    }
    if (tiled) {
      const size_t lds = rescore_tiled_lds(d);
      const unsigned nb = (unsigned)(((int64_t)nf + kRtF - 1) / kRtF);
      auto go = [&](auto kern) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        kern<<<nb, kRtF * kRtS, lds, s>>>(X, n, d, C, B.k_off, fl, (int64_t)nf, labels, nullptr);
      };
      if (d <= 16) go(k_rescore_tiled<16>);
      else if (d <= 32) go(k_rescore_tiled<32>);
      else if (d <= 64) go(k_rescore_tiled<64>);
      else go(k_rescore_tiled<0>);
    } else {
      if (P == 1)
        k_rescore_f64<2><<<grid_for((int64_t)(nf + 1) / 2 * kWave, 256, 65536), 256, 0, s>>>(
            X, n, d, B.Ct, ktot, B.k_off, B.flags, (int64_t)nf, labels, nullptr);
      else
        k_rescore_f64<1><<<grid_for((int64_t)nf * kWave, 256, 65536), 256, 0, s>>>(
            X, n, d, B.Ct, ktot, B.k_off, B.flags, (int64_t)nf, labels, nullptr);
    }
    EAC_LAUNCH_CHECK();
    prof_end(pr2, s);
  }
  if (n_rescored) *n_rescored = (int64_t)nf;
  }  // read-back path
  if (inertia_out && !dists) return fail(EAC_ERR_ARG, "kmeans: inertia needs dists");
  if (dists) {
    const int pgs = (P + kPG - 1) / kPG;
    const int py = pgs < 65535 ? pgs : 65535;
    k_exact_dists<4><<<dim3(grid_for(n, 64, 16384), py), 256, 0, s>>>(X, n, d, C, B.k_off, P,
                                                                        labels, dists);
    EAC_LAUNCH_CHECK();
  }
  if (inertia_out) EAC_TRY(np_sum_rows(dists, n, P, B.csum, inertia_out, s));
  return EAC_OK;
}

size_t kmeans_update_ws_bytes(int64_t n, int d, int64_t ktot, int P) {
  Arena ar(nullptr, 0, true);
  int64_t items = (int64_t)P * n;
  ar.take<int64_t>(P + 1);
  ar.take<uint32_t>(items);
  ar.take<uint32_t>(items);
  ar.take<uint32_t>(items);
  ar.take<uint32_t>(items);
  ar.take<int32_t>(ktot + 1);
  ar.take<int64_t>(ktot + 1);
  ar.take<int32_t>(ktot + 1);
  ar.take<int32_t>(ktot + 1);
  ar.take<uint64_t>(n);
  ar.take<uint64_t>(n);
  size_t a = ar.used;
  size_t sort_a = radix_sort_ws_bytes(items, 4, 4);
  size_t sort_b = radix_sort_ws_bytes(n, 8, 0);
  size_t scan_b = scan_ws_bytes(ktot + 1, 8);
  size_t m = sort_a > sort_b ? sort_a : sort_b;
  m = m > scan_b ? m : scan_b;
  return a + m + 4096;
}

// New centroids for every partition (KMeans.py:709-754); empty clusters take
// the points of smallest (dist, index) in ascending cluster order.
// With g (capturable calls) the empty clusters are reseeded on the device.
int kmeans_update(const float* X, int64_t n, int d, const int32_t* labels, const float* dists,
                  const int64_t* h_koff, int P, float* Cnew, int64_t* n_empty, Arena& ar,
                  hipStream_t s, const int64_t* d_koff = nullptr, const KmGraph* g = nullptr) {
  const int64_t ktot = h_koff[P];
  const int64_t items = (int64_t)P * n;
  if (items > (int64_t)UINT32_MAX || ktot >= (int64_t)UINT32_MAX)
    return fail(EAC_ERR_UNSUPPORTED, "kmeans update: P*n or k too large");
  int64_t* k_off = ar.take<int64_t>(P + 1);
  uint32_t* key = ar.take<uint32_t>(items);
  uint32_t* idx = ar.take<uint32_t>(items);
  uint32_t* skey = ar.take<uint32_t>(items);
  uint32_t* members = ar.take<uint32_t>(items);
  int32_t* count = ar.take<int32_t>(ktot + 1);
  int64_t* mstart = ar.take<int64_t>(ktot + 1);
  int32_t* eflag = ar.take<int32_t>(ktot + 1);
  int32_t* erank = ar.take<int32_t>(ktot + 1);
  uint64_t* dk = ar.take<uint64_t>(n);
  uint64_t* dks = ar.take<uint64_t>(n);
  if (ar.overflow()) return fail(EAC_ERR_WORKSPACE, "kmeans update: workspace too small");
  if (d_koff) {
    k_off = const_cast<int64_t*>(d_koff);  // read only
  } else {
    EAC_HIP_CHECK(hipMemcpyAsync(k_off, h_koff, sizeof(int64_t) * (P + 1),
                                 hipMemcpyHostToDevice, s));
    EAC_HIP_CHECK(hipStreamSynchronize(s));
  }
  EAC_TRY(fill<int32_t>(count, ktot + 1, 0, s));
  {
    // about 8192 blocks in all, each over a run of at least 1024 points
    const int py = P < 65535 ? P : 65535;
    int gx = grid_for(n, 1024, 8192 / py > 1 ? 8192 / py : 1);
    k_label_keys<<<dim3(gx, py), 256, 0, s>>>(labels, n, P, k_off, key, idx, count);
    EAC_LAUNCH_CHECK();
  }
  // stable by centroid: members of each centroid in ascending point order
  const int bits = bits_for((uint64_t)ktot);
  size_t mark = ar.used;
  EAC_TRY((radix_sort<uint32_t, uint32_t>(key, skey, idx, members, items, 0, bits < 1 ? 1 : bits,
                                          ar, s)));
  ar.used = mark;
  EAC_TRY((scan_excl<int32_t, int64_t>(count, mstart, ktot, 0, nullptr, ar, s)));
  ar.used = mark;
  auto narrow = [&](auto kern, int lanes) {
    kern<<<grid_for(ktot, 256 / lanes, 65536), 256, 0, s>>>(X, d, members, mstart, count, ktot,
                                                           Cnew);
  };
  if (d <= 4) narrow(k_centroid_sums_narrow<4>, 4);
  else if (d <= 8) narrow(k_centroid_sums_narrow<8>, 8);
  else if (d <= 16) narrow(k_centroid_sums_narrow<16>, 16);
  else if (d <= 32) narrow(k_centroid_sums_narrow<32>, 32);
  else if (d <= 64) narrow(k_centroid_sums_narrow<64>, 64);
  else
    k_centroid_sums<<<grid_for(ktot, 1, 65536), kWave, 0, s>>>(X, d, members, mstart, count, ktot,
                                                                Cnew);
  EAC_LAUNCH_CHECK();
  // empty clusters
  k_empty_flags<<<grid_for(ktot, 256, 4096), 256, 0, s>>>(count, ktot, eflag);
  EAC_LAUNCH_CHECK();
  EAC_TRY((scan_excl<int32_t, int32_t>(eflag, erank, ktot + 1, 0, nullptr, ar, s)));
  ar.used = mark;
  if (g) {
    k_reseed_dev<<<P < 4096 ? P : 4096, 1024, 0, s>>>(X, d, dists, n, k_off, P, count, erank,
                                                      Cnew, g->err);
    EAC_LAUNCH_CHECK();
    k_fit_tally<<<1, 64, 0, s>>>(g->info, nullptr, erank, ktot, g->any);
    EAC_LAUNCH_CHECK();
    if (n_empty) *n_empty = 0;
    return EAC_OK;
  }
  std::vector<int32_t> er(ktot + 1);
  {
    // host needs only the per-partition empty counts
    std::vector<int32_t> tmp(ktot + 1);
    EAC_TRY(d2h(tmp.data(), erank, sizeof(int32_t) * (ktot + 1), s));
    er.swap(tmp);
  }
  int64_t total_empty = er[ktot];
  if (n_empty) *n_empty = total_empty;
  for (int p = 0; p < P && total_empty > 0; ++p) {
    int64_t e = er[h_koff[p + 1]] - er[h_koff[p]];
    if (e == 0) continue;
    if (e > n) return fail(EAC_ERR_ARG, "kmeans update: more empty clusters than points");
    k_dist_keys<<<grid_for(n, 256, 4096), 256, 0, s>>>(dists + (int64_t)p * n, n, dk);
    EAC_LAUNCH_CHECK();
    EAC_TRY((radix_sort<uint64_t, NoValue>(dk, dks, nullptr, nullptr, n, 0, 64, ar, s)));
    ar.used = mark;
    k_reseed<<<grid_for((h_koff[p + 1] - h_koff[p]) * d, 256, 4096), 256, 0, s>>>(
        X, d, dks, count, erank, h_koff[p], h_koff[p + 1], Cnew);
    EAC_LAUNCH_CHECK();
  }
  return EAC_OK;
}

}  // namespace eac

// ------------------------------------------------------------------ C ABI
using namespace eac;

extern "C" size_t eac_kmeans_workspace_bytes(int64_t n, int32_t d, int64_t k_total, int32_t P) {
  if (n < 1) n = 1;
  if (k_total < 1) k_total = 1;
  if (P < 1) P = 1;
  size_t a = kmeans_label_ws_bytes(n, d, k_total, P);
  size_t b = kmeans_update_ws_bytes(n, d, k_total, P);
  // fit: labels/dists are caller buffers; new centroids and the per-partition
  // state (inertia, previous inertia, offsets, active, iterations, update
  // flags, one flag), each rounded up to 256 bytes
  Arena fa(nullptr, 0, true);
  fa.take<float>((size_t)k_total * d);
  fa.take<float>(P);
  fa.take<float>(P);
  fa.take<int64_t>(P + 1);
  fa.take<int32_t>(P);
  fa.take<int32_t>(P);
  fa.take<int32_t>(P);
  fa.take<int32_t>(1);
  size_t c = fa.used + 4096;
  return (a > b ? a : b) + c;
}

static int km_check(const float* X, int64_t n, int32_t d, const float* C, const int64_t* h_koff,
                    int32_t P) {
  if (n < 1 || d < 1 || P < 1) return fail(EAC_ERR_ARG, "kmeans: need n, d, P >= 1");
  if (!X || !C || !h_koff) return fail(EAC_ERR_ARG, "kmeans: null array");
  if (h_koff[0] != 0) return fail(EAC_ERR_ARG, "kmeans: k_off[0] must be 0");
  for (int p = 0; p < P; ++p)
    if (h_koff[p + 1] <= h_koff[p]) return fail(EAC_ERR_ARG, "kmeans: partition %d has k < 1", p);
  if (n >= (1ll << 40) || P >= (1 << 23)) return fail(EAC_ERR_UNSUPPORTED, "kmeans: n or P too large");
  return EAC_OK;
}

extern "C" int eac_kmeans_label(const float* X, int64_t n, int32_t d, const float* C,
                                const int64_t* h_koff, int32_t P, int32_t mode, int32_t* labels,
                                float* dists, int64_t* h_out, void* ws, size_t ws_bytes,
                                void* stream) {
  EAC_TRY(km_check(X, n, d, C, h_koff, P));
  if (!labels) return fail(EAC_ERR_ARG, "kmeans: null labels");
  size_t need = eac_kmeans_workspace_bytes(n, d, h_koff[P], P);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "kmeans: workspace %zu < %zu bytes", ws_bytes, need);
  Arena ar(ws, ws_bytes);
  hipStream_t s = (hipStream_t)stream;
  int64_t nres = 0;
  EAC_TRY(kmeans_label(X, n, d, C, h_koff, P, mode, labels, dists, nullptr, &nres, ar, s));
  EAC_HIP_CHECK(hipStreamSynchronize(s));
  prof_flush();
  if (h_out) h_out[0] = nres;
  return EAC_OK;
}

extern "C" int eac_kmeans_update(const float* X, int64_t n, int32_t d, const int32_t* labels,
                                 const float* dists, const int64_t* h_koff, int32_t P,
                                 float* C_new, int64_t* h_out, void* ws, size_t ws_bytes,
                                 void* stream) {
  if (!labels || !dists || !C_new) return fail(EAC_ERR_ARG, "kmeans update: null array");
  EAC_TRY(km_check(X, n, d, C_new, h_koff, P));
  size_t need = eac_kmeans_workspace_bytes(n, d, h_koff[P], P);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "kmeans update: workspace %zu < %zu bytes", ws_bytes, need);
  Arena ar(ws, ws_bytes);
  hipStream_t s = (hipStream_t)stream;
  int64_t ne = 0;
  EAC_TRY(kmeans_update(X, n, d, labels, dists, h_koff, P, C_new, &ne, ar, s));
  EAC_HIP_CHECK(hipStreamSynchronize(s));
  if (h_out) h_out[0] = ne;
  return EAC_OK;
}

// ------------------------------------------------- capturable batched fit
namespace eac {
namespace {
// KmGraph buffers (flag count, per-partition counts and cursors, flags
// sorted by partition); info and err are the caller's device arrays.
KmGraph carve_graph(Arena& ar, int64_t n, int P, int64_t* info, unsigned int* err) {
  KmGraph g;
  g.d_nf = ar.take<int64_t>(1);
  g.bcount = ar.take<int32_t>((size_t)P * kSortBlocks);
  g.sorted = ar.take<uint64_t>((size_t)P * n);
  g.info = info;
  g.err = err;
  g.any = nullptr;
  return g;
}
__global__ void k_graph_init(int64_t* info, unsigned int* err) {
  if (threadIdx.x == 0) {
    info[0] = 0;
    info[1] = 0;
    err[0] = 0u;
  }
}
}  // namespace
}  // namespace eac

extern "C" size_t eac_kmeans_fit_graph_workspace_bytes(int64_t n, int32_t d, int64_t k_total,
                                                       int32_t P) {
  if (n < 1) n = 1;
  if (P < 1) P = 1;
  Arena ga(nullptr, 0, true);
  (void)carve_graph(ga, n, P, nullptr, nullptr);
  return eac_kmeans_workspace_bytes(n, d, k_total, P) + ga.used + 4096;
}

// eac_kmeans_fit for capture in a HIP graph: no host read, no host copy, no
// synchronisation. d_koff: the partition offsets already on the device;
// exactly max_iter labelings run (a partition that meets the stopping rule
// keeps its centroids, so its later labelings repeat, and its iteration
// count in d_iters stops); d_info (device int64[2]) receives the flags
// re-scored and the clusters reseeded, d_err (device uint32) error bits
// (1: flag list overflow, 2: more than 2048 empty clusters in one
// partition): the labels are valid when d_err is 0 after the replay.
extern "C" int eac_kmeans_fit_graph(const float* X, int64_t n, int32_t d, float* C_inout,
                                    const int64_t* h_koff, const int64_t* d_koff, int32_t P,
                                    int32_t max_iter, double tol, int32_t mode, int32_t* labels,
                                    float* dists, int32_t* d_iters, int64_t* d_info,
                                    uint32_t* d_err, void* ws, size_t ws_bytes, void* stream) {
  EAC_TRY(km_check(X, n, d, C_inout, h_koff, P));
  if (!labels || !dists || !d_koff || !d_iters || !d_info || !d_err)
    return fail(EAC_ERR_ARG, "kmeans fit graph: null array");
  if (max_iter < 1) return fail(EAC_ERR_ARG, "kmeans fit graph: max_iter < 1");
  const int64_t ktot = h_koff[P];
  size_t need = eac_kmeans_fit_graph_workspace_bytes(n, d, ktot, P);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "kmeans fit graph: workspace %zu < %zu bytes", ws_bytes,
                need);
  hipStream_t s = (hipStream_t)stream;
  Arena ar(ws, ws_bytes);
  KmGraph g = carve_graph(ar, n, P, d_info, d_err);
  float* Cnew = ar.take<float>((size_t)ktot * d);
  float* inertia = ar.take<float>(P);
  float* prev = ar.take<float>(P);
  int32_t* active = ar.take<int32_t>(P);
  int32_t* upd = ar.take<int32_t>(P);
  int32_t* d_any = ar.take<int32_t>(1);
  if (ar.overflow()) return fail(EAC_ERR_WORKSPACE, "kmeans fit graph: workspace too small");
  const size_t mark = ar.used;
  g.any = d_any;
  k_graph_init<<<1, 64, 0, s>>>(d_info, d_err);
  EAC_LAUNCH_CHECK();
  EAC_TRY(fill<int32_t>(d_any, 1, 1, s));  // the first labeling always counts
  k_fit_init<<<grid_for(P, 256, 1024), 256, 0, s>>>(prev, active, d_iters, P);
  EAC_LAUNCH_CHECK();
  for (int it = 1; it <= max_iter; ++it) {
    ar.used = mark;
    int64_t nres = 0;
    EAC_TRY(kmeans_label(X, n, d, C_inout, h_koff, P, mode, labels, dists, inertia, &nres, ar,
                         s, d_koff, &g));
    k_converge<<<1, 1024, 0, s>>>(inertia, prev, active, d_iters, upd, P, it, max_iter,
                                  (float)tol, d_any);
    EAC_LAUNCH_CHECK();
    if (it >= max_iter) break;
    ar.used = mark;
    int64_t ne = 0;
    EAC_TRY(kmeans_update(X, n, d, labels, dists, h_koff, P, Cnew, &ne, ar, s, d_koff, &g));
    const dim3 gc(grid_for(ktot * d / P + 1, 256, 64), P < 65535 ? P : 65535);
    k_copy_running<<<gc, 256, 0, s>>>(C_inout, Cnew, d, d_koff, P, upd);
    EAC_LAUNCH_CHECK();
  }
  return EAC_OK;
}

extern "C" size_t eac_sum_f32_workspace_bytes(int64_t n, int64_t rows) {
  if (n < 0 || rows < 0) return 0;
  return ((size_t)rows * (size_t)np_chunks(n) + 1) * sizeof(float) + 256;
}

// numpy's float32 x[r].sum() of every row (KMeans.py:164 inertia).
extern "C" int eac_sum_f32_rows(const float* x, int64_t n, int64_t rows, float* out, void* ws,
                                size_t ws_bytes, void* stream) {
  if (n < 0 || rows < 0) return fail(EAC_ERR_ARG, "sum_f32: negative size");
  if (rows == 0) return EAC_OK;
  if (!out || (n > 0 && !x)) return fail(EAC_ERR_ARG, "sum_f32: null array");
  if (n >= (1ll << 40)) return fail(EAC_ERR_UNSUPPORTED, "sum_f32: n too large");
  const size_t need = eac_sum_f32_workspace_bytes(n, rows);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "sum_f32: workspace %zu < %zu bytes", ws_bytes, need);
  EAC_TRY(np_sum_rows(x, n, rows, static_cast<float*>(ws), out, (hipStream_t)stream));
  return EAC_OK;
}

// KMeans.py:117-189 per partition: label, count the iteration, stop when
// |sum(dists) - previous| <= tol or iters >= max_iter, otherwise recompute the
// centroids. Partitions stop independently; a stopped partition keeps its
// labels and centroids. labels/dists: [P][n]; C_inout: initial centroids in,
// final centroids out. h_iters[P] (host): iterations run per partition.
// h_out (host): [0] points re-scored in FP64 (all iterations), [1] empty
// clusters re-seeded (all iterations).
extern "C" int eac_kmeans_fit(const float* X, int64_t n, int32_t d, float* C_inout,
                              const int64_t* h_koff, int32_t P, int32_t max_iter, double tol,
                              int32_t mode, int32_t* labels, float* dists, int32_t* h_iters,
                              int64_t* h_out, void* ws, size_t ws_bytes, void* stream) {
  EAC_TRY(km_check(X, n, d, C_inout, h_koff, P));
  if (!labels || !dists) return fail(EAC_ERR_ARG, "kmeans fit: null labels/dists");
  if (max_iter < 1) return fail(EAC_ERR_ARG, "kmeans fit: max_iter < 1");
  const int64_t ktot = h_koff[P];
  size_t need = eac_kmeans_workspace_bytes(n, d, ktot, P);
  if (!ws || ws_bytes < need)
    return fail(EAC_ERR_WORKSPACE, "kmeans fit: workspace %zu < %zu bytes", ws_bytes, need);
  hipStream_t s = (hipStream_t)stream;
  Arena ar(ws, ws_bytes);
  float* Cnew = ar.take<float>((size_t)ktot * d);
  float* inertia = ar.take<float>(P);
  float* prev = ar.take<float>(P);
  int64_t* d_koff = ar.take<int64_t>(P + 1);
  int32_t* active = ar.take<int32_t>(P);
  int32_t* d_iters = ar.take<int32_t>(P);
  int32_t* upd = ar.take<int32_t>(P);
  int32_t* d_any = ar.take<int32_t>(1);
  if (ar.overflow()) return fail(EAC_ERR_WORKSPACE, "kmeans fit: workspace too small");
  const size_t mark = ar.used;
  // the partition offsets go up once; the stopping rule runs on the device,
  // so an iteration costs one read-back (does any partition continue?)
  EAC_TRY(h2d(d_koff, h_koff, sizeof(int64_t) * (P + 1), s));
  k_fit_init<<<grid_for(P, 256, 1024), 256, 0, s>>>(prev, active, d_iters, P);
  EAC_LAUNCH_CHECK();
  int64_t tot_res = 0, tot_empty = 0;
  for (int it = 1; it <= max_iter; ++it) {
    // label every partition (stopped ones relabel identically: unchanged
    // centroids, deterministic kernels)
    int64_t nres = 0;
    ar.used = mark;
    EAC_TRY(kmeans_label(X, n, d, C_inout, h_koff, P, mode, labels, dists, inertia, &nres, ar,
                         s, d_koff));
    tot_res += nres;
    k_converge<<<1, 1024, 0, s>>>(inertia, prev, active, d_iters, upd, P, it, max_iter,
                                  (float)tol, d_any);
    EAC_LAUNCH_CHECK();
    if (it >= max_iter) break;  // every partition stops at max_iter
    int32_t any = 0;
    EAC_TRY(d2h(&any, d_any, sizeof(any), s));
    if (!any) break;
    ar.used = mark;
    int64_t ne = 0;
    EAC_TRY(kmeans_update(X, n, d, labels, dists, h_koff, P, Cnew, &ne, ar, s, d_koff));
    tot_empty += ne;
    const dim3 gc(grid_for(ktot * d / P + 1, 256, 64), P < 65535 ? P : 65535);
    k_copy_running<<<gc, 256, 0, s>>>(C_inout, Cnew, d, d_koff, P, upd);
    EAC_LAUNCH_CHECK();
  }
  if (h_iters) EAC_TRY(d2h(h_iters, d_iters, sizeof(int32_t) * P, s));
  EAC_HIP_CHECK(hipStreamSynchronize(s));
  prof_flush();
  if (h_out) { h_out[0] = tot_res; h_out[1] = tot_empty; }
  return EAC_OK;
}

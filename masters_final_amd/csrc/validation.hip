// Contingency table of two partitions: the shared-sample counts behind the
// reference's clustering accuracy indices.
//
// Reference: MyML/MyML/validation/consistency.py:20-130 (ConsistencyIndex:
// A = clusts1[i] . clusts2[j], B and C the cluster sizes, over the binary
// cluster-by-sample matrices of MyML/MyML/utils/partition.py:17-70) and
// MyML/MyML/validation/hindex.py:8-86 (HungarianIndex: match_matrix =
// clusts1 . clusts2^T). For partitions (every sample in exactly one cluster)
// both products are the contingency table T[i][j] = #{s : a[s] = i, b[s] = j},
// which one pass over the two label arrays counts exactly. The matching on
// top of the table (greedy Jaccard, Hungarian) is host work on ka x kb
// numbers.
#include "prims.hpp"

namespace eac {
namespace {

constexpr int kCtLds = 8192;  // table entries counted in LDS per block (32 KiB)

// Counts in LDS (tables of at most kCtLds entries) or straight in global
// memory; labels outside [a0, a0 + ka) x [b0, b0 + kb) are counted in
// out[ka * kb] so the caller can reject them.
template <bool LDS>
__global__ __launch_bounds__(256) void k_contingency(const int32_t* __restrict__ a,
                                                     const int32_t* __restrict__ b, int64_t n,
                                                     int32_t a0, int32_t ka, int32_t b0,
                                                     int32_t kb,
                                                     unsigned long long* __restrict__ out) {
  __shared__ unsigned int s_t[LDS ? kCtLds : 1];
  const int64_t cells = (int64_t)ka * kb;
  if (LDS) {
    for (int v = threadIdx.x; v < cells; v += blockDim.x) s_t[v] = 0u;
    __syncthreads();
  }
  unsigned int bad = 0;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int32_t i = a[s] - a0, j = b[s] - b0;
    if (i < 0 || i >= ka || j < 0 || j >= kb) {
      ++bad;
      continue;
    }
    const int64_t c = (int64_t)i * kb + j;
    if (LDS)
      atomicAdd(&s_t[c], 1u);
    else
      atomicAdd(&out[c], 1ull);
  }
  bad = wave_sum(bad);
  if (lane_id() == 0 && bad) atomicAdd(&out[cells], (unsigned long long)bad);
  if (LDS) {
    __syncthreads();
    for (int v = threadIdx.x; v < cells; v += blockDim.x)
      if (s_t[v]) atomicAdd(&out[v], (unsigned long long)s_t[v]);
  }
}

}  // namespace
}  // namespace eac

using eac::fail;

extern "C" int eac_contingency(const int32_t* a, const int32_t* b, int64_t n, int32_t a0,
                               int32_t ka, int32_t b0, int32_t kb, int64_t* out,
                               void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n < 0 || ka < 1 || kb < 1)
    return fail(EAC_ERR_ARG, "contingency: need n >= 0, ka, kb >= 1");
  if ((int64_t)ka * kb > ((int64_t)1 << 31))
    return fail(EAC_ERR_UNSUPPORTED, "contingency: table too large");
  if (!out || (n > 0 && (!a || !b))) return fail(EAC_ERR_ARG, "contingency: null array");
  const int64_t cells = (int64_t)ka * kb;
  EAC_TRY(eac::fill<int64_t>(out, cells + 1, 0, s));
  if (n == 0) return EAC_OK;
  unsigned long long* o = reinterpret_cast<unsigned long long*>(out);
  if (cells <= eac::kCtLds) {
    // 16 samples per thread, at most 512 blocks: enough blocks to fill the
    // chip, few enough table flushes (one atomic per nonzero cell per block)
    const int g = eac::grid_for(n, 256 * 16, 512);
    eac::k_contingency<true><<<g, 256, 0, s>>>(a, b, n, a0, ka, b0, kb, o);
  } else {
    eac::k_contingency<false><<<eac::grid_for(n, 256, 4096), 256, 0, s>>>(a, b, n, a0, ka, b0,
                                                                         kb, o);
  }
  EAC_LAUNCH_CHECK();
  return EAC_OK;
}

// Binary cluster-by-sample matrix of a partition.
//
// Reference: MyML/MyML/utils/partition.py:17-35 convertIndexToBin (one row
// per cluster of a list-of-index-arrays partition, N = total listed
// indices, out[i, clust] = 1) and :37-70 convertClusterStringToBin (one row
// per label of a label array, shifted by -1 when the labels start at 1,
// out[label + offset, sample] = 1). Both are a clear of a k x N uint8 matrix
// and a scatter of ones; the host shims work out rows and columns (and the
// reference's ValueError on gapped labels) and call one entry point.
//
// MI355X design: the clear writes 16 bytes per lane (the matrix is the only
// large traffic, k * N bytes); the scatter is one byte store per listed
// entry. Entries outside [0, k) x [0, N) are skipped and counted, so the
// caller can raise numpy's IndexError without the kernel touching memory
// out of bounds.
#include "prims.hpp"

namespace eac {
namespace {

__global__ __launch_bounds__(256) void k_clear16(uint8_t* __restrict__ out, int64_t bytes) {
  const int64_t vec = bytes >> 4;
  uint4* o4 = reinterpret_cast<uint4*>(out);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < vec;
       t += (int64_t)gridDim.x * blockDim.x)
    o4[t] = make_uint4(0u, 0u, 0u, 0u);
  const int64_t tail = (vec << 4) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tail < bytes && tail < (vec << 4) + 16) out[tail] = 0;
}

__global__ __launch_bounds__(256) void k_scatter_ones(const int32_t* __restrict__ rows,
                                                      const int64_t* __restrict__ cols,
                                                      int64_t m, int32_t k, int64_t N,
                                                      uint8_t* __restrict__ out,
                                                      unsigned long long* __restrict__ bad) {
  unsigned int nbad = 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = rows[t];
    const int64_t c = cols ? cols[t] : t;
    if (r < 0 || r >= k || c < 0 || c >= N) {
      ++nbad;
      continue;
    }
    out[(int64_t)r * N + c] = 1;
  }
  nbad = wave_sum(nbad);
  if (bad && lane_id() == 0 && nbad) atomicAdd(bad, (unsigned long long)nbad);
}

}  // namespace
}  // namespace eac

using eac::fail;

extern "C" int eac_partition_to_bin(const int32_t* rows, const int64_t* cols, int64_t m,
                                    int32_t k, int64_t N, uint8_t* out, int64_t* d_bad,
                                    void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (m < 0 || k < 0 || N < 0) return fail(EAC_ERR_ARG, "partition_to_bin: negative size");
  if ((int64_t)k * N > 0 && !out) return fail(EAC_ERR_ARG, "partition_to_bin: null output");
  if (m > 0 && !rows) return fail(EAC_ERR_ARG, "partition_to_bin: null rows");
  if (!cols && m > N) return fail(EAC_ERR_ARG, "partition_to_bin: m > N without columns");
  if (reinterpret_cast<uintptr_t>(out) & 15)
    return fail(EAC_ERR_ARG, "partition_to_bin: output must be 16-byte aligned");
  if (d_bad) EAC_TRY(eac::fill<int64_t>(d_bad, 1, 0, s));
  const int64_t bytes = (int64_t)k * N;
  if (bytes > 0) {
    eac::k_clear16<<<eac::grid_for((bytes >> 4) + 1, 256, 8192), 256, 0, s>>>(out, bytes);
    EAC_LAUNCH_CHECK();
  }
  if (m > 0) {
    eac::k_scatter_ones<<<eac::grid_for(m, 256, 8192), 256, 0, s>>>(
        rows, cols, m, k, N, out, reinterpret_cast<unsigned long long*>(d_bad));
    EAC_LAUNCH_CHECK();
  }
  return EAC_OK;
}

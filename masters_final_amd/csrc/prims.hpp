// Device-wide primitives shared by the graph, co-association and K-Means
// entry points (defined in scan.hip).
#pragma once
#include "common.hpp"

namespace eac {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

// Workspace bytes needed by scan_excl for n elements of elem_bytes each.
size_t scan_ws_bytes(int64_t n, int elem_bytes);

// Exclusive scan: out[i] = init + sum(in[0:i]); d_total (nullable) gets
// init + sum(in). in == out is allowed.
template <typename Tin, typename Tout>
int scan_excl(const Tin* in, Tout* out, int64_t n, Tout init, Tout* d_total,
              Arena& ar, hipStream_t s);

// Fill kernel launcher (used instead of hipMemsetAsync so that callers may
// capture the sequence in a graph).
template <typename T>
int fill(T* p, int64_t n, T value, hipStream_t s);

}  // namespace eac

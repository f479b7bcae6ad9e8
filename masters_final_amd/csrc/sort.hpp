// Stable LSD radix sort (defined in sort.hip).
#pragma once
#include <type_traits>

#include "common.hpp"

namespace eac {

struct NoValue {};

size_t radix_sort_ws_bytes(int64_t n, int key_bytes, int val_bytes);

// Sorts keys (and values, unless V is NoValue) by bits [begin_bit, end_bit),
// stably. Inputs are not modified; kout/vout may not alias kin/vin.
template <typename K, typename V>
int radix_sort(const K* kin, K* kout, const V* vin, V* vout, int64_t n, int begin_bit,
               int end_bit, Arena& ar, hipStream_t s);

}  // namespace eac

// qv_sort.h -- the radix sort configuration of the tree build and the probe
// ordering.
#pragma once
#include <rocprim/device/device_radix_sort.hpp>

namespace qv {

// rocPRIM's radix sort switches from its merge-sort path to onesweep only
// above 2^20 items; for keys of a few bytes' width onesweep (one pass per 8
// key bits) is faster from 2^16 items on: the 21-bit Hilbert keys of the
// walk order take 3 passes.  The 64-bit Morton sort keeps the default
// (8 onesweep passes cost more than its merge path)
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, (1u << 16)>;

}  // namespace qv

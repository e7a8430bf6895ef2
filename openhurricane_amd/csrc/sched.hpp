// sched.hpp -- longest-first cell order for the persistent integrators.
//
// The integrators write each cell's Newton-iteration count into cost[]; the
// next call on the same cells takes them most expensive first, so that the
// long integrations start early and the short ones fill the tail.  The order
// is a bucketed counting sort by cost (exact below 256 iterations, 32
// buckets per octave above), built from the library's own kernels: no
// allocation, no host synchronisation and no memset nodes, so a call can be
// captured into a HIP graph.  Any order gives the same results.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ohcs {

constexpr int kBins = 1024;

// order[0..n) = the cells 0..n-1 by descending cost bucket (ties in no fixed
// order).  bins: 2*kBins unsigned ints of device scratch (counts, offsets).
hipError_t order_by_cost(const int *cost, int *order, unsigned *bins, int64_t n,
                         hipStream_t s);

}  // namespace ohcs

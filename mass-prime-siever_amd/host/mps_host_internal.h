// Host helpers shared by the plain worker (mps_host.cpp) and the batched worker (mps_batch.hip).
#pragma once

#include <vector>

#include "mps_gpu.h"
#include "mps_host.h"

namespace mps {

// Turns the device records of chunk [lo, hi) into relations (p, M, q): q = ord_p(b) for the
// device-decided records, the exact is_prime64 + run_bsgs for the rest.  Sorts recs in place.
int finish_records(const mps_work_item* it, uint64_t lo, uint64_t hi, mps_dev_relation* recs, size_t n,
                   std::vector<mps_relation>& rels, uint64_t* prime_flagged);

}  // namespace mps

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

// finish_records on up to `threads` threads: the sorted records are cut into ranges at changes of
// the number, each range is finished on a thread of its own, and the relations are joined in
// range order, so the result is finish_records' result.  Below `min_records` it runs on the
// calling thread.
int finish_records_parallel(const mps_work_item* it, uint64_t lo, uint64_t hi, mps_dev_relation* recs, size_t n,
                            std::vector<mps_relation>& rels, uint64_t* prime_flagged, int threads,
                            size_t min_records);

// The CPUs this process may run on (at least 1).
int host_cpus();

}  // namespace mps

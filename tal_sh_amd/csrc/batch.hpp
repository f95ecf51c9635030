// Batched contraction: kernel launcher (work items in batch_item.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include "batch_item.hpp"

namespace tal {

// Workgroups of the persistent kernel resident on one CU at once (occupancy).
int batch_blocks_per_cu(int dk, int *blocks);

// Bytes of the work-queue counters of `nq` queues.
size_t batch_counter_bytes(int nq);

// Runs the items of a batch in one launch of `nblocks` workgroups on stream
// s, after a one-block kernel that clears the queue counters (both launches
// are capturable; no memset): queue q holds items [qstart[q], qstart[q + 1]),
// workgroup b starts on queue b % nq.  dk = 4 (R4) or 8 (R8).  counters:
// device memory of batch_counter_bytes(nq) bytes, nq <= 256.
int batch_launch(int dk, const BatchItem *items, const int *qstart, unsigned *counters, int nq, int nblocks,
                 hipStream_t s);

}  // namespace tal

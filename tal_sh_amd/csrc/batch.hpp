// Batched contraction: kernel launcher (work items in batch_item.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include "batch_item.hpp"

namespace tal {

// Waves (independent workers) per workgroup of the persistent kernel.
int batch_workers_per_block();

// Workgroups of the persistent kernel resident on one CU at once (occupancy).
int batch_blocks_per_cu(int dk, int *blocks);

// Runs the items of a batch in one launch of `nblocks` workgroups on stream
// s (capturable): worker w = block * batch_workers_per_block() + wave runs
// items [start[w], start[w + 1]).  dk = 4 (R4) or 8 (R8).
int batch_launch(int dk, const BatchItem *items, const int *start, int nblocks, hipStream_t s);

}  // namespace tal

// Batched contraction: device task descriptor and kernel launcher.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tal {

// One contraction of a batch as the persistent kernel sees it.  Tables at
// tabs[tab ...]: lM[M] lK[K] rK[K] rN[N] dM[M] dN[N] (element offsets).
struct BatchTask {
  const void *L;
  const void *R;
  void *D;
  int M, N, K;
  int tm, tn;            // tile shape (multiples of 16, at most 64)
  int tiles_m, tiles_n;
  int beta_one;
  int amode, bmode;      // thread layout of the A / B gathers: 0 along m / n, 1 along k
  long long tab;
  double alpha;
};

// Clears the work counter and runs all tasks in one persistent launch of
// `grid` workgroups on stream s (both launches are capturable).  Work item
// `it` is tile it - prefix[t] of task t = item_task[it].
int batch_launch(int dk, const BatchTask *tasks, const long long *prefix, const int *item_task, const int *tabs,
                 unsigned long long *counter, long long total, int grid, hipStream_t s);

}  // namespace tal

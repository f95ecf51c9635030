// Internal runtime of libtalsh_amd: devices, buffer pools, tensor images,
// hazard tracking and HIP tasks.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <deque>
#include <vector>

#include "kernels.hpp"
#include "pool.hpp"
#include "talsh.h"

namespace tal {

// ------------------------------------------------------------------ events
struct EvHandle {
  int dev = -1;
  hipEvent_t ev = nullptr;
  ~EvHandle();
};
using EvRef = std::shared_ptr<EvHandle>;

// Per-buffer ordering state: the event that completes the last write and the
// events that complete reads issued since then.  Any stream that touches the
// buffer waits on what it conflicts with, so tasks on different streams of
// one device keep program order on shared tensors.
struct Hazard {
  EvRef writer;
  std::vector<EvRef> readers;
};

// ----------------------------------------------------------------- devices
struct Device {
  int hip_id = -1;
  bool on = false;
  void *base = nullptr;
  size_t size = 0;
  size_t mem_total = 0;
  Arena arena;
  std::vector<hipStream_t> streams;
  // per stream: (completion event, weight) of tasks not yet seen complete, oldest first
  std::vector<std::deque<std::pair<EvRef, double>>> backlog;
  std::vector<hipEvent_t> free_events;
  // statistics (talsh_stats_t of the reference, device_algebra.h:183-193)
  long long tasks_submitted = 0, tasks_completed = 0, tasks_deferred = 0, tasks_failed = 0;
  double flops = 0, traffic_in = 0, traffic_out = 0, time_start = 0;
  int active = 0;
};

struct Runtime {
  bool on = false;
  std::vector<Device> gpus;
  void *host_base = nullptr;
  size_t host_size = 0;
  Arena host_arena;
  int r4mode = R4_FP32;
  int split_method = SPLIT_BF16X6;  // talshAmdSetFastMathSplit
  size_t dev_buffer_request = 0;  // 0: take a fraction of the free memory
  int nstreams = 4;
  size_t xl_mem = 0;  // device bytes a talshTensorContractXL call may use (0: from the free pool)
  int host_policy = 1, host_fallback = 1;  // talshSetMemAllocPolicyHost (MEM_ALLOC_TMP_BUF, fallback on)
  bool log_mem = false, log_ops = false;  // talshMemManagerLogStart / talshTensorOpLogStart
  std::unordered_map<const void *, Hazard> hazards;
  // bumped by every change of the hazard table: a caller that saw the same
  // value before knows every entry is still what it left there
  uint64_t hazard_epoch = 0;
  std::recursive_mutex mu;
};

Runtime &rt();

// ----------------------------------------------------------------- tasks
struct Task {
  int dev = -1;  // GPU index
  hipStream_t stream = nullptr;
  EvRef ev_start, ev_in, ev_comp, ev_done, ev_mm0, ev_mm1;
  int status = TALSH_TASK_EMPTY;
  int error = 0;
  double flops = 0, bytes = 0;
  std::vector<std::function<void()>> on_finalize;  // run once the device work is complete
};

// ----------------------------------------------------------------- helpers
double wall_time();
int dk_size(int dk);
size_t shape_volume(const talsh_tens_shape_t *s);
int canonical_dev(int flat_id);          // normalises kind-3 GPU ids to kind-1; -1 if invalid
int gpu_index(int flat_id);              // GPU index of a (canonical or alias) flat id, -1 if not a GPU
int flat_gpu(int g);                     // canonical flat id of GPU g

EvRef new_event(int g);
void hazard_wait_read(const void *p, hipStream_t s);
void hazard_wait_write(const void *p, hipStream_t s);
void hazard_add_reader(const void *p, const EvRef &e);
void hazard_set_writer(const void *p, const EvRef &e);
void hazard_drain(const void *p);        // blocks until all work on p is complete, forgets p

// Device memory (buffer pool) and host memory (pinned pool, malloc fallback).
int dev_alloc(int g, size_t bytes, void **p, int *handle);
void dev_free(int g, void *p, int handle);
int host_alloc(size_t bytes, void **p, int *handle);
void host_free(void *p, int handle);

// Tensor images.
int image_add(talsh_tens_t *t, int dev_id, void *p, int buf_entry, int attached, int dk);
void image_release(talsh_dev_rsc_t *r);  // frees the body (unless attached)
int image_discard(talsh_tens_t *t, int i);
int image_discard_other(talsh_tens_t *t, int keep);
int image_find(const talsh_tens_t *t, int flat_dev, int dk, bool need_avail);
int tensor_healthy(const talsh_tens_t *t);

// Task lifecycle.
// New task on the stream of GPU g with the least outstanding work; `weight`
// estimates the task's own work in host-link-equivalent bytes.
Task *task_new(int g, double weight = 0.0);
void task_free(Task *t);
int task_finalize_if_done(talsh_task_t *tt, bool block);

}  // namespace tal

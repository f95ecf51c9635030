// Internal runtime of libtalsh_amd: devices, buffer pools, tensor images,
// hazard tracking and HIP tasks.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <utility>
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

// Ordering state of one byte range [start, start + len): the event that
// completes the last write and the events that complete reads issued since
// then.  Any stream that touches memory overlapping the range waits on what it
// conflicts with, so tasks on different streams of one device keep program
// order on shared tensors, also where two views (external bodies, slices of
// one allocation) share bytes without sharing a base address.
struct Hazard {
  EvRef writer;
  std::vector<EvRef> readers;
};
// Ranges are filed by length class c = floor(log2(len)): a range of class c
// that overlaps [p, p + n) starts in (p - 2^(c+1), p + n), so an overlap query
// visits one short key interval per non-empty class.
using HazardKey = std::pair<uintptr_t, size_t>;  // (start, length)
constexpr int kHazardClasses = 64;

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
  std::map<HazardKey, Hazard> hazards[kHazardClasses];
  // unfinalized tasks that name each tensor as an operand: destructing the
  // tensor first completes them (their coherence work still needs it)
  std::unordered_map<const talsh_tens_t *, std::vector<struct Task *>> tens_tasks;
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
  EvRef ev_k0, ev_k1;   // dispatch timestamps of the first / last kernel (KernelClock)
  bool kclock = false;  // ev_k0 / ev_k1 were recorded
  int status = TALSH_TASK_EMPTY;
  int error = 0;  // HIP status of a failed task
  double flops = 0, bytes = 0, exec_time = 0;
  std::vector<std::function<void()>> on_finalize;  // run once the device work is complete
  std::vector<talsh_tens_t *> tens;               // operand tensors the finalizers touch
};

// ----------------------------------------------------------------- helpers
double wall_time();
int dk_size(int dk);
size_t shape_volume(const talsh_tens_shape_t *s);
int canonical_dev(int flat_id);          // normalises kind-3 GPU ids to kind-1; -1 if invalid
int gpu_index(int flat_id);              // GPU index of a (canonical or alias) flat id, -1 if not a GPU
int flat_gpu(int g);                     // canonical flat id of GPU g

EvRef new_event(int g);
// Ordering on the byte range [p, p + n): stream s waits for earlier writes to
// overlapping memory (read) or for earlier reads and writes (write); the
// task's completion event is then filed as a reader or as the new writer.
void hazard_wait_read(const void *p, size_t n, hipStream_t s);
void hazard_wait_write(const void *p, size_t n, hipStream_t s);
void hazard_add_reader(const void *p, size_t n, const EvRef &e);
void hazard_set_writer(const void *p, size_t n, const EvRef &e);
void hazard_drain(const void *p, size_t n);  // blocks until all work on overlapping memory is done, forgets it
size_t image_bytes(const talsh_tens_t *t, int i);  // bytes of image i's body

// Device memory (buffer pool) and host memory (pinned pool, malloc fallback).
int dev_alloc(int g, size_t bytes, void **p, int *handle);
void dev_free(int g, void *p, int handle);
int host_alloc(size_t bytes, void **p, int *handle);
void host_free(void *p, int handle, size_t bytes);

// Tensor images.
int image_add(talsh_tens_t *t, int dev_id, void *p, int buf_entry, int attached, int dk);
void image_release(talsh_dev_rsc_t *r, size_t bytes);  // frees the body (unless attached)
int image_discard(talsh_tens_t *t, int i);
int image_discard_other(talsh_tens_t *t, int keep);
int image_find(const talsh_tens_t *t, int flat_dev, int dk, bool need_avail);
int tensor_healthy(const talsh_tens_t *t);

// Task lifecycle.
// New task on the stream of GPU g with the least outstanding work; `weight`
// estimates the task's own work in host-link-equivalent bytes.
Task *task_new(int g, double weight = 0.0);
void task_free(Task *t);
// Files t as an unfinalized task on each distinct tensor of `tens`.
void task_register(Task *t, const std::vector<talsh_tens_t *> &tens);
// Completion work of a task whose device work is done (block: wait for it):
// runs its finalizers once; returns the task status.
int task_finalize(Task *t, bool block);
int task_finalize_if_done(talsh_task_t *tt, bool block);

}  // namespace tal

// Runtime core of libtalsh_amd: initialisation, devices, pools, hazards,
// tensor images, tasks, and the control/tensor/task parts of the C API.
//
// Reference counterparts: talshInit/talshShutdown (talshc.cpp:546-645) with
// init_gpus/free_gpus (tensor_algebra_gpu_nvidia.cu:2559-2726) and
// arg_buf_allocate (mem_manager.cpp:175-307); tensor construction and image
// bookkeeping (talshc.cpp:1135-1700, :245-318); the task API
// (talshc.cpp:1934-2597) over cuda_task_* (tensor_algebra_gpu_nvidia.cu:
// 3303-4036).
//
// MI355X design: one device buffer pool per GPU carved out of a single
// hipMalloc (best-fit, coalescing, no 1/6-of-buffer cap), a small set of
// default-priority streams per GPU, and per-buffer hazard tracking with HIP
// events instead of the reference's "wait for the previous task" chain
// (LastTask, tensor_algebra_gpu_nvidia.cu:6287-6295), so independent tasks
// overlap while dependent ones stay ordered.
#include "runtime.hpp"
#include "contract.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>

namespace tal {

Runtime &rt() {
  static Runtime r;
  return r;
}

double wall_time() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int dk_size(int dk) {
  switch (dk) {
    case NO_TYPE: return 0;
    case R4: return 4;
    case R8: return 8;
    case C4: return 8;
    case C8: return 16;
    default: return -1;
  }
}

size_t shape_volume(const talsh_tens_shape_t *s) {
  if (!s || s->num_dim < 0) return 0;
  size_t v = 1;
  for (int i = 0; i < s->num_dim; ++i) v *= (size_t)s->dims[i];
  return v;
}

int canonical_dev(int id) {
  if (id < 0 || id >= DEV_MAX) return -1;
  if (id >= 1 + MAX_GPUS_PER_NODE + MAX_MICS_PER_NODE) return id - MAX_GPUS_PER_NODE - MAX_MICS_PER_NODE;
  return id;
}

int gpu_index(int id) {
  const int c = canonical_dev(id);
  if (c < 1 || c > MAX_GPUS_PER_NODE) return -1;
  const int g = c - 1;
  if (g >= (int)rt().gpus.size() || !rt().gpus[g].on) return -1;
  return g;
}

int flat_gpu(int g) { return 1 + g; }

// ------------------------------------------------------------------ events
EvHandle::~EvHandle() {
  if (!ev) return;
  Runtime &r = rt();
  if (dev >= 0 && dev < (int)r.gpus.size() && r.gpus[dev].on)
    r.gpus[dev].free_events.push_back(ev);
  else
    (void)hipEventDestroy(ev);
}

EvRef new_event(int g) {
  Device &d = rt().gpus[g];
  auto h = std::make_shared<EvHandle>();
  h->dev = g;
  if (!d.free_events.empty()) {
    h->ev = d.free_events.back();
    d.free_events.pop_back();
  } else if (hipEventCreate(&h->ev) != hipSuccess) {
    h->ev = nullptr;
  }
  return h;
}

static bool ev_done(const EvRef &e) { return !e || !e->ev || hipEventQuery(e->ev) == hipSuccess; }

namespace {

int len_class(size_t n) { return 63 - __builtin_clzll((unsigned long long)(n | 1)); }

bool hazard_idle(const Hazard &h) {
  if (!ev_done(h.writer)) return false;
  for (auto &e : h.readers)
    if (!ev_done(e)) return false;
  return true;
}

// Calls f(hazard) for every filed range overlapping [p, p + n); ranges whose
// events have all completed are forgotten on the way (they order nothing).
template <class F>
void each_overlap(const void *ptr, size_t n, F &&f) {
  Runtime &r = rt();
  const uintptr_t p = (uintptr_t)ptr, end = p + (n ? n : 1);
  for (int c = 0; c < kHazardClasses; ++c) {
    auto &m = r.hazards[c];
    if (m.empty()) continue;
    const uintptr_t span = (uintptr_t)2 << c;  // class-c ranges are shorter than 2^(c+1)
    const uintptr_t lo = p > span ? p - span : 0;
    for (auto it = m.lower_bound(HazardKey(lo, 0)); it != m.end() && it->first.first < end;) {
      if (it->first.first + it->first.second <= p) {
        ++it;
        continue;
      }
      if (hazard_idle(it->second)) {
        it = m.erase(it);
        r.hazard_epoch++;
        continue;
      }
      f(it->second);
      ++it;
    }
  }
}

Hazard &hazard_entry(const void *p, size_t n) {
  n = n ? n : 1;
  return rt().hazards[len_class(n)][HazardKey((uintptr_t)p, n)];
}

}  // namespace

void hazard_wait_read(const void *p, size_t n, hipStream_t s) {
  each_overlap(p, n, [s](Hazard &h) {
    if (h.writer && !ev_done(h.writer)) (void)hipStreamWaitEvent(s, h.writer->ev, 0);
  });
}

void hazard_wait_write(const void *p, size_t n, hipStream_t s) {
  each_overlap(p, n, [s](Hazard &h) {
    if (h.writer && !ev_done(h.writer)) (void)hipStreamWaitEvent(s, h.writer->ev, 0);
    for (auto &e : h.readers)
      if (!ev_done(e)) (void)hipStreamWaitEvent(s, e->ev, 0);
  });
}

void hazard_add_reader(const void *p, size_t n, const EvRef &e) {
  rt().hazard_epoch++;
  Hazard &h = hazard_entry(p, n);
  for (auto &x : h.readers)
    if (x == e) return;  // the same event (re-recorded by its owner) is listed once
  if (h.readers.size() >= 16) {  // prune completed readers
    std::vector<EvRef> keep;
    for (auto &x : h.readers)
      if (!ev_done(x)) keep.push_back(x);
    h.readers.swap(keep);
  }
  h.readers.push_back(e);
}

// The writer has waited (hazard_wait_write) on every read of this range, so
// the range's readers are superseded.  Other overlapping ranges keep their
// own state: queries visit them too.
void hazard_set_writer(const void *p, size_t n, const EvRef &e) {
  rt().hazard_epoch++;
  Hazard &h = hazard_entry(p, n);
  h.writer = e;
  h.readers.clear();
}

void hazard_drain(const void *ptr, size_t n) {
  Runtime &r = rt();
  const uintptr_t p = (uintptr_t)ptr, end = p + (n ? n : 1);
  for (int c = 0; c < kHazardClasses; ++c) {
    auto &m = r.hazards[c];
    if (m.empty()) continue;
    const uintptr_t span = (uintptr_t)2 << c;
    const uintptr_t lo = p > span ? p - span : 0;
    for (auto it = m.lower_bound(HazardKey(lo, 0)); it != m.end() && it->first.first < end;) {
      if (it->first.first + it->first.second <= p) {
        ++it;
        continue;
      }
      Hazard &h = it->second;
      if (h.writer && h.writer->ev) (void)hipEventSynchronize(h.writer->ev);
      for (auto &e : h.readers)
        if (e && e->ev) (void)hipEventSynchronize(e->ev);
      it = m.erase(it);
      r.hazard_epoch++;
    }
  }
}

size_t image_bytes(const talsh_tens_t *t, int i) {
  if (!t || !t->shape_p || i < 0 || i >= t->ndev) return 0;
  const int s = dk_size(t->data_kind[i]);
  return shape_volume(t->shape_p) * (size_t)(s > 0 ? s : 1);
}

// ------------------------------------------------------------------ memory
int dev_alloc(int g, size_t bytes, void **p, int *handle) {
  const int e = rt().gpus[g].arena.alloc(bytes, p, handle);
  if (rt().log_mem)
    std::printf("#TALSH MEM: GPU %d allocate %zu bytes -> %s %p (entry %d)\n", g, bytes, e ? "failed" : "ok",
                e ? nullptr : *p, e ? -1 : *handle);
  return e;
}

void dev_free(int g, void *p, int handle) {
  hazard_drain(p, rt().gpus[g].arena.length(handle));
  rt().gpus[g].arena.release(handle);
  if (rt().log_mem) std::printf("#TALSH MEM: GPU %d release %p (entry %d)\n", g, p, handle);
}

int host_alloc(size_t bytes, void **p, int *handle) {
  Runtime &r = rt();
  if (r.host_policy != MEM_ALLOC_REGULAR) {
    if (r.host_size > 0 && r.host_arena.alloc(bytes, p, handle) == 0) return 0;
    if (!r.host_fallback) return TRY_LATER;
  }
  // regular memory (policy MEM_ALLOC_REGULAR, or the fallback of talsh.h:20)
  *p = std::malloc(bytes ? bytes : 1);
  *handle = -1;
  return *p ? 0 : TRY_LATER;
}

void host_free(void *p, int handle, size_t bytes) {
  hazard_drain(p, handle >= 0 ? rt().host_arena.length(handle) : bytes);
  if (handle >= 0)
    rt().host_arena.release(handle);
  else
    std::free(p);
}

// ------------------------------------------------------------------ images
int image_add(talsh_tens_t *t, int dev_id, void *p, int buf_entry, int attached, int dk) {
  if (t->ndev >= t->dev_rsc_len) return -1;
  const int i = t->ndev++;
  t->dev_rsc[i].dev_id = canonical_dev(dev_id);
  t->dev_rsc[i].gmem_p = p;
  t->dev_rsc[i].buf_entry = buf_entry;
  t->dev_rsc[i].mem_attached = attached;
  t->data_kind[i] = dk;
  t->avail[i] = YEP;
  return i;
}

void image_release(talsh_dev_rsc_t *r, size_t bytes) {
  if (r->gmem_p && !r->mem_attached) {
    const int g = gpu_index(r->dev_id);
    if (r->dev_id == DEV_HOST)
      host_free(r->gmem_p, r->buf_entry, bytes);
    else if (g >= 0)
      dev_free(g, r->gmem_p, r->buf_entry);
  }
  // Attached (caller-owned) memory is not drained: the tasks that name this
  // tensor were finalized before its images go, and other tensors' tasks on
  // overlapping memory keep their hazard entries, which still order later
  // views of it.  (Draining them made a destruct wait for unrelated tasks
  // that read the same caller buffer.)
  r->dev_id = DEV_NULL;
  r->gmem_p = nullptr;
  r->buf_entry = -1;
  r->mem_attached = 0;
}

int image_discard(talsh_tens_t *t, int i) {
  if (i < 0 || i >= t->ndev) return -1;
  image_release(&t->dev_rsc[i], image_bytes(t, i));
  for (int j = i; j + 1 < t->ndev; ++j) {
    t->dev_rsc[j] = t->dev_rsc[j + 1];
    t->data_kind[j] = t->data_kind[j + 1];
    t->avail[j] = t->avail[j + 1];
  }
  t->ndev--;
  t->dev_rsc[t->ndev] = talsh_dev_rsc_t{DEV_NULL, nullptr, -1, 0};
  t->data_kind[t->ndev] = NO_TYPE;
  t->avail[t->ndev] = NOPE;
  return 0;
}

int image_discard_other(talsh_tens_t *t, int keep) {
  if (keep < 0 || keep >= t->ndev) return -1;
  for (int i = t->ndev - 1; i > keep; --i) image_discard(t, i);
  for (int i = keep - 1; i >= 0; --i) image_discard(t, i);
  return 0;
}

int image_find(const talsh_tens_t *t, int flat_dev, int dk, bool need_avail) {
  const int c = canonical_dev(flat_dev);
  for (int i = 0; i < t->ndev; ++i)
    if (t->dev_rsc[i].dev_id == c && (dk == NO_TYPE || t->data_kind[i] == dk) && (!need_avail || t->avail[i] == YEP))
      return i;
  return -1;
}

int tensor_healthy(const talsh_tens_t *t) {
  if (!t || !t->shape_p || t->shape_p->num_dim < 0 || !t->dev_rsc || !t->data_kind || !t->avail) return NOPE;
  if (t->ndev < 0 || t->ndev > t->dev_rsc_len) return NOPE;
  return YEP;
}

// ------------------------------------------------------------------ tasks
// Fixed cost of a task in host-link-equivalent bytes (about 70 us at 57 GB/s:
// launches and the short kernels of a small contraction).
constexpr double kTaskOverheadBytes = 4.0e6;

Task *task_new(int g, double weight) {
  Task *t = new Task();
  t->dev = g;
  Device &d = rt().gpus[g];
  // Least outstanding work first: a large task does not hold up the small ones
  // that would otherwise queue behind it on a round-robin stream.
  size_t best = 0;
  double best_w = 0.0;
  for (size_t i = 0; i < d.streams.size(); ++i) {
    auto &q = d.backlog[i];
    while (!q.empty() && hipEventQuery(q.front().first->ev) != hipErrorNotReady) q.pop_front();
    double w = 0.0;
    for (auto &e : q) w += e.second;
    if (i == 0 || w < best_w) {
      best = i;
      best_w = w;
    }
  }
  t->stream = d.streams[best];
  t->ev_start = new_event(g);
  t->ev_in = new_event(g);
  t->ev_comp = new_event(g);
  t->ev_done = new_event(g);
  d.backlog[best].emplace_back(t->ev_done, weight + kTaskOverheadBytes);
  t->status = TALSH_TASK_SCHEDULED;
  return t;
}

void task_free(Task *t) { delete t; }

void task_register(Task *t, const std::vector<talsh_tens_t *> &tens) {
  for (talsh_tens_t *x : tens) {
    if (!x) continue;
    bool seen = false;
    for (talsh_tens_t *y : t->tens) seen = seen || y == x;
    if (seen) continue;
    t->tens.push_back(x);
    rt().tens_tasks[x].push_back(t);
  }
}

int task_finalize(Task *t, bool block) {
  if (t->status == TALSH_TASK_COMPLETED || t->status == TALSH_TASK_ERROR) return t->status;
  // a blocking wait polls the completion event, as talshTaskWait polls
  // talshTaskComplete (talshc.cpp:2481): the runtime's own blocking wait
  // wakes the thread some 10-20 us after a short task ends
  hipError_t q = hipEventQuery(t->ev_done->ev);
  while (block && q == hipErrorNotReady) q = hipEventQuery(t->ev_done->ev);
  if (q == hipErrorNotReady) return t->status;
  Runtime &r = rt();
  Device &d = r.gpus[t->dev];
  auto fin = std::move(t->on_finalize);
  t->on_finalize.clear();
  for (auto &f : fin) f();
  for (talsh_tens_t *x : t->tens) {
    auto it = r.tens_tasks.find(x);
    if (it == r.tens_tasks.end()) continue;
    auto &v = it->second;
    for (size_t i = 0; i < v.size(); ++i)
      if (v[i] == t) {
        v.erase(v.begin() + (long)i);
        break;
      }
    if (v.empty()) r.tens_tasks.erase(it);
  }
  t->tens.clear();
  d.active--;
  if (q == hipSuccess) {
    t->status = TALSH_TASK_COMPLETED;
    d.tasks_completed++;
    d.flops += t->flops;
    float ms = 0;
    if (hipEventElapsedTime(&ms, t->ev_start->ev, t->ev_done->ev) == hipSuccess) t->exec_time = ms * 1e-3;
  } else {
    t->status = TALSH_TASK_ERROR;
    t->error = (int)q;
    d.tasks_failed++;
  }
  return t->status;
}

// Finalises a scheduled task once its device work is done; returns the status.
int task_finalize_if_done(talsh_task_t *tt, bool block) {
  Task *t = static_cast<Task *>(tt->task_p);
  if (!t) return TALSH_TASK_EMPTY;
  const int st = task_finalize(t, block);
  if (st == TALSH_TASK_COMPLETED) {
    tt->task_error = 0;
    tt->exec_time = t->exec_time;
  } else if (st == TALSH_TASK_ERROR) {
    tt->task_error = t->error;
  }
  return st;
}

}  // namespace tal

using namespace tal;

// =========================================================================
// C API: helpers and control (talsh.h:149-229)
// =========================================================================
extern "C" {

int talshValidDataKind(int datk, int *datk_size) {
  const int s = dk_size(datk);
  if (datk_size) *datk_size = s < 0 ? 0 : s;
  return s < 0 ? NOPE : YEP;
}

int talshInit(size_t *host_buf_size, int *host_arg_max, int ngpus, int gpu_list[], int nmics, int mic_list[],
              int namds, int amd_list[]) {
  (void)mic_list;
  Runtime &r = rt();
  std::lock_guard<std::recursive_mutex> lk(r.mu);
  if (r.on) return TALSH_ALREADY_INITIALIZED;
  if (nmics > 0) return TALSH_NOT_AVAILABLE;
  if (ngpus < 0 || namds < 0 || ngpus > MAX_GPUS_PER_NODE || namds > MAX_AMDS_PER_NODE) return TALSH_INVALID_ARGS;
  // default Host allocation policy (talshc.cpp:602, talsh.h:19-20)
  r.host_policy = MEM_ALLOC_TMP_BUF;
  r.host_fallback = 1;
  // GPUs may be listed under either device kind; both lists name HIP device ordinals.
  std::vector<int> list;
  for (int i = 0; i < ngpus; ++i) list.push_back(gpu_list ? gpu_list[i] : i);
  for (int i = 0; i < namds; ++i) list.push_back(amd_list ? amd_list[i] : i);
  int nvis = 0;
  if (!list.empty() && hipGetDeviceCount(&nvis) != hipSuccess) return TALSH_NOT_AVAILABLE;
  r.gpus.assign(MAX_GPUS_PER_NODE, Device());
  for (int id : list) {
    if (id < 0 || id >= nvis || id >= MAX_GPUS_PER_NODE) return TALSH_INVALID_ARGS;
    Device &d = r.gpus[id];
    if (d.on) continue;
    if (hipSetDevice(id) != hipSuccess) return TALSH_FAILURE;
    // address of this GPU's copy of the allocation-free entry points' scratch
    if (noalloc_scratch_setup(id) != 0) return TALSH_FAILURE;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return TALSH_FAILURE;
    d.hip_id = id;
    d.mem_total = tot;
    size_t want = r.dev_buffer_request ? r.dev_buffer_request : (size_t)(0.80 * (double)fr);
    want = std::min(want, fr);
    void *base = nullptr;
    while (want >= (64u << 20) && hipMalloc(&base, want) != hipSuccess) want = want / 10 * 9;
    if (!base) return TALSH_FAILURE;
    d.base = base;
    d.size = want;
    d.arena.init(base, want, 256);
    d.streams.resize(r.nstreams);
    for (auto &s : d.streams)
      if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return TALSH_FAILURE;
    d.backlog.assign(d.streams.size(), {});
    d.time_start = wall_time();
    d.on = true;
  }
  // pinned host argument buffer (HAB)
  size_t hsz = host_buf_size ? *host_buf_size : 0;
  if (hsz > 0) {
    void *hb = nullptr;
    while (hsz >= (1u << 20) && hipHostMalloc(&hb, hsz, hipHostMallocDefault) != hipSuccess) hsz = hsz / 10 * 9;
    if (hb) {
      r.host_base = hb;
      r.host_size = hsz;
      r.host_arena.init(hb, hsz, 256);
    } else {
      hsz = 0;
    }
  }
  if (host_buf_size) *host_buf_size = r.host_size;
  if (host_arg_max) *host_arg_max = r.host_size ? (int)std::min<size_t>(r.host_size / 4096, 1 << 20) : 0;
  r.on = true;
  return TALSH_SUCCESS;
}

int talshShutdown(void) {
  Runtime &r = rt();
  std::lock_guard<std::recursive_mutex> lk(r.mu);
  if (!r.on) return TALSH_NOT_INITIALIZED;
  for (auto &d : r.gpus) {
    if (!d.on) continue;
    (void)hipSetDevice(d.hip_id);
    (void)hipDeviceSynchronize();
  }
  for (auto &m : r.hazards) m.clear();
  r.tens_tasks.clear();
  r.hazard_epoch++;
  ttgt_cache_clear();
  perm_tune_clear();
  for (auto &d : r.gpus) {
    if (!d.on) continue;
    d.backlog.clear();
    for (auto e : d.free_events) (void)hipEventDestroy(e);
    d.free_events.clear();
    for (auto s : d.streams) (void)hipStreamDestroy(s);
    d.streams.clear();
    (void)hipFree(d.base);
    d.on = false;
  }
  if (r.host_base) (void)hipHostFree(r.host_base);
  r.host_base = nullptr;
  r.host_size = 0;
  r.host_arena.init(nullptr, 0, 256);
  r.gpus.clear();
  r.on = false;
  return TALSH_SUCCESS;
}

int talshDeviceCount(int dev_kind, int *dev_count) {
  if (!dev_count) return TALSH_INVALID_ARGS;
  *dev_count = 0;
  switch (dev_kind) {
    case DEV_HOST: *dev_count = 1; return TALSH_SUCCESS;
    case DEV_NVIDIA_GPU:
    case DEV_AMD_GPU: {
      int n = 0;
      if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
      *dev_count = std::min(n, MAX_GPUS_PER_NODE);
      return TALSH_SUCCESS;
    }
    case DEV_INTEL_MIC: return TALSH_SUCCESS;
    default: return TALSH_INVALID_ARGS;
  }
}

int talshFlatDevId(int dev_kind, int dev_num) {
  switch (dev_kind) {
    case DEV_HOST: return dev_num == 0 ? 0 : DEV_MAX;
    case DEV_NVIDIA_GPU: return (dev_num >= 0 && dev_num < MAX_GPUS_PER_NODE) ? 1 + dev_num : DEV_MAX;
    case DEV_INTEL_MIC: return (dev_num >= 0 && dev_num < MAX_MICS_PER_NODE) ? 1 + MAX_GPUS_PER_NODE + dev_num : DEV_MAX;
    case DEV_AMD_GPU:
      return (dev_num >= 0 && dev_num < MAX_AMDS_PER_NODE) ? 1 + MAX_GPUS_PER_NODE + MAX_MICS_PER_NODE + dev_num
                                                            : DEV_MAX;
    default: return DEV_MAX;
  }
}

int talshKindDevId(int dev_id, int *dev_kind) {
  int k = DEV_NULL, n = -1;
  if (dev_id == 0) { k = DEV_HOST; n = 0; }
  else if (dev_id >= 1 && dev_id <= MAX_GPUS_PER_NODE) { k = DEV_NVIDIA_GPU; n = dev_id - 1; }
  else if (dev_id > MAX_GPUS_PER_NODE && dev_id <= MAX_GPUS_PER_NODE + MAX_MICS_PER_NODE) {
    k = DEV_INTEL_MIC; n = dev_id - 1 - MAX_GPUS_PER_NODE;
  } else if (dev_id > MAX_GPUS_PER_NODE + MAX_MICS_PER_NODE && dev_id < DEV_MAX) {
    k = DEV_AMD_GPU; n = dev_id - 1 - MAX_GPUS_PER_NODE - MAX_MICS_PER_NODE;
  }
  if (dev_kind) *dev_kind = k;
  return n;
}

static int resolve_flat(int dev_num, int dev_kind) {
  return dev_kind == DEV_NULL ? dev_num : talshFlatDevId(dev_kind, dev_num);
}

int talshDeviceState(int dev_num, int dev_kind) {
  Runtime &r = rt();
  if (!r.on) return DEV_OFF;
  const int id = resolve_flat(dev_num, dev_kind);
  if (id == DEV_HOST) return DEV_ON;
  return gpu_index(id) >= 0 ? DEV_ON : DEV_OFF;
}
int talshDeviceState_(int dev_num, int dev_kind) { return talshDeviceState(dev_num, dev_kind); }

int talshDeviceBusyLeast(int dev_kind) {
  Runtime &r = rt();
  int best = -1;
  for (int g = 0; g < (int)r.gpus.size(); ++g)
    if (r.gpus[g].on && (best < 0 || r.gpus[g].active < r.gpus[best].active)) best = g;
  if (best < 0) return dev_kind == DEV_NULL ? DEV_HOST : -1;
  if (dev_kind == DEV_NULL) return flat_gpu(best);
  return best;
}
int talshDeviceBusyLeast_(int dev_kind) { return talshDeviceBusyLeast(dev_kind); }

size_t talshDeviceMemorySize(int dev_num, int dev_kind) {
  const int id = resolve_flat(dev_num, dev_kind);
  const int g = gpu_index(id);
  if (g >= 0) return rt().gpus[g].mem_total;
  if (id == DEV_HOST) return rt().host_size;
  return 0;
}
size_t talshDeviceMemorySize_(int dev_num, int dev_kind) { return talshDeviceMemorySize(dev_num, dev_kind); }

size_t talshDeviceBufferSize(int dev_num, int dev_kind) {
  const int id = resolve_flat(dev_num, dev_kind);
  const int g = gpu_index(id);
  if (g >= 0) return rt().gpus[g].size;
  if (id == DEV_HOST) return rt().host_size;
  return 0;
}
size_t talshDeviceBufferSize_(int dev_num, int dev_kind) { return talshDeviceBufferSize(dev_num, dev_kind); }

void *talshDeviceBufferBasePtr(int dev_num, int dev_kind) {
  const int id = resolve_flat(dev_num, dev_kind);
  const int g = gpu_index(id);
  if (g >= 0 && rt().gpus[g].on) return rt().gpus[g].base;
  if (id == DEV_HOST) return rt().host_base;
  return nullptr;
}
void *talshDeviceBufferBasePtr_(int dev_num, int dev_kind) { return talshDeviceBufferBasePtr(dev_num, dev_kind); }

// Execution device for up to three operands (talshc.cpp:366-488): the device
// that already holds available images of the most operand elements, so the
// fewest elements have to move.  Every operation of this library runs on a
// GPU, so when the Host wins (or nothing is resident anywhere) the least busy
// GPU is returned instead; the Host only when no GPU is initialised.
int talshDetermineOptimalDevice(const talsh_tens_t *tens0, const talsh_tens_t *tens1, const talsh_tens_t *tens2) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!tens0) return DEV_NULL;
  std::map<int, double> resident;  // flat device id -> elements with an available image there
  for (const talsh_tens_t *t : {tens0, tens1, tens2}) {
    if (!t || tensor_healthy(t) != YEP || !t->shape_p) continue;
    const double v = (double)shape_volume(t->shape_p);
    std::set<int> seen;
    for (int i = 0; i < t->ndev; ++i)
      if (t->avail[i] == YEP && seen.insert(t->dev_rsc[i].dev_id).second) resident[t->dev_rsc[i].dev_id] += v;
  }
  int best = DEV_NULL;
  double best_v = -1.0;
  for (const auto &kv : resident)
    if (gpu_index(kv.first) >= 0 && rt().gpus[gpu_index(kv.first)].on && kv.second > best_v) {
      best = kv.first;
      best_v = kv.second;
    }
  if (best != DEV_NULL) return best;
  const int g = talshDeviceBusyLeast(DEV_NVIDIA_GPU);
  if (g >= 0) return talshFlatDevId(DEV_NVIDIA_GPU, g);
  return talshFlatDevId(DEV_HOST, 0);
}

// Logging switches (talshc.cpp:1057-1075): buffer-pool allocations and tensor
// operations are printed to stdout while on.
void talshMemManagerLogStart(void) { rt().log_mem = true; }
void talshMemManagerLogFinish(void) { rt().log_mem = false; }
int talsh_tens_no_init(const talsh_tens_data_t *tens_data, const talsh_tens_shape_t *tens_shape,
                       const talsh_tens_signature_t *tens_signature) {
  (void)tens_data, (void)tens_shape, (void)tens_signature;
  return 0;
}

void talshSetMemAllocPolicyHost(int mem_policy, int fallback, int *ierr) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  const bool ok = mem_policy == MEM_ALLOC_REGULAR || mem_policy == MEM_ALLOC_TMP_BUF || mem_policy == MEM_ALLOC_ALL_BUF;
  if (ok) {
    rt().host_policy = mem_policy;
    rt().host_fallback = fallback != 0;
  }
  if (ierr) *ierr = ok ? 0 : 1;
}

void talshTensorOpLogStart(void) { rt().log_ops = true; }
void talshTensorOpLogFinish(void) { rt().log_ops = false; }

size_t talshDeviceTensorSize(int dev_num, int dev_kind) {
  // no per-tensor cap beyond the buffer itself (the reference caps at 1/6, mem_manager.cpp:626-631)
  return talshDeviceBufferSize(dev_num, dev_kind);
}
size_t talshDeviceTensorSize_(int dev_num, int dev_kind) { return talshDeviceTensorSize(dev_num, dev_kind); }

size_t talshDeviceBufferFreeSize(int dev_num, int dev_kind) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  const int id = resolve_flat(dev_num, dev_kind);
  const int g = gpu_index(id);
  if (g >= 0) return rt().gpus[g].arena.free_bytes();
  if (id == DEV_HOST) return rt().host_arena.free_bytes();
  return 0;
}
size_t talshDeviceBufferFreeSize_(int dev_num, int dev_kind) { return talshDeviceBufferFreeSize(dev_num, dev_kind); }

double talshDeviceGetFlops(int dev_kind, int dev_id) {
  Runtime &r = rt();
  double f = 0;
  if (dev_kind == DEV_DEFAULT && dev_id == DEV_DEFAULT) {
    for (auto &d : r.gpus) f += d.flops;
    return f;
  }
  const int g = gpu_index(resolve_flat(dev_id, dev_kind));
  return g >= 0 ? r.gpus[g].flops : 0.0;
}

int talshStats(int dev_id, int dev_kind) {
  Runtime &r = rt();
  if (!r.on) return TALSH_NOT_INITIALIZED;
  for (int g = 0; g < (int)r.gpus.size(); ++g) {
    if (!r.gpus[g].on) continue;
    if (dev_id >= 0 && gpu_index(resolve_flat(dev_id, dev_kind)) != g) continue;
    const Device &d = r.gpus[g];
    std::printf("#MSG(TAL-SH::MI355X): Statistics on GPU #%d:\n", g);
    std::printf(" Time active (sec): %f\n", wall_time() - d.time_start);
    std::printf(" Number of tasks submitted: %lld\n", d.tasks_submitted);
    std::printf(" Number of tasks completed: %lld\n", d.tasks_completed);
    std::printf(" Number of tasks deferred : %lld\n", d.tasks_deferred);
    std::printf(" Number of tasks failed   : %lld\n", d.tasks_failed);
    std::printf(" Number of Flops processed: %e\n", d.flops);
    std::printf(" Number of Bytes to GPU   : %e\n", d.traffic_in);
    std::printf(" Number of Bytes from GPU : %e\n", d.traffic_out);
    std::printf(" Buffer pool free / size  : %zu / %zu\n", d.arena.free_bytes(), d.size);
  }
  return TALSH_SUCCESS;
}
int talshStats_(int dev_id, int dev_kind) { return talshStats(dev_id, dev_kind); }

// =========================================================================
// C API: tensor blocks (talsh.h:232-311)
// =========================================================================
int talshTensorCreate(talsh_tens_t **tens_block) {
  if (!tens_block) return TALSH_INVALID_ARGS;
  *tens_block = (talsh_tens_t *)std::malloc(sizeof(talsh_tens_t));
  if (!*tens_block) return TRY_LATER;
  return talshTensorClean(*tens_block);
}

int talshTensorClean(talsh_tens_t *t) {
  if (!t) return TALSH_INVALID_ARGS;
  t->shape_p = nullptr;
  t->dev_rsc = nullptr;
  t->data_kind = nullptr;
  t->avail = nullptr;
  t->dev_rsc_len = 0;
  t->ndev = 0;
  return TALSH_SUCCESS;
}

int talshTensorIsEmpty(const talsh_tens_t *t) {
  if (!t) return TALSH_INVALID_ARGS;
  return t->shape_p == nullptr ? YEP : NOPE;
}

int talshTensorConstruct(talsh_tens_t *t, int data_kind, int tens_rank, const int tens_dims[], int dev_id,
                         void *ext_mem, int in_hab, talsh_tens_init_i init_method, double init_val_real,
                         double init_val_imag) {
  Runtime &r = rt();
  std::lock_guard<std::recursive_mutex> lk(r.mu);
  if (!r.on) return TALSH_NOT_INITIALIZED;
  (void)in_hab;
  if (!t) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(t) != YEP) return TALSH_OBJECT_NOT_EMPTY;
  int dks = 0;
  if (talshValidDataKind(data_kind, &dks) != YEP) return TALSH_INVALID_ARGS;
  if (tens_rank < 0 || tens_rank > MAX_TENSOR_RANK || (tens_rank > 0 && !tens_dims)) return TALSH_INVALID_ARGS;
  for (int i = 0; i < tens_rank; ++i)
    if (tens_dims[i] <= 0) return TALSH_INVALID_ARGS;
  const int cdev = canonical_dev(dev_id);
  if (cdev < 0) return TALSH_INVALID_ARGS;
  const int g = gpu_index(cdev);
  if (cdev != DEV_HOST && g < 0) return TALSH_INVALID_ARGS;
  // shape
  t->shape_p = (talsh_tens_shape_t *)std::malloc(sizeof(talsh_tens_shape_t));
  t->dev_rsc = (talsh_dev_rsc_t *)std::malloc(sizeof(talsh_dev_rsc_t) * TALSH_MAX_DEV_PRESENT);
  t->data_kind = (int *)std::malloc(sizeof(int) * TALSH_MAX_DEV_PRESENT);
  t->avail = (int *)std::malloc(sizeof(int) * TALSH_MAX_DEV_PRESENT);
  if (!t->shape_p || !t->dev_rsc || !t->data_kind || !t->avail) {
    std::free(t->shape_p); std::free(t->dev_rsc); std::free(t->data_kind); std::free(t->avail);
    talshTensorClean(t);
    return TRY_LATER;
  }
  t->shape_p->num_dim = tens_rank;
  t->shape_p->dims = tens_rank > 0 ? (int *)std::malloc(sizeof(int) * tens_rank) : nullptr;
  t->shape_p->divs = nullptr;
  t->shape_p->grps = nullptr;
  for (int i = 0; i < tens_rank; ++i) t->shape_p->dims[i] = tens_dims[i];
  t->dev_rsc_len = TALSH_MAX_DEV_PRESENT;
  t->ndev = 0;
  for (int i = 0; i < TALSH_MAX_DEV_PRESENT; ++i) {
    t->dev_rsc[i] = talsh_dev_rsc_t{DEV_NULL, nullptr, -1, 0};
    t->data_kind[i] = NO_TYPE;
    t->avail[i] = NOPE;
  }
  const size_t vol = shape_volume(t->shape_p);
  if (ext_mem) {  // external body: attached, no initialisation
    image_add(t, cdev, ext_mem, -1, 1, data_kind);
    return TALSH_SUCCESS;
  }
  if (data_kind == NO_TYPE) return TALSH_SUCCESS;
  const size_t bytes = vol * (size_t)dks;
  void *p = nullptr;
  int h = -1, err;
  if (cdev == DEV_HOST)
    err = host_alloc(bytes, &p, &h);
  else
    err = dev_alloc(g, bytes, &p, &h);
  if (err) {
    talshTensorDestruct(t);
    return err;
  }
  image_add(t, cdev, p, h, 0, data_kind);
  if (cdev == DEV_HOST) {
    if (init_method) {
      talsh_tens_data_t td{p, vol, data_kind};
      if (init_method(&td, t->shape_p, nullptr)) return NOT_CLEAN;
      return TALSH_SUCCESS;
    }
    for (size_t l = 0; l < vol; ++l) {
      switch (data_kind) {
        case R4: ((float *)p)[l] = (float)init_val_real; break;
        case R8: ((double *)p)[l] = init_val_real; break;
        case C4: ((float *)p)[2 * l] = (float)init_val_real; ((float *)p)[2 * l + 1] = (float)init_val_imag; break;
        case C8: ((double *)p)[2 * l] = init_val_real; ((double *)p)[2 * l + 1] = init_val_imag; break;
      }
    }
    return TALSH_SUCCESS;
  }
  // device body: initialise on the device (blocking, as the reference does), unless asked not to
  if (init_method == talsh_tens_no_init) return TALSH_SUCCESS;
  if (talshTensorInit(t, init_val_real, init_val_imag, cdev, DEV_NULL, COPY_M, nullptr) != TALSH_SUCCESS)
    return NOT_CLEAN;
  return TALSH_SUCCESS;
}

int talshTensorConstruct_(talsh_tens_t *t, int data_kind, int tens_rank, const int tens_dims[], int dev_id,
                          void *ext_mem, int in_hab, talsh_tens_init_i init_method, double init_val_real,
                          double init_val_imag) {
  return talshTensorConstruct(t, data_kind, tens_rank, tens_dims, dev_id, ext_mem, in_hab, init_method,
                              init_val_real, init_val_imag);
}

int talshTensorImportData(talsh_tens_t *t, int data_kind, const void *ext_data) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (!t || !ext_data) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(t) == YEP) return TALSH_OBJECT_IS_EMPTY;
  const int i = image_find(t, DEV_HOST, data_kind, true);
  if (i < 0) return TALSH_NOT_FOUND;
  hazard_drain(t->dev_rsc[i].gmem_p, image_bytes(t, i));
  std::memcpy(t->dev_rsc[i].gmem_p, ext_data, shape_volume(t->shape_p) * (size_t)dk_size(data_kind));
  return TALSH_SUCCESS;
}

int talshTensorDestruct(talsh_tens_t *t) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!t) return TALSH_INVALID_ARGS;
  // Tasks that still name the tensor finish first: their coherence work
  // (images, buffers they own) needs it.  Their handles then report COMPLETED.
  for (auto it = rt().tens_tasks.find(t); it != rt().tens_tasks.end(); it = rt().tens_tasks.find(t)) {
    Task *task = it->second.front();
    (void)task_finalize(task, true);
    it = rt().tens_tasks.find(t);
    if (it != rt().tens_tasks.end() && it->second.front() == task) {  // still listed: drop the link
      it->second.erase(it->second.begin());
      if (it->second.empty()) rt().tens_tasks.erase(it);
    }
  }
  if (t->dev_rsc) {
    for (int i = 0; i < t->ndev && i < t->dev_rsc_len; ++i) image_release(&t->dev_rsc[i], image_bytes(t, i));
    std::free(t->dev_rsc);
  }
  if (t->shape_p) {
    std::free(t->shape_p->dims);
    std::free(t->shape_p);
  }
  std::free(t->data_kind);
  std::free(t->avail);
  talshTensorClean(t);
  return TALSH_SUCCESS;
}

int talshTensorDestroy(talsh_tens_t *t) {
  if (!t) return TALSH_INVALID_ARGS;
  const int e = talshTensorDestruct(t);
  std::free(t);
  return e;
}

int talshTensorRank(const talsh_tens_t *t) { return (t && t->shape_p) ? t->shape_p->num_dim : -1; }

size_t talshTensorVolume(const talsh_tens_t *t) { return (t && t->shape_p) ? shape_volume(t->shape_p) : 0; }

size_t talshTensorSizeAllImages(const talsh_tens_t *t, int *num_images) {
  size_t s = 0;
  if (num_images) *num_images = 0;
  if (!t || !t->shape_p) return 0;
  const size_t v = shape_volume(t->shape_p);
  for (int i = 0; i < t->ndev; ++i) s += v * (size_t)std::max(0, dk_size(t->data_kind[i]));
  if (num_images) *num_images = t->ndev;
  return s;
}

const int *talshTensorDimExtents(const talsh_tens_t *t, int *rank) {
  if (!t || !t->shape_p) {
    if (rank) *rank = -1;
    return nullptr;
  }
  if (rank) *rank = t->shape_p->num_dim;
  return t->shape_p->dims;
}

int talshTensorShape(const talsh_tens_t *t, talsh_tens_shape_t *s) {
  if (!t || !s) return TALSH_INVALID_ARGS;
  if (!t->shape_p) return TALSH_OBJECT_IS_EMPTY;
  const int n = t->shape_p->num_dim;
  s->num_dim = n;
  s->dims = n > 0 ? (int *)std::malloc(sizeof(int) * n) : nullptr;
  s->divs = nullptr;
  s->grps = nullptr;
  for (int i = 0; i < n; ++i) s->dims[i] = t->shape_p->dims[i];
  return TALSH_SUCCESS;
}

int talshTensorDataKind(const talsh_tens_t *t, int *num_images, int *data_kinds) {
  if (!t || !num_images || !data_kinds) return TALSH_INVALID_ARGS;
  *num_images = t->ndev;
  for (int i = 0; i < t->ndev; ++i) data_kinds[i] = t->data_kind[i];
  return TALSH_SUCCESS;
}

int talshTensorReshape(talsh_tens_t *t, int tens_rank, const int tens_dims[]) {
  if (!t || !t->shape_p || tens_rank < 0 || tens_rank > MAX_TENSOR_RANK) return TALSH_INVALID_ARGS;
  size_t v = 1;
  for (int i = 0; i < tens_rank; ++i) v *= (size_t)tens_dims[i];
  if (v != shape_volume(t->shape_p)) return TALSH_INVALID_ARGS;
  int *nd = tens_rank > 0 ? (int *)std::malloc(sizeof(int) * tens_rank) : nullptr;
  for (int i = 0; i < tens_rank; ++i) nd[i] = tens_dims[i];
  std::free(t->shape_p->dims);
  t->shape_p->dims = nd;
  t->shape_p->num_dim = tens_rank;
  return TALSH_SUCCESS;
}

int talshTensorInUse(const talsh_tens_t *t) {
  if (!t || !t->avail) return NOPE;
  for (int i = 0; i < t->ndev; ++i)
    if (t->avail[i] != YEP) return YEP;
  return NOPE;
}

int talshTensorPresence(const talsh_tens_t *t, int *ncopies, int copies[], int data_kinds[], int dev_kind,
                        int dev_id) {
  if (!t || !ncopies) return TALSH_INVALID_ARGS;
  *ncopies = 0;
  for (int i = 0; i < t->ndev; ++i) {
    int k;
    const int n = talshKindDevId(t->dev_rsc[i].dev_id, &k);
    if (dev_kind != DEV_NULL && k != dev_kind && !(dev_kind == DEV_AMD_GPU && k == DEV_NVIDIA_GPU)) continue;
    if (dev_kind != DEV_NULL && dev_id >= 0 && n != dev_id) continue;
    if (copies) copies[*ncopies] = t->dev_rsc[i].dev_id;
    if (data_kinds) data_kinds[*ncopies] = t->data_kind[i];
    (*ncopies)++;
  }
  return TALSH_SUCCESS;
}
int talshTensorPresence_(const talsh_tens_t *t, int *ncopies, int copies[], int data_kinds[], int dev_kind,
                         int dev_id) {
  return talshTensorPresence(t, ncopies, copies, data_kinds, dev_kind, dev_id);
}

int talshTensorGetBodyAccess(talsh_tens_t *t, void **body_p, int data_kind, int dev_id, int dev_kind) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!t || !body_p) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(t) == YEP) return TALSH_OBJECT_IS_EMPTY;
  const int id = resolve_flat(dev_id, dev_kind);
  const int i = image_find(t, id, data_kind, false);
  if (i < 0) return TALSH_NOT_FOUND;
  if (t->avail[i] != YEP) return TALSH_NOT_AVAILABLE;
  hazard_drain(t->dev_rsc[i].gmem_p, image_bytes(t, i));  // the caller may touch it directly
  *body_p = t->dev_rsc[i].gmem_p;
  return TALSH_SUCCESS;
}
int talshTensorGetBodyAccess_(talsh_tens_t *t, void **body_p, int data_kind, int dev_id, int dev_kind) {
  return talshTensorGetBodyAccess(t, body_p, data_kind, dev_id, dev_kind);
}

int talshTensorGetBodyAccessConst(const talsh_tens_t *t, const void **body_p, int data_kind, int dev_id,
                                  int dev_kind) {
  return talshTensorGetBodyAccess(const_cast<talsh_tens_t *>(t), const_cast<void **>(body_p), data_kind, dev_id,
                                  dev_kind);
}

int talshTensorGetScalar(talsh_tens_t *t, double *scalar_real, double *scalar_imag) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!t || !scalar_real || !scalar_imag) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(t) == YEP) return TALSH_OBJECT_IS_EMPTY;
  if (shape_volume(t->shape_p) != 1) return TALSH_INVALID_ARGS;
  for (int i = 0; i < t->ndev; ++i) {
    if (t->avail[i] != YEP) continue;
    const int dk = t->data_kind[i];
    unsigned char buf[16] = {0};
    const int g = gpu_index(t->dev_rsc[i].dev_id);
    hazard_drain(t->dev_rsc[i].gmem_p, image_bytes(t, i));
    if (t->dev_rsc[i].dev_id == DEV_HOST)
      std::memcpy(buf, t->dev_rsc[i].gmem_p, dk_size(dk));
    else if (g >= 0) {
      (void)hipSetDevice(rt().gpus[g].hip_id);
      if (hipMemcpy(buf, t->dev_rsc[i].gmem_p, dk_size(dk), hipMemcpyDeviceToHost) != hipSuccess)
        return TALSH_FAILURE;
    } else
      continue;
    switch (dk) {
      case R4: *scalar_real = *(float *)buf; *scalar_imag = 0; break;
      case R8: *scalar_real = *(double *)buf; *scalar_imag = 0; break;
      case C4: *scalar_real = ((float *)buf)[0]; *scalar_imag = ((float *)buf)[1]; break;
      case C8: *scalar_real = ((double *)buf)[0]; *scalar_imag = ((double *)buf)[1]; break;
      default: return TALSH_FAILURE;
    }
    return TALSH_SUCCESS;
  }
  return TALSH_NOT_AVAILABLE;
}

void talshTensorPrint(const talsh_tens_t *t) {
  if (!t || !t->shape_p) {
    std::printf("#MSG(TAL-SH::talsh_tens_t): Empty\n");
    return;
  }
  std::printf("#MSG(TAL-SH::talsh_tens_t): Shape (");
  for (int i = 0; i < t->shape_p->num_dim; ++i) std::printf(i ? ",%d" : "%d", t->shape_p->dims[i]);
  std::printf(")\n");
}

void talshTensorPrintInfo(const talsh_tens_t *t) {
  talshTensorPrint(t);
  if (!t || !t->shape_p) return;
  for (int i = 0; i < t->ndev; ++i)
    std::printf(" Image %d: device %d, data kind %d, available %d, ptr %p\n", i, t->dev_rsc[i].dev_id,
                t->data_kind[i], t->avail[i], t->dev_rsc[i].gmem_p);
}

// Elements of the Host image with |x| >= thresh and their multi-indices
// (talshc.cpp:1753-1850): one line "value i1 i2 ..." per element.
void talshTensorPrintBody(const talsh_tens_t *t, double thresh) {
  std::printf("#MESSAGE: Printing tensor body:");
  if (!t || !t->shape_p || t->shape_p->num_dim < 0) {
    std::printf(" Empty tensor\n");
    return;
  }
  const int i = image_find(t, DEV_HOST, NO_TYPE, true);
  if (i < 0) {
    std::printf(" No Host image\n");
    return;
  }
  hazard_drain(t->dev_rsc[i].gmem_p, image_bytes(t, i));
  const int nd = t->shape_p->num_dim;
  const int *dims = t->shape_p->dims;
  const size_t v = shape_volume(t->shape_p);
  const void *p = t->dev_rsc[i].gmem_p;
  for (size_t l = 0; l < v; ++l) {
    double re = 0, im = 0;
    switch (t->data_kind[i]) {
      case R4: re = ((const float *)p)[l]; break;
      case R8: re = ((const double *)p)[l]; break;
      case C4: re = ((const float *)p)[2 * l]; im = ((const float *)p)[2 * l + 1]; break;
      case C8: re = ((const double *)p)[2 * l]; im = ((const double *)p)[2 * l + 1]; break;
    }
    const bool cplx = t->data_kind[i] == C4 || t->data_kind[i] == C8;
    if (std::hypot(re, im) < thresh) continue;
    if (cplx) std::printf("\n(%E,%E)", re, im);
    else std::printf("\n%E", re);
    size_t q = l;
    for (int k = 0; k < nd; ++k) {
      std::printf(" %zu", q % (size_t)dims[k]);
      q /= (size_t)dims[k];
    }
  }
  std::printf("\n#END OF MESSAGE\n");
}

double talshTensorImageNorm1_cpu(const talsh_tens_t *t) {
  // Debugging accessor over an existing Host image (talshc.cpp, talsh.h:593).
  if (!t || !t->shape_p) return -1.0;
  const int i = image_find(t, DEV_HOST, NO_TYPE, false);
  if (i < 0) return -1.0;
  hazard_drain(t->dev_rsc[i].gmem_p, image_bytes(t, i));
  const size_t v = shape_volume(t->shape_p);
  double s = 0;
  const void *p = t->dev_rsc[i].gmem_p;
  for (size_t l = 0; l < v; ++l) {
    switch (t->data_kind[i]) {
      case R4: s += std::abs(((const float *)p)[l]); break;
      case R8: s += std::abs(((const double *)p)[l]); break;
      case C4: s += std::hypot(((const float *)p)[2 * l], ((const float *)p)[2 * l + 1]); break;
      case C8: s += std::hypot(((const double *)p)[2 * l], ((const double *)p)[2 * l + 1]); break;
    }
  }
  return s;
}

// =========================================================================
// C API: tasks (talsh.h:332-372)
// =========================================================================
int talshTaskCreate(talsh_task_t **tt) {
  if (!tt) return TALSH_INVALID_ARGS;
  *tt = (talsh_task_t *)std::malloc(sizeof(talsh_task_t));
  if (!*tt) return TRY_LATER;
  return talshTaskClean(*tt);
}

int talshTaskClean(talsh_task_t *tt) {
  if (!tt) return TALSH_INVALID_ARGS;
  tt->task_p = nullptr;
  tt->task_error = -1;
  tt->dev_kind = DEV_NULL;
  tt->data_kind = NO_TYPE;
  tt->coherence = -1;
  tt->num_args = 0;
  for (int i = 0; i < MAX_TENSOR_OPERANDS; ++i) {
    tt->tens_args[i].tens_p = nullptr;
    tt->tens_args[i].source_image = -1;
  }
  tt->data_vol = 0;
  tt->flops = 0;
  tt->exec_time = 0;
  return TALSH_SUCCESS;
}

int talshTaskIsEmpty(const talsh_task_t *tt) {
  if (!tt) return TALSH_INVALID_ARGS;
  return (tt->dev_kind == DEV_NULL && tt->task_p == nullptr) ? YEP : NOPE;
}

int talshTaskDestruct(talsh_task_t *tt) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!tt) return TALSH_INVALID_ARGS;
  Task *t = static_cast<Task *>(tt->task_p);
  if (t) {
    const int st = task_finalize_if_done(tt, false);
    if (st != TALSH_TASK_COMPLETED && st != TALSH_TASK_ERROR) return TALSH_IN_PROGRESS;
    task_free(t);
  }
  return talshTaskClean(tt);
}

int talshTaskDestroy(talsh_task_t *tt) {
  if (!tt) return TALSH_INVALID_ARGS;
  const int e = talshTaskDestruct(tt);
  if (e == TALSH_IN_PROGRESS) return e;
  std::free(tt);
  return e;
}

int talshTaskDevId(talsh_task_t *tt, int *dev_kind) {
  if (!tt) return DEV_NULL;
  Task *t = static_cast<Task *>(tt->task_p);
  if (!t) {
    if (dev_kind) *dev_kind = DEV_NULL;
    return DEV_NULL;
  }
  if (dev_kind) {
    *dev_kind = DEV_NVIDIA_GPU;
    return t->dev;
  }
  return flat_gpu(t->dev);
}
int talshTaskDevId_(talsh_task_t *tt, int *dev_kind) { return talshTaskDevId(tt, dev_kind); }

int talshTaskArgCoherence(const talsh_task_t *tt) { return tt ? tt->coherence : -1; }

const talshTensArg_t *talshTaskTensArgs(const talsh_task_t *tt, int *num_args) {
  if (!tt) return nullptr;
  if (num_args) *num_args = tt->num_args;
  return tt->tens_args;
}

int talshTaskStatus(talsh_task_t *tt) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!tt) return TALSH_INVALID_ARGS;
  if (!tt->task_p) return TALSH_TASK_EMPTY;
  return task_finalize_if_done(tt, false);
}

int talshTaskComplete(talsh_task_t *tt, int *stats, int *ierr) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (ierr) *ierr = TALSH_SUCCESS;
  if (!tt) {
    if (ierr) *ierr = TALSH_INVALID_ARGS;
    return NOPE;
  }
  if (!tt->task_p) {
    if (stats) *stats = TALSH_TASK_EMPTY;
    if (ierr) *ierr = TALSH_OBJECT_IS_EMPTY;
    return NOPE;
  }
  const int st = task_finalize_if_done(tt, false);
  if (stats) *stats = st;
  if (st == TALSH_TASK_ERROR && ierr) *ierr = TALSH_TASK_ERROR;
  return (st == TALSH_TASK_COMPLETED || st == TALSH_TASK_ERROR) ? YEP : NOPE;
}

int talshTaskWait(talsh_task_t *tt, int *stats) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!tt) return TALSH_INVALID_ARGS;
  if (!tt->task_p) {
    if (stats) *stats = TALSH_TASK_EMPTY;
    return TALSH_OBJECT_IS_EMPTY;
  }
  const int st = task_finalize_if_done(tt, true);
  if (stats) *stats = st;
  return TALSH_SUCCESS;
}

int talshTasksWait(int ntasks, talsh_task_t talsh_tasks[], int stats[]) {
  if (ntasks < 0 || (ntasks > 0 && (!talsh_tasks || !stats))) return TALSH_INVALID_ARGS;
  for (int i = 0; i < ntasks; ++i) {
    const int e = talshTaskWait(&talsh_tasks[i], &stats[i]);
    if (e != TALSH_SUCCESS && e != TALSH_OBJECT_IS_EMPTY) return e;
  }
  return TALSH_SUCCESS;
}

int talshTaskTime(talsh_task_t *tt, double *total, double *comput, double *input, double *output, double *mmul) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!tt || !total) return TALSH_INVALID_ARGS;
  Task *t = static_cast<Task *>(tt->task_p);
  if (!t) return TALSH_OBJECT_IS_EMPTY;
  const int st = task_finalize_if_done(tt, false);
  if (st != TALSH_TASK_COMPLETED) return TALSH_IN_PROGRESS;
  auto el = [](const EvRef &a, const EvRef &b) {
    float ms = 0;
    if (!a || !b || !a->ev || !b->ev) return 0.0;
    if (hipEventElapsedTime(&ms, a->ev, b->ev) != hipSuccess) return 0.0;
    return (double)ms * 1e-3;
  };
  *total = el(t->ev_start, t->ev_done);
  if (input) *input = el(t->ev_start, t->ev_in);
  // compute: the kernels' own span when they carried dispatch timestamps
  if (comput) *comput = t->kclock ? el(t->ev_k0, t->ev_k1) : el(t->ev_in, t->ev_comp);
  if (output) *output = el(t->ev_comp, t->ev_done);
  if (mmul) *mmul = (t->ev_mm0 && t->ev_mm1) ? el(t->ev_mm0, t->ev_mm1) : 0.0;
  return TALSH_SUCCESS;
}
int talshTaskTime_(talsh_task_t *tt, double *total, double *comput, double *input, double *output, double *mmul) {
  return talshTaskTime(tt, total, comput, input, output, mmul);
}

void talshTaskPrint(const talsh_task_t *tt) {
  if (!tt) return;
  std::printf("#MSG(TAL-SH::talsh_task): dev_kind %d, data_kind %d, coherence %d, args %d, error %d, flops %e\n",
              tt->dev_kind, tt->data_kind, tt->coherence, tt->num_args, tt->task_error, tt->flops);
}

}  // extern "C"

// Tensor operations of the C API: Init, Scale, Copy, Add, Contract, Norm2,
// Place, Discard.  Each call validates and plans on the host, acquires all
// device buffers up front (rolling back on TRY_LATER / DEVICE_UNABLE, which
// leave the task empty, as the reference does), then enqueues transfers and
// kernels on one HIP stream of the execution GPU and returns; completion and
// coherence bookkeeping happen when the task is observed complete
// (talshTaskComplete / talshTaskWait), or before returning for a NULL task.
//
// Reference counterparts: talshTensorInit/Scale (talshc.cpp:3562-4529),
// talshTensorCopy (:4530-4787), talshTensorAdd (:4795-5066),
// talshTensorContract (:5073-5385) and their GPU back ends
// gpu_tensor_block_{init,scale,copy,add,contract_dlf,place}
// (tensor_algebra_gpu_nvidia.cu:4089-7188); coherence letters as in
// DOC/TALSH_manual.txt:719-750 (D discard, M move, T temporary, K keep).
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "contract.hpp"
#include "plan.hpp"
#include "runtime.hpp"
#include "talsh_amd.h"

namespace tal {

namespace {

int coh_letter(int coh_ctrl, int nargs, int arg) { return (coh_ctrl >> (2 * (nargs - 1 - arg))) & 3; }

struct Operand {
  talsh_tens_t *t = nullptr;
  int coh = COPY_T;
  bool output = false;
  bool read = true;       // the operation reads the operand's value
  int img = -1;           // source image index at scheduling time
  void *src_p = nullptr;  // source image body
  int src_dev = -1;       // canonical flat id of the source image
  bool resident = false;  // source image is on the execution GPU
  void *dptr = nullptr;   // body used by the kernels
  int dhandle = -1;       // pool handle when a buffer was allocated on the GPU
  int dk = NO_TYPE;
  size_t bytes = 0;
};

class Op {
 public:
  Op(int g, talsh_task_t *user_task, const char *name) : g_(g), user_(user_task), name_(name) {}
  ~Op() {
    kernel_clock() = KernelClock();
    if (!committed_) rollback();
  }

  int add(talsh_tens_t *t, int coh, bool output, bool read) {
    if (!t || talshTensorIsEmpty(t) == YEP || tensor_healthy(t) != YEP) return TALSH_INVALID_ARGS;
    Operand o;
    o.t = t;
    o.coh = coh;
    o.output = output;
    o.read = read;
    const int exec = flat_gpu(g_);
    int pick = -1;
    for (int i = 0; i < t->ndev; ++i)
      if (t->avail[i] == YEP && t->dev_rsc[i].dev_id == exec) { pick = i; break; }
    if (pick < 0)
      for (int i = 0; i < t->ndev; ++i)
        if (t->avail[i] == YEP && t->dev_rsc[i].dev_id == DEV_HOST) { pick = i; break; }
    if (pick < 0)
      for (int i = 0; i < t->ndev; ++i)
        if (t->avail[i] == YEP && gpu_index(t->dev_rsc[i].dev_id) >= 0) { pick = i; break; }
    if (pick < 0) return TALSH_NOT_AVAILABLE;
    o.img = pick;
    o.src_p = t->dev_rsc[pick].gmem_p;
    o.src_dev = t->dev_rsc[pick].dev_id;
    o.resident = (o.src_dev == exec);
    o.dk = t->data_kind[pick];
    o.bytes = shape_volume(t->shape_p) * (size_t)dk_size(o.dk);
    ops_.push_back(o);
    return TALSH_SUCCESS;
  }

  Operand &operand(int i) { return ops_[i]; }

  // Device buffers: one per non-resident operand plus the requested temporaries.
  int acquire(const std::vector<size_t> &temp_bytes) {
    for (auto &o : ops_) {
      if (o.resident) {
        o.dptr = o.src_p;
        continue;
      }
      const int e = dev_alloc(g_, o.bytes, &o.dptr, &o.dhandle);
      if (e) return e;
    }
    for (size_t b : temp_bytes) {
      void *p = nullptr;
      int h = -1;
      const int e = dev_alloc(g_, b, &p, &h);
      if (e) return e;
      temps_.push_back(p);
      temp_handles_.push_back(h);
      temp_bytes_.push_back(b);
    }
    return TALSH_SUCCESS;
  }
  void *temp(int i) const { return temps_[i]; }

  // Starts the task: stream, ordering against earlier work, input transfers.
  // clock: the kernels launched until end() carry dispatch timestamps, which
  // talshTaskTime reports as the compute time.
  int begin(int data_kind, int coh_ctrl, bool clock = false) {
    Device &d = rt().gpus[g_];
    (void)hipSetDevice(d.hip_id);
    // stream choice weight: host-link bytes, plus device bytes at the HBM/link speed ratio
    double weight = 0.0;
    for (auto &o : ops_) {
      const bool xfer_in = !o.resident && o.read;
      const bool xfer_out = !o.resident && o.output && (o.coh == COPY_T || o.coh == COPY_K);
      weight += (double)o.bytes * ((xfer_in ? 1.0 : 0.0) + (xfer_out ? 1.0 : 0.0) + 1.0 / 64.0);
    }
    for (size_t i = 0; i < temps_.size(); ++i) weight += 2.0 * (double)temp_bytes_[i] / 64.0;
    task_ = task_new(g_, weight);
    s_ = task_->stream;
    (void)hipEventRecord(task_->ev_start->ev, s_);
    bool moved = false;
    for (auto &o : ops_) {
      if (o.resident) {
        if (o.output) hazard_wait_write(o.dptr, o.bytes, s_);
        else hazard_wait_read(o.dptr, o.bytes, s_);
      } else {
        hazard_wait_read(o.src_p, o.bytes, s_);
        if (o.read) {
          if (hipMemcpyAsync(o.dptr, o.src_p, o.bytes, hipMemcpyDefault, s_) != hipSuccess) return TALSH_FAILURE;
          task_->bytes += (double)o.bytes;
          d.traffic_in += (double)o.bytes;
          moved = true;
        }
      }
    }
    for (size_t i = 0; i < temps_.size(); ++i) hazard_wait_write(temps_[i], temp_bytes_[i], s_);
    // with no input transfer the input stage is empty: its end is the start
    // (one event record less on the launch path of a short task)
    if (moved) (void)hipEventRecord(task_->ev_in->ev, s_);
    else task_->ev_in = task_->ev_start;
    if (clock) {
      task_->ev_k0 = new_event(g_);
      task_->ev_k1 = new_event(g_);
      KernelClock &c = kernel_clock();
      c.on = true;
      c.started = false;
      c.t0 = task_->ev_k0->ev;
      c.t1 = task_->ev_k1->ev;
    }
    data_kind_ = data_kind;
    coh_ctrl_ = coh_ctrl;
    return TALSH_SUCCESS;
  }

  hipStream_t stream() const { return s_; }
  Task *task() { return task_; }

  // Events that bracket the GEMM (talshTaskTime's mmul); the enqueuer records them.
  void gemm_events() {
    task_->ev_mm0 = new_event(g_);
    task_->ev_mm1 = new_event(g_);
  }

  // Ends the task: output transfers, completion event, coherence bookkeeping.
  int end(double flops) {
    Device &d = rt().gpus[g_];
    KernelClock &clk = kernel_clock();
    task_->kclock = clk.on && clk.started;
    clk = KernelClock();
    bool back = false;
    for (auto &o : ops_) back = back || copied_back(o);
    // likewise an empty output stage starts where the task ends
    if (back) (void)hipEventRecord(task_->ev_comp->ev, s_);
    else task_->ev_comp = task_->ev_done;
    for (auto &o : ops_) {
      if (copied_back(o)) {
        hazard_wait_write(o.src_p, o.bytes, s_);
        if (hipMemcpyAsync(o.src_p, o.dptr, o.bytes, hipMemcpyDefault, s_) != hipSuccess) return TALSH_FAILURE;
        task_->bytes += (double)o.bytes;
        d.traffic_out += (double)o.bytes;
      }
    }
    (void)hipEventRecord(task_->ev_done->ev, s_);
    if (hipGetLastError() != hipSuccess) return TALSH_FAILURE;
    // ordering state of every buffer touched; a non-resident output's source
    // image is written only when the result is copied back to it
    for (auto &o : ops_) {
      if (o.output) hazard_set_writer(o.dptr, o.bytes, task_->ev_done);
      else hazard_add_reader(o.dptr, o.bytes, task_->ev_done);
      if (o.resident) continue;
      if (copied_back(o)) hazard_set_writer(o.src_p, o.bytes, task_->ev_done);
      else hazard_add_reader(o.src_p, o.bytes, task_->ev_done);
    }
    for (size_t i = 0; i < temps_.size(); ++i) hazard_set_writer(temps_[i], temp_bytes_[i], task_->ev_done);
    // output images other than the source become stale
    for (auto &o : ops_)
      if (o.output) {
        image_discard_other(o.t, o.img);
        o.img = 0;
      }
    // images in use
    for (auto &o : ops_) {
      const int i = find_image(o.t, o.src_p);
      if (i < 0) continue;
      if (o.output || o.coh == COPY_D || (o.coh == COPY_M && !o.resident)) o.t->avail[i] = NOPE;
    }
    // deferred bookkeeping
    const int g = g_;
    std::vector<Operand> ops = ops_;
    std::vector<void *> temps = temps_;
    std::vector<int> th = temp_handles_;
    task_->on_finalize.push_back([g, ops, temps, th]() {
      for (size_t i = 0; i < temps.size(); ++i) dev_free(g, temps[i], th[i]);
      for (const auto &o : ops) finalize_operand(g, o);
    });
    std::vector<talsh_tens_t *> tens;
    for (auto &o : ops_) tens.push_back(o.t);
    task_register(task_, tens);
    task_->flops = flops;
    d.tasks_submitted++;
    d.active++;
    committed_ = true;
    return TALSH_SUCCESS;
  }

  // Publishes the task through the user's handle, or waits for it when the
  // call is blocking.
  int publish(int nargs) {
    talsh_task_t local;
    talsh_task_t *tt = user_ ? user_ : &local;
    if (!user_) talshTaskClean(&local);
    tt->task_p = task_;
    tt->task_error = -1;
    tt->dev_kind = DEV_NVIDIA_GPU;
    tt->data_kind = data_kind_;
    tt->coherence = coh_ctrl_;
    tt->num_args = nargs;
    for (int i = 0; i < nargs && i < MAX_TENSOR_OPERANDS; ++i) {
      tt->tens_args[i].tens_p = ops_[i].t;
      tt->tens_args[i].source_image = ops_[i].img;
    }
    tt->flops = task_->flops;
    tt->data_vol = task_->bytes;
    if (rt().log_ops)
      std::printf("#TALSH OP: %s on GPU %d: %d operands, %.6g flop, %.6g bytes moved\n", name_, g_, nargs,
                  task_->flops, task_->bytes);
    task_ = nullptr;
    if (user_) return TALSH_SUCCESS;
    const int st = task_finalize_if_done(tt, true);
    talshTaskDestruct(tt);
    return st == TALSH_TASK_COMPLETED ? TALSH_SUCCESS : TALSH_TASK_ERROR;
  }

 private:
  static bool copied_back(const Operand &o) {
    return o.output && !o.resident && (o.coh == COPY_T || o.coh == COPY_K);
  }

  static int find_image(const talsh_tens_t *t, const void *p) {
    for (int i = 0; i < t->ndev; ++i)
      if (t->dev_rsc[i].gmem_p == p) return i;
    return -1;
  }

  // Coherence work of one operand.  The tensor is alive here: destructing a
  // tensor first finalizes the tasks that name it.  A buffer the task owns is
  // released even if the tensor was emptied by other means meanwhile.
  static void finalize_operand(int g, const Operand &o) {
    talsh_tens_t *t = o.t;
    const bool owns = !o.resident && o.dptr && o.dhandle >= 0;
    if (talshTensorIsEmpty(t) == YEP || tensor_healthy(t) != YEP) {
      if (owns) dev_free(g, o.dptr, o.dhandle);
      return;
    }
    const int si = find_image(t, o.src_p);
    if (o.resident) {
      if (si >= 0) {
        if (o.coh == COPY_D) image_discard(t, si);
        else t->avail[si] = YEP;
      }
      return;
    }
    switch (o.coh) {
      case COPY_D:
        dev_free(g, o.dptr, o.dhandle);
        if (si >= 0) image_discard(t, si);
        break;
      case COPY_M:
        if (si >= 0) image_discard(t, si);
        image_add(t, flat_gpu(g), o.dptr, o.dhandle, 0, o.dk);
        break;
      case COPY_T:
        dev_free(g, o.dptr, o.dhandle);
        if (si >= 0) t->avail[si] = YEP;
        break;
      default:  // COPY_K
        if (si >= 0) t->avail[si] = YEP;
        image_add(t, flat_gpu(g), o.dptr, o.dhandle, 0, o.dk);
        break;
    }
  }

  void rollback() {
    if (s_) (void)hipStreamSynchronize(s_);  // transfers already queued must land before buffers are reused
    for (auto &o : ops_)
      if (!o.resident && o.dptr && o.dhandle >= 0) rt().gpus[g_].arena.release(o.dhandle);
    for (size_t i = 0; i < temps_.size(); ++i) rt().gpus[g_].arena.release(temp_handles_[i]);
    if (task_) task_free(task_);
    task_ = nullptr;
  }

  int g_;
  talsh_task_t *user_;
  const char *name_;
  std::vector<Operand> ops_;
  std::vector<void *> temps_;
  std::vector<int> temp_handles_;
  std::vector<size_t> temp_bytes_;
  Task *task_ = nullptr;
  hipStream_t s_ = nullptr;
  bool committed_ = false;
  int data_kind_ = NO_TYPE, coh_ctrl_ = 0;
};

}  // namespace

// Execution GPU for (dev_id, dev_kind); operands are consulted when the
// caller leaves the choice to the runtime.  Host execution is not provided.
int exec_gpu(int dev_id, int dev_kind, std::initializer_list<const talsh_tens_t *> tens, int *g) {
  Runtime &r = rt();
  int flat;
  if (dev_kind == DEV_DEFAULT) {
    if (dev_id == DEV_DEFAULT) {
      for (const talsh_tens_t *t : tens) {
        if (!t || !t->dev_rsc) continue;
        for (int i = 0; i < t->ndev; ++i) {
          const int gi = gpu_index(t->dev_rsc[i].dev_id);
          if (gi >= 0) { *g = gi; return TALSH_SUCCESS; }
        }
      }
      const int b = talshDeviceBusyLeast(DEV_NVIDIA_GPU);
      if (b < 0) return TALSH_NOT_AVAILABLE;
      *g = b;
      return TALSH_SUCCESS;
    }
    flat = dev_id;
  } else if (dev_kind == DEV_HOST) {
    return TALSH_NOT_AVAILABLE;
  } else if (dev_kind == DEV_NVIDIA_GPU || dev_kind == DEV_AMD_GPU) {
    if (dev_id == DEV_DEFAULT) {
      const int b = talshDeviceBusyLeast(DEV_NVIDIA_GPU);
      if (b < 0) return TALSH_NOT_AVAILABLE;
      *g = b;
      return TALSH_SUCCESS;
    }
    flat = talshFlatDevId(dev_kind, dev_id);
  } else {
    return TALSH_NOT_AVAILABLE;
  }
  if (canonical_dev(flat) == DEV_HOST) return TALSH_NOT_AVAILABLE;
  const int gi = gpu_index(flat);
  if (gi < 0) return TALSH_INVALID_ARGS;
  (void)r;
  *g = gi;
  return TALSH_SUCCESS;
}

namespace {

int check_task(talsh_task_t *tt) {
  if (tt && talshTaskIsEmpty(tt) != YEP) return TALSH_OBJECT_NOT_EMPTY;
  return TALSH_SUCCESS;
}

int deferred(int g, int e) {
  if (e == TRY_LATER || e == DEVICE_UNABLE) rt().gpus[g].tasks_deferred++;
  return e;
}

std::vector<int> dims_of(const talsh_tens_t *t) {
  return std::vector<int>(t->shape_p->dims, t->shape_p->dims + t->shape_p->num_dim);
}

}  // namespace

namespace {

// Whether the GEMM writes a permuted D in place through a C map instead of
// transposing D into a temporary and back (the planner's d_in_place: D's
// stride-1 dim leads D' with an extent of at least 16, so that the GEMM's and
// the K-split summing pass's stores, runs along m, stay contiguous; no batch),
// and the leading dims of D' multiply to exactly M and the next ones to N.
// Sets the number of D' dims in each group.
bool d_in_place(const ContractPlan &plan, int *nm, int *nn) {
  if (!plan.d_in_place || plan.d_o2n.empty() || plan.dprime_dims.empty()) return false;
  const int r = (int)plan.dprime_dims.size();
  if (r > kMaxMapDims) return false;
  int t = 0;
  int64_t pm = 1, pn = 1;
  while (t < r && pm < plan.M) pm *= plan.dprime_dims[t++];
  *nm = t;
  while (t < r && pn < plan.N) pn *= plan.dprime_dims[t++];
  *nn = t - *nm;
  return pm == plan.M && pn == plan.N;
}

// C maps of in-place GEMMs, kept on each device for shapes that recur: a map
// depends only on D's extents and the plan, so a repeated contraction reads
// the stored table instead of launching offset_map_kernel (a launch and a
// few microseconds on a short contraction's critical path).  A shape's
// second use writes its table with that kernel into a per-device chunk, on
// the task's stream, and records an event that later users on other streams
// wait for; shapes seen once (most of a mixed serving stream) cost nothing.
// The chunk is bounded (further shapes take the per-call kernel); freed at
// shutdown.
constexpr size_t kCmapChunkBytes = (size_t)64 << 20;
using CmapKey = std::pair<int, std::vector<int64_t>>;
struct CmapEntry {
  int64_t *d = nullptr;
  hipEvent_t ready = nullptr;
};
struct CmapChunk {
  char *base = nullptr;
  size_t used = 0;
  bool failed = false;
};
struct CmapCache {
  std::mutex mu;
  std::map<CmapKey, int> seen;  // uses of shapes not stored yet
  std::map<CmapKey, CmapEntry> maps;
  std::map<int, CmapChunk> chunks;  // per HIP device
};
CmapCache &cmap_cache() {
  static CmapCache c;
  return c;
}

const int64_t *cached_cmap(const OffsetMap &map, hipStream_t s) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  const int nt = map.nm + map.nn;
  CmapKey k;
  k.first = dev;
  k.second.reserve(2 + 2 * (size_t)nt);
  k.second.push_back(map.nm);
  k.second.push_back(map.nn);
  k.second.insert(k.second.end(), map.ext, map.ext + nt);
  k.second.insert(k.second.end(), map.str, map.str + nt);
  CmapCache &c = cmap_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.maps.find(k);
  if (it != c.maps.end()) {
    if (hipEventQuery(it->second.ready) == hipErrorNotReady) (void)hipStreamWaitEvent(s, it->second.ready, 0);
    return it->second.d;
  }
  if (c.seen.size() >= 4096) c.seen.clear();
  if (++c.seen[k] < 2) return nullptr;
  CmapChunk &ch = c.chunks[dev];
  const size_t bytes = ((size_t)(map.M + map.N) * sizeof(int64_t) + 255) / 256 * 256;
  if (!ch.base && !ch.failed && hipMalloc((void **)&ch.base, kCmapChunkBytes) != hipSuccess) {
    ch.base = nullptr;
    ch.failed = true;
  }
  if (!ch.base || ch.used + bytes > kCmapChunkBytes) return nullptr;
  CmapEntry en;
  en.d = (int64_t *)(ch.base + ch.used);
  if (hipEventCreateWithFlags(&en.ready, hipEventDisableTiming) != hipSuccess) return nullptr;
  if (launch_offset_map(map, en.d, s) || hipEventRecord(en.ready, s) != hipSuccess) {
    (void)hipEventDestroy(en.ready);
    return nullptr;
  }
  ch.used += bytes;
  c.seen.erase(k);
  c.maps.emplace(std::move(k), en);
  return en.d;
}

}  // namespace

void ttgt_cache_clear() {
  CmapCache &c = cmap_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  for (auto &m : c.maps) (void)hipEventDestroy(m.second.ready);
  for (auto &ch : c.chunks)
    if (ch.second.base) (void)hipFree(ch.second.base);
  c.maps.clear();
  c.seen.clear();
  c.chunks.clear();
}

TtgtTemps ttgt_temp_bytes(const ContractPlan &plan, int dk, size_t lbytes, size_t rbytes, size_t dbytes) {
  TtgtTemps t;
  if (plan.a.transpose) t.a = plan.swap ? rbytes : lbytes;
  if (plan.b.transpose) t.b = plan.swap ? lbytes : rbytes;
  int nm = 0, nn = 0;
  if (d_in_place(plan, &nm, &nn)) t.d = (size_t)(plan.M + plan.N) * sizeof(int64_t);  // the C map
  else if (plan.d_transpose) t.d = dbytes;
  if (dk == R4 && rt().r4mode == R4_SPLIT_BF16) {
    t.split = rt().split_method;
    t.w = gemm_split_workspace(t.split, plan.M, plan.N, plan.K);
  } else {
    t.s = gemm_tail_workspace(dk, plan.M, plan.N, plan.K, plan.batch);
  }
  return t;
}

int enqueue_ttgt(const DigitalPattern &p, const ContractPlan &plan, int dk, bool acc, double ar, double ai,
                 void *D, const void *L, const void *R, const int *ddims, const int *ldims, const int *rdims,
                 const TtgtBuffers &tmp, hipStream_t s, hipEvent_t gemm_start, hipEvent_t gemm_end) {
  const bool cplx = (dk == C4 || dk == C8);
  const void *X = plan.swap ? R : L, *Y = plan.swap ? L : R;
  const int *xd = plan.swap ? rdims : ldims, *yd = plan.swap ? ldims : rdims;
  const int xr = plan.swap ? p.rrank : p.lrank, yr = plan.swap ? p.lrank : p.rrank;
  const void *A = X, *B = Y;
  void *C = D;
  if (plan.a.transpose && plan.b.transpose) {  // both operand transposes in one launch where they can
    if (!tmp.a || !tmp.b ||
        launch_permute_pair(dk, X, tmp.a, xr, xd, plan.a.o2n.data(), Y, tmp.b, yr, yd, plan.b.o2n.data(), s))
      return TALSH_FAILURE;
    A = tmp.a;
    B = tmp.b;
  } else if (plan.a.transpose) {
    if (!tmp.a || launch_permute(dk, X, tmp.a, xr, xd, plan.a.o2n.data(), PERM_COPY, 1, 0, 0, s))
      return TALSH_FAILURE;
    A = tmp.a;
  } else if (plan.b.transpose) {
    if (!tmp.b || launch_permute(dk, Y, tmp.b, yr, yd, plan.b.o2n.data(), PERM_COPY, 1, 0, 0, s))
      return TALSH_FAILURE;
    B = tmp.b;
  }
  int nm = 0, nn = 0;
  const bool in_place = d_in_place(plan, &nm, &nn);
  const int64_t *cmap = nullptr;
  if (in_place) {
    // D' dim t is D dim i with d_o2n[i] == t; its D stride is the product of the extents before i
    if (!tmp.d || (int)plan.d_o2n.size() != p.drank) return TALSH_FAILURE;
    OffsetMap map;
    map.nm = nm;
    map.nn = nn;
    map.M = plan.M;
    map.N = plan.N;
    int64_t stride = 1;
    for (int i = 0; i < p.drank; ++i) {
      const int t = plan.d_o2n[i];
      if (t < nm + nn) {
        map.ext[t] = ddims[i];
        map.str[t] = stride;
      }
      stride *= ddims[i];
    }
    cmap = cached_cmap(map, s);
    if (!cmap) {
      if (launch_offset_map(map, (int64_t *)tmp.d, s)) return TALSH_FAILURE;
      cmap = (const int64_t *)tmp.d;
    }
  } else if (plan.d_transpose) {
    if (!tmp.d) return TALSH_FAILURE;
    C = tmp.d;
    if (acc && launch_permute(dk, D, C, p.drank, ddims, plan.d_o2n.data(), PERM_COPY, 1, 0, 0, s))
      return TALSH_FAILURE;
  }
  int lconj = (p.conj >> 1) & 1, rconj = (p.conj >> 2) & 1;
  if (p.conj & 1) { lconj ^= 1; rconj ^= 1; }
  if (!cplx) lconj = rconj = 0;
  if (gemm_start) (void)hipEventRecord(gemm_start, s);
  const bool split = tmp.split >= 0;
  // hyper indices: plan.batch products of M x N x K, each operand strided by one matrix
  const int64_t sa = plan.M * plan.K, sb = plan.K * plan.N, sc = plan.M * plan.N;
  if (split) {
    if (!tmp.w) return TALSH_FAILURE;
    for (int64_t q = 0; q < plan.batch; ++q)  // the planes workspace is reused in stream order
      if (launch_gemm_split(tmp.split, plan.a.op, plan.b.op, plan.M, plan.N, plan.K, ar,
                            (const float *)A + q * sa, plan.a.ld,
                            (const float *)B + q * sb, plan.b.ld, acc ? 1 : 0, (float *)C + q * sc,
                            std::max<int64_t>(1, plan.M), tmp.w, s, cmap))
        return TALSH_FAILURE;
  } else if (launch_gemm(dk, R4_FP32, plan.a.op, plan.b.op, plan.swap ? rconj : lconj, plan.swap ? lconj : rconj,
                         plan.M, plan.N, plan.K, ar, cplx ? ai : 0.0, A, plan.a.ld, B, plan.b.ld, acc ? 1 : 0, C,
                         std::max<int64_t>(1, plan.M), s, plan.batch, sa, sb, sc, tmp.s, tmp.s_bytes, cmap)) {
    return TALSH_FAILURE;
  }
  if (gemm_end) (void)hipEventRecord(gemm_end, s);
  if (plan.d_transpose && !in_place &&
      launch_permute(dk, C, D, (int)plan.dprime_dims.size(), plan.dprime_dims.data(), plan.d_back_o2n.data(),
                     PERM_COPY, 1, 0, 0, s))
    return TALSH_FAILURE;
  return TALSH_SUCCESS;
}

}  // namespace tal

using namespace tal;

extern "C" {

int talshTensorInit(talsh_tens_t *dtens, double val_real, double val_imag, int dev_id, int dev_kind, int copy_ctrl,
                    talsh_task_t *talsh_task) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (int e = check_task(talsh_task)) return e;
  int g;
  if (int e = exec_gpu(dev_id, dev_kind, {dtens}, &g)) return e;
  Op op(g, talsh_task, __func__);
  if (int e = op.add(dtens, coh_letter(copy_ctrl, 1, 0), true, false)) return e;
  if (int e = op.acquire({})) return deferred(g, e);
  Operand &d = op.operand(0);
  if (int e = op.begin(d.dk, copy_ctrl, true)) return e;
  if (launch_init(d.dk, d.dptr, shape_volume(dtens->shape_p), val_real, val_imag, op.stream())) return TALSH_FAILURE;
  if (int e = op.end(0.0)) return e;
  return op.publish(1);
}
int talshTensorInit_(talsh_tens_t *dtens, double val_real, double val_imag, int dev_id, int dev_kind,
                     int copy_ctrl, talsh_task_t *talsh_task) {
  return talshTensorInit(dtens, val_real, val_imag, dev_id, dev_kind, copy_ctrl, talsh_task);
}

int talshTensorScale(talsh_tens_t *dtens, double val_real, double val_imag, int dev_id, int dev_kind,
                     int copy_ctrl, talsh_task_t *talsh_task) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (int e = check_task(talsh_task)) return e;
  int g;
  if (int e = exec_gpu(dev_id, dev_kind, {dtens}, &g)) return e;
  Op op(g, talsh_task, __func__);
  if (int e = op.add(dtens, coh_letter(copy_ctrl, 1, 0), true, true)) return e;
  if (int e = op.acquire({})) return deferred(g, e);
  Operand &d = op.operand(0);
  if (int e = op.begin(d.dk, copy_ctrl, true)) return e;
  const size_t vol = shape_volume(dtens->shape_p);
  if (launch_scale(d.dk, d.dptr, vol, val_real, val_imag, op.stream())) return TALSH_FAILURE;
  if (int e = op.end((d.dk == C4 || d.dk == C8) ? 6.0 * vol : 1.0 * vol)) return e;
  return op.publish(1);
}
int talshTensorScale_(talsh_tens_t *dtens, double val_real, double val_imag, int dev_id, int dev_kind,
                      int copy_ctrl, talsh_task_t *talsh_task) {
  return talshTensorScale(dtens, val_real, val_imag, dev_id, dev_kind, copy_ctrl, talsh_task);
}

// Unary pattern -> 0-based O2N permutation of L into D, with shape checks.
static int unary_perm(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, std::vector<int> *o2n, int *conj) {
  DigitalPattern p;
  if (parse_pattern(cptrn, &p) != 0) return TALSH_INVALID_ARGS;
  if (p.rrank != 0 || p.drank != p.lrank) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(dtens) == YEP || talshTensorIsEmpty(ltens) == YEP) return TALSH_OBJECT_IS_EMPTY;
  if (dtens->shape_p->num_dim != p.drank || ltens->shape_p->num_dim != p.lrank) return TALSH_INVALID_ARGS;
  o2n->resize(p.lrank);
  for (int i = 0; i < p.lrank; ++i) {
    const int t = p.dig[i];
    if (t < 1 || t > p.drank || ltens->shape_p->dims[i] != dtens->shape_p->dims[t - 1]) return TALSH_INVALID_ARGS;
    (*o2n)[i] = t - 1;
  }
  *conj = ((p.conj >> 1) & 1) ^ (p.conj & 1);
  return TALSH_SUCCESS;
}

int talshTensorCopy(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, int dev_id, int dev_kind,
                    int copy_ctrl, talsh_task_t *talsh_task) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (!cptrn || !dtens || !ltens) return TALSH_INVALID_ARGS;
  if (int e = check_task(talsh_task)) return e;
  std::vector<int> o2n;
  int conj = 0;
  if (int e = unary_perm(cptrn, dtens, ltens, &o2n, &conj)) return e;
  int g;
  if (int e = exec_gpu(dev_id, dev_kind, {dtens, ltens}, &g)) return e;
  Op op(g, talsh_task, __func__);
  if (int e = op.add(dtens, coh_letter(copy_ctrl, 2, 0), true, false)) return e;
  if (int e = op.add(ltens, coh_letter(copy_ctrl, 2, 1), false, true)) return e;
  if (op.operand(0).dk != op.operand(1).dk) return TALSH_INVALID_ARGS;
  if (int e = op.acquire({})) return deferred(g, e);
  const int dk = op.operand(0).dk;
  if (int e = op.begin(dk, copy_ctrl, true)) return e;
  std::vector<int> ld = dims_of(ltens);
  if (launch_permute(dk, op.operand(1).dptr, op.operand(0).dptr, (int)ld.size(), ld.data(), o2n.data(), PERM_COPY,
                     1.0, 0.0, conj, op.stream()))
    return TALSH_FAILURE;
  if (int e = op.end(0.0)) return e;
  return op.publish(2);
}
int talshTensorCopy_(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, int dev_id, int dev_kind,
                     int copy_ctrl, talsh_task_t *talsh_task) {
  return talshTensorCopy(cptrn, dtens, ltens, dev_id, dev_kind, copy_ctrl, talsh_task);
}

int talshTensorAdd(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, double scale_real,
                   double scale_imag, int dev_id, int dev_kind, int copy_ctrl, talsh_task_t *talsh_task) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (!cptrn || !dtens || !ltens) return TALSH_INVALID_ARGS;
  if (int e = check_task(talsh_task)) return e;
  std::vector<int> o2n;
  int conj = 0;
  if (int e = unary_perm(cptrn, dtens, ltens, &o2n, &conj)) return e;
  int g;
  if (int e = exec_gpu(dev_id, dev_kind, {dtens, ltens}, &g)) return e;
  Op op(g, talsh_task, __func__);
  if (int e = op.add(dtens, coh_letter(copy_ctrl, 2, 0), true, true)) return e;
  if (int e = op.add(ltens, coh_letter(copy_ctrl, 2, 1), false, true)) return e;
  if (op.operand(0).dk != op.operand(1).dk) return TALSH_INVALID_ARGS;
  if (int e = op.acquire({})) return deferred(g, e);
  const int dk = op.operand(0).dk;
  if (int e = op.begin(dk, copy_ctrl, true)) return e;
  std::vector<int> ld = dims_of(ltens);
  if (launch_permute(dk, op.operand(1).dptr, op.operand(0).dptr, (int)ld.size(), ld.data(), o2n.data(), PERM_ADD,
                     scale_real, scale_imag, conj, op.stream()))
    return TALSH_FAILURE;
  const double vol = (double)shape_volume(dtens->shape_p);
  if (int e = op.end((dk == C4 || dk == C8) ? 8.0 * vol : 2.0 * vol)) return e;
  return op.publish(2);
}
int talshTensorAdd_(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, double scale_real,
                    double scale_imag, int dev_id, int dev_kind, int copy_ctrl, talsh_task_t *talsh_task) {
  return talshTensorAdd(cptrn, dtens, ltens, scale_real, scale_imag, dev_id, dev_kind, copy_ctrl, talsh_task);
}

// Slice (mode 0: dtens (+)= ltens[offsets + i]) and insert (mode 1:
// dtens[offsets + i] (+)= ltens): talshTensorSlice / talshTensorInsert
// (talshc.cpp:4008-4524) with the reference's CPU semantics
// (tensor_algebra_cpu.F90:2315-2626, 6821-7053); its GPU back ends
// (tensor_algebra_gpu_nvidia.cu:4788-4804) are unimplemented.  An insert
// writes only part of dtens, so dtens is always read (uploaded if needed).
static int slice_insert(int mode, talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets, int dev_id,
                        int dev_kind, int copy_ctrl, int accumulative, talsh_task_t *talsh_task, const char *name) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (!dtens || !ltens) return TALSH_INVALID_ARGS;
  if (int e = check_task(talsh_task)) return e;
  if (talshTensorIsEmpty(dtens) == YEP || talshTensorIsEmpty(ltens) == YEP) return TALSH_OBJECT_IS_EMPTY;
  const int rank = dtens->shape_p->num_dim;
  if (ltens->shape_p->num_dim != rank || (rank > 0 && !offsets)) return TALSH_INVALID_ARGS;
  const talsh_tens_t *small = mode == 0 ? dtens : ltens, *big = mode == 0 ? ltens : dtens;
  for (int k = 0; k < rank; ++k)
    if (offsets[k] < 0 || offsets[k] + small->shape_p->dims[k] > big->shape_p->dims[k]) return TALSH_INVALID_ARGS;
  const bool acc = accumulative != NOPE;
  int g;
  if (int e = exec_gpu(dev_id, dev_kind, {dtens, ltens}, &g)) return e;
  Op op(g, talsh_task, name);
  if (int e = op.add(dtens, coh_letter(copy_ctrl, 2, 0), true, acc || mode == 1)) return e;
  if (int e = op.add(ltens, coh_letter(copy_ctrl, 2, 1), false, true)) return e;
  if (op.operand(0).dk != op.operand(1).dk) return TALSH_INVALID_ARGS;
  if (int e = op.acquire({})) return deferred(g, e);
  const int dk = op.operand(0).dk;
  if (int e = op.begin(dk, copy_ctrl, true)) return e;
  std::vector<int> sd = dims_of(small), bd = dims_of(big);
  if (launch_slice(dk, mode, op.operand(1).dptr, op.operand(0).dptr, rank, sd.data(), bd.data(), offsets,
                   acc ? 1 : 0, op.stream()))
    return TALSH_FAILURE;
  const double vol = (double)shape_volume(small->shape_p);
  if (int e = op.end(acc ? ((dk == C4 || dk == C8) ? 2.0 : 1.0) * vol : 0.0)) return e;
  return op.publish(2);
}

int talshTensorSlice(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets, int dev_id, int dev_kind,
                     int copy_ctrl, int accumulative, talsh_task_t *talsh_task) {
  return slice_insert(0, dtens, ltens, offsets, dev_id, dev_kind, copy_ctrl, accumulative, talsh_task, __func__);
}
int talshTensorSlice_(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets, int dev_id, int dev_kind,
                      int copy_ctrl, int accumulative, talsh_task_t *talsh_task) {
  return talshTensorSlice(dtens, ltens, offsets, dev_id, dev_kind, copy_ctrl, accumulative, talsh_task);
}
int talshTensorInsert(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets, int dev_id, int dev_kind,
                      int copy_ctrl, int accumulative, talsh_task_t *talsh_task) {
  return slice_insert(1, dtens, ltens, offsets, dev_id, dev_kind, copy_ctrl, accumulative, talsh_task, __func__);
}
int talshTensorInsert_(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets, int dev_id, int dev_kind,
                       int copy_ctrl, int accumulative, talsh_task_t *talsh_task) {
  return talshTensorInsert(dtens, ltens, offsets, dev_id, dev_kind, copy_ctrl, accumulative, talsh_task);
}

// Decompositions (talshc.cpp:5650-6290) are Host LAPACK work outside the
// contraction path: arguments are checked as the reference does, then
// TALSH_NOT_IMPLEMENTED is returned.
static int decomposition_args(const char *cptrn, std::initializer_list<const talsh_tens_t *> tens) {
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (!cptrn) return TALSH_INVALID_ARGS;
  for (const talsh_tens_t *t : tens) {
    if (!t) return TALSH_INVALID_ARGS;
    if (talshTensorIsEmpty(t) == YEP) return TALSH_OBJECT_IS_EMPTY;
  }
  return TALSH_NOT_IMPLEMENTED;
}
int talshTensorDecomposeSVD(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                            talsh_tens_t *stens, const char absorb, int dev_id, int dev_kind) {
  (void)absorb, (void)dev_id, (void)dev_kind;
  if (!stens) return TALSH_INVALID_ARGS;
  return decomposition_args(cptrn, {dtens, ltens, rtens});
}
int talshTensorDecomposeSVDL(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                             int dev_id, int dev_kind) {
  (void)dev_id, (void)dev_kind;
  return decomposition_args(cptrn, {dtens, ltens, rtens});
}
int talshTensorDecomposeSVDR(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                             int dev_id, int dev_kind) {
  (void)dev_id, (void)dev_kind;
  return decomposition_args(cptrn, {dtens, ltens, rtens});
}
int talshTensorDecomposeSVDLR(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                              int dev_id, int dev_kind) {
  (void)dev_id, (void)dev_kind;
  return decomposition_args(cptrn, {dtens, ltens, rtens});
}
int talshTensorOrthogonalizeSVD(const char *cptrn, talsh_tens_t *dtens, int dev_id, int dev_kind) {
  (void)dev_id, (void)dev_kind;
  return decomposition_args(cptrn, {dtens});
}
int talshTensorOrthogonalizeMGS(talsh_tens_t *dtens, int num_iso_dims, int *iso_dims, int dev_id, int dev_kind) {
  (void)dev_id, (void)dev_kind;
  if (num_iso_dims < 0 || (num_iso_dims > 0 && !iso_dims)) return TALSH_INVALID_ARGS;
  return decomposition_args("", {dtens});
}

int talshTensorContract(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                        double scale_real, double scale_imag, int dev_id, int dev_kind, int copy_ctrl,
                        int accumulative, talsh_task_t *talsh_task) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (!cptrn || !dtens || !ltens || !rtens) return TALSH_INVALID_ARGS;
  // D is written while L and R are read (in place, when the GEMM writes D through offsets): no aliasing
  if (dtens == ltens || dtens == rtens) return TALSH_INVALID_ARGS;
  if (int e = check_task(talsh_task)) return e;
  if (talshTensorIsEmpty(dtens) == YEP || talshTensorIsEmpty(ltens) == YEP || talshTensorIsEmpty(rtens) == YEP)
    return TALSH_OBJECT_IS_EMPTY;
  DigitalPattern p;
  if (parse_pattern(cptrn, &p) != 0) return TALSH_INVALID_ARGS;
  if (p.drank != dtens->shape_p->num_dim || p.lrank != ltens->shape_p->num_dim || p.rrank != rtens->shape_p->num_dim)
    return TALSH_INVALID_ARGS;
  ContractPlan plan;
  const int pe = plan_contraction(p, dtens->shape_p->dims, ltens->shape_p->dims, rtens->shape_p->dims,
                                  accumulative != NOPE, &plan);
  if (pe != 0) return TALSH_INVALID_ARGS;
  int g;
  if (int e = exec_gpu(dev_id, dev_kind, {dtens, ltens, rtens}, &g)) return e;
  Op op(g, talsh_task, __func__);
  const bool acc = accumulative != NOPE;
  if (int e = op.add(dtens, coh_letter(copy_ctrl, 3, 0), true, acc)) return e;
  if (int e = op.add(ltens, coh_letter(copy_ctrl, 3, 1), false, true)) return e;
  if (int e = op.add(rtens, coh_letter(copy_ctrl, 3, 2), false, true)) return e;
  const int dk = op.operand(0).dk;
  if (op.operand(1).dk != dk || op.operand(2).dk != dk) return TALSH_INVALID_ARGS;
  const bool cplx = (dk == C4 || dk == C8);
  // temporaries: [A'], [B'], [D'], [split-BF16 planes]
  const TtgtTemps tb = ttgt_temp_bytes(plan, dk, op.operand(1).bytes, op.operand(2).bytes, op.operand(0).bytes);
  std::vector<size_t> temps;
  int ia = -1, ib = -1, id = -1, iw = -1, is = -1;
  if (tb.a) { ia = (int)temps.size(); temps.push_back(tb.a); }
  if (tb.b) { ib = (int)temps.size(); temps.push_back(tb.b); }
  if (tb.d) { id = (int)temps.size(); temps.push_back(tb.d); }
  if (tb.w) { iw = (int)temps.size(); temps.push_back(tb.w); }
  if (tb.s) { is = (int)temps.size(); temps.push_back(tb.s); }
  if (int e = op.acquire(temps)) return deferred(g, e);
  if (int e = op.begin(dk, copy_ctrl)) return e;
  TtgtBuffers bufs;
  bufs.a = ia >= 0 ? op.temp(ia) : nullptr;
  bufs.b = ib >= 0 ? op.temp(ib) : nullptr;
  bufs.d = id >= 0 ? op.temp(id) : nullptr;
  bufs.w = iw >= 0 ? op.temp(iw) : nullptr;
  bufs.s = is >= 0 ? op.temp(is) : nullptr;
  bufs.s_bytes = tb.s;
  bufs.split = tb.split;
  op.gemm_events();
  if (int e = enqueue_ttgt(p, plan, dk, acc, scale_real, scale_imag, op.operand(0).dptr, op.operand(1).dptr,
                           op.operand(2).dptr, dtens->shape_p->dims, ltens->shape_p->dims, rtens->shape_p->dims,
                           bufs, op.stream(), op.task()->ev_mm0->ev, op.task()->ev_mm1->ev))
    return e;
  const double flops = (cplx ? 8.0 : 2.0) * (double)plan.M * (double)plan.N * (double)plan.K * (double)plan.batch;
  if (int e = op.end(flops)) return e;
  return op.publish(3);
}
int talshTensorContract_(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                         double scale_real, double scale_imag, int dev_id, int dev_kind, int copy_ctrl,
                         int accumulative, talsh_task_t *talsh_task) {
  return talshTensorContract(cptrn, dtens, ltens, rtens, scale_real, scale_imag, dev_id, dev_kind, copy_ctrl,
                             accumulative, talsh_task);
}

int talshTensorNorm2(talsh_tens_t *dtens, talsh_tens_t *ltens, int dev_id, int dev_kind, int copy_ctrl,
                     talsh_task_t *talsh_task) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (!dtens || !ltens) return TALSH_INVALID_ARGS;
  if (int e = check_task(talsh_task)) return e;
  if (talshTensorIsEmpty(dtens) == YEP || talshTensorIsEmpty(ltens) == YEP) return TALSH_OBJECT_IS_EMPTY;
  if (shape_volume(dtens->shape_p) != 1) return TALSH_INVALID_ARGS;
  int g;
  if (int e = exec_gpu(dev_id, dev_kind, {ltens, dtens}, &g)) return e;
  Op op(g, talsh_task, __func__);
  if (int e = op.add(dtens, coh_letter(copy_ctrl, 2, 0), true, false)) return e;
  if (int e = op.add(ltens, coh_letter(copy_ctrl, 2, 1), false, true)) return e;
  const int dk = op.operand(1).dk;
  if (op.operand(0).dk != dk) return TALSH_INVALID_ARGS;
  if (int e = op.acquire({norm2_scratch_elems() * sizeof(double)})) return deferred(g, e);
  if (int e = op.begin(dk, copy_ctrl, true)) return e;
  const size_t vol = shape_volume(ltens->shape_p);
  if (launch_norm2(dk, op.operand(1).dptr, vol, (double *)op.temp(0), op.operand(0).dptr, op.stream()))
    return TALSH_FAILURE;
  if (int e = op.end((dk == C4 || dk == C8) ? 4.0 * vol : 2.0 * vol)) return e;
  return op.publish(2);
}

// Places a body image on a device (talshc.cpp:3318-3480): COPY_M moves it
// (source image discarded at completion), COPY_K keeps both.
int talshTensorPlace(talsh_tens_t *tens, int dev_id, int dev_kind, void *dev_mem, int copy_ctrl,
                     talsh_task_t *talsh_task) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (!tens) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(tens) == YEP) return TALSH_OBJECT_IS_EMPTY;
  if (tensor_healthy(tens) != YEP) return TALSH_FAILURE;
  if (copy_ctrl != COPY_M && copy_ctrl != COPY_K) return TALSH_INVALID_ARGS;
  if (int e = check_task(talsh_task)) return e;
  const int target = canonical_dev(dev_kind == DEV_NULL ? dev_id : talshFlatDevId(dev_kind, dev_id));
  if (target < 0) return TALSH_INVALID_ARGS;
  const int tg = gpu_index(target);
  if (target != DEV_HOST && tg < 0) return TALSH_INVALID_ARGS;
  // already there?
  int src = -1;
  for (int i = 0; i < tens->ndev; ++i)
    if (tens->avail[i] == YEP && tens->dev_rsc[i].dev_id == target) return TALSH_SUCCESS;
  for (int i = 0; i < tens->ndev; ++i)
    if (tens->avail[i] == YEP && (src < 0 || tens->dev_rsc[i].dev_id == DEV_HOST)) src = i;
  if (src < 0) return TALSH_NOT_FOUND;
  const int sg = gpu_index(tens->dev_rsc[src].dev_id);
  const int g = tg >= 0 ? tg : sg;  // GPU that drives the transfer
  if (g < 0) return TALSH_FAILURE;
  const int dk = tens->data_kind[src];
  const size_t bytes = shape_volume(tens->shape_p) * (size_t)dk_size(dk);
  void *dst = dev_mem;
  int h = -1, attached = dev_mem ? 1 : 0;
  if (!dst) {
    const int e = (target == DEV_HOST) ? host_alloc(bytes, &dst, &h) : dev_alloc(tg, bytes, &dst, &h);
    if (e) return deferred(g, e);
  }
  Device &d = rt().gpus[g];
  (void)hipSetDevice(d.hip_id);
  Task *task = task_new(g, (double)bytes);
  hipStream_t s = task->stream;
  void *sp = tens->dev_rsc[src].gmem_p;
  (void)hipEventRecord(task->ev_start->ev, s);
  hazard_wait_read(sp, bytes, s);
  hazard_wait_write(dst, bytes, s);
  (void)hipEventRecord(task->ev_in->ev, s);
  if (hipMemcpyAsync(dst, sp, bytes, hipMemcpyDefault, s) != hipSuccess) {
    task_free(task);
    if (!attached) {
      if (target == DEV_HOST) host_free(dst, h, bytes);
      else dev_free(tg, dst, h);
    }
    return TALSH_FAILURE;
  }
  (void)hipEventRecord(task->ev_comp->ev, s);
  (void)hipEventRecord(task->ev_done->ev, s);
  hazard_add_reader(sp, bytes, task->ev_done);
  hazard_set_writer(dst, bytes, task->ev_done);
  if (target == DEV_HOST) d.traffic_out += (double)bytes;
  else d.traffic_in += (double)bytes;
  task->bytes = (double)bytes;
  if (copy_ctrl == COPY_M) tens->avail[src] = NOPE;
  task->on_finalize.push_back([tens, sp, dst, h, attached, target, tg, bytes, dk, copy_ctrl]() {
    if (talshTensorIsEmpty(tens) == YEP) {  // the new image has no owner: release it
      if (!attached) {
        if (target == DEV_HOST) host_free(dst, h, bytes);
        else dev_free(tg, dst, h);
      }
      return;
    }
    int si = -1;
    for (int i = 0; i < tens->ndev; ++i)
      if (tens->dev_rsc[i].gmem_p == sp) si = i;
    if (si >= 0) {
      if (copy_ctrl == COPY_M) image_discard(tens, si);
      else tens->avail[si] = YEP;
    }
    image_add(tens, target, dst, h, attached, dk);
  });
  task_register(task, {tens});
  d.tasks_submitted++;
  d.active++;
  talsh_task_t local;
  talsh_task_t *tt = talsh_task ? talsh_task : &local;
  if (!talsh_task) talshTaskClean(&local);
  tt->task_p = task;
  tt->task_error = -1;
  tt->dev_kind = DEV_NVIDIA_GPU;
  tt->data_kind = dk;
  tt->coherence = copy_ctrl;
  tt->num_args = 1;
  tt->tens_args[0].tens_p = tens;
  tt->tens_args[0].source_image = src;
  tt->data_vol = (double)bytes;
  if (talsh_task) return TALSH_SUCCESS;
  const int st = task_finalize_if_done(tt, true);
  talshTaskDestruct(tt);
  return st == TALSH_TASK_COMPLETED ? TALSH_SUCCESS : TALSH_TASK_ERROR;
}
int talshTensorPlace_(talsh_tens_t *tens, int dev_id, int dev_kind, void *dev_mem, int copy_ctrl,
                      talsh_task_t *talsh_task) {
  return talshTensorPlace(tens, dev_id, dev_kind, dev_mem, copy_ctrl, talsh_task);
}

int talshTensorDiscard(talsh_tens_t *tens, int dev_id, int dev_kind) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (!tens) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(tens) == YEP) return TALSH_OBJECT_IS_EMPTY;
  const int target = canonical_dev(dev_kind == DEV_NULL ? dev_id : talshFlatDevId(dev_kind, dev_id));
  for (int i = tens->ndev - 1; i >= 0; --i)
    if (tens->dev_rsc[i].dev_id == target) {
      if (tens->avail[i] != YEP) return TALSH_NOT_AVAILABLE;
      image_discard(tens, i);
    }
  return TALSH_SUCCESS;
}
int talshTensorDiscard_(talsh_tens_t *tens, int dev_id, int dev_kind) {
  return talshTensorDiscard(tens, dev_id, dev_kind);
}

int talshTensorDiscardOther(talsh_tens_t *tens, int dev_id, int dev_kind) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (!tens) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(tens) == YEP) return TALSH_OBJECT_IS_EMPTY;
  const int target = canonical_dev(dev_kind == DEV_NULL ? dev_id : talshFlatDevId(dev_kind, dev_id));
  int keep = -1;
  for (int i = 0; i < tens->ndev; ++i)
    if (tens->dev_rsc[i].dev_id == target) keep = i;
  if (keep < 0) return TALSH_NOT_FOUND;
  for (int i = 0; i < tens->ndev; ++i)
    if (i != keep && tens->avail[i] != YEP) return TALSH_NOT_AVAILABLE;
  image_discard_other(tens, keep);
  return TALSH_SUCCESS;
}
int talshTensorDiscardOther_(talsh_tens_t *tens, int dev_id, int dev_kind) {
  return talshTensorDiscardOther(tens, dev_id, dev_kind);
}

// Fast math (talsh.h:167-174): on this build it selects the split MFMA GEMM
// (method: talshAmdSetFastMathSplit) for R4 contractions; C4/R8/C8 are
// unaffected.
int talshEnableFastMath(int dev_kind, int dev_id) {
  (void)dev_kind;
  (void)dev_id;
  if (!gemm_split_bf16_available()) return TALSH_NOT_AVAILABLE;
  std::lock_guard<std::recursive_mutex> lk(rt().mu);  // not while an operation plans
  rt().r4mode = R4_SPLIT_BF16;
  return TALSH_SUCCESS;
}
int talshDisableFastMath(int dev_kind, int dev_id) {
  (void)dev_kind;
  (void)dev_id;
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  rt().r4mode = R4_FP32;
  return TALSH_SUCCESS;
}
int talshQueryFastMath(int dev_kind, int dev_id) {
  (void)dev_kind;
  (void)dev_id;
  return rt().r4mode == R4_SPLIT_BF16 ? YEP : NOPE;
}

// ---- extensions (talsh_amd.h) ----
static_assert(SPLIT_BF16X6 == TALSH_AMD_SPLIT_BF16X6 && SPLIT_FP16X3 == TALSH_AMD_SPLIT_FP16X3, "split ids");
int talshAmdSetFastMathSplit(int method) {
  if (method != SPLIT_BF16X6 && method != SPLIT_FP16X3) return TALSH_INVALID_ARGS;
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  rt().split_method = method;
  return TALSH_SUCCESS;
}
int talshAmdGetFastMathSplit(void) { return rt().split_method; }
int talshAmdSetDeviceBufferSize(size_t bytes) {
  if (rt().on) return TALSH_ALREADY_INITIALIZED;
  rt().dev_buffer_request = bytes;
  return TALSH_SUCCESS;
}
int talshAmdSetNumStreams(int n) {
  if (rt().on) return TALSH_ALREADY_INITIALIZED;
  if (n < 1 || n > 32) return TALSH_INVALID_ARGS;
  rt().nstreams = n;
  return TALSH_SUCCESS;
}
int talshAmdDeviceSynchronize(int gpu) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  const int g = gpu_index(flat_gpu(gpu));
  if (g < 0) return TALSH_INVALID_ARGS;
  (void)hipSetDevice(rt().gpus[g].hip_id);
  return hipDeviceSynchronize() == hipSuccess ? TALSH_SUCCESS : TALSH_FAILURE;
}
// talshTensorNorm2's kernels on raw device pointers with the per-GPU static
// scratch: no pool buffer, no task, no event, no allocation of any kind.
int talshAmdTensorNorm2NoAlloc(int data_kind, const void *body, size_t volume, void *result, int gpu,
                               void *stream) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  const int g = gpu_index(flat_gpu(gpu));
  if (g < 0 || !rt().gpus[g].on) return TALSH_INVALID_ARGS;
  if (data_kind != R4 && data_kind != R8 && data_kind != C4 && data_kind != C8) return TALSH_INVALID_ARGS;
  if (!body || !result || volume == 0) return TALSH_INVALID_ARGS;
  // the reduction's partials never exceed the fixed scratch (one double per stage-1 block)
  if (norm2_scratch_elems() * sizeof(double) > norm2_noalloc_scratch_bytes()) return TALSH_INVALID_ARGS;
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) return TALSH_FAILURE;
  const int dev = rt().gpus[g].hip_id;
  if (prev != dev && hipSetDevice(dev) != hipSuccess) return TALSH_FAILURE;
  const int e = launch_norm2_noalloc(dev, data_kind, body, volume, result, (hipStream_t)stream);
  if (prev != dev) (void)hipSetDevice(prev);
  return e == 0 ? TALSH_SUCCESS : TALSH_FAILURE;
}

int talshAmdBuildInfo(char *buf, int len) {
  const char *s = "libtalsh_amd gfx950 MFMA f64/f32 TTGT";
  if (!buf || len <= 0) return TALSH_INVALID_ARGS;
  std::strncpy(buf, s, (size_t)len - 1);
  buf[len - 1] = 0;
  return TALSH_SUCCESS;
}

}  // extern "C"

// ---- pattern / plan queries ----
extern "C" {

// Symbolic -> digital pattern, same signature and meaning as the reference's
// C-callable converter (talshf.F90:619-667, declared in talshc.cpp:109).
int talsh_get_contr_ptrn_str2dig(const char *c_str, int *dig_ptrn, int *drank, int *lrank, int *rrank,
                                 int *conj_bits) {
  if (!c_str || !dig_ptrn || !drank || !lrank || !rrank || !conj_bits) return -1;
  DigitalPattern p;
  *drank = *lrank = *rrank = -1;
  *conj_bits = 0;
  const int e = parse_pattern(c_str, &p);
  if (e) return e;
  *drank = p.drank;
  *lrank = p.lrank;
  *rrank = p.rrank;
  *conj_bits = p.conj;
  for (int i = 0; i < p.lrank + p.rrank; ++i) dig_ptrn[i] = p.dig[i];
  return 0;
}

// Digital -> symbolic pattern (tensor_algebra.h:265-266): D gets the
// uncontracted labels a, b, ... in order, each contracted index of L a new
// label after them, and R reuses L's label for the index it contracts with.
void get_contr_pattern_sym(const int *rank_left, const int *rank_right, const int *conj_bits, const int *cptrn_dig,
                           char *cptrn_sym, int *cpl, int *ierr) {
  if (!rank_left || !rank_right || !conj_bits || !cptrn_sym || !cpl || !ierr) return;
  *cpl = 0;
  *ierr = 0;
  const int lr = *rank_left, rr = *rank_right, cj = *conj_bits;
  if (lr < 0 || rr < 0 || lr > MAX_TENSOR_RANK || rr > MAX_TENSOR_RANK || (lr + rr > 0 && !cptrn_dig)) {
    *ierr = 1;
    return;
  }
  std::string out;
  int nu = 0;
  for (int i = 0; i < lr + rr; ++i) nu += cptrn_dig[i] > 0;
  auto tensor = [&](const char *name, int bit, const std::vector<char> &labels) {
    out += name;
    if (cj & bit) out += '+';
    out += '(';
    for (size_t i = 0; i < labels.size(); ++i) {
      if (i) out += ',';
      out += labels[i];
    }
    out += ')';
  };
  std::vector<char> dl, ll, rl, lcon(lr > 0 ? lr : 1, '?');
  for (int i = 0; i < nu; ++i) dl.push_back((char)('a' + i));
  char next = (char)('a' + nu);
  for (int i = 0; i < lr; ++i) {
    const int d = cptrn_dig[i];
    if (d > 0) {
      ll.push_back((char)('a' + d - 1));
    } else {
      lcon[i] = next++;
      ll.push_back(lcon[i]);
    }
  }
  for (int i = 0; i < rr; ++i) {
    const int d = cptrn_dig[lr + i];
    if (d > 0) rl.push_back((char)('a' + d - 1));
    else if (-d >= 1 && -d <= lr) rl.push_back(lcon[-d - 1]);
    else {
      *ierr = 1;
      return;
    }
  }
  tensor("D", 1, dl);
  out += "+=";
  tensor("L", 2, ll);
  out += '*';
  tensor("R", 4, rl);
  std::memcpy(cptrn_sym, out.c_str(), out.size() + 1);
  *cpl = (int)out.size();
}

// Contraction plan chosen for a pattern and shapes (see talsh_amd.h).
int talshAmdPlanContraction(const char *pattern, const int *ddims, const int *ldims, const int *rdims,
                            int accumulative, long long *mnk, int *flags, char *ops, int *perms, double *moved) {
  DigitalPattern p;
  if (!pattern || parse_pattern(pattern, &p) != 0) return TALSH_INVALID_ARGS;
  ContractPlan c;
  const int e = plan_contraction(p, ddims, ldims, rdims, accumulative, &c);
  if (e) return TALSH_INVALID_ARGS;
  if (mnk) { mnk[0] = c.M; mnk[1] = c.N; mnk[2] = c.K; }
  if (flags) { flags[0] = c.swap; flags[1] = c.a.transpose; flags[2] = c.b.transpose; flags[3] = c.d_transpose; }
  if (ops) { ops[0] = c.a.op; ops[1] = c.b.op; }
  if (perms) {
    // [A o2n (rank of A)] [B o2n] [D->D' o2n]; -1 fills unused entries
    int *q = perms;
    const int ra = c.swap ? p.rrank : p.lrank, rb = c.swap ? p.lrank : p.rrank;
    for (int i = 0; i < ra; ++i) *q++ = c.a.transpose ? c.a.o2n[i] : -1;
    for (int i = 0; i < rb; ++i) *q++ = c.b.transpose ? c.b.o2n[i] : -1;
    for (int i = 0; i < p.drank; ++i) *q++ = c.d_transpose ? c.d_o2n[i] : -1;
  }
  if (moved) *moved = c.moved_elems;
  return TALSH_SUCCESS;
}

}  // extern "C"

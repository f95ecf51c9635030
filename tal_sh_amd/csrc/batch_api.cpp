// Batched contractions: host API (extension, include/talsh_amd.h).
//
// A batch collects many talshTensorContract-style operations on GPU-resident
// tensors and runs them as one persistent kernel launch captured in a
// hipGraph.  The first execution plans the batch (index tables, tile shapes,
// graph); later executions of an unchanged batch replay the graph.
//
// Tile shapes come from an autotuner: tasks are grouped into shape classes
// by (log2 M, log2 N) and, for every class not tuned yet in this process,
// each candidate tile shape is timed on a sample of the class's own tasks
// (writing into scratch outputs so user data is untouched); the fastest
// shape is cached per (data kind, class) for the rest of the process.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "batch.hpp"
#include "plan.hpp"
#include "runtime.hpp"
#include "talsh_amd.h"

struct talshAmdBatch {
  struct Entry {
    std::string pattern;
    talsh_tens_t *d, *l, *r;
    double ar, ai;
    int acc;
  };
  std::vector<Entry> entries;
  bool built = false;
  int g = -1, dk = NO_TYPE, grid = 0, ntasks = 0;
  long long total = 0;
  double flops = 0.0;
  std::vector<const void *> sig;
  std::vector<size_t> sig_bytes;  // body bytes of each sig entry
  void *dev = nullptr;
  int dev_handle = -1;
  tal::BatchTask *d_tasks = nullptr;
  long long *d_prefix = nullptr;
  int *d_item_task = nullptr;
  unsigned long long *d_counter = nullptr;
  int *d_tabs = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int tuned_classes = 0, replays = 0;
  double tune_ms = 0.0;
  // Ordering state shared by every operand: the hazard entries of all its
  // tensors point at `hz`, re-recorded after each execution.  `epoch` is the
  // hazard-table epoch right after the last execution; while it is unchanged
  // nothing else has touched any operand (or moved a body) since.
  tal::EvRef hz;
  uint64_t epoch = 0;
  int policy = 0;  // tile policy chosen by whole-batch timing
};

namespace tal {
namespace {

struct TileShape {
  int tm, tn;
};
const TileShape kCandidates[] = {{16, 16}, {16, 32}, {32, 16}, {32, 32}, {16, 64},
                                 {64, 16}, {32, 64}, {64, 32}, {64, 64}};

std::map<std::pair<int, int>, TileShape> &tuned_tiles() {
  static std::map<std::pair<int, int>, TileShape> t;
  return t;
}

int bucket(int x) {
  int b = 0;
  while ((1 << b) < x && b < 12) ++b;
  return b;
}
int shape_class(int dk, int M, int N) { return dk * 10000 + bucket(M) * 100 + bucket(N); }

// Host description of one contraction after planning.
struct HostTask {
  BatchTask t;
  std::vector<int> tab;  // lM lK rK rN dM dN
  int cls = 0;
  size_t dbytes = 0;
};

// Mixed-radix offsets: for every multi-index over `ext` (first fastest),
// sum_k idx_k * stride_a[k] and sum_k idx_k * stride_b[k].
void odometer(const std::vector<int> &ext, const std::vector<long long> &sa, const std::vector<long long> &sb,
              std::vector<int> *oa, std::vector<int> *ob) {
  long long n = 1;
  for (int e : ext) n *= e;
  std::vector<int> idx(ext.size(), 0);
  long long a = 0, b = 0;
  for (long long i = 0; i < n; ++i) {
    oa->push_back((int)a);
    if (ob) ob->push_back((int)b);
    for (size_t k = 0; k < ext.size(); ++k) {
      if (++idx[k] < ext[k]) {
        a += sa[k];
        b += sb[k];
        break;
      }
      a -= sa[k] * (ext[k] - 1);
      b -= sb[k] * (ext[k] - 1);
      idx[k] = 0;
    }
  }
}

std::vector<long long> strides(const talsh_tens_t *t) {
  std::vector<long long> s(t->shape_p->num_dim);
  long long acc = 1;
  for (int i = 0; i < t->shape_p->num_dim; ++i) {
    s[i] = acc;
    acc *= t->shape_p->dims[i];
  }
  return s;
}

void *gpu_body(talsh_tens_t *t, int g, int *dk) {
  const int i = image_find(t, flat_gpu(g), NO_TYPE, true);
  if (i < 0) return nullptr;
  *dk = t->data_kind[i];
  return t->dev_rsc[i].gmem_p;
}

int plan_task(const talshAmdBatch::Entry &e, int g, HostTask *h) {
  DigitalPattern p;
  if (parse_pattern(e.pattern.c_str(), &p) != 0) return TALSH_INVALID_ARGS;
  const talsh_tens_t *d = e.d, *l = e.l, *r = e.r;
  if (p.drank != d->shape_p->num_dim || p.lrank != l->shape_p->num_dim || p.rrank != r->shape_p->num_dim)
    return TALSH_INVALID_ARGS;
  if (shape_volume(d->shape_p) >= (1ull << 31) || shape_volume(l->shape_p) >= (1ull << 31) ||
      shape_volume(r->shape_p) >= (1ull << 31))
    return DEVICE_UNABLE;  // 32-bit offset tables: large contractions belong to the TTGT path
  ContractPlan cp;  // validates the pattern against the shapes
  const int pe = plan_contraction(p, d->shape_p->dims, l->shape_p->dims, r->shape_p->dims, e.acc, &cp);
  if (pe) return TALSH_INVALID_ARGS;
  if (cp.batch > 1) return DEVICE_UNABLE;  // hyper indices: the TTGT path batches them
  const std::vector<long long> ls = strides(l), rs = strides(r), ds = strides(d);
  std::vector<int> mext, next, kext;
  std::vector<long long> ml, md, nr, nd, kl, kr;
  for (int i = 0; i < p.lrank; ++i) {
    const int c = p.dig[i];
    if (c > 0) {
      mext.push_back(l->shape_p->dims[i]);
      ml.push_back(ls[i]);
      md.push_back(ds[c - 1]);
    } else {
      kext.push_back(l->shape_p->dims[i]);
      kl.push_back(ls[i]);
      kr.push_back(rs[-c - 1]);
    }
  }
  for (int j = 0; j < p.rrank; ++j) {
    const int c = p.dig[p.lrank + j];
    if (c > 0) {
      next.push_back(r->shape_p->dims[j]);
      nr.push_back(rs[j]);
      nd.push_back(ds[c - 1]);
    }
  }
  // Contracted indices run in L's storage order unless only R has one as its
  // contiguous index: then in R's order, so the B gathers can run along k.
  const bool l_kfast = p.lrank > 0 && p.dig[0] < 0, r_kfast = p.rrank > 0 && p.dig[p.lrank] < 0;
  if (r_kfast && !l_kfast) {
    std::vector<size_t> ord(kext.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return kr[x] < kr[y]; });
    std::vector<int> e2;
    std::vector<long long> l2, r2;
    for (size_t i : ord) {
      e2.push_back(kext[i]);
      l2.push_back(kl[i]);
      r2.push_back(kr[i]);
    }
    kext.swap(e2);
    kl.swap(l2);
    kr.swap(r2);
  }
  std::vector<int> lM, dM, lK, rK, rN, dN;
  odometer(mext, ml, md, &lM, &dM);
  odometer(kext, kl, kr, &lK, &rK);
  odometer(next, nr, nd, &rN, &dN);
  // gather layouts: along k when k (not m / n) is unit-stride in the operand
  auto unit = [](const std::vector<int> &v) { return v.size() >= 2 && v[1] - v[0] == 1; };
  h->t.amode = (unit(lK) && !unit(lM)) ? 1 : 0;
  h->t.bmode = (unit(rK) && !unit(rN)) ? 1 : 0;
  int dk[3];
  h->t.D = gpu_body(e.d, g, &dk[0]);
  h->t.L = gpu_body(e.l, g, &dk[1]);
  h->t.R = gpu_body(e.r, g, &dk[2]);
  if (!h->t.D || !h->t.L || !h->t.R) return TALSH_NOT_AVAILABLE;
  if (dk[0] != dk[1] || dk[0] != dk[2] || (dk[0] != R4 && dk[0] != R8)) return TALSH_NOT_IMPLEMENTED;
  h->t.M = (int)lM.size();
  h->t.K = (int)lK.size();
  h->t.N = (int)rN.size();
  h->t.alpha = e.ar;
  h->t.beta_one = e.acc ? 1 : 0;
  h->tab.reserve(2 * (h->t.M + h->t.N + h->t.K));
  for (auto *v : {&lM, &lK, &rK, &rN, &dM, &dN}) h->tab.insert(h->tab.end(), v->begin(), v->end());
  h->cls = shape_class(dk[0], h->t.M, h->t.N);
  h->dbytes = shape_volume(d->shape_p) * (size_t)dk_size(dk[0]);
  return TALSH_SUCCESS;
}

TileShape default_tile(int M, int N) {
  auto fit = [](int x) { return x <= 16 ? 16 : x <= 32 ? 32 : 64; };
  return TileShape{fit(M), fit(N)};
}

void set_tile(HostTask &h, TileShape s) {
  h.t.tm = s.tm;
  h.t.tn = s.tn;
  h.t.tiles_m = (h.t.M + s.tm - 1) / s.tm;
  h.t.tiles_n = (h.t.N + s.tn - 1) / s.tn;
}

// Device image of a task list: [tasks][prefix][counter][item -> task][tables].
struct DeviceBatch {
  void *base = nullptr;
  int handle = -1;
  BatchTask *tasks = nullptr;
  long long *prefix = nullptr;
  unsigned long long *counter = nullptr;
  int *item_task = nullptr;
  int *tabs = nullptr;
  long long total = 0;
};

int upload(int g, std::vector<HostTask> &hts, DeviceBatch *db, hipStream_t s) {
  const int n = (int)hts.size();
  std::vector<BatchTask> tasks(n);
  std::vector<long long> prefix(n + 1, 0);
  std::vector<int> tabs;
  for (int i = 0; i < n; ++i) {
    hts[i].t.tab = (long long)tabs.size();
    tabs.insert(tabs.end(), hts[i].tab.begin(), hts[i].tab.end());
    tasks[i] = hts[i].t;
    prefix[i + 1] = prefix[i] + (long long)tasks[i].tiles_m * tasks[i].tiles_n;
  }
  if (tabs.empty()) tabs.push_back(0);
  std::vector<int> item_task((size_t)std::max<long long>(1, prefix[n]));
  for (int i = 0; i < n; ++i)
    for (long long it = prefix[i]; it < prefix[i + 1]; ++it) item_task[(size_t)it] = i;
  const size_t bt = ((sizeof(BatchTask) * n + 255) / 256) * 256, bp = ((8 * (n + 1) + 255) / 256) * 256;
  const size_t bi = ((4 * item_task.size() + 255) / 256) * 256;
  const size_t bytes = bt + bp + 256 + bi + 4 * tabs.size();
  int e = dev_alloc(g, bytes, &db->base, &db->handle);
  if (e) return e;
  char *b = (char *)db->base;
  db->tasks = (BatchTask *)b;
  db->prefix = (long long *)(b + bt);
  db->counter = (unsigned long long *)(b + bt + bp);
  db->item_task = (int *)(b + bt + bp + 256);
  db->tabs = (int *)(b + bt + bp + 256 + bi);
  db->total = prefix[n];
  if (hipMemcpyAsync(db->tasks, tasks.data(), sizeof(BatchTask) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(db->prefix, prefix.data(), 8 * (n + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(db->item_task, item_task.data(), 4 * item_task.size(), hipMemcpyHostToDevice, s) !=
          hipSuccess ||
      hipMemcpyAsync(db->tabs, tabs.data(), 4 * tabs.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    dev_free(g, db->base, db->handle);
    return TALSH_FAILURE;
  }
  return TALSH_SUCCESS;
}

// Workgroups of the persistent launch: as many as are resident on the CUs at
// once (4 per CU: 4 waves per SIMD by registers, 35 KB of LDS each for R8).
int persistent_grid(int g, int dk) {
  hipDeviceProp_t prop;
  const int per_cu = 4;
  (void)dk;
  if (hipGetDeviceProperties(&prop, rt().gpus[g].hip_id) != hipSuccess) return 256 * per_cu;
  return std::max(1, prop.multiProcessorCount) * per_cu;
}

// Times every candidate tile shape on up to 32 tasks of each untuned class.
int autotune(int g, int dk, std::vector<HostTask> &hts, hipStream_t s, int *nclasses, double *ms_spent) {
  std::map<int, std::vector<int>> classes;
  for (int i = 0; i < (int)hts.size(); ++i) classes[hts[i].cls].push_back(i);
  const int grid = persistent_grid(g, dk);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return TALSH_FAILURE;
  const double t0 = wall_time();
  int err = TALSH_SUCCESS;
  for (auto &kv : classes) {
    const auto key = std::make_pair(dk, kv.first);
    if (tuned_tiles().count(key)) continue;
    std::vector<HostTask> probe;
    size_t scratch = 0;
    for (int i = 0; i < (int)kv.second.size() && i < 32; ++i) {
      probe.push_back(hts[kv.second[i]]);
      scratch += (probe.back().dbytes + 255) / 256 * 256;
    }
    void *sp = nullptr;
    int sh = -1;
    if ((err = dev_alloc(g, scratch, &sp, &sh)) != 0) break;
    size_t off = 0;
    for (auto &p : probe) {  // write into scratch outputs, never into user tensors
      p.t.D = (char *)sp + off;
      p.t.beta_one = 0;
      off += (p.dbytes + 255) / 256 * 256;
    }
    float best = 1e30f;
    TileShape pick = default_tile(probe[0].t.M, probe[0].t.N);
    for (const TileShape &c : kCandidates) {
      for (auto &p : probe) set_tile(p, c);
      DeviceBatch db;
      if ((err = upload(g, probe, &db, s)) != 0) break;
      float ms = 0;
      for (int rep = 0; rep < 2; ++rep) {  // first launch warms caches
        (void)hipEventRecord(e0, s);
        batch_launch(dk, db.tasks, db.prefix, db.item_task, db.tabs, db.counter, db.total, grid, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
      }
      dev_free(g, db.base, db.handle);
      if (ms < best) {
        best = ms;
        pick = c;
      }
    }
    dev_free(g, sp, sh);
    if (err) break;
    tuned_tiles()[key] = pick;
    (*nclasses)++;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_spent += 1e3 * (wall_time() - t0);
  return err;
}

// Second tuning stage, on the whole batch: a class timed alone favours small
// tiles (they expose parallelism a 32-task probe lacks), while the full batch
// has work items to spare and pays a fixed cost per tile.  Each tile policy
// below is timed on one launch of the entire batch writing scratch outputs;
// the fastest is applied.  Skipped (policy 0) when the scratch outputs would
// not fit in a quarter of the free pool.
enum TilePolicy { POLICY_CLASS_TUNED = 0, POLICY_FIT = 1, POLICY_FIT_MIN32 = 2, POLICY_MAX = 3, NPOLICY = 4 };

TileShape policy_tile(int policy, int dk, const HostTask &h) {
  const TileShape fit = default_tile(h.t.M, h.t.N);
  switch (policy) {
    case POLICY_FIT: return fit;
    case POLICY_FIT_MIN32: return TileShape{std::max(32, fit.tm), std::max(32, fit.tn)};
    case POLICY_MAX: return TileShape{64, 64};
    default: {
      auto it = tuned_tiles().find(std::make_pair(dk, h.cls));
      return it != tuned_tiles().end() ? it->second : fit;
    }
  }
}

int choose_policy(int g, int dk, std::vector<HostTask> hts, hipStream_t s, int *policy, double *ms_spent) {
  *policy = POLICY_CLASS_TUNED;
  size_t scratch = 0;
  for (auto &h : hts) scratch += (h.dbytes + 255) / 256 * 256;
  if (scratch == 0 || scratch > rt().gpus[g].arena.free_bytes() / 4) return TALSH_SUCCESS;
  void *sp = nullptr;
  int sh = -1;
  if (dev_alloc(g, scratch, &sp, &sh) != 0) return TALSH_SUCCESS;
  size_t off = 0;
  for (auto &h : hts) {  // scratch outputs: user tensors are never written while tuning
    h.t.D = (char *)sp + off;
    h.t.beta_one = 0;
    off += (h.dbytes + 255) / 256 * 256;
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    dev_free(g, sp, sh);
    return TALSH_FAILURE;
  }
  const double t0 = wall_time();
  const int grid = persistent_grid(g, dk);
  float best = 1e30f;
  int err = TALSH_SUCCESS;
  for (int pol = 0; pol < NPOLICY && !err; ++pol) {
    for (auto &h : hts) set_tile(h, policy_tile(pol, dk, h));
    DeviceBatch db;
    if ((err = upload(g, hts, &db, s)) != 0) break;
    float ms = 0;
    for (int rep = 0; rep < 2; ++rep) {  // first launch warms caches
      (void)hipEventRecord(e0, s);
      batch_launch(dk, db.tasks, db.prefix, db.item_task, db.tabs, db.counter, db.total, grid, s);
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      (void)hipEventElapsedTime(&ms, e0, e1);
    }
    dev_free(g, db.base, db.handle);
    if (ms < best) {
      best = ms;
      *policy = pol;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(g, sp, sh);
  *ms_spent += 1e3 * (wall_time() - t0);
  return err;
}

void release_build(talshAmdBatch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  if (b->dev && b->g >= 0) dev_free(b->g, b->dev, b->dev_handle);
  b->dev = nullptr;
  b->built = false;
}

std::vector<const void *> signature(const talshAmdBatch *b, int g) {
  std::vector<const void *> s;
  for (auto &e : b->entries) {
    int dk;
    s.push_back(gpu_body(e.d, g, &dk));
    s.push_back(gpu_body(e.l, g, &dk));
    s.push_back(gpu_body(e.r, g, &dk));
  }
  return s;
}

int build(talshAmdBatch *b, int g) {
  release_build(b);
  std::vector<HostTask> hts(b->entries.size());
  for (size_t i = 0; i < b->entries.size(); ++i)
    if (int e = plan_task(b->entries[i], g, &hts[i])) return e;
  const int dk = hts.empty() ? R8 : (hts[0].t.D ? hts[0].cls / 10000 : R8);
  for (auto &h : hts)
    if (h.cls / 10000 != dk) return TALSH_NOT_IMPLEMENTED;  // one data kind per batch
  // outputs must not overlap one another nor any input of the batch (byte ranges)
  const size_t esz = (size_t)dk_size(dk);
  std::vector<std::pair<uintptr_t, uintptr_t>> outs, ins;
  for (size_t i = 0; i < hts.size(); ++i) {
    const auto &e = b->entries[i];
    const uintptr_t d = (uintptr_t)hts[i].t.D, l = (uintptr_t)hts[i].t.L, r = (uintptr_t)hts[i].t.R;
    outs.emplace_back(d, d + shape_volume(e.d->shape_p) * esz);
    ins.emplace_back(l, l + shape_volume(e.l->shape_p) * esz);
    ins.emplace_back(r, r + shape_volume(e.r->shape_p) * esz);
  }
  std::sort(outs.begin(), outs.end());
  for (size_t i = 1; i < outs.size(); ++i)
    if (outs[i].first < outs[i - 1].second) return TALSH_INVALID_ARGS;
  for (const auto &in : ins) {  // the output starting last before in's end is the only candidate
    auto it = std::lower_bound(outs.begin(), outs.end(), std::make_pair(in.second, (uintptr_t)0));
    if (it != outs.begin() && std::prev(it)->second > in.first) return TALSH_INVALID_ARGS;
  }
  Device &dev = rt().gpus[g];
  hipStream_t s = dev.streams[0];
  if (int e = autotune(g, dk, hts, s, &b->tuned_classes, &b->tune_ms)) return e;
  int policy = POLICY_CLASS_TUNED;
  if (int e = choose_policy(g, dk, hts, s, &policy, &b->tune_ms)) return e;
  b->policy = policy;
  b->flops = 0;
  for (auto &h : hts) {
    set_tile(h, policy_tile(policy, dk, h));
    b->flops += 2.0 * h.t.M * (double)h.t.N * h.t.K;
  }
  DeviceBatch db;
  if (int e = upload(g, hts, &db, s)) return e;
  b->dev = db.base;
  b->dev_handle = db.handle;
  b->d_tasks = db.tasks;
  b->d_prefix = db.prefix;
  b->d_item_task = db.item_task;
  b->d_counter = db.counter;
  b->d_tabs = db.tabs;
  b->total = db.total;
  b->ntasks = (int)hts.size();
  b->grid = persistent_grid(g, dk);
  b->g = g;
  b->dk = dk;
  // capture [clear counter, persistent kernel] once; executions replay it
  hipStream_t cs;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return TALSH_FAILURE;
  int err = TALSH_SUCCESS;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) err = TALSH_FAILURE;
  if (!err && batch_launch(dk, b->d_tasks, b->d_prefix, b->d_item_task, b->d_tabs, b->d_counter, b->total, b->grid, cs))
    err = TALSH_FAILURE;
  if (hipStreamEndCapture(cs, &b->graph) != hipSuccess) err = TALSH_FAILURE;
  if (!err && hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) err = TALSH_FAILURE;
  (void)hipStreamDestroy(cs);
  if (err) {
    release_build(b);
    return err;
  }
  b->sig = signature(b, g);
  b->sig_bytes.clear();
  for (auto &e : b->entries)
    for (const talsh_tens_t *t : {e.d, e.l, e.r}) b->sig_bytes.push_back(shape_volume(t->shape_p) * esz);
  b->built = true;
  return TALSH_SUCCESS;
}

}  // namespace
}  // namespace tal

using namespace tal;

extern "C" {

int talshAmdBatchCreate(talshAmdBatch **batch) {
  if (!batch) return TALSH_INVALID_ARGS;
  *batch = new talshAmdBatch();
  return TALSH_SUCCESS;
}

int talshAmdBatchAdd(talshAmdBatch *b, const char *pattern, talsh_tens_t *dtens, talsh_tens_t *ltens,
                     talsh_tens_t *rtens, double scale_real, double scale_imag, int accumulative) {
  if (!b || !pattern || !dtens || !ltens || !rtens) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(dtens) == YEP || talshTensorIsEmpty(ltens) == YEP || talshTensorIsEmpty(rtens) == YEP)
    return TALSH_OBJECT_IS_EMPTY;
  if (scale_imag != 0.0) return TALSH_NOT_IMPLEMENTED;
  b->entries.push_back({pattern, dtens, ltens, rtens, scale_real, scale_imag, accumulative != NOPE});
  b->built = false;
  return TALSH_SUCCESS;
}

int talshAmdBatchSize(const talshAmdBatch *b) { return b ? (int)b->entries.size() : -1; }

int talshAmdBatchExecute(talshAmdBatch *b, int dev_id, int dev_kind, talsh_task_t *talsh_task) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  if (!rt().on) return TALSH_NOT_INITIALIZED;
  if (!b) return TALSH_INVALID_ARGS;
  if (talsh_task && talshTaskIsEmpty(talsh_task) != YEP) return TALSH_OBJECT_NOT_EMPTY;
  if (b->entries.empty()) return TALSH_SUCCESS;
  int g = -1;
  if (dev_kind == DEV_DEFAULT && dev_id == DEV_DEFAULT) g = talshDeviceBusyLeast(DEV_NVIDIA_GPU);
  else g = gpu_index(dev_kind == DEV_DEFAULT ? dev_id : talshFlatDevId(dev_kind, dev_id));
  if (g < 0) return TALSH_NOT_AVAILABLE;
  Device &dev = rt().gpus[g];
  (void)hipSetDevice(dev.hip_id);
  // Replay fast path: same GPU, graph built, hazard table unchanged since the
  // last execution, so every operand's ordering state is exactly `hz` and one
  // wait on it covers all of them (no per-tensor lookups).
  const bool fast = b->built && b->g == g && b->hz && rt().hazard_epoch == b->epoch;
  if (!fast && (!b->built || b->g != g || signature(b, g) != b->sig)) {
    const int e = build(b, g);
    if (e) return e;
  }
  Task *task = task_new(g, 3.0 * (double)b->entries.size() * 4096.0);
  hipStream_t s = task->stream;
  (void)hipEventRecord(task->ev_start->ev, s);
  if (fast) {
    (void)hipStreamWaitEvent(s, b->hz->ev, 0);
  } else {
    for (size_t i = 0; i < b->entries.size(); ++i) {
      hazard_wait_write(b->sig[3 * i], b->sig_bytes[3 * i], s);
      hazard_wait_read(b->sig[3 * i + 1], b->sig_bytes[3 * i + 1], s);
      hazard_wait_read(b->sig[3 * i + 2], b->sig_bytes[3 * i + 2], s);
    }
  }
  (void)hipEventRecord(task->ev_in->ev, s);
  task->ev_mm0 = new_event(g);
  (void)hipEventRecord(task->ev_mm0->ev, s);
  if (hipGraphLaunch(b->exec, s) != hipSuccess) {
    task_free(task);
    return TALSH_FAILURE;
  }
  task->ev_mm1 = new_event(g);
  (void)hipEventRecord(task->ev_mm1->ev, s);
  (void)hipEventRecord(task->ev_comp->ev, s);
  (void)hipEventRecord(task->ev_done->ev, s);
  if (!fast) {
    if (!b->hz) b->hz = new_event(g);
    for (size_t i = 0; i < b->entries.size(); ++i) {
      hazard_set_writer(b->sig[3 * i], b->sig_bytes[3 * i], b->hz);
      hazard_add_reader(b->sig[3 * i + 1], b->sig_bytes[3 * i + 1], b->hz);
      hazard_add_reader(b->sig[3 * i + 2], b->sig_bytes[3 * i + 2], b->hz);
      talsh_tens_t *d = b->entries[i].d;  // other images of each output become stale
      const int gi = image_find(d, flat_gpu(g), NO_TYPE, true);
      if (gi >= 0) image_discard_other(d, gi);
    }
  }
  (void)hipEventRecord(b->hz->ev, s);
  b->epoch = rt().hazard_epoch;
  task->flops = b->flops;
  dev.tasks_submitted++;
  dev.active++;
  b->replays++;
  talsh_task_t local;
  talsh_task_t *tt = talsh_task ? talsh_task : &local;
  if (!talsh_task) talshTaskClean(&local);
  tt->task_p = task;
  tt->task_error = -1;
  tt->dev_kind = DEV_NVIDIA_GPU;
  tt->data_kind = b->dk;
  tt->coherence = 0;
  tt->num_args = 0;
  tt->flops = b->flops;
  if (talsh_task) return TALSH_SUCCESS;
  const int st = task_finalize_if_done(tt, true);
  talshTaskDestruct(tt);
  return st == TALSH_TASK_COMPLETED ? TALSH_SUCCESS : TALSH_TASK_ERROR;
}

int talshAmdBatchStats(const talshAmdBatch *b, double *flops, long long *work_items, int *tuned_classes,
                       double *tune_ms, int *executions) {
  if (!b) return TALSH_INVALID_ARGS;
  if (flops) *flops = b->flops;
  if (work_items) *work_items = b->total;
  if (tuned_classes) *tuned_classes = b->tuned_classes;
  if (tune_ms) *tune_ms = b->tune_ms;
  if (executions) *executions = b->replays;
  return TALSH_SUCCESS;
}

int talshAmdBatchDestroy(talshAmdBatch *b) {
  if (!b) return TALSH_INVALID_ARGS;
  {
    std::lock_guard<std::recursive_mutex> lk(rt().mu);
    if (b->built && rt().on && b->g >= 0) {
      for (size_t i = 0; i < b->sig.size(); ++i) hazard_drain(b->sig[i], b->sig_bytes[i]);
      release_build(b);
    }
  }
  delete b;
  return TALSH_SUCCESS;
}

}  // extern "C"

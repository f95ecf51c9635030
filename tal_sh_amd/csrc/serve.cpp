// Contraction serving engine (extension, include/talsh_amd.h).
//
// Requests (talshTensorContract operations) arrive at any time.  The engine
// keeps them in arrival order and admits them onto the GPU as soon as both a
// slot (at most max_inflight requests in flight) and device memory (the
// request's transient HBM need fits the remaining budget) are free; admitted
// requests are scheduled asynchronously on the GPU's streams, so their
// transfers and kernels overlap with those of requests already running.
// Completions are observed by polling (talshAmdServerProgress), which frees
// the request's budget and immediately admits more work: continuous
// batching, not a queue in front of a serial executor.  Operands that are
// already resident on the GPU are used in place and cost no budget.
//
// Fairness: the queue head may be overtaken by smaller requests while it
// waits for memory, but only kMaxSkips times; after that nothing behind it
// is admitted until it fits, so large requests cannot starve.
//
// Transient need of a request = bodies it must bring to the GPU (operands
// without an available image there) + the temporaries of its TTGT plan
// (transposed operands), computed from the same planner the contraction uses.
#include <algorithm>
#include <deque>
#include <string>
#include <vector>

#include "contract.hpp"
#include "plan.hpp"
#include "runtime.hpp"
#include "talsh_amd.h"

struct talshAmdServer {
  struct Req {
    talshAmdRequest r;
    std::string pattern;
    double t_arrive = 0, t_admit = 0;
    size_t need = 0;
    double flops = 0;
    int skips = 0;
    talsh_task_t task;
  };
  int g = -1;
  size_t budget = 0, in_use = 0, peak = 0;
  int max_inflight = 1;
  std::deque<Req *> queue;
  std::vector<Req *> inflight;
  std::vector<double> latencies, queue_waits;
  std::vector<std::pair<long long, int>> rejected;  // (tag, error) refused at admission, reported by Progress
  long long submitted = 0, completed = 0, failed = 0, deferred = 0;
  double flops_done = 0, t_first = -1, t_last = 0;
};

namespace {

constexpr int kMaxSkips = 64;

size_t body_bytes(const talsh_tens_t *t) {
  for (int i = 0; i < t->ndev; ++i)
    if (t->avail[i] == YEP) return tal::shape_volume(t->shape_p) * (size_t)tal::dk_size(t->data_kind[i]);
  return 0;
}

bool resident(const talsh_tens_t *t, int g) { return tal::image_find(t, tal::flat_gpu(g), NO_TYPE, true) >= 0; }

// Transient device bytes and FLOP of one request; 0 on success.
int estimate(talshAmdServer *s, talshAmdServer::Req *q) {
  tal::DigitalPattern p;
  const talshAmdRequest &r = q->r;
  if (!r.dtens || !r.ltens || !r.rtens || talshTensorIsEmpty(r.dtens) == YEP || talshTensorIsEmpty(r.ltens) == YEP ||
      talshTensorIsEmpty(r.rtens) == YEP)
    return TALSH_INVALID_ARGS;
  if (tal::parse_pattern(q->pattern.c_str(), &p) != 0) return TALSH_INVALID_ARGS;
  tal::ContractPlan plan;
  const int e = tal::plan_contraction(p, r.dtens->shape_p->dims, r.ltens->shape_p->dims, r.rtens->shape_p->dims,
                                      r.accumulative != NOPE, &plan);
  if (e) return TALSH_INVALID_ARGS;
  int dk = NO_TYPE;
  for (int i = 0; i < r.dtens->ndev; ++i)
    if (r.dtens->avail[i] == YEP) dk = r.dtens->data_kind[i];
  const size_t bd = body_bytes(r.dtens), bl = body_bytes(r.ltens), br = body_bytes(r.rtens);
  size_t need = 0;
  if (!resident(r.dtens, s->g)) need += bd;
  if (!resident(r.ltens, s->g)) need += bl;
  if (!resident(r.rtens, s->g)) need += br;
  // the TTGT temporaries talshTensorContract will take from the pool
  const tal::TtgtTemps t = tal::ttgt_temp_bytes(plan, dk, bl, br, bd);
  need += t.a + t.b + t.d + t.w + t.s;
  q->need = need;
  q->flops = ((dk == C4 || dk == C8) ? 8.0 : 2.0) * (double)plan.M * (double)plan.N * (double)plan.K *
             (double)plan.batch;
  return TALSH_SUCCESS;
}

void admit(talshAmdServer *s) {
  const double now = tal::wall_time();
  bool head_blocked = false;
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    talshAmdServer::Req *q = *it;
    if ((int)s->inflight.size() >= s->max_inflight) break;
    bool fits = s->in_use + q->need <= s->budget;
    // The need depends on where the operands are now (a request that reads
    // another's output was estimated while that output was in flight) and on
    // the fast-math mode: it is recomputed before the request is charged, and
    // for the head whenever it is blocked.
    if (fits || it == s->queue.begin()) {
      if (estimate(s, q) != TALSH_SUCCESS) {
        s->failed++;
        s->rejected.emplace_back(q->r.tag, TALSH_INVALID_ARGS);
        it = s->queue.erase(it);
        delete q;
        continue;
      }
      fits = s->in_use + q->need <= s->budget;
    }
    if (!fits) {
      if (it == s->queue.begin()) {
        head_blocked = true;
        if (q->skips >= kMaxSkips) break;  // let the head fill the budget as it drains
      }
      ++it;
      continue;
    }
    talshTaskClean(&q->task);
    const int e = talshTensorContract(q->pattern.c_str(), q->r.dtens, q->r.ltens, q->r.rtens, q->r.scale_real,
                                      q->r.scale_imag, s->g, DEV_NVIDIA_GPU, q->r.copy_ctrl, q->r.accumulative,
                                      &q->task);
    if ((e == TRY_LATER || e == TALSH_NOT_AVAILABLE) && !s->inflight.empty()) {
      // device pool fragmented by requests in flight, or an operand is the output of a
      // request still in flight: retry once some of them complete
      s->deferred++;
      if (it == s->queue.begin()) head_blocked = true;
      ++it;
      continue;
    }
    if (head_blocked && it != s->queue.begin()) s->queue.front()->skips++;
    it = s->queue.erase(it);
    if (e != TALSH_SUCCESS) {
      // cannot run even on an idle device (or invalid): reported as failed by the next Progress call
      s->failed++;
      s->rejected.emplace_back(q->r.tag, e);
      talshTaskDestruct(&q->task);
      delete q;
      continue;
    }
    q->t_admit = now;
    s->in_use += q->need;
    s->peak = std::max(s->peak, s->in_use);
    s->inflight.push_back(q);
  }
}

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double x = p * (double)(v.size() - 1);
  const size_t i = (size_t)x;
  const double f = x - (double)i;
  return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

int talshAmdServerCreate(talshAmdServer **srv, int gpu, size_t hbm_budget, int max_inflight) {
  if (!srv || max_inflight < 1 || hbm_budget == 0) return TALSH_INVALID_ARGS;
  std::lock_guard<std::recursive_mutex> lk(tal::rt().mu);
  if (!tal::rt().on) return TALSH_NOT_INITIALIZED;
  const int g = tal::gpu_index(tal::flat_gpu(gpu));
  if (g < 0) return TALSH_INVALID_ARGS;
  talshAmdServer *s = new talshAmdServer();
  s->g = g;
  s->budget = std::min(hbm_budget, tal::rt().gpus[g].size);
  s->max_inflight = max_inflight;
  *srv = s;
  return TALSH_SUCCESS;
}

int talshAmdServerSubmit(talshAmdServer *s, const talshAmdRequest *req) {
  if (!s || !req || !req->pattern) return TALSH_INVALID_ARGS;
  std::lock_guard<std::recursive_mutex> lk(tal::rt().mu);
  auto *q = new talshAmdServer::Req();
  q->r = *req;
  q->pattern = req->pattern;
  q->r.pattern = nullptr;
  q->t_arrive = tal::wall_time();
  const int e = estimate(s, q);
  if (e) {
    delete q;
    return e;
  }
  if (q->need > s->budget) {
    delete q;
    return DEVICE_UNABLE;  // can never be admitted under this budget
  }
  if (s->t_first < 0) s->t_first = q->t_arrive;
  s->submitted++;
  s->queue.push_back(q);
  admit(s);
  return TALSH_SUCCESS;
}

int talshAmdServerProgress(talshAmdServer *s, int max_out, long long *tags, double *latency_s, int *status,
                           int *nout) {
  if (!s || !nout) return TALSH_INVALID_ARGS;
  std::lock_guard<std::recursive_mutex> lk(tal::rt().mu);
  *nout = 0;
  while (!s->rejected.empty() && *nout < max_out) {
    if (tags) tags[*nout] = s->rejected.front().first;
    if (latency_s) latency_s[*nout] = 0.0;
    if (status) status[*nout] = TALSH_TASK_ERROR;
    (*nout)++;
    s->rejected.erase(s->rejected.begin());
  }
  for (auto it = s->inflight.begin(); it != s->inflight.end() && *nout < max_out;) {
    talshAmdServer::Req *q = *it;
    int st = 0, ierr = 0;
    if (talshTaskComplete(&q->task, &st, &ierr) != YEP) {
      ++it;
      continue;
    }
    const double now = tal::wall_time();
    if (st == TALSH_TASK_COMPLETED) {
      s->completed++;
      s->flops_done += q->flops;
      s->latencies.push_back(now - q->t_arrive);
      s->queue_waits.push_back(q->t_admit - q->t_arrive);
    } else {
      s->failed++;
    }
    s->t_last = now;
    s->in_use -= q->need;
    if (tags) tags[*nout] = q->r.tag;
    if (latency_s) latency_s[*nout] = now - q->t_arrive;
    if (status) status[*nout] = st;
    (*nout)++;
    talshTaskDestruct(&q->task);
    delete q;
    it = s->inflight.erase(it);
  }
  admit(s);
  return TALSH_SUCCESS;
}

int talshAmdServerPending(const talshAmdServer *s, int *queued, int *inflight) {
  if (!s) return TALSH_INVALID_ARGS;
  if (queued) *queued = (int)s->queue.size();
  if (inflight) *inflight = (int)s->inflight.size();
  return TALSH_SUCCESS;
}

int talshAmdServerStats(const talshAmdServer *s, double *out, int nout) {
  // out = {submitted, completed, failed, deferred, p50 latency s, p99 latency s, mean latency s,
  //        p50 queue wait s, p99 queue wait s, TFLOP/s (completed FLOP / first arrival..last completion),
  //        peak transient bytes, budget bytes}
  if (!s || !out || nout < 12) return TALSH_INVALID_ARGS;
  double mean = 0;
  for (double x : s->latencies) mean += x;
  if (!s->latencies.empty()) mean /= (double)s->latencies.size();
  const double span = (s->t_first >= 0 && s->t_last > s->t_first) ? s->t_last - s->t_first : 0.0;
  const double v[12] = {(double)s->submitted, (double)s->completed, (double)s->failed, (double)s->deferred,
                        percentile(s->latencies, 0.50), percentile(s->latencies, 0.99), mean,
                        percentile(s->queue_waits, 0.50), percentile(s->queue_waits, 0.99),
                        span > 0 ? s->flops_done / span / 1e12 : 0.0, (double)s->peak, (double)s->budget};
  for (int i = 0; i < 12; ++i) out[i] = v[i];
  return TALSH_SUCCESS;
}

int talshAmdServerDestroy(talshAmdServer *s) {
  if (!s) return TALSH_INVALID_ARGS;
  std::lock_guard<std::recursive_mutex> lk(tal::rt().mu);
  for (auto *q : s->inflight) {
    int st;
    talshTaskWait(&q->task, &st);
    talshTaskDestruct(&q->task);
    delete q;
  }
  for (auto *q : s->queue) delete q;
  delete s;
  return TALSH_SUCCESS;
}

}  // extern "C"

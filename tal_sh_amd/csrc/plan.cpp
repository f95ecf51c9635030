// Pattern parsing and contraction planning.
//
// parse_pattern follows the semantics of talsh_get_contr_ptrn_str2dig
// (talshf.F90:619-667) and get_contr_pattern_dig (tensor_algebra_cpu.F90:
// 5307-5531): arguments NAME[+](idx,...), index separators ',' or '|', a
// trailing '+' on an index marks contravariance (ignored for the digits), a
// '+' after an argument name conjugates it, a trailing scalar factor "*2" is
// dropped, and a unary pattern gets a rank-0 right operand.
//
// plan_contraction replaces the reference GPU matricization
// get_contr_permutations (tensor_algebra_cpu.F90:5613-5755, called from
// tensor_algebra_gpu_nvidia.cu:6083-6095), which always produces the TN form
// and transposes an operand unless the fixed TN order happens to match.  Here
// both operand roles (L as A, or R as A), both storage forms of each operand
// ('N' or 'T') and three orderings for the M, N and K index groups are
// enumerated, and the plan that copies the fewest elements is kept; operands
// already in a usable layout go straight into the GEMM.  Hyper indices (in D,
// L and R; the reference GPU path refuses them, its CPU path batches them,
// tensor_algebra_cpu.F90:5757-5956 and :3983-3985) become the batch
// dimension of a strided batched GEMM, placed after M, N and K in every
// operand.
#include "plan.hpp"

#include <algorithm>
#include <cctype>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

namespace tal {

namespace {

bool label_char(char c) { return std::isalnum(static_cast<unsigned char>(c)) != 0; }

struct Arg {
  std::vector<std::string> idx;
  bool conj = false;
};

int split_args(const std::string &s, std::vector<Arg> *args) {
  size_t i = 0;
  while (i < s.size()) {
    size_t open = s.find('(', i);
    if (open == std::string::npos) break;
    if (args->size() >= 3) return 1;
    Arg a;
    if (open >= 2 && s[open - 1] == '+' && (label_char(s[open - 2]) || s[open - 2] == '_')) a.conj = true;
    size_t close = s.find(')', open + 1);
    if (close == std::string::npos) return 5;
    std::string body = s.substr(open + 1, close - open - 1);
    if (!body.empty()) {
      size_t start = 0;
      while (true) {
        size_t sep = body.find_first_of(",|", start);
        std::string lab = body.substr(start, sep == std::string::npos ? std::string::npos : sep - start);
        if (lab.empty()) return 2;
        // alphanumeric label with an optional trailing '+' (contravariant)
        if (lab.size() > 1 && lab.back() == '+') lab.pop_back();
        for (char c : lab)
          if (!label_char(c)) return 4;
        a.idx.push_back(lab);
        if (sep == std::string::npos) break;
        start = sep + 1;
      }
    }
    args->push_back(a);
    i = close + 1;
  }
  return 0;
}

}  // namespace

int parse_pattern(const char *pattern, DigitalPattern *out) {
  if (!pattern || !out) return 16;
  std::string s(pattern);
  if (s.empty()) return 16;
  if (s.size() > 1024) return 17;
  // drop a trailing scalar factor "...)*2.0"; add a rank-0 right operand to unary patterns
  size_t star = s.find('*');
  if (star != std::string::npos && s.find('(', star) == std::string::npos) {
    s.resize(star);
    if (s.empty() || s.back() != ')') return 3;
    star = std::string::npos;
  }
  if (star == std::string::npos) s += "*R()";
  std::vector<Arg> args;
  int err = split_args(s, &args);
  if (err) return err;
  if (args.size() != 3) return 15;
  const int nd = (int)args[0].idx.size(), nl = (int)args[1].idx.size(), nr = (int)args[2].idx.size();
  std::vector<int> dig(nl + nr, 0);
  // every label must occur exactly twice (D-L, D-R or L-R), or D-L-R for a hyper index
  std::vector<std::pair<int, int>> occ;  // (arg, pos)
  std::vector<std::vector<bool>> done(3);
  for (int a = 0; a < 3; ++a) done[a].assign(args[a].idx.size(), false);
  for (int a = 0; a < 3; ++a)
    for (int x = 0; x < (int)args[a].idx.size(); ++x) {
      if (done[a][x]) continue;
      occ.clear();
      occ.emplace_back(a, x);
      for (int b = a; b < 3; ++b)
        for (int y = (b == a ? x + 1 : 0); y < (int)args[b].idx.size(); ++y)
          if (args[b].idx[y] == args[a].idx[x]) occ.emplace_back(b, y);
      if (occ.size() < 2) return 11;
      if (occ.size() > 3) return 10;
      for (auto &o : occ) done[o.first][o.second] = true;
      const int a0 = occ[0].first, p0 = occ[0].second, a1 = occ[1].first, p1 = occ[1].second;
      if (occ.size() == 3) {
        if (!(a0 == 0 && a1 == 1 && occ[2].first == 2)) return 9;
        dig[p1] = p0 + 1;
        dig[nl + occ[2].second] = p0 + 1;
      } else if (a0 == 0 && a1 == 1) {
        dig[p1] = p0 + 1;
      } else if (a0 == 0 && a1 == 2) {
        dig[nl + p1] = p0 + 1;
      } else if (a0 == 1 && a1 == 2) {
        dig[p0] = -(p1 + 1);
        dig[nl + p1] = -(p0 + 1);
      } else {
        return 10;
      }
    }
  out->drank = nd;
  out->lrank = nl;
  out->rrank = nr;
  out->conj = (args[0].conj ? 1 : 0) | (args[1].conj ? 2 : 0) | (args[2].conj ? 4 : 0);
  out->dig = dig;
  return 0;
}

namespace {

constexpr int kContr = 1000;  // symbol of a contracted pair = kContr + its L position

std::vector<int> filter(const std::vector<int> &seq, bool contracted) {
  std::vector<int> r;
  for (int s : seq)
    if ((s >= kContr) == contracted) r.push_back(s);
  return r;
}

std::vector<int> cat(const std::vector<int> &a, const std::vector<int> &b) {
  std::vector<int> r(a);
  r.insert(r.end(), b.begin(), b.end());
  return r;
}

// O2N map from `seq` (current order) to `target` (0-based positions).
std::vector<int> o2n_of(const std::vector<int> &seq, const std::vector<int> &target) {
  std::vector<int> o(seq.size());
  for (size_t i = 0; i < seq.size(); ++i)
    o[i] = (int)(std::find(target.begin(), target.end(), seq[i]) - target.begin());
  return o;
}

double vol_of(const int *dims, int n) {
  double v = 1;
  for (int i = 0; i < n; ++i) v *= dims[i];
  return v;
}

int plan_uncached(const DigitalPattern &p, const int *ddims, const int *ldims, const int *rdims, int accumulative,
                  ContractPlan *plan);

// Plans already made, keyed by ranks, digits, extents and the accumulate
// flag (the plan depends on nothing else).  The search over matricizations
// costs ~16 us per call on the host, as much as a small contraction's GPU
// work; repeated shapes (iterative solvers, benchmarks, serving templates)
// then plan in well under a microsecond.  Cleared when full.
constexpr size_t kPlanCacheMax = 4096;
std::mutex g_plan_mu;
std::map<std::vector<int>, ContractPlan> g_plans;

}  // namespace

int plan_contraction(const DigitalPattern &p, const int *ddims, const int *ldims, const int *rdims,
                     int accumulative, ContractPlan *plan) {
  const int nd = p.drank, nl = p.lrank, nr = p.rrank;
  if (nd < 0 || nl < 0 || nr < 0 || (int)p.dig.size() != nl + nr || !plan) return 1;
  std::vector<int> key;
  key.reserve(4 + 2 * (size_t)(nl + nr) + nd);
  key.push_back(nd);
  key.push_back(nl);
  key.push_back(nr);
  key.push_back(accumulative ? 1 : 0);
  key.insert(key.end(), p.dig.begin(), p.dig.end());
  key.insert(key.end(), ddims, ddims + nd);
  key.insert(key.end(), ldims, ldims + nl);
  key.insert(key.end(), rdims, rdims + nr);
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
      *plan = it->second;
      return 0;
    }
  }
  const int e = plan_uncached(p, ddims, ldims, rdims, accumulative, plan);
  if (e == 0) {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    if (g_plans.size() >= kPlanCacheMax) g_plans.clear();
    g_plans.emplace(std::move(key), *plan);
  }
  return e;
}

namespace {

int plan_uncached(const DigitalPattern &p, const int *ddims, const int *ldims, const int *rdims, int accumulative,
                  ContractPlan *plan) {
  const int nd = p.drank, nl = p.lrank, nr = p.rrank;
  if (nd < 0 || nl < 0 || nr < 0 || (int)p.dig.size() != nl + nr) return 1;
  std::vector<int> sl(nl), sr(nr), sd(nd);
  std::vector<int> dcount(nd, 0);
  std::vector<int64_t> ext_of(kContr + nl + 1, 0);  // extent per symbol
  for (int i = 0; i < nd; ++i) {
    sd[i] = i;
    ext_of[i] = ddims[i];
  }
  for (int i = 0; i < nl; ++i) {
    const int c = p.dig[i];
    if (c > 0) {
      if (c > nd || ldims[i] != ddims[c - 1]) return 1;
      sl[i] = c - 1;
      dcount[c - 1]++;
    } else if (c < 0) {
      const int j = -c - 1;
      if (j >= nr || p.dig[nl + j] != -(i + 1) || ldims[i] != rdims[j]) return 1;
      sl[i] = kContr + i;
      ext_of[kContr + i] = ldims[i];
    } else {
      return 1;
    }
  }
  for (int j = 0; j < nr; ++j) {
    const int c = p.dig[nl + j];
    if (c > 0) {
      if (c > nd || rdims[j] != ddims[c - 1]) return 1;
      sr[j] = c - 1;
      dcount[c - 1]++;
    } else if (c < 0) {
      const int i = -c - 1;
      if (i >= nl || p.dig[i] != -(j + 1)) return 1;
      sr[j] = kContr + i;
    } else {
      return 1;
    }
  }
  // a D index present in both inputs is a hyper index: it becomes a batch
  // dimension, the slowest of every operand of a strided batched GEMM
  std::vector<int> hyper;
  for (int i = 0; i < nd; ++i) {
    if (dcount[i] == 0) return 1;
    if (dcount[i] > 1) hyper.push_back(i);
  }
  auto is_hyper = [&](int s) { return s < kContr && dcount[s] > 1; };
  auto drop_hyper = [&](const std::vector<int> &v) {
    std::vector<int> r;
    for (int s : v)
      if (!is_hyper(s)) r.push_back(s);
    return r;
  };
  int64_t batch = 1;
  for (int s : hyper) batch *= ext_of[s];
  const double vd = vol_of(ddims, nd), vl = vol_of(ldims, nl), vr = vol_of(rdims, nr);

  bool have = false;
  ContractPlan best;
  int best_ntrans = 99;
  bool best_d0_first = false;
  for (int swap = 0; swap < 2; ++swap) {
    const std::vector<int> &sx = swap ? sr : sl, &sy = swap ? sl : sr;
    const int *xdims = swap ? rdims : ldims, *ydims = swap ? ldims : rdims;
    const int nx = swap ? nr : nl, ny = swap ? nl : nr;
    const double vx = swap ? vr : vl, vy = swap ? vl : vr;
    const std::vector<int> mx = drop_hyper(filter(sx, false)), ny_ = drop_hyper(filter(sy, false));
    std::vector<int> m_sorted(mx), n_sorted(ny_);
    std::sort(m_sorted.begin(), m_sorted.end());
    std::sort(n_sorted.begin(), n_sorted.end());
    const std::vector<int> kx = filter(sx, true), ky = filter(sy, true);
    for (int mo = 0; mo < 2; ++mo)
      for (int no = 0; no < 2; ++no)
        for (int ko = 0; ko < 2; ++ko) {
          const std::vector<int> &M = mo ? mx : m_sorted, &N = no ? ny_ : n_sorted, &K = ko ? ky : kx;
          ContractPlan c;
          c.swap = swap != 0;
          const std::vector<int> dseq = cat(cat(M, N), hyper);
          c.d_transpose = (dseq != sd);
          // operand X (A): 'T' form = K++M, 'N' form = M++K
          const std::vector<int> xt = cat(cat(K, M), hyper), xn = cat(cat(M, K), hyper);
          if (sx == xt) { c.a.op = 'T'; }
          else if (sx == xn) { c.a.op = 'N'; }
          else {
            // a copy is needed anyway: produce the K-contiguous form, which
            // every GEMM kernel stages into LDS without scattering
            c.a.transpose = true;
            c.a.op = 'T';
            c.a.o2n = o2n_of(sx, xt);
          }
          // operand Y (B): 'N' form = K++N, 'T' form = N++K
          const std::vector<int> yn = cat(cat(K, N), hyper), yt = cat(cat(N, K), hyper);
          if (sy == yn) { c.b.op = 'N'; }
          else if (sy == yt) { c.b.op = 'T'; }
          else {
            c.b.transpose = true;
            c.b.op = 'N';
            c.b.o2n = o2n_of(sy, yn);
          }
          c.M = 1;
          for (int s : M) c.M *= ext_of[s];
          c.N = 1;
          for (int s : N) c.N *= ext_of[s];
          c.K = 1;
          for (int s : K) c.K *= ext_of[s];
          c.batch = batch;
          // the GEMM stages K-contiguous operand tiles into LDS without
          // scattering: for long-K products an M-contiguous A ('N') or an
          // N-contiguous B ('T') costs more than transposing that operand
          // once the other GEMM dimension is large (5136^3 matrix product:
          // 48.5 TFLOP/s on the operands as given, 58.5 with L transposed).
          // Operands above 2^28 elements are left as given: their temporary
          // would take a large share of the device pool.
          constexpr double kMaxForcedCopy = 268435456.0;
          if (batch == 1 && c.K >= 1024) {
            if (!c.a.transpose && c.a.op == 'N' && c.N >= 1024 && vx <= kMaxForcedCopy) {
              c.a.transpose = true;
              c.a.op = 'T';
              c.a.o2n = o2n_of(sx, xt);
            }
            if (!c.b.transpose && c.b.op == 'T' && c.M >= 1024 && vy <= kMaxForcedCopy) {
              c.b.transpose = true;
              c.b.op = 'N';
              c.b.o2n = o2n_of(sy, yn);
            }
          }
          c.a.ld = std::max<int64_t>(1, c.a.op == 'T' ? c.K : c.M);
          c.b.ld = std::max<int64_t>(1, c.b.op == 'N' ? c.K : c.N);
          if (c.d_transpose) {
            c.d_o2n = o2n_of(sd, dseq);
            c.d_back_o2n = o2n_of(dseq, sd);
            c.dprime_dims.resize(nd);
            for (int i = 0; i < nd; ++i) c.dprime_dims[i] = (int)ext_of[dseq[i]];
            c.d_in_place = batch == 1 && c.d_o2n[0] == 0 && c.dprime_dims[0] >= 16;
          }
          // a D transpose done in place by the GEMM moves nothing extra
          c.moved_elems = (c.a.transpose ? vx : 0) + (c.b.transpose ? vy : 0) +
                          (c.d_transpose && !c.d_in_place ? vd * (accumulative ? 2 : 1) : 0);
          const int ntrans = (int)c.a.transpose + (int)c.b.transpose + (int)c.d_transpose;
          // among equal plans, one whose D' starts with D's first (stride-1) index lets the GEMM
          // write D in place with m-contiguous stores
          const bool d0_first = !c.d_transpose || c.d_o2n[0] == 0;
          if (!have || c.moved_elems < best.moved_elems ||
              (c.moved_elems == best.moved_elems &&
               (ntrans < best_ntrans || (ntrans == best_ntrans && d0_first && !best_d0_first)))) {
            best = c;
            best_ntrans = ntrans;
            best_d0_first = d0_first;
            have = true;
          }
          (void)nx;
          (void)ny;
          (void)xdims;
          (void)ydims;
        }
  }
  *plan = best;
  return 0;
}

}  // namespace

}  // namespace tal

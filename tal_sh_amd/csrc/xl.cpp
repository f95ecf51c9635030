// Extra-large contraction, talshTensorContractXL (talshc.cpp:5394-5640).
//
// The reference halves the operation recursively (talshTensorOpDecompose2,
// talshc.cpp:3150-3317) until each piece fits a device buffer entry, extracts
// every piece's operand slices on the host, contracts them on the GPU
// (COPY_TTT) and accumulates each piece into the host image of D.
//
// Here the index space is tiled instead.  Every index of the pattern gets a
// tile size; tiles are halved one index at a time, always the index whose
// halving shrinks a piece the most, until a piece -- the D, L and R boxes of
// one tile plus the TTGT temporaries -- fits half of the device memory granted
// to the call (two pieces are in flight).  Pieces are grouped by D box: a
// group brings the D box to the GPU once (accumulative calls), contracts all of
// its contracted-index tiles into it there, one after the other, and writes
// the box back once, so D crosses the host link at most twice whatever the
// split of the contracted indices.  Operand boxes are cut from the image on the
// execution GPU when there is one (slice kernel), otherwise moved from the host
// image by strided DMA copies; an operand that one tile covers and that already
// lives on the GPU is used in place.  Consecutive groups alternate between two
// streams, so one group's copies overlap the other's kernels.  As in the
// reference, D ends with a single, up-to-date host image.  Hyper indices are
// D indices that L and R share: their tiles cut all three operands alike.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "contract.hpp"
#include "plan.hpp"
#include "runtime.hpp"
#include "talsh_amd.h"

namespace tal {

// Copies the box [off, off + sd) of a dlf tensor with extents bd between a
// host body and a packed device buffer.  The box is cut into rows of
// contiguous bytes (the leading dims it spans in full plus the first one it
// cuts); consecutive rows over the next such run of dims form one 2-D copy;
// the dims beyond are enumerated here, one 2-D copy per position.
int box_dma(bool h2d, char *host, char *dev, int rank, const int *bd, const int *off, const int *sd, size_t esz,
            hipStream_t s) {
  std::vector<size_t> stride(rank + 1, 1);
  for (int k = 0; k < rank; ++k) stride[k + 1] = stride[k] * (size_t)bd[k];
  size_t base = 0;
  for (int k = 0; k < rank; ++k) base += (size_t)off[k] * stride[k];
  int k = 0;
  size_t row = 1;
  while (k < rank) {
    row *= (size_t)sd[k];
    const bool cut = sd[k] != bd[k];
    ++k;
    if (cut) break;
  }
  const int k2 = k;
  size_t height = 1;
  while (k < rank) {
    height *= (size_t)sd[k];
    const bool cut = sd[k] != bd[k];
    ++k;
    if (cut) break;
  }
  const int j = k;  // dims [j, rank) are enumerated
  const size_t pitch = stride[std::min(k2, rank)] * esz, width = row * esz;
  const hipMemcpyKind kind = h2d ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  std::vector<int> idx(rank, 0);
  size_t packed = 0;  // element offset of the current 2-D block in the device buffer
  while (true) {
    size_t ho = base;
    for (int m = j; m < rank; ++m) ho += (size_t)idx[m] * stride[m];
    char *hp = host + ho * esz, *dp = dev + packed * esz;
    hipError_t e;
    if (height == 1) {
      e = h2d ? hipMemcpyAsync(dp, hp, width, kind, s) : hipMemcpyAsync(hp, dp, width, kind, s);
    } else if (pitch < (size_t(1) << 31)) {
      e = h2d ? hipMemcpy2DAsync(dp, width, hp, pitch, width, height, kind, s)
              : hipMemcpy2DAsync(hp, pitch, dp, width, width, height, kind, s);
    } else {  // very long pitches: one copy per row
      e = hipSuccess;
      for (size_t r = 0; r < height && e == hipSuccess; ++r)
        e = h2d ? hipMemcpyAsync(dp + r * width, hp + r * pitch, width, kind, s)
                : hipMemcpyAsync(hp + r * pitch, dp + r * width, width, kind, s);
    }
    if (e != hipSuccess) return TALSH_FAILURE;
    packed += row * height;
    int m = j;
    while (m < rank && ++idx[m] == sd[m]) idx[m++] = 0;
    if (m == rank) break;
  }
  return TALSH_SUCCESS;
}

namespace {

// Where an operand's body is read from (or, for D, kept) during the call.
struct Home {
  char *p = nullptr;
  bool gpu = false;  // on the execution GPU, else host memory
};

int pick_home(talsh_tens_t *t, int g, int dk, Home *h) {
  int i = image_find(t, flat_gpu(g), dk, true);
  if (i >= 0) {
    h->p = (char *)t->dev_rsc[i].gmem_p;
    h->gpu = true;
    return TALSH_SUCCESS;
  }
  i = image_find(t, DEV_HOST, dk, true);
  if (i < 0) {  // only on another GPU: bring a host copy first
    if (int e = talshTensorPlace(t, 0, DEV_HOST, nullptr, COPY_K, nullptr)) return e;
    i = image_find(t, DEV_HOST, dk, true);
    if (i < 0) return TALSH_NOT_AVAILABLE;
  }
  h->p = (char *)t->dev_rsc[i].gmem_p;
  h->gpu = false;
  return TALSH_SUCCESS;
}

struct Buf {
  void *p = nullptr;
  int h = -1;
};

// Device buffers of one group, released once its completion event has passed.
struct Slot {
  std::vector<Buf> bufs;
  hipEvent_t done = nullptr;
  bool busy = false;
};

void slot_release(int g, Slot &sl) {
  if (sl.busy) (void)hipEventSynchronize(sl.done);
  for (Buf &b : sl.bufs) dev_free(g, b.p, b.h);
  sl.bufs.clear();
  sl.busy = false;
}

void *slot_alloc(int g, Slot &sl, size_t bytes, int *err) {
  Buf b;
  *err = dev_alloc(g, std::max<size_t>(bytes, 1), &b.p, &b.h);
  if (*err) return nullptr;
  sl.bufs.push_back(b);
  return b.p;
}

}  // namespace
}  // namespace tal

using namespace tal;

int talshAmdSetContractXLMemory(size_t bytes) {
  std::lock_guard<std::recursive_mutex> lk(rt().mu);
  rt().xl_mem = bytes;
  return TALSH_SUCCESS;
}

int talshTensorContractXL(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                          double scale_real, double scale_imag, int dev_id, int dev_kind, int accumulative) {
  Runtime &r = rt();
  std::lock_guard<std::recursive_mutex> lk(r.mu);
  if (!r.on) return TALSH_NOT_INITIALIZED;
  if (!cptrn || !dtens || !ltens || !rtens) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(dtens) != NOPE || talshTensorIsEmpty(ltens) != NOPE || talshTensorIsEmpty(rtens) != NOPE)
    return TALSH_OBJECT_IS_EMPTY;
  if (tensor_healthy(dtens) != YEP || tensor_healthy(ltens) != YEP || tensor_healthy(rtens) != YEP)
    return TALSH_OBJECT_BROKEN;
  if (dtens == ltens || dtens == rtens) return TALSH_INVALID_ARGS;
  DigitalPattern p;
  if (parse_pattern(cptrn, &p) != 0) return TALSH_INVALID_ARGS;
  if (p.drank != dtens->shape_p->num_dim || p.lrank != ltens->shape_p->num_dim || p.rrank != rtens->shape_p->num_dim)
    return TALSH_INVALID_ARGS;
  const int *dd = dtens->shape_p->dims, *ld = ltens->shape_p->dims, *rd = rtens->shape_p->dims;
  const bool acc = accumulative != NOPE;
  {
    ContractPlan whole;
    if (plan_contraction(p, dd, ld, rd, acc, &whole) != 0) return TALSH_INVALID_ARGS;
  }
  int g;
  if (int e = exec_gpu(dev_id, dev_kind, {dtens, ltens, rtens}, &g)) return e;
  Device &dev = r.gpus[g];
  (void)hipSetDevice(dev.hip_id);

  // execution data kind: D's (talshc.cpp:5466-5471)
  int dk = NO_TYPE;
  for (int i = 0; i < dtens->ndev; ++i)
    if (dtens->avail[i] == YEP) { dk = dtens->data_kind[i]; break; }
  if (dk == NO_TYPE) return TALSH_FAILURE;
  Home hd, hl, hr;
  if (int e = pick_home(dtens, g, dk, &hd)) return e;
  if (int e = pick_home(ltens, g, dk, &hl)) return e == TALSH_NOT_AVAILABLE ? TALSH_INVALID_ARGS : e;
  if (int e = pick_home(rtens, g, dk, &hr)) return e == TALSH_NOT_AVAILABLE ? TALSH_INVALID_ARGS : e;
  // the call owns the bodies from here on: wait for earlier work on any image
  for (talsh_tens_t *t : {dtens, ltens, rtens})
    for (int i = 0; i < t->ndev; ++i) hazard_drain(t->dev_rsc[i].gmem_p, image_bytes(t, i));

  // ---- index space: D indices [0, drank), then one id per contracted pair
  const int nd = p.drank;
  std::vector<int> lid(p.lrank), rid(p.rrank), ext(nd);
  for (int k = 0; k < nd; ++k) ext[k] = dd[k];
  for (int i = 0; i < p.lrank; ++i)
    if (p.dig[i] > 0) lid[i] = p.dig[i] - 1;
  for (int j = 0; j < p.rrank; ++j)
    if (p.dig[p.lrank + j] > 0) rid[j] = p.dig[p.lrank + j] - 1;
  for (int i = 0; i < p.lrank; ++i)
    if (p.dig[i] < 0) {
      const int id = (int)ext.size();
      ext.push_back(ld[i]);
      lid[i] = id;
      rid[-p.dig[i] - 1] = id;
    }
  const int nid = (int)ext.size();
  // slowness of each index: sum of its dim positions (halving slow dims keeps DMA rows long)
  std::vector<int> slow(nid, 0);
  for (int k = 0; k < nd; ++k) slow[k] += k;
  for (int i = 0; i < p.lrank; ++i) slow[lid[i]] += i;
  for (int j = 0; j < p.rrank; ++j) slow[rid[j]] += j;

  const size_t esz = (size_t)dk_size(dk);
  auto vol_of = [&](const std::vector<int> &ids, const std::vector<int> &t) {
    size_t v = 1;
    for (int id : ids) v *= (size_t)t[id];
    return v;
  };
  std::vector<int> did(nd);
  for (int k = 0; k < nd; ++k) did[k] = k;
  auto covers = [&](const std::vector<int> &ids, const std::vector<int> &t) {
    for (int id : ids)
      if (t[id] != ext[id]) return false;
    return true;
  };
  // device bytes of one piece: buffers for the boxes that are not used in place,
  // plus temporaries bounded by one more copy of each operand and the split planes
  auto need = [&](const std::vector<int> &t) {
    const size_t dv = vol_of(did, t) * esz, lv = vol_of(lid, t) * esz, rv = vol_of(rid, t) * esz;
    size_t b = dv + lv + rv;
    if (covers(did, t) && hd.gpu) b -= dv;
    if (covers(lid, t) && hl.gpu) b -= lv;
    if (covers(rid, t) && hr.gpu) b -= rv;
    b += dv + lv + rv;
    int64_t M = 1, N = 1, K = 1;  // a bound: hyper indices count in M and N
    for (int i = 0; i < p.lrank; ++i)
      if (lid[i] < nd) M *= t[lid[i]];
      else K *= t[lid[i]];
    for (int j = 0; j < p.rrank; ++j)
      if (rid[j] < nd) N *= t[rid[j]];
    if (dk == R4 && r.r4mode == R4_SPLIT_BF16) b += gemm_split_workspace(r.split_method, M, N, K);
    else b += gemm_tail_workspace(dk, M, N, K, 1);
    return b;
  };
  size_t budget = dev.arena.largest_free() / 10 * 9;
  if (r.xl_mem > 0) budget = std::min(budget, r.xl_mem);
  std::vector<int> tile = ext;
  auto groups_of = [&](const std::vector<int> &t) {
    size_t n = 1;
    for (int k = 0; k < nd; ++k) n *= (size_t)((ext[k] + t[k] - 1) / t[k]);
    return n;
  };
  while (need(tile) * (groups_of(tile) > 1 ? 2 : 1) > budget) {
    int best = -1;
    size_t best_need = 0;
    for (int id = 0; id < nid; ++id) {
      if (tile[id] <= 1) continue;
      std::vector<int> t = tile;
      t[id] = (t[id] + 1) / 2;
      const size_t n = need(t);
      const bool better = best < 0 || n < best_need ||
                          (n == best_need && ((id < nd) != (best < nd) ? id < nd : slow[id] > slow[best]));
      if (better) { best = id; best_need = n; }
    }
    if (best < 0) return DEVICE_UNABLE;
    tile[best] = (tile[best] + 1) / 2;
  }
  // With host operands and a single piece nothing would overlap the host link
  // with the GPU: cut the larger host-resident input along one of its D
  // indices (the slowest, so DMA rows stay long) into kStreamBoxes boxes, so
  // that one box crosses the link while another is contracted; the other input
  // then crosses once, whole.
  constexpr int kStreamBoxes = 4;
  if (groups_of(tile) == 1 && covers(lid, tile) && covers(rid, tile) && (!hl.gpu || !hr.gpu)) {
    const size_t lvol = vol_of(lid, ext), rvol = vol_of(rid, ext);
    const bool use_l = !hl.gpu && (hr.gpu || lvol >= rvol);
    int best = -1;
    for (int id : use_l ? lid : rid)
      if (id < nd && ext[id] >= kStreamBoxes &&
          (best < 0 || slow[id] > slow[best] || (slow[id] == slow[best] && ext[id] > ext[best])))
        best = id;
    if (best >= 0) {
      std::vector<int> t = tile;
      t[best] = (ext[best] + kStreamBoxes - 1) / kStreamBoxes;
      if (2 * need(t) <= budget) tile = t;
    }
  }
  std::vector<int> ntile(nid);
  size_t npieces = 1;
  for (int id = 0; id < nid; ++id) {
    ntile[id] = (ext[id] + tile[id] - 1) / tile[id];
    npieces *= (size_t)ntile[id];
  }
  const size_t ngroups = groups_of(tile);
  const size_t nk = npieces / ngroups;
  if (r.log_ops)
    std::printf("#TALSH OP: talshTensorContractXL on GPU %d: %zu pieces (%zu D boxes x %zu contracted tiles)\n", g,
                npieces, ngroups, nk);

  const bool d_in_place = covers(did, tile) && hd.gpu;
  bool l_in_place = covers(lid, tile) && hl.gpu;
  bool r_in_place = covers(rid, tile) && hr.gpu;
  hipStream_t streams[2] = {dev.streams[0], dev.streams[dev.streams.size() > 1 ? 1 : 0]};
  Slot slots[2], whole;  // whole: inputs moved to the GPU once for the whole call
  for (Slot &sl : slots)
    if (hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) return TALSH_FAILURE;
  if (hipEventCreateWithFlags(&whole.done, hipEventDisableTiming) != hipSuccess) return TALSH_FAILURE;
  int err = TALSH_SUCCESS;
  double flops = 0.0, moved_in = 0.0, moved_out = 0.0;
  // an input that one tile covers but that lives on the host crosses the link
  // once, up front, into a buffer that every piece reads in place
  auto upload_whole = [&](Home &h, bool &in_place, const std::vector<int> &ids, int rank, const int *dims) {
    if (err || in_place || h.gpu || !covers(ids, tile)) return;
    const size_t bytes = vol_of(ids, ext) * esz;
    void *buf = slot_alloc(g, whole, bytes, &err);
    if (err) return;
    std::vector<int> zero(rank, 0);
    if (box_dma(true, h.p, (char *)buf, rank, dims, zero.data(), dims, esz, streams[0])) {
      err = TALSH_FAILURE;
      return;
    }
    moved_in += (double)bytes;
    h.p = (char *)buf;
    h.gpu = true;
    in_place = true;
    whole.busy = true;
  };
  upload_whole(hl, l_in_place, lid, p.lrank, ld);
  upload_whole(hr, r_in_place, rid, p.rrank, rd);
  if (!err && whole.busy) {
    if (hipEventRecord(whole.done, streams[0]) != hipSuccess ||
        hipStreamWaitEvent(streams[1], whole.done, 0) != hipSuccess)
      err = TALSH_FAILURE;
  }
  std::vector<int> pos(nid, 0);  // tile coordinates of the current piece
  std::vector<int> pd(nd), pl(p.lrank), pr(p.rrank), od(nd), ol(p.lrank), orr(p.rrank);
  auto piece_dims = [&]() {
    for (int k = 0; k < nd; ++k) {
      od[k] = pos[k] * tile[k];
      pd[k] = std::min(tile[k], ext[k] - od[k]);
    }
    for (int i = 0; i < p.lrank; ++i) {
      const int id = lid[i];
      ol[i] = pos[id] * tile[id];
      pl[i] = std::min(tile[id], ext[id] - ol[i]);
    }
    for (int j = 0; j < p.rrank; ++j) {
      const int id = rid[j];
      orr[j] = pos[id] * tile[id];
      pr[j] = std::min(tile[id], ext[id] - orr[j]);
    }
  };
  // fetch a box into a packed device buffer (or hand out the body in place)
  auto fetch = [&](const Home &h, bool in_place, int rank, const int *bd, const int *off, const int *sd, void *buf,
                   hipStream_t s) -> const void * {
    if (in_place) return h.p;
    if (h.gpu) {
      if (launch_slice(dk, 0, h.p, buf, rank, sd, bd, off, 0, s)) return nullptr;
    } else {
      if (box_dma(true, h.p, (char *)buf, rank, bd, off, sd, esz, s)) return nullptr;
      size_t v = 1;
      for (int k = 0; k < rank; ++k) v *= (size_t)sd[k];
      moved_in += (double)(v * esz);
    }
    return buf;
  };
  const bool cplx = (dk == C4 || dk == C8);
  for (size_t gi = 0; gi < ngroups && err == TALSH_SUCCESS; ++gi) {
    Slot &sl = slots[gi & 1];
    hipStream_t s = streams[gi & 1];
    slot_release(g, sl);
    // D box of this group; contracted tiles start at 0
    {
      size_t q = gi;
      for (int k = 0; k < nd; ++k) {
        pos[k] = (int)(q % (size_t)ntile[k]);
        q /= (size_t)ntile[k];
      }
      for (int id = nd; id < nid; ++id) pos[id] = 0;
    }
    piece_dims();
    // the group's first piece is its largest: size the buffers for it
    const size_t dv = vol_of(did, tile) * esz, lv = vol_of(lid, tile) * esz, rv = vol_of(rid, tile) * esz;
    void *dbuf = d_in_place ? hd.p : slot_alloc(g, sl, dv, &err);
    void *lbuf = (l_in_place || err) ? nullptr : slot_alloc(g, sl, lv, &err);
    void *rbuf = (r_in_place || err) ? nullptr : slot_alloc(g, sl, rv, &err);
    if (err) break;
    if (acc && !d_in_place) {
      if (!fetch(hd, false, nd, dd, od.data(), pd.data(), dbuf, s)) { err = TALSH_FAILURE; break; }
    }
    std::vector<int> pd_box = pd, od_box = od;
    TtgtBuffers tmp;
    TtgtTemps have;
    for (size_t kt = 0; kt < nk && err == TALSH_SUCCESS; ++kt) {
      {
        size_t q = kt;
        for (int id = nd; id < nid; ++id) {
          pos[id] = (int)(q % (size_t)ntile[id]);
          q /= (size_t)ntile[id];
        }
      }
      piece_dims();
      ContractPlan plan;
      if (plan_contraction(p, pd.data(), pl.data(), pr.data(), acc || kt > 0, &plan) != 0) { err = TALSH_FAILURE; break; }
      size_t pdv = esz, plv = esz, prv = esz;
      for (int k = 0; k < nd; ++k) pdv *= (size_t)pd[k];
      for (int i = 0; i < p.lrank; ++i) plv *= (size_t)pl[i];
      for (int j = 0; j < p.rrank; ++j) prv *= (size_t)pr[j];
      const TtgtTemps want = ttgt_temp_bytes(plan, dk, plv, prv, pdv);
      // temporaries grow on demand (sized for the group's full tile)
      auto grow = [&](size_t want_b, size_t &have_b, void *&ptr, size_t full) {
        if (want_b == 0 || want_b <= have_b) return;
        const size_t b = std::max(want_b, full);
        ptr = slot_alloc(g, sl, b, &err);
        have_b = b;
      };
      const size_t fa = plan.swap ? rv : lv, fb = plan.swap ? lv : rv;
      grow(want.a, have.a, tmp.a, fa);
      grow(want.b, have.b, tmp.b, fb);
      grow(want.d, have.d, tmp.d, want.d == pdv ? dv : 0);  // a D' temporary, or the GEMM's small offset table
      grow(want.w, have.w, tmp.w, 0);
      grow(want.s, have.s, tmp.s, 0);
      tmp.s_bytes = have.s;
      tmp.split = want.split;
      if (err) break;
      const void *lp = fetch(hl, l_in_place, p.lrank, ld, ol.data(), pl.data(), lbuf, s);
      const void *rp = fetch(hr, r_in_place, p.rrank, rd, orr.data(), pr.data(), rbuf, s);
      if (!lp || !rp) { err = TALSH_FAILURE; break; }
      if (int e = enqueue_ttgt(p, plan, dk, acc || kt > 0, scale_real, scale_imag, dbuf, lp, rp, pd.data(),
                               pl.data(), pr.data(), tmp, s, nullptr, nullptr)) {
        err = e;
        break;
      }
      flops += (cplx ? 8.0 : 2.0) * (double)plan.M * (double)plan.N * (double)plan.K * (double)plan.batch;
    }
    if (err) break;
    if (!d_in_place) {
      if (hd.gpu) {
        if (launch_slice(dk, 1, dbuf, hd.p, nd, pd_box.data(), dd, od_box.data(), 0, s)) { err = TALSH_FAILURE; break; }
      } else {
        if (box_dma(false, hd.p, (char *)dbuf, nd, dd, od_box.data(), pd_box.data(), esz, s)) {
          err = TALSH_FAILURE;
          break;
        }
        moved_out += (double)(vol_of(did, tile) * esz);
      }
    }
    if (hipEventRecord(sl.done, s) != hipSuccess) { err = TALSH_FAILURE; break; }
    sl.busy = true;
  }
  if (err) (void)hipStreamSynchronize(streams[0]), (void)hipStreamSynchronize(streams[1]);
  for (Slot &sl : slots) {
    slot_release(g, sl);
    (void)hipEventDestroy(sl.done);
  }
  slot_release(g, whole);  // every slot's work, which reads these, has completed
  (void)hipEventDestroy(whole.done);
  if (err) {
    dev.tasks_failed++;
    return err == TRY_LATER || err == DEVICE_UNABLE ? err : TALSH_FAILURE;
  }
  if (r.log_ops)
    std::printf("#TALSH OP: talshTensorContractXL done: %.6g flop, %.6g bytes in, %.6g bytes out\n", flops,
                moved_in, moved_out);
  dev.flops += flops;
  dev.traffic_in += moved_in;
  dev.traffic_out += moved_out;
  dev.tasks_submitted++;
  dev.tasks_completed++;
  // D keeps one image, on the host (talshc.cpp:5460-5461)
  const int keep = image_find(dtens, hd.gpu ? flat_gpu(g) : DEV_HOST, dk, true);
  if (keep < 0) return TALSH_FAILURE;
  image_discard_other(dtens, keep);
  if (hd.gpu) return talshTensorPlace(dtens, 0, DEV_HOST, nullptr, COPY_M, nullptr);
  return TALSH_SUCCESS;
}

int talshTensorContractXL_(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                           double scale_real, double scale_imag, int dev_id, int dev_kind, int accumulative) {
  return talshTensorContractXL(cptrn, dtens, ltens, rtens, scale_real, scale_imag, dev_id, dev_kind, accumulative);
}

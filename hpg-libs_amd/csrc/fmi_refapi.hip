// fmi_refapi.hip -- the reference's C surface of the search path
// (include/hpg_search.h) over the engine, and the general per-task kernels
// it needs: exact search and interval vectors from any starting interval and
// window, single LF steps, the anchors of BWSearch1CPU, and
// BWSearch1VectorHelper driven by the caller's vectors.
//
// A per-read call is a batch of one on a stream of the calling thread: the
// read and the inputs go to a device staging buffer of that thread, the
// batch runs, the outputs come back and are written into the caller's
// result / results_list the way the reference's function writes them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fmi.h"
#include "../../include/hpg_search.h"
#include "fmi_device.h"
#include "fmi_index.h"
#include "fmi_internal.h"
#include "fmi_kernels.h"
#include "fmi_vhelper.h"

static_assert(sizeof(result) == 112 && offsetof(result, k) == 8 && offsetof(result, start) == 24 &&
                  offsetof(result, num_mismatches) == 30 && offsetof(result, err_kind) == 31 &&
                  offsetof(result, err_base) == 51 && offsetof(result, err_pos) == 72,
              "result must keep the layout of search/results.h:13-21");

namespace fmi {

// --------------------------------------------------------------------------
// kernels
// --------------------------------------------------------------------------

// BWExactSearchBackward / Forward (search/search.h:9-65) from each task's own
// record: in[t].k, .l, .start, .pos, .end (k, l as the reference's intmax_t,
// given here as rows 0..n+1 and -1..n); out[t] gets k, l and pos, anchor[t]
// (may be NULL) the last non-empty interval -- (0, 0) when the first step
// already empties -- with bounds [i + 1, pos] (backward) or [pos, i - 1]
// (forward) and pos = i, i being where the loop stopped.
__global__ __launch_bounds__(256) void gen_exact_kernel(DevIndex ix, uint32_t dsel, uint32_t comp,
                                                        uint32_t direction, const uint8_t *reads,
                                                        const uint64_t *offsets, const fmi_hit *in,
                                                        uint32_t ntasks, fmi_hit *out,
                                                        fmi_hit *anchor) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntasks) return;
  const DevDir &D = dir_of(ix, dsel);
  const uint8_t *W = reads + offsets[t];
  const fmi_hit r = in[t];
  uint32_t k2 = (uint32_t)r.k, l2 = (uint32_t)r.l, k = 0u, l = 0u;
  int32_t i = r.pos;
  if (direction == 0) {
    for (; i >= r.start; --i) {
      const uint32_t s = W[i];
      if (s > 3u) { k2 = ix.n + 1u; l2 = ix.n; }
      else { uint32_t nb; lf_one(D, k2, l2, comp ? 3u - s : s, k2, l2, nb); }
      if (k2 > l2) break;
      k = k2; l = l2;
    }
    if (anchor)
      store_hit(anchor + t, k, l, (int16_t)(i + 1), (int16_t)i, r.pos, 0, 0, 0, 0, 0, 0);
  } else {
    for (; i <= r.end; ++i) {
      const uint32_t s = W[i];
      if (s > 3u) { k2 = ix.n + 1u; l2 = ix.n; }
      else { uint32_t nb; lf_one(D, k2, l2, comp ? 3u - s : s, k2, l2, nb); }
      if (k2 > l2) break;
      k = k2; l = l2;
    }
    if (anchor)
      store_hit(anchor + t, k, l, r.pos, (int16_t)i, (int16_t)(i - 1), 0, 1, 0, 0, 0, 0);
  }
  store_hit(out + t, k2, l2, r.start, (int16_t)i, r.end, 0, r.dir, 0, 0, 0, 0);
}

// BWExactSearchVectorBackward / Forward (search/search.c:1407-1489) from
// (k, l) over W[start..end] of each task: vk/vl[voff[t] + j] for j = 0 ..
// end - start, last_good[t] the return value.  The reference returns 0
// without writing when k > l or start > end.
__global__ __launch_bounds__(256) void gen_vector_kernel(DevIndex ix, uint32_t dsel, uint32_t comp,
                                                         uint32_t direction, const uint8_t *reads,
                                                         const uint64_t *offsets,
                                                         const int64_t *kl, const int16_t *se,
                                                         const uint64_t *voff, uint32_t ntasks,
                                                         int64_t *vk, int64_t *vl,
                                                         int16_t *last_good) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntasks) return;
  const DevDir &D = dir_of(ix, dsel);
  const uint8_t *W = reads + offsets[t];
  const int64_t k = kl[2 * t], l = kl[2 * t + 1];
  const int32_t start = se[2 * t], end = se[2 * t + 1];
  if (k > l || start > end) { last_good[t] = 0; return; }
  int64_t *ok = vk + voff[t], *ol = vl + voff[t];
  uint32_t k2 = (uint32_t)k, l2 = (uint32_t)l;
  int32_t good = 0;
  auto step = [&](int32_t i) {
    const uint32_t s = W[i];
    if (s > 3u) { k2 = ix.n + 1u; l2 = ix.n; return; }
    uint32_t nb;
    lf_one(D, k2, l2, comp ? 3u - s : s, k2, l2, nb);
  };
  if (direction == 0) {
    int32_t i = end, j = end - start;
    for (; i >= start; --i, --j) {
      step(i);
      ok[j] = k2; ol[j] = l2;
      if (k2 > l2) { --i; --j; break; }
      ++good;
    }
    for (; i >= start; --i, --j) { ok[j] = k2; ol[j] = l2; }
  } else {
    int32_t i = start, j = 0;
    for (; i <= end; ++i, ++j) {
      step(i);
      ok[j] = k2; ol[j] = l2;
      if (k2 > l2) { ++i; ++j; break; }
      ++good;
    }
    for (; i <= end; ++i, ++j) { ok[j] = k2; ol[j] = l2; }
  }
  last_good[t] = (int16_t)good;
}

// BWiteration (search/runtime.h:38-42) on each task's (k, l, b).
__global__ __launch_bounds__(256) void lf_batch_kernel(DevIndex ix, uint32_t dsel, uint32_t comp,
                                                       const int64_t *k, const int64_t *l,
                                                       const uint8_t *b, uint32_t ntasks,
                                                       int64_t *ko, int64_t *lo) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntasks) return;
  const uint32_t s = b[t];
  if (s > 3u) { ko[t] = ix.n + 1u; lo[t] = ix.n; return; }
  uint32_t a, c, nb;
  lf_one(dir_of(ix, dsel), (uint32_t)k[t], (uint32_t)l[t], comp ? 3u - s : s, a, c, nb);
  ko[t] = a;
  lo[t] = c;
}

// The two anchors of BWSearch1CPU (search/search.c:1876-1933) per (read,
// strand): pass 1 is BWExactSearchBackward of W[half..len-1] on the pair's
// first index, pass 2 BWExactSearchForward of W[0..half-1] on its second
// (half = len / 2), both from [0, n]; anchors[4 * read + 2 * strand + pass].
__global__ __launch_bounds__(256) void search1_anchor_kernel(DevIndex ix, uint32_t strand_mask,
                                                             const uint8_t *reads,
                                                             const uint64_t *offsets,
                                                             uint32_t nreads, fmi_hit *anchors) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4u * nreads) return;
  const uint32_t rd = t >> 2, strand = (t >> 1) & 1u, pass = t & 1u;
  if (!(strand_mask & (1u << strand))) return;
  const uint8_t *W = reads + offsets[rd];
  const int32_t len = (int32_t)(offsets[rd + 1] - offsets[rd]);
  const int32_t half = len / 2;
  // + strand: (backward, forward); - strand: (forward_rev, backward_rev)
  const uint32_t dsel = (pass == 0) == (strand == 0) ? 0u : 1u;
  const DevDir &D = dir_of(ix, dsel);
  uint32_t k2 = 0u, l2 = ix.n, k = 0u, l = 0u;
  auto step = [&](int32_t i) {
    const uint32_t s = W[i];
    if (s > 3u) { k2 = ix.n + 1u; l2 = ix.n; return; }
    uint32_t nb;
    lf_one(D, k2, l2, strand ? 3u - s : s, k2, l2, nb);
  };
  if (pass == 0) {
    int32_t i = len - 1;
    for (; i >= half; --i) {
      step(i);
      if (k2 > l2) break;
      k = k2; l = l2;
    }
    store_hit(anchors + t, k, l, (int16_t)(i + 1), (int16_t)i, (int16_t)(len - 1), 0, 0, 0, 0, 0,
              (uint8_t)strand);
  } else {
    const int32_t end = half - 1;      // search.c:1912-1914: half-- then [start, half]
    int32_t i = 0;
    for (; i <= end; ++i) {
      step(i);
      if (k2 > l2) break;
      k = k2; l = l2;
    }
    store_hit(anchors + t, k, l, 0, (int16_t)i, (int16_t)(i - 1), 0, 1, 0, 0, 0, (uint8_t)strand);
  }
}

// BWSearch1VectorHelper on one window per task with the caller's vectors
// (vk/vl: the backward vector, vki/vli: the forward vector, n + 1 entries
// each at voff[t], n = end - start); lanes stride over the tasks with a
// scratch slice of kMaxReadLen + 3 intervals each.
struct VhGivenParams {
  DevIndex ix;
  const uint8_t *reads;
  const uint64_t *offsets;
  const int16_t *se;
  const uint64_t *voff;
  const int64_t *vk, *vl, *vki, *vli;
  const uint8_t *strand;
  fmi_hit *hits;
  uint32_t *counts;
  uint2 *scratch;
  uint32_t ntasks, cap;
};

__global__ __launch_bounds__(256) void vhelper_given_kernel(VhGivenParams P) {
  const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lanes = gridDim.x * blockDim.x;
  uint2 *vb = P.scratch + (uint64_t)lane * (kMaxReadLen + 3u);
  for (uint32_t t = lane; t < P.ntasks; t += lanes) {
    const int32_t start = P.se[2 * t], end = P.se[2 * t + 1];
    const int32_t n = end - start, half = n / 2;
    const uint64_t v = P.voff[t];
    for (int32_t j = 0; j <= half; ++j)
      vb[j] = make_uint2((uint32_t)P.vk[v + j], (uint32_t)P.vl[v + j]);
    uint2 *vf = vb + (half + 1);
    for (int32_t j = half - 1; j <= n; ++j)
      vf[j - (half - 1)] = make_uint2((uint32_t)P.vki[v + j], (uint32_t)P.vli[v + j]);
    VhSink S{P.hits + (uint64_t)t * P.cap, P.cap, 0u, 0u, P.strand[t], (int16_t)start,
             (int16_t)end};
    vh_replay(P.ix, VhBytes{P.reads + P.offsets[t]}, start, end, P.strand[t], 0u, vb, S);
    P.counts[t] = S.cnt;
  }
}

}  // namespace fmi

// --------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------

namespace {

thread_local int g_status = FMI_OK;

int status(int rc) {
  g_status = rc;
  return rc;
}

int launch_rc(hipError_t e) { return e == hipSuccess ? FMI_OK : FMI_EDEVICE; }

// Per-thread, per-device stream and staging buffers of the per-read calls.
struct Ctx {
  hipStream_t s = nullptr;
  void *d = nullptr;
  size_t dcap = 0;
};

// Kept until the thread ends (no teardown: a thread-exit destructor could
// run after the HIP runtime has gone at process exit).
struct ThreadCtx {
  std::map<int, Ctx> by_dev;
};
thread_local ThreadCtx g_ctx;

// Makes the index's device current and returns the thread's context there,
// with at least `bytes` of device staging.
int ctx_for(const fmi_index *ix, size_t bytes, Ctx **out) {
  if (!ix || !ix->on_device) return FMI_ESTATE;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return FMI_EDEVICE;
  if (cur != ix->device && hipSetDevice(ix->device) != hipSuccess) return FMI_EDEVICE;
  Ctx &c = g_ctx.by_dev[ix->device];
  if (!c.s && hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking) != hipSuccess) return FMI_EDEVICE;
  if (c.dcap < bytes) {
    if (c.d) {
      if (hipStreamSynchronize(c.s) != hipSuccess) return FMI_EDEVICE;
      (void)hipFree(c.d);
      c.d = nullptr;
      c.dcap = 0;
    }
    const size_t want = std::max<size_t>(bytes, 1u << 20);
    if (hipMalloc(&c.d, want) != hipSuccess) return FMI_ENOMEM;
    c.dcap = want;
  }
  *out = &c;
  return FMI_OK;
}

// Carves aligned pieces out of a staging buffer.
struct Carve {
  char *base;
  size_t off = 0;
  template <class T>
  T *take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T *p = reinterpret_cast<T *>(base + off);
    off += std::max<size_t>(count, 1) * sizeof(T);
    return p;
  }
};
size_t carve_size(std::initializer_list<size_t> bytes) {
  size_t s = 0;
  for (size_t b : bytes) s += ((std::max<size_t>(b, 1) + 255) & ~size_t(255));
  return s + 256;
}

bool view_ok(const bwt_index *v) {
  return v && v->fmi && v->view >= FMI_VIEW_BACKWARD && v->view <= FMI_VIEW_FORWARD_REV;
}
uint32_t dsel_of(int view) { return (view == FMI_VIEW_FORWARD || view == FMI_VIEW_FORWARD_REV) ? 1u : 0u; }
uint32_t comp_of(int view) { return view >= FMI_VIEW_BACKWARD_REV ? 1u : 0u; }

// The strand of a (backward, forward) pair as the mapping layer forms them
// (bwt.c:1690-1701): 0 for (BACKWARD, FORWARD), 1 for (FORWARD_REV,
// BACKWARD_REV), -1 otherwise.
int pair_strand(const bwt_index *b, const bwt_index *f) {
  if (!view_ok(b) || !view_ok(f) || b->fmi != f->fmi) return -1;
  if (b->view == FMI_VIEW_BACKWARD && f->view == FMI_VIEW_FORWARD) return 0;
  if (b->view == FMI_VIEW_FORWARD_REV && f->view == FMI_VIEW_BACKWARD_REV) return 1;
  return -1;
}

// Writes an engine record into a reference result as copy_result does: the
// fields of the record and its error entries, nothing past them.
void to_result(const fmi_hit &h, result *r) {
  r->dir = h.dir != 0;
  r->num_mismatches = 0;
  r->start = h.start;
  r->end = h.end;
  r->k = h.k;
  r->l = h.l;
  r->pos = h.pos;
  if (h.num_mismatches) {
    r->err_kind[0] = h.err_kind;
    r->err_base[0] = h.err_base;
    r->err_pos[0] = h.err_pos;
    r->num_mismatches = 1;
  }
}

void to_result_k(const fmi_hitk &h, result *r) {
  r->dir = h.dir != 0;
  r->start = h.start;
  r->end = h.end;
  r->k = h.k;
  r->l = h.l;
  r->pos = h.pos;
  r->num_mismatches = 0;
  for (uint32_t e = 0; e < h.num_mismatches && e < FMI_KMAX; ++e) {
    r->err_kind[e] = h.err_kind[e];
    r->err_base[e] = h.err_base[e];
    r->err_pos[e] = h.err_pos[e];
    r->num_mismatches++;
  }
}

bool list_push(results_list *rl, const result &r) {
  if (rl->num_results >= rl->max_results) return false;
  result *d = &rl->list[rl->num_results++];
  d->dir = r.dir;
  d->num_mismatches = 0;
  d->start = r.start;
  d->end = r.end;
  d->k = r.k;
  d->l = r.l;
  d->pos = r.pos;
  for (uint8_t e = 0; e < r.num_mismatches; ++e) {
    d->err_kind[e] = r.err_kind[e];
    d->err_base[e] = r.err_base[e];
    d->err_pos[e] = r.err_pos[e];
    d->num_mismatches++;
  }
  return true;
}

// Replays BWSearch1CPU's list operations (search/search.c:1876-1933) for one
// (read, strand) from its records in the reference's order (at least
// free + 1 of them stored, or all) and its two anchors: pass-1 records have
// dir 0 and come first, pass-2 records dir 1.  add_result returns true from
// BWSearch1CPU itself for the anchors and for the match after pass 1; a full
// list inside BWSimpleSearch1Backward returns to BWSearch1CPU with res
// unchanged and still non-empty, so the match's add_result returns true
// there too; a full list inside BWSimpleSearch1Forward is ignored.
bool replay_search1(const fmi_hit *recs, uint32_t stored, const fmi_hit *anc,
                    results_list *r_list, results_list *r_list_anchor) {
  result r;
  std::memset(&r, 0, sizeof r);
  if (r_list_anchor) {
    r_list_anchor->num_results = 0;
    to_result(anc[0], &r);
    if (!list_push(r_list_anchor, r)) return true;
  }
  uint32_t j = 0;
  for (; j < stored && recs[j].dir == 0; ++j) {
    to_result(recs[j], &r);
    if (!list_push(r_list, r)) return true;
  }
  if (r_list_anchor) {
    to_result(anc[1], &r);
    if (!list_push(r_list_anchor, r)) return true;
  }
  for (; j < stored; ++j) {
    to_result(recs[j], &r);
    if (!list_push(r_list, r)) return false;
  }
  return false;
}

uint64_t list_free(const results_list *rl) {
  return rl->num_results < rl->max_results ? rl->max_results - rl->num_results : 0;
}

std::mutex g_dirs_mu;
struct DirEntry {
  fmi_index *ix = nullptr;
  int refs = 0;
};
std::map<std::string, DirEntry> g_dirs;

}  // namespace

namespace fmi {
hipError_t launch_search1_anchors(const DevIndex &ix, uint32_t strand_mask, const uint8_t *reads,
                                  const uint64_t *offsets, uint32_t nreads, fmi_hit *anchors,
                                  hipStream_t s) {
  const uint64_t grid = (4ull * nreads + 255) / 256;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(search1_anchor_kernel, dim3((uint32_t)grid), dim3(256), 0, s, ix, strand_mask,
                     reads, offsets, nreads, anchors);
  return hipGetLastError();
}
}  // namespace fmi

extern "C" {

int hpg_search_status(void) { return g_status; }

int hpg_bwt_index_view(fmi_index *ix, int view, bwt_index *out) {
  if (!ix || !out || view < FMI_VIEW_BACKWARD || view > FMI_VIEW_FORWARD_REV) return status(FMI_EINVAL);
  out->fmi = ix;
  out->view = view;
  out->siz = ix->n + 1;
  out->owner = nullptr;
  return status(FMI_OK);
}

void load_bwt_index(bwt_index *index_rev, bwt_index *index, const char *directory, int direction,
                    bool inverse_sa, bwt_config_t config) {
  (void)inverse_sa;
  (void)config;
  if (index) std::memset(index, 0, sizeof *index);
  if (index_rev) std::memset(index_rev, 0, sizeof *index_rev);
  if (!index || !directory) { status(FMI_EINVAL); return; }
  std::lock_guard<std::mutex> lk(g_dirs_mu);
  DirEntry &e = g_dirs[directory];
  if (!e.ix) {
    fmi_index *ix = nullptr;
    int rc = fmi_index_import_reference(directory, &ix);
    if (rc == FMI_OK) rc = fmi_index_to_device(ix);
    if (rc != FMI_OK) {
      if (ix) fmi_index_free(ix);
      g_dirs.erase(directory);
      status(rc);
      return;
    }
    e.ix = ix;
  }
  e.refs++;
  const int v = direction ? FMI_VIEW_BACKWARD : FMI_VIEW_FORWARD;
  hpg_bwt_index_view(e.ix, v, index);
  index->owner = &e;
  if (index_rev) {
    hpg_bwt_index_view(e.ix, v + 2, index_rev);
    index_rev->owner = &e;
  }
  status(FMI_OK);
}

void free_bwt_index(bwt_index *index_rev, bwt_index *index, bool inverse_sa) {
  (void)inverse_sa;
  if (!index || !index->owner) return;
  std::lock_guard<std::mutex> lk(g_dirs_mu);
  for (auto it = g_dirs.begin(); it != g_dirs.end(); ++it) {
    if (&it->second != index->owner) continue;
    if (--it->second.refs == 0) {
      fmi_index_free(it->second.ix);
      g_dirs.erase(it);
    }
    break;
  }
  std::memset(index, 0, sizeof *index);
  if (index_rev) std::memset(index_rev, 0, sizeof *index_rev);
}

uint64_t hpg_size_sa(const bwt_index *index) { return index ? index->siz : 0; }

void hpg_bw_iteration(intmax_t k_in, intmax_t l_in, intmax_t *k_out, intmax_t *l_out, uint8_t b,
                      const bwt_index *index) {
  if (!view_ok(index) || !k_out || !l_out) { status(FMI_EINVAL); return; }
  const fmi_index *ix = index->fmi;
  const intmax_t n = (intmax_t)ix->n;
  if (k_in < 0 || k_in > n + 1 || l_in < -1 || l_in > n) {
    *k_out = 1;
    *l_out = 0;
    status(FMI_ERANGE);
    return;
  }
  Ctx *c = nullptr;
  int rc = ctx_for(ix, carve_size({8, 8, 1, 8, 8}), &c);
  if (rc) { status(rc); return; }
  Carve cv{static_cast<char *>(c->d)};
  int64_t *dk = cv.take<int64_t>(1), *dl = cv.take<int64_t>(1), *dko = cv.take<int64_t>(1),
          *dlo = cv.take<int64_t>(1);
  uint8_t *db = cv.take<uint8_t>(1);
  const int64_t in[2] = {(int64_t)k_in, (int64_t)l_in};
  int64_t outv[2] = {1, 0};
  hipError_t e = hipMemcpyAsync(dk, &in[0], 8, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dl, &in[1], 8, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(db, &b, 1, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(fmi::lf_batch_kernel, dim3(1), dim3(64), 0, c->s, ix->dev,
                       dsel_of(index->view), comp_of(index->view), dk, dl, db, 1u, dko, dlo);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&outv[0], dko, 8, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(&outv[1], dlo, 8, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  if (e != hipSuccess) { status(FMI_EDEVICE); return; }
  *k_out = outv[0];
  *l_out = outv[1];
  status(FMI_OK);
}

uintmax_t hpg_get_sa(uintmax_t m, const bwt_index *index) {
  if (!view_ok(index)) { status(FMI_EINVAL); return 0; }
  const fmi_index *ix = index->fmi;
  if (m > ix->n) { status(FMI_ERANGE); return 0; }
  Ctx *c = nullptr;
  int rc = ctx_for(ix, carve_size({8, 8, 8, 4}), &c);
  if (rc) { status(rc); return 0; }
  Carve cv{static_cast<char *>(c->d)};
  int64_t *dk = cv.take<int64_t>(1);
  uint64_t *dp = cv.take<uint64_t>(1);
  uint32_t *dn = cv.take<uint32_t>(1);
  const int64_t kk = (int64_t)m;
  uint64_t pos = 0;
  if (hipMemcpyAsync(dk, &kk, 8, hipMemcpyHostToDevice, c->s) != hipSuccess) { status(FMI_EDEVICE); return 0; }
  rc = fmi_locate_batch(ix, (int)dsel_of(index->view), dk, dk, 1, 1, dp, dn, c->s);
  if (rc) { status(rc); return 0; }
  if (hipMemcpyAsync(&pos, dp, 8, hipMemcpyDeviceToHost, c->s) != hipSuccess ||
      hipStreamSynchronize(c->s) != hipSuccess) {
    status(FMI_EDEVICE);
    return 0;
  }
  status(FMI_OK);
  return pos;
}

uintmax_t hpg_get_isa(uintmax_t m, const bwt_index *index) {
  if (!view_ok(index)) { status(FMI_EINVAL); return 0; }
  const fmi_index *ix = index->fmi;
  if (m > ix->n) { status(FMI_ERANGE); return 0; }
  Ctx *c = nullptr;
  int rc = ctx_for(ix, carve_size({8, 8}), &c);
  if (rc) { status(rc); return 0; }
  Carve cv{static_cast<char *>(c->d)};
  uint64_t *dq = cv.take<uint64_t>(1), *dr = cv.take<uint64_t>(1);
  const uint64_t q = m;
  uint64_t row = 0;
  if (hipMemcpyAsync(dq, &q, 8, hipMemcpyHostToDevice, c->s) != hipSuccess) { status(FMI_EDEVICE); return 0; }
  rc = fmi_inverse_locate_batch(ix, (int)dsel_of(index->view), dq, 1, dr, c->s);
  if (rc) { status(rc); return 0; }
  if (hipMemcpyAsync(&row, dr, 8, hipMemcpyDeviceToHost, c->s) != hipSuccess ||
      hipStreamSynchronize(c->s) != hipSuccess) {
    status(FMI_EDEVICE);
    return 0;
  }
  status(FMI_OK);
  return row;
}

uintmax_t binsearch(uintmax_t *array, uintmax_t size, uintmax_t key) {
  if (!array) return 0;
  return (uintmax_t)(std::upper_bound(array, array + size, key) - array);
}

int hpg_exact_array(const uint8_t *reads, const uint64_t *offsets, uint32_t nreads,
                    bwt_index *index, int direction, result *r, result *anchors) {
  if (!view_ok(index) || (direction != 0 && direction != 1)) return status(FMI_EINVAL);
  if (nreads == 0) return status(FMI_OK);
  if (!reads || !offsets || !r) return status(FMI_EINVAL);
  const fmi_index *ix = index->fmi;
  const intmax_t n = (intmax_t)ix->n;
  std::vector<fmi_hit> in(nreads);
  for (uint32_t t = 0; t < nreads; ++t) {
    const result &q = r[t];
    const uint64_t len = offsets[t + 1] - offsets[t];
    if (offsets[t + 1] < offsets[t] || len > FMI_MAX_READ_LEN) return status(FMI_ERANGE);
    if (q.k < 0 || q.k > n + 1 || q.l < -1 || q.l > n) return status(FMI_ERANGE);
    // the loop reads W[pos .. start] or W[pos .. end]: both inside the read
    const int32_t lo = direction ? q.pos : q.start, hi = direction ? q.end : q.pos;
    if (lo <= hi && (lo < 0 || hi >= (int32_t)len)) return status(FMI_ERANGE);
    std::memset(&in[t], 0, sizeof(fmi_hit));
    in[t].k = q.k;
    in[t].l = q.l;
    in[t].start = q.start;
    in[t].pos = q.pos;
    in[t].end = q.end;
    in[t].dir = q.dir ? 1 : 0;
  }
  const uint64_t bases = offsets[nreads] - offsets[0];
  Ctx *c = nullptr;
  int rc = ctx_for(ix, carve_size({bases, 8ull * (nreads + 1), 32ull * nreads, 32ull * nreads,
                                   32ull * nreads}),
                   &c);
  if (rc) return status(rc);
  Carve cv{static_cast<char *>(c->d)};
  uint8_t *dw = cv.take<uint8_t>(bases);
  uint64_t *doff = cv.take<uint64_t>(nreads + 1);
  fmi_hit *din = cv.take<fmi_hit>(nreads), *dout = cv.take<fmi_hit>(nreads),
          *danc = cv.take<fmi_hit>(nreads);
  std::vector<uint64_t> off(offsets, offsets + nreads + 1);
  for (auto &o : off) o -= offsets[0];
  std::vector<fmi_hit> out(nreads), anc(anchors ? nreads : 0);
  hipError_t e = hipMemcpyAsync(dw, reads + offsets[0], bases, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(doff, off.data(), 8ull * (nreads + 1), hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(din, in.data(), 32ull * nreads, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(fmi::gen_exact_kernel, dim3((nreads + 255) / 256), dim3(256), 0, c->s,
                       ix->dev, dsel_of(index->view), comp_of(index->view), (uint32_t)direction,
                       dw, doff, din, nreads, dout, anchors ? danc : nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out.data(), dout, 32ull * nreads, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess && anchors)
    e = hipMemcpyAsync(anc.data(), danc, 32ull * nreads, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  if (e != hipSuccess) return status(launch_rc(e));
  for (uint32_t t = 0; t < nreads; ++t) {
    // change_result(r, k2, l2, i): the other fields of r stay as they were
    r[t].k = out[t].k;
    r[t].l = out[t].l;
    r[t].pos = out[t].pos;
    if (anchors) {
      // init_result, bound_result, change_result: err_* are not touched
      anchors[t].dir = anc[t].dir != 0;
      anchors[t].num_mismatches = 0;
      anchors[t].start = anc[t].start;
      anchors[t].end = anc[t].end;
      anchors[t].k = anc[t].k;
      anchors[t].l = anc[t].l;
      anchors[t].pos = anc[t].pos;
    }
  }
  return status(FMI_OK);
}

static void exact_one(uint8_t *W, bwt_index *index, result *r, result *r_anchor, int direction) {
  if (!W || !r) { status(FMI_EINVAL); return; }
  // the read is W[0 .. the last position the loop can touch]
  const int32_t last = direction ? r->end : r->pos;
  const uint64_t len = last >= 0 ? (uint64_t)last + 1 : 0;
  const uint64_t off[2] = {0, len};
  hpg_exact_array(W, off, 1, index, direction, r, r_anchor);
}

void BWExactSearchBackward(uint8_t *W, bwt_index *index, result *r, result *r_anchor) {
  exact_one(W, index, r, r_anchor, 0);
}

void BWExactSearchForward(uint8_t *W, bwt_index *index, result *r, result *r_anchor) {
  exact_one(W, index, r, r_anchor, 1);
}

static int16_t vector_one(uint8_t *W, int16_t start, int16_t end, intmax_t k, intmax_t l,
                          intmax_t *vec_k, intmax_t *vec_l, bwt_index *index, int direction) {
  if (!view_ok(index)) { status(FMI_EINVAL); return 0; }
  if (k > l || start > end) { status(FMI_OK); return 0; }   // search.c:1411-1412
  const fmi_index *ix = index->fmi;
  const intmax_t n = (intmax_t)ix->n;
  if (!W || !vec_k || !vec_l || start < 0 || end >= (int16_t)FMI_MAX_READ_LEN || k < 0 ||
      l > n)
    { status(FMI_ERANGE); return 0; }
  const uint32_t len = (uint32_t)end + 1, m = (uint32_t)(end - start + 1);
  Ctx *c = nullptr;
  int rc = ctx_for(ix, carve_size({len, 16, 16, 4, 8, 8ull * m, 8ull * m, 2}), &c);
  if (rc) { status(rc); return 0; }
  Carve cv{static_cast<char *>(c->d)};
  uint8_t *dw = cv.take<uint8_t>(len);
  uint64_t *doff = cv.take<uint64_t>(2);
  int64_t *dkl = cv.take<int64_t>(2);
  int16_t *dse = cv.take<int16_t>(2);
  uint64_t *dvo = cv.take<uint64_t>(1);
  int64_t *dvk = cv.take<int64_t>(m), *dvl = cv.take<int64_t>(m);
  int16_t *dg = cv.take<int16_t>(1);
  const uint64_t off[2] = {0, len}, vo = 0;
  const int64_t kl[2] = {(int64_t)k, (int64_t)l};
  const int16_t se[2] = {start, end};
  int16_t good = 0;
  std::vector<int64_t> vk(m), vl(m);
  hipError_t e = hipMemcpyAsync(dw, W, len, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(doff, off, 16, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dkl, kl, 16, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dse, se, 4, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dvo, &vo, 8, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(fmi::gen_vector_kernel, dim3(1), dim3(64), 0, c->s, ix->dev,
                       dsel_of(index->view), comp_of(index->view), (uint32_t)direction, dw, doff,
                       dkl, dse, dvo, 1u, dvk, dvl, dg);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(vk.data(), dvk, 8ull * m, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(vl.data(), dvl, 8ull * m, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(&good, dg, 2, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  if (e != hipSuccess) { status(FMI_EDEVICE); return 0; }
  for (uint32_t j = 0; j < m; ++j) {
    vec_k[j] = vk[j];
    vec_l[j] = vl[j];
  }
  status(FMI_OK);
  return good;
}

int16_t BWExactSearchVectorBackward(uint8_t *W, int16_t start, int16_t end, intmax_t k,
                                    intmax_t l, intmax_t *vec_k, intmax_t *vec_l,
                                    bwt_index *index) {
  return vector_one(W, start, end, k, l, vec_k, vec_l, index, 0);
}

int16_t BWExactSearchVectorForward(uint8_t *W, int16_t start, int16_t end, intmax_t k,
                                   intmax_t l, intmax_t *vec_k, intmax_t *vec_l,
                                   bwt_index *index) {
  return vector_one(W, start, end, k, l, vec_k, vec_l, index, 1);
}

int hpg_search1_array(const uint8_t *reads, const uint64_t *offsets, uint32_t nreads,
                      bwt_index *backward, bwt_index *forward, results_list *r_lists,
                      results_list *r_list_anchors, bool *ret) {
  const int strand = pair_strand(backward, forward);
  if (strand < 0) return status(FMI_EINVAL);
  if (nreads == 0) return status(FMI_OK);
  if (!reads || !offsets || !r_lists || !ret) return status(FMI_EINVAL);
  const fmi_index *ix = backward->fmi;
  uint64_t cap = 1;
  for (uint32_t t = 0; t < nreads; ++t) {
    if (offsets[t + 1] < offsets[t] || offsets[t + 1] - offsets[t] > FMI_MAX_READ_LEN)
      return status(FMI_ERANGE);
    if (!r_lists[t].list && r_lists[t].max_results) return status(FMI_EINVAL);
    if (r_list_anchors && !r_list_anchors[t].list && r_list_anchors[t].max_results)
      return status(FMI_EINVAL);
    cap = std::max<uint64_t>(cap, list_free(&r_lists[t]) + 1);
  }
  // at most 2^24 stored records (512 MiB) per call: callers split larger arrays
  if (cap > 65536 || (uint64_t)nreads * 2 * cap > (1ull << 24)) return status(FMI_ERANGE);
  const uint64_t bases = offsets[nreads] - offsets[0];
  const uint64_t slots = 2ull * nreads;
  Ctx *c = nullptr;
  int rc = ctx_for(ix, carve_size({bases, 8ull * (nreads + 1), 32ull * slots * cap, 4ull * slots,
                                   32ull * 4 * nreads}),
                   &c);
  if (rc) return status(rc);
  Carve cv{static_cast<char *>(c->d)};
  uint8_t *dw = cv.take<uint8_t>(bases);
  uint64_t *doff = cv.take<uint64_t>(nreads + 1);
  fmi_hit *dh = cv.take<fmi_hit>(slots * cap);
  uint32_t *dc = cv.take<uint32_t>(slots);
  fmi_hit *da = cv.take<fmi_hit>(4ull * nreads);
  std::vector<uint64_t> off(offsets, offsets + nreads + 1);
  for (auto &o : off) o -= offsets[0];
  hipError_t e = hipMemcpyAsync(dw, reads + offsets[0], bases, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(doff, off.data(), 8ull * (nreads + 1), hipMemcpyHostToDevice, c->s);
  if (e != hipSuccess) return status(FMI_EDEVICE);
  const int mask = 1 << strand;
  rc = fmi_search1_batch(ix, FMI_MODE_EDIT, mask, dw, doff, nreads, (uint32_t)cap, dh, dc, c->s);
  if (rc) return status(rc);
  if (r_list_anchors) {
    e = fmi::launch_search1_anchors(ix->dev, (uint32_t)mask, dw, doff, nreads, da, c->s);
    if (e != hipSuccess) return status(FMI_EDEVICE);
  }
  std::vector<fmi_hit> hits(slots * cap), anc(r_list_anchors ? 4ull * nreads : 0);
  std::vector<uint32_t> cnt(slots);
  e = hipMemcpyAsync(hits.data(), dh, 32ull * slots * cap, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), dc, 4ull * slots, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess && r_list_anchors)
    e = hipMemcpyAsync(anc.data(), da, 32ull * 4 * nreads, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  if (e != hipSuccess) return status(FMI_EDEVICE);
  for (uint32_t t = 0; t < nreads; ++t) {
    const uint64_t slot = 2ull * t + (uint64_t)strand;
    const uint32_t stored = std::min<uint32_t>(cnt[slot], (uint32_t)cap);
    ret[t] = replay_search1(hits.data() + slot * cap, stored,
                            r_list_anchors ? anc.data() + 4ull * t + 2ull * strand : nullptr,
                            &r_lists[t], r_list_anchors ? &r_list_anchors[t] : nullptr);
  }
  return status(FMI_OK);
}

bool BWSearch1CPU(uint8_t *W, bwt_index *backward, bwt_index *forward, result *res,
                  results_list *r_list, results_list *r_list_anchor, uint8_t nA) {
  if (!W || !res || !r_list || nA != 4 || pair_strand(backward, forward) < 0) {
    status(FMI_EINVAL);
    return false;
  }
  // the mapping layer's res: the whole read from the whole index
  // (bwt.c:1686-1701); the search uses absolute positions from start 0
  if (res->k != 0 || res->l != (intmax_t)backward->siz - 1 || res->start != 0 || res->end < -1) {
    status(FMI_EINVAL);
    return false;
  }
  const uint64_t off[2] = {0, (uint64_t)(res->end + 1)};
  bool ret = false;
  hpg_search1_array(W, off, 1, backward, forward, r_list, r_list_anchor, &ret);
  return ret;
}

bool BWSearch1VectorHelper(uint8_t *W, int16_t start, int16_t end, intmax_t *vec_k,
                           intmax_t *vec_l, intmax_t *vec_ki, intmax_t *vec_li,
                           bwt_index *backward, bwt_index *forward, results_list *r_list,
                           uint8_t nA) {
  const int strand = pair_strand(backward, forward);
  if (!W || !r_list || nA != 4 || strand < 0 || !vec_k || !vec_l || !vec_ki || !vec_li) {
    status(FMI_EINVAL);
    return false;
  }
  if (start < 0 || end >= (int16_t)FMI_MAX_READ_LEN) { status(FMI_ERANGE); return false; }
  if (end - start < 2) { status(FMI_OK); return false; }    // no records (see the header)
  const fmi_index *ix = backward->fmi;
  const int32_t n = end - start, m = n + 1;
  for (int32_t j = 0; j < m; ++j) {
    for (intmax_t v : {vec_k[j], vec_l[j], vec_ki[j], vec_li[j]}) {
      if (v < 0 || v > (intmax_t)ix->n + 1) { status(FMI_ERANGE); return false; }
    }
  }
  const uint64_t cap = list_free(r_list) + 1;
  if (cap > (1u << 20)) { status(FMI_ERANGE); return false; }
  const uint32_t len = (uint32_t)end + 1;
  const uint64_t scratch = (uint64_t)256 * (fmi::kMaxReadLen + 3u) * sizeof(uint2);
  Ctx *c = nullptr;
  int rc = ctx_for(ix, carve_size({len, 16, 4, 8, 8ull * m, 8ull * m, 8ull * m, 8ull * m, 1,
                                   32ull * cap, 4, scratch}),
                   &c);
  if (rc) { status(rc); return false; }
  Carve cv{static_cast<char *>(c->d)};
  uint8_t *dw = cv.take<uint8_t>(len);
  uint64_t *doff = cv.take<uint64_t>(2);
  int16_t *dse = cv.take<int16_t>(2);
  uint64_t *dvo = cv.take<uint64_t>(1);
  int64_t *dvk = cv.take<int64_t>(m), *dvl = cv.take<int64_t>(m), *dvki = cv.take<int64_t>(m),
          *dvli = cv.take<int64_t>(m);
  uint8_t *dst = cv.take<uint8_t>(1);
  fmi_hit *dh = cv.take<fmi_hit>(cap);
  uint32_t *dc = cv.take<uint32_t>(1);
  uint2 *dscr = cv.take<uint2>(256ull * (fmi::kMaxReadLen + 3u));
  const uint64_t off[2] = {0, len}, vo = 0;
  const int16_t se[2] = {start, end};
  const uint8_t st = (uint8_t)strand;
  std::vector<int64_t> tmp(4ull * m);
  for (int32_t j = 0; j < m; ++j) {
    tmp[j] = vec_k[j];
    tmp[m + j] = vec_l[j];
    tmp[2 * m + j] = vec_ki[j];
    tmp[3 * m + j] = vec_li[j];
  }
  hipError_t e = hipMemcpyAsync(dw, W, len, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(doff, off, 16, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dse, se, 4, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dvo, &vo, 8, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dvk, tmp.data(), 8ull * m, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dvl, tmp.data() + m, 8ull * m, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dvki, tmp.data() + 2 * m, 8ull * m, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dvli, tmp.data() + 3 * m, 8ull * m, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dst, &st, 1, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) {
    fmi::VhGivenParams P{ix->dev, dw, doff, dse, dvo, dvk, dvl, dvki, dvli, dst, dh, dc, dscr,
                         1u, (uint32_t)cap};
    hipLaunchKernelGGL(fmi::vhelper_given_kernel, dim3(1), dim3(256), 0, c->s, P);
    e = hipGetLastError();
  }
  std::vector<fmi_hit> hits(cap);
  uint32_t count = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(hits.data(), dh, 32ull * cap, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(&count, dc, 4, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  if (e != hipSuccess) { status(FMI_EDEVICE); return false; }
  status(FMI_OK);
  result r;
  std::memset(&r, 0, sizeof r);
  const uint32_t stored = std::min<uint32_t>(count, (uint32_t)cap);
  for (uint32_t j = 0; j < stored; ++j) {
    to_result(hits[j], &r);
    if (!list_push(r_list, r)) return true;      // add_result inside the helper
  }
  return false;
}

bool BWSearchCPU(uint8_t *W, uint64_t nW, bwt_index *backward, bwt_index *forward,
                 results_list *rl_prev, results_list *rl_next, results_list *rl_prev_i,
                 results_list *rl_next_i, results_list *rl_final, results_list *rl_anchor,
                 int16_t fragsize, bool type, uint8_t nA) {
  const int strand = pair_strand(backward, forward);
  if (!W || nA != 4 || strand < 0 || (type ? 1 : 0) != (strand == 0 ? 1 : 0) || !rl_prev ||
      !rl_next || !rl_prev_i || !rl_next_i || !rl_final) {
    status(FMI_EINVAL);
    return false;
  }
  const uint64_t cap = rl_final->max_results;
  if (rl_final->num_results != 0 || rl_prev->max_results != cap || rl_next->max_results != cap ||
      rl_prev_i->max_results != cap || rl_next_i->max_results != cap || cap == 0 || cap > 65536 ||
      !rl_final->list || (rl_anchor && (rl_anchor->max_results == 0 ||
                                        rl_anchor->max_results > 4096 || !rl_anchor->list))) {
    status(FMI_EINVAL);
    return false;
  }
  if (nW > FMI_MAX_READ_LEN || fragsize <= 0 || (int64_t)nW / fragsize - 1 > FMI_KMAX) {
    status(FMI_ERANGE);
    return false;
  }
  const fmi_index *ix = backward->fmi;
  const uint32_t acap = rl_anchor ? (uint32_t)rl_anchor->max_results : 1u;
  Ctx *c = nullptr;
  int rc = ctx_for(ix, carve_size({nW, 16, 2, 64ull * 2 * cap, 8, 64ull * 2 * acap, 8, 2}), &c);
  if (rc) { status(rc); return false; }
  Carve cv{static_cast<char *>(c->d)};
  uint8_t *dw = cv.take<uint8_t>(nW);
  uint64_t *doff = cv.take<uint64_t>(2);
  int16_t *dfs = cv.take<int16_t>(1);
  fmi_hitk *dfin = cv.take<fmi_hitk>(2 * cap);
  uint32_t *dnf = cv.take<uint32_t>(2);
  fmi_hitk *danc = cv.take<fmi_hitk>(2ull * acap);
  uint32_t *dna = cv.take<uint32_t>(2);
  uint8_t *dfull = cv.take<uint8_t>(2);
  const uint64_t off[2] = {0, nW};
  hipError_t e = hipMemcpyAsync(dw, W, nW, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(doff, off, 16, hipMemcpyHostToDevice, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dfs, &fragsize, 2, hipMemcpyHostToDevice, c->s);
  if (e != hipSuccess) { status(FMI_EDEVICE); return false; }
  rc = fmi_search_k_batch(ix, 1 << strand, dw, doff, dfs, 1, (uint32_t)cap,
                          rl_anchor ? acap : 0u, dfin, dnf, rl_anchor ? danc : nullptr,
                          rl_anchor ? dna : nullptr, dfull, c->s);
  if (rc) { status(rc); return false; }
  std::vector<fmi_hitk> fin(cap), anc(acap);
  uint32_t nf = 0, na = 0;
  uint8_t full = 0;
  e = hipMemcpyAsync(fin.data(), dfin + (uint64_t)strand * cap, 64ull * cap, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(&nf, dnf + strand, 4, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess && rl_anchor)
    e = hipMemcpyAsync(anc.data(), danc + (uint64_t)strand * acap, 64ull * acap, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess && rl_anchor) e = hipMemcpyAsync(&na, dna + strand, 4, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipMemcpyAsync(&full, dfull + strand, 1, hipMemcpyDeviceToHost, c->s);
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  if (e != hipSuccess) { status(FMI_EDEVICE); return false; }
  if (full > 1) { status(FMI_ERANGE); return false; }
  status(FMI_OK);
  rl_prev->num_results = rl_next->num_results = rl_prev_i->num_results = rl_next_i->num_results = 0;
  result r;
  std::memset(&r, 0, sizeof r);
  for (uint32_t j = 0; j < std::min<uint64_t>(nf, cap); ++j) {
    to_result_k(fin[j], &r);
    list_push(rl_final, r);
  }
  if (rl_anchor) {
    rl_anchor->num_results = 0;
    for (uint32_t j = 0; j < std::min<uint32_t>(na, acap); ++j) {
      to_result_k(anc[j], &r);
      list_push(rl_anchor, r);
    }
  }
  return full == 1;
}

int fmi_search1_anchors_batch(const fmi_index *ix, int strand_mask, const uint8_t *d_reads,
                              const uint64_t *d_offsets, uint32_t nreads, fmi_hit *d_anchors,
                              fmi_stream s) {
  if (!ix) return FMI_EINVAL;
  if (!ix->on_device) return FMI_ESTATE;
  if (strand_mask < 1 || strand_mask > 3) return FMI_EINVAL;
  if (nreads == 0) return FMI_OK;
  if (!d_reads || !d_offsets || !d_anchors) return FMI_EINVAL;
  if (nreads > (1u << 28)) return FMI_ERANGE;
  return launch_rc(fmi::launch_search1_anchors(ix->dev, (uint32_t)strand_mask, d_reads, d_offsets,
                                               nreads, d_anchors, static_cast<hipStream_t>(s)));
}

}  // extern "C"

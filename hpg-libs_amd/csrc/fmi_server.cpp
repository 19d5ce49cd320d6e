// fmi_server.cpp -- continuous-batching request loop over a device-resident
// index (fmi_server_* in include/fmi.h).
//
// Host memory is pinned and allocated once: an input staging area, where
// fmi_server_submit copies a request's reads and offsets (their only host
// copy) on the caller's thread, and an output mirror of the device arena.
// One scheduler thread owns the device side; a launch:
//   1. takes pending requests first-come first-served while their reads
//      and results fit max_batch_reads and the device arena;
//   2. copies each request's reads and request-relative offsets from its
//      staging into the launch's device region (one copy each; a request
//      submitted in the fastq_batch_t layout is copied as it is and encoded
//      there by a kernel), then a rebase kernel turns the offsets into
//      launch offsets;
//   3. runs the persistent 1-error search and compacts the stored records of
//      every (read, strand) slot into a dense array (a scan of the counts);
//   4. copies the counts and the record total into the output mirror and
//      records an event; once it has passed, the requests' input staging is
//      released and exactly the dense records are copied back;
//   5. once those have landed, a copy worker hands each request its counts
//      and records and releases the launch's region.
// Up to max_inflight launches run at once on separate (default-priority)
// streams.  Submission is bounded: a request that finds max_pending requests
// queued, or no room in the input staging, is refused with FMI_EAGAIN (the
// caller retries; the reference's producer blocks at max_num_work_items,
// commons/workflow_scheduler.c:252-271).  A queued request can be cancelled;
// one already on the GPU is dropped when its launch completes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fmi.h"
#include "fmi_index.h"
#include "fmi_kernels.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint64_t kAlign = 256;
inline uint64_t align_up(uint64_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }
constexpr uint32_t kCopyWorkers = 2;
constexpr uint32_t kDefaultMaxPending = 4096;

// A launch's region of the arena (device memory and its pinned host
// mirror), held until every request of the launch has fetched its results
// from the mirror or been cancelled.
struct Region {
  uint64_t off = 0, size = 0;
  uint32_t refs = 0;
};

struct Request {
  uint64_t id = 0;
  uint32_t nreads = 0;
  uint64_t bases = 0;
  bool fastq = false;                  // staged as fastq_batch_t seq + data_indices
  uint64_t raw = 0;                    // fastq: seq bytes (bases + nreads NULs)
  uint64_t in_off = 0, in_bytes = 0;   // staging: reads, then nreads + 1 offsets
  bool staged = false;                 // holds input staging
  // results, in the launch's output mirror: 2 * nreads counts and nrec
  // dense records, valid while the request holds its region
  std::shared_ptr<Region> region;
  const uint32_t *counts = nullptr;
  const fmi_hit *records = nullptr;
  uint64_t nrec = 0;
  Clock::time_point submit, done;
  bool finished = false;
  bool cancelled = false;
  int status = FMI_OK;
};

// Device layout of one launch.
struct Layout {
  uint64_t o_reads, o_offs, o_table, o_raw, o_hits, o_counts, o_total, o_compact, o_scratch, size;
};

// raw: staging bytes of the launch's fastq requests (copied up as they are,
// then encoded on the device)
Layout layout_for(uint64_t bases, uint64_t nreads, uint64_t nreq, uint32_t cap, uint64_t raw) {
  const uint64_t slots = 2 * nreads;
  Layout L;
  L.o_reads = 0;
  L.o_offs = align_up(bases ? bases : 1);
  L.o_table = L.o_offs + align_up((nreads + 1) * 8);
  L.o_raw = L.o_table + align_up(2 * nreq * 8);
  L.o_hits = L.o_raw + align_up(raw);
  L.o_counts = L.o_hits + align_up(slots * (uint64_t)cap * sizeof(fmi_hit));
  L.o_total = L.o_counts + align_up(slots * 4);
  L.o_compact = L.o_total + kAlign;
  L.o_scratch = L.o_compact + align_up(slots * (uint64_t)cap * sizeof(fmi_hit));
  L.size = L.o_scratch + align_up(fmi::compact_scratch_bytes((uint32_t)slots));
  return L;
}

uint64_t input_bytes(uint64_t bases, uint64_t nreads) {
  return align_up(bases ? bases : 1) + align_up((nreads + 1) * 8);
}

// staging of a fastq request: seq bytes, then nreads + 1 int32 indices
uint64_t fastq_bytes(uint64_t raw, uint64_t nreads) {
  return align_up(raw ? raw : 1) + align_up((nreads + 1) * 4);
}

// First-fit allocator over [0, size) in kAlign units.
class Arena {
 public:
  void reset(uint64_t size) {
    free_.clear();
    free_[0] = size;
  }
  bool alloc(uint64_t bytes, uint64_t *off) {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const uint64_t rest = it->second - bytes, start = it->first + bytes;
        free_.erase(it);
        if (rest) free_[start] = rest;
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    bytes = align_up(bytes);
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<uint64_t, uint64_t> free_;
};

struct Launch {
  std::vector<std::shared_ptr<Request>> reqs;
  std::vector<uint32_t> first_read;   // each request's first read in the launch
  uint64_t off = 0;
  Layout lay{};
  uint32_t nreads = 0;
  int slot = -1;
  int phase = 1;                      // 1: counts on the way, 2: records on the way
  uint64_t total = 0;
};

// Launch offsets from request-relative ones: table holds, per request, its
// first read in the launch and the launch position of its first base.
// Entry i belongs to the last request whose first read is <= i (a request's
// closing entry was overwritten by the next request's opening 0).
__global__ void rebase_kernel(uint64_t *offs, uint32_t n1, const uint64_t *table, uint32_t nreq) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n1) return;
  uint32_t lo = 0, hi = nreq;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (table[2 * mid] <= i) lo = mid;
    else hi = mid;
  }
  offs[i] += table[2 * lo + 1];
}

}  // namespace

struct fmi_server {
  const fmi_index *ix = nullptr;
  fmi_server_config cfg{};
  int device = 0;
  char *d_arena = nullptr;
  uint8_t *h_mirror = nullptr;   // output mirror of the device arena (pinned)
  uint8_t *h_in = nullptr;       // input staging (pinned)
  uint64_t in_size = 0;
  Arena arena, in_arena;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<bool> slot_busy;

  std::mutex mu;
  std::condition_variable cv_submit, cv_done, cv_work;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
  std::vector<Launch> inflight;       // touched by the scheduler thread only
  std::deque<std::function<void()>> work;   // result hand-out tasks
  uint64_t next_id = 1;
  bool stop = false, workers_stop = false;
  fmi_server_stats stats{};
  std::string fault;
  std::thread thr;
  std::vector<std::thread> workers;

  void run();
  void worker();
  int start_launch(Launch &L);
  int fetch_records(Launch &L);
  void hand_out(Launch L, int status);
  void release_inputs(Launch &L);
};

int fmi_server::start_launch(Launch &L) {
  const Layout &lay = L.lay;
  hipStream_t s = streams[L.slot];
  char *d = d_arena + L.off;
  // request table of the rebase kernel, staged in the output mirror
  uint64_t *table = reinterpret_cast<uint64_t *>(h_mirror + L.off + lay.o_table);
  uint64_t base = 0, raw_off = 0;
  uint32_t r = 0;
  for (size_t q = 0; q < L.reqs.size(); ++q) {
    Request &rq = *L.reqs[q];
    L.first_read.push_back(r);
    table[2 * q] = r;
    table[2 * q + 1] = base;
    const uint8_t *src = h_in + rq.in_off;
    if (rq.fastq) {
      // seq and indices as staged, encoded into the launch's codes and
      // request-relative offsets on the device
      char *dst = d + lay.o_raw + raw_off;
      if (hipMemcpyAsync(dst, src, rq.in_bytes, hipMemcpyHostToDevice, s) != hipSuccess)
        return FMI_EDEVICE;
      if (fmi::launch_fastq_encode(dst, reinterpret_cast<const int32_t *>(dst + align_up(rq.raw ? rq.raw : 1)),
                                   rq.nreads, rq.raw, base, r,
                                   reinterpret_cast<uint8_t *>(d + lay.o_reads),
                                   reinterpret_cast<uint64_t *>(d + lay.o_offs), s) != hipSuccess)
        return FMI_EDEVICE;
      raw_off += rq.in_bytes;
    } else {
      if (rq.bases &&
          hipMemcpyAsync(d + lay.o_reads + base, src, rq.bases, hipMemcpyHostToDevice, s) != hipSuccess)
        return FMI_EDEVICE;
      // request-relative offsets; the next request's opening entry (0)
      // overwrites this request's closing one, the rebase kernel adds bases
      if (hipMemcpyAsync(d + lay.o_offs + 8ull * r, src + align_up(rq.bases ? rq.bases : 1),
                         8ull * (rq.nreads + 1), hipMemcpyHostToDevice, s) != hipSuccess)
        return FMI_EDEVICE;
    }
    base += rq.bases;
    r += rq.nreads;
  }
  const uint32_t nreq = (uint32_t)L.reqs.size();
  if (hipMemcpyAsync(d + lay.o_table, table, 16ull * nreq, hipMemcpyHostToDevice, s) != hipSuccess)
    return FMI_EDEVICE;
  hipLaunchKernelGGL(rebase_kernel, dim3((L.nreads + 1 + 255) / 256), dim3(256), 0, s,
                     reinterpret_cast<uint64_t *>(d + lay.o_offs), L.nreads + 1,
                     reinterpret_cast<const uint64_t *>(d + lay.o_table), nreq);
  if (hipGetLastError() != hipSuccess) return FMI_EDEVICE;
  const uint32_t slots = 2 * L.nreads;
  int rc = fmi_search1_batch(ix, cfg.mode, cfg.strand_mask, reinterpret_cast<const uint8_t *>(d),
                             reinterpret_cast<const uint64_t *>(d + lay.o_offs), L.nreads, cfg.cap,
                             reinterpret_cast<fmi_hit *>(d + lay.o_hits),
                             reinterpret_cast<uint32_t *>(d + lay.o_counts), s);
  if (rc != FMI_OK) return rc;
  if (fmi::launch_compact(reinterpret_cast<const uint32_t *>(d + lay.o_counts),
                          reinterpret_cast<const fmi_hit *>(d + lay.o_hits), slots, cfg.cap,
                          reinterpret_cast<fmi_hit *>(d + lay.o_compact),
                          reinterpret_cast<unsigned long long *>(d + lay.o_total),
                          d + lay.o_scratch, s) != hipSuccess)
    return FMI_EDEVICE;
  // counts and the record total are adjacent in the layout: one copy
  if (hipMemcpyAsync(h_mirror + L.off + lay.o_counts, d + lay.o_counts,
                     lay.o_compact - lay.o_counts, hipMemcpyDeviceToHost, s) != hipSuccess)
    return FMI_EDEVICE;
  if (hipEventRecord(events[L.slot], s) != hipSuccess) return FMI_EDEVICE;
  L.phase = 1;
  return FMI_OK;
}

void fmi_server::release_inputs(Launch &L) {
  std::lock_guard<std::mutex> lk(mu);
  for (auto &rp : L.reqs) {
    if (rp->staged) {
      in_arena.release(rp->in_off, rp->in_bytes);
      rp->staged = false;
    }
  }
}

int fmi_server::fetch_records(Launch &L) {
  L.total = *reinterpret_cast<const unsigned long long *>(h_mirror + L.off + L.lay.o_total);
  hipStream_t s = streams[L.slot];
  if (L.total &&
      hipMemcpyAsync(h_mirror + L.off + L.lay.o_compact, d_arena + L.off + L.lay.o_compact,
                     L.total * sizeof(fmi_hit), hipMemcpyDeviceToHost, s) != hipSuccess)
    return FMI_EDEVICE;
  if (hipEventRecord(events[L.slot], s) != hipSuccess) return FMI_EDEVICE;
  L.phase = 2;
  return FMI_OK;
}

// Points each request at its slice of the launch's output mirror (on a copy
// worker: the record counts are summed per request) and frees the stream
// slot; the launch's region is released by the last fetch or cancel of its
// requests (at once when none of them will fetch).
void fmi_server::hand_out(Launch L, int status) {
  auto region = std::make_shared<Region>();
  region->off = L.off;
  region->size = L.lay.size;
  std::vector<uint64_t> nrec(L.reqs.size(), 0);
  const uint32_t *cnt = reinterpret_cast<const uint32_t *>(h_mirror + L.off + L.lay.o_counts);
  const fmi_hit *rec = reinterpret_cast<const fmi_hit *>(h_mirror + L.off + L.lay.o_compact);
  std::vector<uint64_t> first_rec(L.reqs.size(), 0);
  if (status == FMI_OK) {
    uint64_t o = 0;
    for (size_t q = 0; q < L.reqs.size(); ++q) {
      const uint64_t s0 = 2ull * L.first_read[q], ns = 2ull * L.reqs[q]->nreads;
      uint64_t m = 0;
      for (uint64_t i = 0; i < ns; ++i) m += std::min<uint32_t>(cnt[s0 + i], cfg.cap);
      first_rec[q] = o;
      nrec[q] = m;
      o += m;
    }
  }
  const auto now = Clock::now();
  std::lock_guard<std::mutex> lk(mu);
  for (size_t q = 0; q < L.reqs.size(); ++q) {
    Request &r = *L.reqs[q];
    if (r.staged) {
      in_arena.release(r.in_off, r.in_bytes);
      r.staged = false;
    }
    r.done = now;
    r.finished = true;
    r.status = r.cancelled ? FMI_ECANCELED : status;
    if (status == FMI_OK && !r.cancelled) {
      stats.requests_done++;
      stats.reads_done += r.nreads;
      r.region = region;
      r.counts = cnt + 2ull * L.first_read[q];
      r.records = rec + first_rec[q];
      r.nrec = nrec[q];
      region->refs++;
    }
    if (r.cancelled) live.erase(r.id);
  }
  if (region->refs == 0) arena.release(region->off, region->size);
  slot_busy[L.slot] = false;
  cv_done.notify_all();
  cv_submit.notify_all();   // the scheduler may admit into the freed memory
}

// Drops request r's hold on its launch's region (mu held); the last one
// frees the region.
void drop_region(fmi_server *srv, Request &r) {
  if (!r.region) return;
  if (--r.region->refs == 0) {
    srv->arena.release(r.region->off, r.region->size);
    srv->cv_submit.notify_all();
  }
  r.region.reset();
  r.counts = nullptr;
  r.records = nullptr;
}

void fmi_server::worker() {
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    cv_work.wait(lk, [&] { return workers_stop || !work.empty(); });
    if (work.empty()) return;
    auto task = std::move(work.front());
    work.pop_front();
    lk.unlock();
    task();
    lk.lock();
  }
}

void fmi_server::run() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    bool progressed = false;
    // launches whose current copy has landed
    for (size_t i = 0; i < inflight.size();) {
      Launch &L = inflight[i];
      if (hipEventQuery(events[L.slot]) != hipSuccess) { ++i; continue; }
      progressed = true;
      lk.unlock();
      int rc = FMI_OK;
      bool done = true;
      if (L.phase == 1) {
        release_inputs(L);          // the reads are on the device
        rc = fetch_records(L);
        done = rc != FMI_OK;
      }
      lk.lock();
      if (done) {
        Launch fin = std::move(inflight[i]);
        inflight.erase(inflight.begin() + (long)i);
        work.emplace_back([this, fin, rc]() mutable { hand_out(std::move(fin), rc); });
        cv_work.notify_one();
      } else {
        ++i;
      }
    }
    // admission: FIFO, as many requests as fit one launch and the arena
    while (!pending.empty() && inflight.size() < cfg.max_inflight) {
      int slot = -1;
      for (uint32_t s = 0; s < cfg.max_inflight; ++s)
        if (!slot_busy[s]) { slot = (int)s; break; }
      if (slot < 0) break;
      Launch L;
      uint64_t bases = 0, nr = 0, raw = 0;
      for (auto &rp : pending) {
        if (nr + rp->nreads > cfg.max_batch_reads) break;
        L.reqs.push_back(rp);
        bases += rp->bases;
        nr += rp->nreads;
        if (rp->fastq) raw += rp->in_bytes;
      }
      while (!L.reqs.empty()) {
        L.lay = layout_for(bases, nr, L.reqs.size(), cfg.cap, raw);
        if (arena.alloc(L.lay.size, &L.off)) break;
        bases -= L.reqs.back()->bases;
        nr -= L.reqs.back()->nreads;
        if (L.reqs.back()->fastq) raw -= L.reqs.back()->in_bytes;
        L.reqs.pop_back();
      }
      if (L.reqs.empty()) break;   // wait for memory to free up
      for (size_t q = 0; q < L.reqs.size(); ++q) pending.pop_front();
      slot_busy[slot] = true;
      L.slot = slot;
      L.nreads = (uint32_t)nr;
      stats.launches++;
      if (nr > stats.max_batch_reads_seen) stats.max_batch_reads_seen = nr;
      lk.unlock();
      const int rc = start_launch(L);
      if (rc != FMI_OK) {
        (void)hipStreamSynchronize(streams[slot]);
        {
          std::lock_guard<std::mutex> g(mu);
          fault = fmi_last_error();
        }
        hand_out(std::move(L), rc);
        lk.lock();
      } else {
        lk.lock();
        inflight.push_back(std::move(L));
        if (inflight.size() > stats.max_inflight_seen) stats.max_inflight_seen = (uint32_t)inflight.size();
      }
      progressed = true;
    }
    stats.pending = (uint32_t)pending.size();
    if (stop && pending.empty() && inflight.empty()) break;
    if (!progressed) {
      if (inflight.empty()) cv_submit.wait_for(lk, std::chrono::milliseconds(2));
      else cv_submit.wait_for(lk, std::chrono::microseconds(20));
    }
  }
}

extern "C" {

int fmi_server_create(const fmi_index *ix, const fmi_server_config *cfg, fmi_server **out) {
  if (!ix || !cfg || !out) return FMI_EINVAL;
  *out = nullptr;
  if (!ix->on_device) return FMI_ESTATE;
  if (cfg->max_inflight < 1 || cfg->max_inflight > 8 || cfg->max_batch_reads < 1 ||
      cfg->max_batch_reads > (1u << 24) || cfg->cap > 4096 ||
      (uint64_t)cfg->max_batch_reads * 2 * cfg->cap >= (1ull << 32) ||
      (cfg->mode != FMI_MODE_EDIT && cfg->mode != FMI_MODE_MISMATCH) || cfg->strand_mask < 1 ||
      cfg->strand_mask > 3 || cfg->device_budget_bytes < (1u << 20) ||
      cfg->max_pending > (1u << 24) ||
      (cfg->staging_bytes && cfg->staging_bytes < (1u << 20)))
    return FMI_EINVAL;
  auto *srv = new (std::nothrow) fmi_server();
  if (!srv) return FMI_ENOMEM;
  srv->ix = ix;
  srv->cfg = *cfg;
  if (!srv->cfg.max_pending) srv->cfg.max_pending = kDefaultMaxPending;
  srv->in_size = cfg->staging_bytes ? cfg->staging_bytes : 2 * cfg->device_budget_bytes;
  srv->cfg.staging_bytes = srv->in_size;
  srv->device = ix->device;
  void *d = nullptr;
  if (hipSetDevice(ix->device) != hipSuccess || hipMalloc(&d, cfg->device_budget_bytes) != hipSuccess) {
    delete srv;
    return FMI_ENOMEM;
  }
  srv->d_arena = static_cast<char *>(d);
  void *h = nullptr, *hi = nullptr;
  if (hipHostMalloc(&h, cfg->device_budget_bytes, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&hi, srv->in_size, hipHostMallocDefault) != hipSuccess) {
    if (h) (void)hipHostFree(h);
    (void)hipFree(srv->d_arena);
    delete srv;
    return FMI_ENOMEM;
  }
  srv->h_mirror = static_cast<uint8_t *>(h);
  srv->h_in = static_cast<uint8_t *>(hi);
  srv->arena.reset(cfg->device_budget_bytes);
  srv->in_arena.reset(srv->in_size);
  for (uint32_t s = 0; s < cfg->max_inflight; ++s) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      fmi_server_destroy(srv);
      return FMI_EDEVICE;
    }
    srv->streams.push_back(st);
    srv->events.push_back(ev);
    srv->slot_busy.push_back(false);
  }
  for (uint32_t w = 0; w < kCopyWorkers; ++w) srv->workers.emplace_back([srv] { srv->worker(); });
  srv->thr = std::thread([srv] { srv->run(); });
  *out = srv;
  return FMI_OK;
}

int fmi_server_submit(fmi_server *srv, const uint8_t *reads, const uint64_t *offsets,
                      uint32_t nreads, uint64_t *request_id) {
  if (!srv || !offsets || !request_id || (nreads && !reads)) return FMI_EINVAL;
  if (nreads > srv->cfg.max_batch_reads) return FMI_ERANGE;
  for (uint32_t i = 0; i < nreads; ++i) {
    if (offsets[i + 1] < offsets[i] || offsets[i + 1] - offsets[i] > FMI_MAX_READ_LEN)
      return FMI_EINVAL;
  }
  const uint64_t bases = offsets[nreads] - offsets[0];
  const uint64_t in_bytes = input_bytes(bases, nreads);
  if (layout_for(bases, nreads, 1, srv->cfg.cap, 0).size > srv->cfg.device_budget_bytes ||
      in_bytes > srv->in_size)
    return FMI_ERANGE;
  auto r = std::make_shared<Request>();
  r->nreads = nreads;
  r->bases = bases;
  r->in_bytes = in_bytes;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return FMI_ESTATE;
    if (nreads && (srv->pending.size() >= srv->cfg.max_pending ||
                   !srv->in_arena.alloc(in_bytes, &r->in_off))) {
      srv->stats.refused++;
      return FMI_EAGAIN;     // backpressure: the queue or the staging is full
    }
    r->staged = nreads != 0;
    r->id = srv->next_id++;
  }
  if (nreads) {
    // the request's only host copy: reads and request-relative offsets
    // straight into its pinned staging, on the caller's thread
    uint8_t *dst = srv->h_in + r->in_off;
    std::memcpy(dst, reads + offsets[0], bases);
    uint64_t *o = reinterpret_cast<uint64_t *>(dst + align_up(bases ? bases : 1));
    const uint64_t o0 = offsets[0];
    for (uint32_t i = 0; i <= nreads; ++i) o[i] = offsets[i] - o0;
  }
  r->submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->live[r->id] = r;
    if (nreads == 0) {
      r->finished = true;
      r->done = r->submit;
    } else {
      srv->pending.push_back(r);
    }
    *request_id = r->id;
  }
  srv->cv_submit.notify_one();
  return FMI_OK;
}

int fmi_server_submit_fastq(fmi_server *srv, const char *seq, const int32_t *data_indices,
                            uint32_t num_reads, uint64_t *request_id) {
  if (!srv || !data_indices || !request_id || (num_reads && !seq)) return FMI_EINVAL;
  if (num_reads > srv->cfg.max_batch_reads) return FMI_ERANGE;
  for (uint32_t i = 0; i < num_reads; ++i) {
    const int64_t len = (int64_t)data_indices[i + 1] - data_indices[i] - 1;
    if (len < 0 || len > (int64_t)FMI_MAX_READ_LEN) return FMI_EINVAL;
  }
  const uint64_t raw = (uint64_t)((int64_t)data_indices[num_reads] - data_indices[0]);
  const uint64_t bases = raw - num_reads;
  const uint64_t in_bytes = fastq_bytes(raw, num_reads);
  if (layout_for(bases, num_reads, 1, srv->cfg.cap, in_bytes).size > srv->cfg.device_budget_bytes ||
      in_bytes > srv->in_size)
    return FMI_ERANGE;
  auto r = std::make_shared<Request>();
  r->nreads = num_reads;
  r->bases = bases;
  r->fastq = true;
  r->raw = raw;
  r->in_bytes = in_bytes;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return FMI_ESTATE;
    if (num_reads && (srv->pending.size() >= srv->cfg.max_pending ||
                      !srv->in_arena.alloc(in_bytes, &r->in_off))) {
      srv->stats.refused++;
      return FMI_EAGAIN;
    }
    r->staged = num_reads != 0;
    r->id = srv->next_id++;
  }
  if (num_reads) {
    // the request's only host copy: seq as it is and indices from 0
    uint8_t *dst = srv->h_in + r->in_off;
    std::memcpy(dst, seq + data_indices[0], raw);
    int32_t *o = reinterpret_cast<int32_t *>(dst + align_up(raw ? raw : 1));
    const int32_t d0 = data_indices[0];
    for (uint32_t i = 0; i <= num_reads; ++i) o[i] = data_indices[i] - d0;
  }
  r->submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->live[r->id] = r;
    if (num_reads == 0) {
      r->finished = true;
      r->done = r->submit;
    } else {
      srv->pending.push_back(r);
    }
    *request_id = r->id;
  }
  srv->cv_submit.notify_one();
  return FMI_OK;
}

int fmi_server_cancel(fmi_server *srv, uint64_t request_id) {
  if (!srv) return FMI_EINVAL;
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->live.find(request_id);
  if (it == srv->live.end()) return FMI_EINVAL;
  std::shared_ptr<Request> r = it->second;
  if (r->finished) {
    // its results were ready: drop them
    drop_region(srv, *r);
    srv->live.erase(it);
    return FMI_OK;
  }
  r->cancelled = true;
  srv->stats.cancelled++;
  auto p = std::find(srv->pending.begin(), srv->pending.end(), r);
  if (p != srv->pending.end()) {
    // still queued: never launched
    srv->pending.erase(p);
    if (r->staged) {
      srv->in_arena.release(r->in_off, r->in_bytes);
      r->staged = false;
    }
    r->finished = true;
    r->status = FMI_ECANCELED;
    r->done = Clock::now();
    srv->live.erase(request_id);
    srv->cv_done.notify_all();
  }
  // else on the GPU: its launch completes and the results are dropped
  return FMI_OK;
}

int fmi_server_wait(fmi_server *srv, uint64_t request_id, int64_t timeout_us, double *latency_ms,
                    uint64_t *nrecords) {
  if (!srv) return FMI_EINVAL;
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->live.find(request_id);
  if (it == srv->live.end()) return FMI_EINVAL;
  std::shared_ptr<Request> r = it->second;
  auto ready = [&] { return r->finished; };
  if (timeout_us < 0) {
    srv->cv_done.wait(lk, ready);
  } else if (!srv->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), ready)) {
    return FMI_EAGAIN;
  }
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r->done - r->submit).count();
  if (nrecords) *nrecords = r->nrec;
  return r->status;
}

int fmi_server_fetch(fmi_server *srv, uint64_t request_id, uint32_t *counts, fmi_hit *records) {
  if (!srv) return FMI_EINVAL;
  std::shared_ptr<Request> r;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->live.find(request_id);
    if (it == srv->live.end()) return FMI_EINVAL;
    r = it->second;
    if (!r->finished) return FMI_EAGAIN;
    srv->live.erase(it);
  }
  // the request still holds its region: the mirror is not reused meanwhile
  if (r->status == FMI_OK) {
    if (counts && r->counts) std::memcpy(counts, r->counts, 2ull * r->nreads * 4);
    if (records && r->nrec) std::memcpy(records, r->records, r->nrec * sizeof(fmi_hit));
  }
  std::lock_guard<std::mutex> lk(srv->mu);
  drop_region(srv, *r);
  return r->status;
}

int fmi_server_get_stats(fmi_server *srv, fmi_server_stats *stats) {
  if (!srv || !stats) return FMI_EINVAL;
  std::lock_guard<std::mutex> lk(srv->mu);
  *stats = srv->stats;
  stats->pending = (uint32_t)srv->pending.size();
  return FMI_OK;
}

int fmi_server_destroy(fmi_server *srv) {
  if (!srv) return FMI_EINVAL;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_submit.notify_all();
  if (srv->thr.joinable()) srv->thr.join();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->workers_stop = true;
  }
  srv->cv_work.notify_all();
  for (auto &w : srv->workers)
    if (w.joinable()) w.join();
  for (auto s : srv->streams) {
    (void)fmi::release_stream_state(s);   // the search's counters on that stream
    (void)hipStreamDestroy(s);
  }
  for (auto e : srv->events) (void)hipEventDestroy(e);
  if (srv->h_mirror) (void)hipHostFree(srv->h_mirror);
  if (srv->h_in) (void)hipHostFree(srv->h_in);
  if (srv->d_arena) (void)hipFree(srv->d_arena);
  delete srv;
  return FMI_OK;
}

}  // extern "C"

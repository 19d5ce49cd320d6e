// fmi_server.cpp -- continuous-batching request loop over a device-resident
// index (fmi_server_* in include/fmi.h).
//
// One scheduler thread owns the device side: a device arena of
// device_budget_bytes and a pinned host staging area of the same size share
// one first-fit allocator, so a launch's region has the same offsets in
// both.  A launch:
//   1. packs the admitted requests' reads and rebased offsets into staging
//      and copies them up;
//   2. runs the persistent 1-error search, then compacts the stored records
//      of every (read, strand) slot into a dense array (scan of the counts);
//   3. copies the counts and the record total back and records an event;
//   4. once that event has completed, copies exactly the dense records back
//      and records a second event;
//   5. once that one has completed, hands each request its slice.
// Up to max_inflight launches run at once on separate (default-priority)
// streams, so requests that arrive while others are on the GPU are admitted
// as soon as memory and a stream free up.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
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

struct Request {
  uint64_t id = 0;
  std::vector<uint8_t> reads;
  std::vector<uint64_t> offsets;   // rebased to 0, nreads + 1 entries
  uint32_t nreads = 0;
  std::vector<uint32_t> counts;    // results (2 * nreads)
  std::vector<fmi_hit> records;    // results (dense)
  Clock::time_point submit, done;
  bool finished = false;
  int status = FMI_OK;
};

struct Layout {
  uint64_t o_reads, o_offs, o_hits, o_counts, o_total, o_compact, o_scratch, size;
};

Layout layout_for(uint64_t bases, uint64_t nreads, uint32_t cap) {
  const uint64_t slots = 2 * nreads;
  Layout L;
  L.o_reads = 0;
  L.o_offs = align_up(bases ? bases : 1);
  L.o_hits = L.o_offs + align_up((nreads + 1) * 8);
  L.o_counts = L.o_hits + align_up(slots * (uint64_t)cap * sizeof(fmi_hit));
  L.o_total = L.o_counts + align_up(slots * 4);
  L.o_compact = L.o_total + kAlign;
  L.o_scratch = L.o_compact + align_up(slots * (uint64_t)cap * sizeof(fmi_hit));
  L.size = L.o_scratch + align_up(fmi::compact_scratch_bytes((uint32_t)slots));
  return L;
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

}  // namespace

struct fmi_server {
  const fmi_index *ix = nullptr;
  fmi_server_config cfg{};
  int device = 0;
  char *d_arena = nullptr;
  uint8_t *h_stage = nullptr;
  Arena arena;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<bool> slot_busy;

  std::mutex mu;
  std::condition_variable cv_submit, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
  std::vector<Launch> inflight;       // touched by the scheduler thread only
  uint64_t next_id = 1;
  bool stop = false;
  fmi_server_stats stats{};
  std::string fault;
  std::thread thr;

  void run();
  int start_launch(Launch &L);
  int fetch_records(Launch &L);
  void finish_launch(Launch &L, int status);
};

int fmi_server::start_launch(Launch &L) {
  uint8_t *st = h_stage + L.off;
  const Layout &lay = L.lay;
  uint64_t *offs = reinterpret_cast<uint64_t *>(st + lay.o_offs);
  uint64_t base = 0;
  uint32_t r = 0;
  offs[0] = 0;
  for (auto &rp : L.reqs) {
    L.first_read.push_back(r);
    std::memcpy(st + base, rp->reads.data(), rp->reads.size());
    for (uint32_t i = 0; i < rp->nreads; ++i) offs[r + i + 1] = base + rp->offsets[i + 1];
    base += rp->reads.size();
    r += rp->nreads;
  }
  hipStream_t s = streams[L.slot];
  char *d = d_arena + L.off;
  const uint32_t slots = 2 * L.nreads;
  if (hipMemcpyAsync(d, st, lay.o_hits, hipMemcpyHostToDevice, s) != hipSuccess) return FMI_EDEVICE;
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
  if (hipMemcpyAsync(st + lay.o_counts, d + lay.o_counts, lay.o_compact - lay.o_counts,
                     hipMemcpyDeviceToHost, s) != hipSuccess)
    return FMI_EDEVICE;
  if (hipEventRecord(events[L.slot], s) != hipSuccess) return FMI_EDEVICE;
  L.phase = 1;
  return FMI_OK;
}

int fmi_server::fetch_records(Launch &L) {
  L.total = *reinterpret_cast<const unsigned long long *>(h_stage + L.off + L.lay.o_total);
  hipStream_t s = streams[L.slot];
  if (L.total &&
      hipMemcpyAsync(h_stage + L.off + L.lay.o_compact, d_arena + L.off + L.lay.o_compact,
                     L.total * sizeof(fmi_hit), hipMemcpyDeviceToHost, s) != hipSuccess)
    return FMI_EDEVICE;
  if (hipEventRecord(events[L.slot], s) != hipSuccess) return FMI_EDEVICE;
  L.phase = 2;
  return FMI_OK;
}

void fmi_server::finish_launch(Launch &L, int status) {
  // Hand out results outside the lock: only this thread writes them.
  if (status == FMI_OK) {
    const uint32_t *cnt = reinterpret_cast<const uint32_t *>(h_stage + L.off + L.lay.o_counts);
    const fmi_hit *rec = reinterpret_cast<const fmi_hit *>(h_stage + L.off + L.lay.o_compact);
    uint64_t o = 0;
    for (size_t q = 0; q < L.reqs.size(); ++q) {
      Request &r = *L.reqs[q];
      const uint64_t s0 = 2ull * L.first_read[q], ns = 2ull * r.nreads;
      r.counts.assign(cnt + s0, cnt + s0 + ns);
      uint64_t m = 0;
      for (uint64_t i = 0; i < ns; ++i) m += std::min<uint32_t>(cnt[s0 + i], cfg.cap);
      r.records.assign(rec + o, rec + o + m);
      o += m;
    }
  }
  const auto now = Clock::now();
  std::lock_guard<std::mutex> lk(mu);
  for (auto &rp : L.reqs) {
    rp->done = now;
    rp->finished = true;
    rp->status = status;
    if (status == FMI_OK) {
      stats.requests_done++;
      stats.reads_done += rp->nreads;
    }
  }
  arena.release(L.off, L.lay.size);
  slot_busy[L.slot] = false;
  cv_done.notify_all();
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
        rc = fetch_records(L);
        done = rc != FMI_OK;
      }
      if (done) finish_launch(L, rc);
      lk.lock();
      if (done) inflight.erase(inflight.begin() + (long)i);
      else ++i;
    }
    // admission: FIFO, as many requests as fit one launch and the arena
    while (!pending.empty() && inflight.size() < cfg.max_inflight) {
      int slot = -1;
      for (uint32_t s = 0; s < cfg.max_inflight; ++s)
        if (!slot_busy[s]) { slot = (int)s; break; }
      if (slot < 0) break;
      Launch L;
      uint64_t bases = 0, nr = 0;
      for (auto &rp : pending) {
        if (nr + rp->nreads > cfg.max_batch_reads) break;
        L.reqs.push_back(rp);
        bases += rp->reads.size();
        nr += rp->nreads;
      }
      while (!L.reqs.empty()) {
        L.lay = layout_for(bases, nr, cfg.cap);
        if (arena.alloc(L.lay.size, &L.off)) break;
        bases -= L.reqs.back()->reads.size();
        nr -= L.reqs.back()->nreads;
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
        finish_launch(L, rc);
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
      cfg->strand_mask > 3 || cfg->device_budget_bytes < (1u << 20))
    return FMI_EINVAL;
  auto *srv = new (std::nothrow) fmi_server();
  if (!srv) return FMI_ENOMEM;
  srv->ix = ix;
  srv->cfg = *cfg;
  srv->device = ix->device;
  void *d = nullptr;
  if (hipSetDevice(ix->device) != hipSuccess || hipMalloc(&d, cfg->device_budget_bytes) != hipSuccess) {
    delete srv;
    return FMI_ENOMEM;
  }
  srv->d_arena = static_cast<char *>(d);
  void *h = nullptr;
  if (hipHostMalloc(&h, cfg->device_budget_bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(srv->d_arena);
    delete srv;
    return FMI_ENOMEM;
  }
  srv->h_stage = static_cast<uint8_t *>(h);
  srv->arena.reset(cfg->device_budget_bytes);
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
  if (layout_for(bases, nreads, srv->cfg.cap).size > srv->cfg.device_budget_bytes) return FMI_ERANGE;
  auto r = std::make_shared<Request>();
  r->reads.assign(reads + offsets[0], reads + offsets[nreads]);
  r->offsets.resize(nreads + 1);
  for (uint32_t i = 0; i <= nreads; ++i) r->offsets[i] = offsets[i] - offsets[0];
  r->nreads = nreads;
  r->submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return FMI_ESTATE;
    r->id = srv->next_id++;
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
  if (nrecords) *nrecords = r->records.size();
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
  if (r->status != FMI_OK) return r->status;
  if (counts && !r->counts.empty()) std::memcpy(counts, r->counts.data(), r->counts.size() * 4);
  if (records && !r->records.empty())
    std::memcpy(records, r->records.data(), r->records.size() * sizeof(fmi_hit));
  return FMI_OK;
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
  for (auto s : srv->streams) {
    (void)fmi::release_stream_state(s);   // the search's counters on that stream
    (void)hipStreamDestroy(s);
  }
  for (auto e : srv->events) (void)hipEventDestroy(e);
  if (srv->h_stage) (void)hipHostFree(srv->h_stage);
  if (srv->d_arena) (void)hipFree(srv->d_arena);
  delete srv;
  return FMI_OK;
}

}  // extern "C"

// fmi_api.cpp -- the extern "C" boundary of libhpgfm.so (declared in
// include/fmi.h).  Validates arguments on the host, owns the device copy of
// the index and per-stream work counters, and times every launch with HIP
// events on the stream it runs on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/fmi.h"
#include "fmi_index.h"
#include "fmi_kernels.h"

namespace {

thread_local std::string g_err;
thread_local hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
thread_local bool g_timed = false;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return FMI_EDEVICE;
}

#define HIP_TRY(expr, what)                         \
  do {                                              \
    hipError_t e_ = (expr);                         \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

// Defaults of the persistent launch (the autotuner overrides them).
constexpr uint32_t kDefaultWgPerCu = 8;
constexpr uint32_t kDefaultChunk = 64;
constexpr uint32_t kDefaultRefill = 16;

// An index whose device arrays exceed this spreads the search's accesses
// over enough memory that issuing only the loads a step needs pays
// (FMI_LOADS_AUTO).
constexpr uint64_t kOwnLoadsBytes = 16ull << 30;

struct Tuning {
  uint32_t wg_per_cu = 0;   // workgroups of 256 per CU; 0 = default
  uint32_t chunk = kDefaultChunk;            // tasks a wave takes per atomic
  uint32_t refill_min = kDefaultRefill;      // idle lanes before a wave refills
  uint32_t loads = FMI_LOADS_AUTO;           // fmi_set_load_mode
  bool staging = true;                       // fmi_set_staging
};
Tuning g_tuning;
std::mutex g_mu;

// Per-(device, stream) state: the persistent launches' work counters and
// the scratch of the vector helper and the k-error search, so that launches
// on different streams never share a queue or a scratch slice.  A graph
// captured on a stream holds these addresses: while one is alive
// (fmi_graph_end .. fmi_graph_destroy) the scratch may not move, and none of
// it may be created while the stream is capturing.
struct StreamState {
  uint32_t *queue = nullptr;              // 256 B: queue counter, then stats
  unsigned long long *stats = nullptr;
  void *scratch = nullptr;
  uint64_t scratch_bytes = 0;
  uint32_t graphs = 0;                    // live graphs captured on the stream
};
std::unordered_map<std::string, StreamState> g_streams;
// graph -> the keys of the streams whose counters and scratch it holds (the
// capturing stream and the streams that joined its capture)
std::unordered_map<void *, std::vector<std::string>> g_graph_stream;
// capture id -> the streams that joined it through fmi_stream_wait
std::unordered_map<unsigned long long, std::vector<std::string>> g_capture_members;

// The device memory of fmi_search1_batch_noalloc: the persistent launch's
// work-queue counter (bytes 0..3) and its work counters (64..103), the same
// layout as a stream's workspace.  One copy per GPU, held until the process
// ends; in host code `scratch` is only a handle for hipGetSymbolAddress, the
// kernels get the device address kept in g_scratch_addr.
constexpr size_t kNoAllocScratchBytes = 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) on the first no-alloc call there.
int noalloc_scratch(void **out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(FMI_EDEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return FMI_OK;
}

int stream_key(hipStream_t s, std::string *key) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  char buf[64];
  snprintf(buf, sizeof buf, "%d:%p", dev, (void *)s);
  *key = buf;
  return FMI_OK;
}

bool capturing(hipStream_t s);

struct Workspace {
  uint32_t *queue = nullptr;
  unsigned long long *stats = nullptr;
};

int workspace_for(hipStream_t s, Workspace *out) {
  std::string key;
  int rc = stream_key(s, &key);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  StreamState &st = g_streams[key];
  if (!st.queue) {
    if (capturing(s))
      return fail(FMI_ESTATE, "run the call once on this stream before capturing it");
    void *p = nullptr;
    HIP_TRY(hipMalloc(&p, 256), "hipMalloc(workspace)");
    st.queue = static_cast<uint32_t *>(p);
    st.stats = reinterpret_cast<unsigned long long *>(static_cast<char *>(p) + 64);
  }
  out->queue = st.queue;
  out->stats = st.stats;
  return FMI_OK;
}

// Scratch that grows to the largest request made on the stream; an older,
// smaller one is freed after the stream drains.  Growing is refused while
// the stream captures or a graph captured on it is alive.
int scratch_for(hipStream_t s, uint64_t bytes, void **out) {
  std::string key;
  int rc = stream_key(s, &key);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  StreamState &st = g_streams[key];
  if (st.scratch_bytes < bytes) {
    if (capturing(s) || st.graphs)
      return fail(FMI_ESTATE, "the call needs " + std::to_string(bytes) +
                                  " bytes of stream scratch, more than the " +
                                  std::to_string(st.scratch_bytes) +
                                  " a graph captured on this stream holds");
    if (st.scratch) {
      HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
      HIP_TRY(hipFree(st.scratch), "hipFree(scratch)");
      st.scratch = nullptr;
      st.scratch_bytes = 0;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(FMI_ENOMEM, "device scratch of " + std::to_string(bytes) + " bytes");
    st.scratch = p;
    st.scratch_bytes = bytes;
  }
  *out = st.scratch;
  return FMI_OK;
}

int cu_count(int *cus) {
  static int cached[64] = {0};
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  if (dev >= 0 && dev < 64 && cached[dev]) { *cus = cached[dev]; return FMI_OK; }
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  *cus = prop.multiProcessorCount;
  if (dev >= 0 && dev < 64) cached[dev] = *cus;
  return FMI_OK;
}

// Timing region (fmi_timing_begin/end): one event pair per launch, taken
// from a per-thread pool that grows as needed and is reused across regions.
thread_local bool g_region = false;
thread_local std::vector<std::pair<hipEvent_t, hipEvent_t>> g_pool;
thread_local size_t g_used = 0;

thread_local hipEvent_t g_single0 = nullptr, g_single1 = nullptr;

// A stream being captured into a graph (fmi_graph_begin) records no timing
// events: the launches are not run then.
bool capturing(hipStream_t s) {
  if (!s) return false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

int begin_timed(hipStream_t s) {
  if (capturing(s)) return FMI_OK;
  if (g_region) {
    if (g_used == g_pool.size()) {
      hipEvent_t a, b;
      HIP_TRY(hipEventCreate(&a), "hipEventCreate");
      HIP_TRY(hipEventCreate(&b), "hipEventCreate");
      g_pool.emplace_back(a, b);
    }
    g_ev0 = g_pool[g_used].first;
    g_ev1 = g_pool[g_used].second;
    g_used++;
  } else {
    if (!g_single0) {
      HIP_TRY(hipEventCreate(&g_single0), "hipEventCreate");
      HIP_TRY(hipEventCreate(&g_single1), "hipEventCreate");
    }
    g_ev0 = g_single0;
    g_ev1 = g_single1;
  }
  HIP_TRY(hipEventRecord(g_ev0, s), "hipEventRecord");
  return FMI_OK;
}

int end_timed(hipStream_t s, hipError_t launch_err, const char *what) {
  if (launch_err != hipSuccess) return hip_fail(launch_err, what);
  if (capturing(s)) return FMI_OK;
  HIP_TRY(hipEventRecord(g_ev1, s), "hipEventRecord");
  g_timed = true;
  return FMI_OK;
}

int need_device(const fmi_index *ix) {
  if (!ix) return fail(FMI_EINVAL, "index is NULL");
  if (!ix->on_device) return fail(FMI_ESTATE, "index is not on the device (call fmi_index_to_device)");
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  if (dev != ix->device) return fail(FMI_ESTATE, "index lives on another device");
  return FMI_OK;
}

int view_params(int view, int direction, uint32_t *dsel, uint32_t *comp) {
  if (view < 0 || view > 3) return fail(FMI_EINVAL, "view must be 0..3");
  if (direction != 0 && direction != 1) return fail(FMI_EINVAL, "direction must be 0 or 1");
  *dsel = (view == FMI_VIEW_FORWARD || view == FMI_VIEW_FORWARD_REV) ? 1u : 0u;
  *comp = view >= FMI_VIEW_BACKWARD_REV ? 1u : 0u;
  return FMI_OK;
}

int search1_common(const fmi_index *ix, int mode, int strand_mask, const uint8_t *d_reads,
                   const uint64_t *d_offsets, uint32_t nreads, uint32_t cap, fmi_hit *d_hits,
                   uint32_t *d_counts, fmi_stream s, bool persistent) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (mode != FMI_MODE_EDIT && mode != FMI_MODE_MISMATCH) return fail(FMI_EINVAL, "mode must be EDIT or MISMATCH");
  if (strand_mask < 1 || strand_mask > 3) return fail(FMI_EINVAL, "strand_mask must be 1, 2 or 3");
  if (nreads == 0) return FMI_OK;
  if (!d_reads || !d_offsets || !d_counts) return fail(FMI_EINVAL, "NULL buffer");
  if (cap > 0 && !d_hits) return fail(FMI_EINVAL, "d_hits is NULL with cap > 0");
  if (nreads > (1u << 30)) return fail(FMI_ERANGE, "more than 2^30 reads in one batch");
  hipStream_t st = static_cast<hipStream_t>(s);
  Workspace w;
  rc = workspace_for(st, &w);
  if (rc) return rc;
  uint32_t grid = 0;
  Tuning t;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    t = g_tuning;
  }
  if (persistent) {
    int cus = 0;
    rc = cu_count(&cus);
    if (rc) return rc;
    grid = (uint32_t)cus * (t.wg_per_cu ? t.wg_per_cu : kDefaultWgPerCu);
    HIP_TRY(fmi::launch_zero(w.queue, 1, st), "zero queue");
  }
  void *staged = nullptr;
  if (persistent && t.staging && nreads <= fmi::kStagedMaxTasks) {
    rc = scratch_for(st, fmi::search1_staged_scratch_bytes(nreads), &staged);
    if (rc) return rc;
  }
  rc = begin_timed(st);
  if (rc) return rc;
  const bool own = t.loads == FMI_LOADS_OWN ||
                   (t.loads == FMI_LOADS_AUTO && fmi_index_device_bytes(ix) > kOwnLoadsBytes);
  hipError_t e = fmi::launch_search1(ix->dev, mode == FMI_MODE_MISMATCH, persistent, false,
                                     (uint32_t)strand_mask, d_reads, d_offsets, nreads, cap,
                                     d_hits, d_counts, w.queue, w.stats, grid, t.chunk,
                                     t.refill_min, own, st, staged);
  return end_timed(st, e, "search1 launch");
}

}  // namespace

namespace fmi {

int release_stream_state(hipStream_t s) {
  std::string key;
  int rc = stream_key(s, &key);
  if (rc) return rc;
  StreamState st;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_streams.find(key);
    if (it == g_streams.end()) return FMI_OK;
    if (it->second.graphs)
      return fail(FMI_ESTATE, "destroy the graphs captured on this stream first");
    st = it->second;
    g_streams.erase(it);
  }
  HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (st.queue) HIP_TRY(hipFree(st.queue), "hipFree(workspace)");
  if (st.scratch) HIP_TRY(hipFree(st.scratch), "hipFree(scratch)");
  return FMI_OK;
}

}  // namespace fmi

extern "C" {

const char *fmi_status_str(int status) {
  switch (status) {
    case FMI_OK: return "ok";
    case FMI_EINVAL: return "invalid argument";
    case FMI_ENOMEM: return "out of memory";
    case FMI_EIO: return "i/o error";
    case FMI_EFORMAT: return "bad index format";
    case FMI_EDEVICE: return "device error";
    case FMI_ERANGE: return "size out of range";
    case FMI_ESTATE: return "call out of order";
    case FMI_EAGAIN: return "not now (time-out or full queue)";
    case FMI_ECANCELED: return "cancelled";
    default: return "unknown status";
  }
}

const char *fmi_last_error(void) { return g_err.c_str(); }

int fmi_index_build(const uint8_t *text, uint64_t n, uint32_t sa_ratio, fmi_index **out) {
  if (!out) return fail(FMI_EINVAL, "out is NULL");
  *out = nullptr;
  fmi_index *ix = new (std::nothrow) fmi_index();
  if (!ix) return fail(FMI_ENOMEM, "out of host memory");
  std::string err;
  int rc = fmi::build_index(text, n, sa_ratio, ix, &err);
  if (rc) { delete ix; return fail(rc, err); }
  *out = ix;
  return FMI_OK;
}

int fmi_index_build_gpu(const uint8_t *text, uint64_t n, uint32_t sa_ratio, fmi_index **out) {
  if (!out) return fail(FMI_EINVAL, "out is NULL");
  *out = nullptr;
  fmi_index *ix = new (std::nothrow) fmi_index();
  if (!ix) return fail(FMI_ENOMEM, "out of host memory");
  std::string err;
  int rc = fmi::build_index_gpu(text, n, sa_ratio, ix, &err);
  if (rc) { delete ix; return fail(rc, err); }
  *out = ix;
  return FMI_OK;
}

int fmi_index_save(const fmi_index *ix, const char *path) {
  if (!ix || !path) return fail(FMI_EINVAL, "NULL argument");
  std::string err;
  int rc = fmi::save_index(ix, path, &err);
  return rc ? fail(rc, err) : FMI_OK;
}

int fmi_index_load(const char *path, fmi_index **out) {
  if (!path || !out) return fail(FMI_EINVAL, "NULL argument");
  *out = nullptr;
  fmi_index *ix = new (std::nothrow) fmi_index();
  if (!ix) return fail(FMI_ENOMEM, "out of host memory");
  std::string err;
  int rc = fmi::load_index(path, ix, &err);
  if (rc) { delete ix; return fail(rc, err); }
  *out = ix;
  return FMI_OK;
}

int fmi_index_import_reference(const char *dir, fmi_index **out) {
  if (!dir || !out) return fail(FMI_EINVAL, "NULL argument");
  *out = nullptr;
  fmi_index *ix = new (std::nothrow) fmi_index();
  if (!ix) return fail(FMI_ENOMEM, "out of host memory");
  std::string err;
  int rc = fmi::import_reference_dir(dir, ix, &err);
  if (rc) { delete ix; return fail(rc, err); }
  *out = ix;
  return FMI_OK;
}

void fmi_index_free(fmi_index *ix) {
  if (!ix) return;
  if (ix->on_device) {
    int cur = -1;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(ix->device);
    (void)hipFree(ix->d_mem);
    (void)hipFree(ix->d_full);
    (void)hipFree(ix->d_kmer);
    (void)hipFree(ix->d_koffset);
    (void)hipFree(ix->d_kstart);
    if (cur >= 0) (void)hipSetDevice(cur);
  }
  delete ix;
}

uint64_t fmi_index_size_sa(const fmi_index *ix) { return ix ? ix->n + 1 : 0; }

uint64_t fmi_index_sample_count(const fmi_index *ix) { return ix ? ix->h[0].nsamp : 0; }

int fmi_index_get_samples(const fmi_index *ix, int forward, uint32_t *S, uint32_t *R) {
  if (!ix) return fail(FMI_EINVAL, "index is NULL");
  const fmi::HostDir &h = ix->h[forward ? 1 : 0];
  if (S) std::memcpy(S, h.S, h.nsamp * 4);
  if (R) std::memcpy(R, h.R, h.nsamp * 4);
  return FMI_OK;
}
uint32_t fmi_index_sa_ratio(const fmi_index *ix) { return ix ? ix->ratio : 0; }

uint64_t fmi_index_device_bytes(const fmi_index *ix) {
  if (!ix) return 0;
  uint64_t b = 0;
  for (int d = 0; d < 2; ++d)
    b += ix->h[d].nblocks * sizeof(fmi::OccBlock) + 2 * ((ix->h[d].nsamp * 4 + 255) & ~255ull);
  return b + ix->full_bytes + ix->kmer_bytes;
}

int fmi_index_device_layout(const fmi_index *ix, uint64_t *out) {
  if (!ix || !out) return fail(FMI_EINVAL, "NULL argument");
  out[0] = fmi_index_device_bytes(ix) - ix->full_bytes - ix->kmer_bytes;
  out[1] = ix->full_bytes;
  out[2] = ix->kmer_bytes;
  return FMI_OK;
}

int fmi_index_expand(fmi_index *ix) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (ix->d_full) return FMI_OK;
  if (ix->ratio == 0) return fail(FMI_ESTATE, "the index has no samples");
  const uint64_t n1 = ix->n + 1;
  const uint64_t arr = (n1 * 4 + 255) & ~255ull;
  const uint64_t tbytes = (((ix->n + 15) / 16) * 4 + 255) & ~255ull;
  const uint64_t bytes = 4 * arr + tbytes;
  void *mem = nullptr;
  hipError_t e = hipMalloc(&mem, bytes);
  if (e != hipSuccess)
    return fail(FMI_ENOMEM, "hipMalloc(full SA / ISA / text, " + std::to_string(bytes) + " bytes)");
  char *p = static_cast<char *>(mem);
  uint32_t *sa[2], *isa[2];
  for (int d = 0; d < 2; ++d) {
    sa[d] = reinterpret_cast<uint32_t *>(p); p += arr;
    isa[d] = reinterpret_cast<uint32_t *>(p); p += arr;
  }
  uint32_t *text = reinterpret_cast<uint32_t *>(p);
  for (int d = 0; d < 2 && e == hipSuccess; ++d)
    e = fmi::launch_expand(ix->dev.d[d], (uint32_t)ix->n, ix->ratio, sa[d], isa[d], nullptr);
  if (e == hipSuccess) e = fmi::launch_pack_text(ix->dev.d[0], (uint32_t)ix->n, isa[0], text, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(mem);
    return hip_fail(e, "expand launch");
  }
  for (int d = 0; d < 2; ++d) {
    ix->dev.d[d].sa = sa[d];
    ix->dev.d[d].isa = isa[d];
  }
  ix->dev.text = text;
  ix->d_full = mem;
  ix->full_bytes = bytes;
  return FMI_OK;
}

int fmi_index_kmer_table(fmi_index *ix, uint32_t k) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (k != 0 && (k < 4 || k > 15)) return fail(FMI_EINVAL, "k must be 0 or 4..15");
  if (ix->d_kmer) {
    HIP_TRY(hipDeviceSynchronize(), "drain before freeing the table");
    (void)hipFree(ix->d_kmer);
    ix->d_kmer = nullptr;
    ix->kmer_bytes = 0;
    ix->dev.kmer[0] = ix->dev.kmer[1] = nullptr;
    ix->dev.kmer_k = 0;
    ix->dev.kmer2[0] = ix->dev.kmer2[1] = nullptr;
    ix->dev.kmer2_k = 0;
  }
  if (k == 0) return FMI_OK;
  // the tables of k-strings and, from k = 8 on, of (k - 4)-strings
  const uint32_t k2 = k >= 8 ? k - 4 : 0;
  const uint64_t per = (1ull << (2 * k)) * sizeof(uint64_t);
  const uint64_t per2 = k2 ? (1ull << (2 * k2)) * sizeof(uint64_t) : 0;
  void *mem = nullptr;
  hipError_t e = hipMalloc(&mem, 2 * (per + per2));
  if (e != hipSuccess)
    return fail(FMI_ENOMEM, "hipMalloc(k-mer table, " + std::to_string(2 * (per + per2)) + " bytes)");
  char *base = static_cast<char *>(mem);
  uint64_t *t[2] = {reinterpret_cast<uint64_t *>(base), reinterpret_cast<uint64_t *>(base + per)};
  uint64_t *t2[2] = {k2 ? reinterpret_cast<uint64_t *>(base + 2 * per) : nullptr,
                     k2 ? reinterpret_cast<uint64_t *>(base + 2 * per + per2) : nullptr};
  for (int d = 0; d < 2 && e == hipSuccess; ++d) {
    e = fmi::launch_kmer_table(ix->dev.d[d], (uint32_t)ix->n, k, t[d], nullptr);
    if (e == hipSuccess && k2) e = fmi::launch_kmer_table(ix->dev.d[d], (uint32_t)ix->n, k2, t2[d], nullptr);
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(mem);
    return hip_fail(e, "k-mer table launch");
  }
  ix->d_kmer = mem;
  ix->kmer_bytes = 2 * (per + per2);
  ix->dev.kmer[0] = t[0];
  ix->dev.kmer[1] = t[1];
  ix->dev.kmer_k = k;
  ix->dev.kmer2[0] = t2[0];
  ix->dev.kmer2[1] = t2[1];
  ix->dev.kmer2_k = k2;
  return FMI_OK;
}

int fmi_index_get_kmer_table(const fmi_index *ix, int forward, uint64_t *out) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (!out) return fail(FMI_EINVAL, "NULL argument");
  if (!ix->d_kmer) return fail(FMI_ESTATE, "call fmi_index_kmer_table first");
  const uint64_t per = (1ull << (2 * ix->dev.kmer_k)) * sizeof(uint64_t);
  HIP_TRY(hipMemcpy(out, ix->dev.kmer[forward ? 1 : 0], per, hipMemcpyDeviceToHost),
          "copy k-mer table");
  return FMI_OK;
}

int fmi_index_get_full(const fmi_index *ix, int forward, uint32_t *sa, uint32_t *isa,
                       uint32_t *text) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (!ix->d_full) return fail(FMI_ESTATE, "call fmi_index_expand first");
  const int d = forward ? 1 : 0;
  const uint64_t n1 = ix->n + 1;
  if (sa) HIP_TRY(hipMemcpy(sa, ix->dev.d[d].sa, n1 * 4, hipMemcpyDeviceToHost), "copy SA");
  if (isa) HIP_TRY(hipMemcpy(isa, ix->dev.d[d].isa, n1 * 4, hipMemcpyDeviceToHost), "copy ISA");
  if (text)
    HIP_TRY(hipMemcpy(text, ix->dev.text, ((ix->n + 15) / 16) * 4, hipMemcpyDeviceToHost), "copy text");
  return FMI_OK;
}

int fmi_index_to_device(fmi_index *ix) {
  if (!ix) return fail(FMI_EINVAL, "index is NULL");
  if (ix->on_device) return FMI_OK;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev), "hipGetDevice (no usable GPU?)");
  const uint64_t bytes = fmi_index_device_bytes(ix);
  void *mem = nullptr;
  hipError_t e = hipMalloc(&mem, bytes);
  if (e != hipSuccess) return fail(FMI_ENOMEM, std::string("hipMalloc(index): ") + hipGetErrorString(e));
  char *p = static_cast<char *>(mem);
  fmi::DevIndex dv{};
  for (int d = 0; d < 2; ++d) {
    const fmi::HostDir &h = ix->h[d];
    const uint64_t bb = h.nblocks * sizeof(fmi::OccBlock);
    const uint64_t sb = (h.nsamp * 4 + 255) & ~255ull;
    HIP_TRY(hipMemcpy(p, h.blk, bb, hipMemcpyHostToDevice), "copy blocks");
    dv.d[d].blk = reinterpret_cast<const fmi::OccBlock *>(p);
    p += bb;
    HIP_TRY(hipMemcpy(p, h.S, h.nsamp * 4, hipMemcpyHostToDevice), "copy S");
    dv.d[d].S = reinterpret_cast<const uint32_t *>(p);
    p += sb;
    HIP_TRY(hipMemcpy(p, h.R, h.nsamp * 4, hipMemcpyHostToDevice), "copy R");
    dv.d[d].R = reinterpret_cast<const uint32_t *>(p);
    p += sb;
    for (int b = 0; b < 4; ++b) dv.d[d].C[b] = h.C[b];
    dv.d[d].dollar = (uint32_t)h.dollar;
    dv.d[d].dollar_blk = (uint32_t)(h.dollar / fmi::kBlockRows);
    dv.d[d].dollar_off = (uint32_t)(h.dollar % fmi::kBlockRows);
  }
  dv.n = (uint32_t)ix->n;
  dv.ratio = ix->ratio;
  const uint64_t kc = ix->kstart.size();
  void *ko = nullptr, *ks = nullptr;
  HIP_TRY(hipMalloc(&ko, (kc + 1) * 8), "hipMalloc(karyotype)");
  HIP_TRY(hipMalloc(&ks, (kc ? kc : 1) * 8), "hipMalloc(karyotype)");
  HIP_TRY(hipMemcpy(ko, ix->koffset.data(), (kc + 1) * 8, hipMemcpyHostToDevice), "copy karyotype");
  if (kc) HIP_TRY(hipMemcpy(ks, ix->kstart.data(), kc * 8, hipMemcpyHostToDevice), "copy karyotype");
  ix->d_mem = mem;
  ix->d_koffset = static_cast<uint64_t *>(ko);
  ix->d_kstart = static_cast<uint64_t *>(ks);
  ix->dev = dv;
  ix->device = dev;
  ix->device_bytes = bytes;
  ix->on_device = true;
  return FMI_OK;
}

int fmi_index_get_bwt(const fmi_index *ix, int forward, uint8_t *out, uint64_t *dollar) {
  if (!ix || !out || !dollar) return fail(FMI_EINVAL, "NULL argument");
  const fmi::HostDir &h = ix->h[forward ? 1 : 0];
  // rows [0, dollar) land at out[r], rows past the '$' row at out[r - 1];
  // threads take contiguous row ranges
  const uint64_t rows = ix->n + 1;
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(nt, 16u));
  if (rows < (1u << 20)) nt = 1;
  const uint64_t per = (rows + nt - 1) / nt;
  auto work = [&](unsigned t) {
    const uint64_t r0 = t * per, r1 = std::min(rows, r0 + per);
    for (uint64_t r = r0; r < r1; ++r) {
      if (r == h.dollar) continue;
      out[r < h.dollar ? r : r - 1] = (uint8_t)fmi::host_bwt_sym_raw(h, r);
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(work, t);
  work(0);
  for (auto &x : th) x.join();
  *dollar = h.dollar;
  return FMI_OK;
}

int fmi_index_set_karyotype(fmi_index *ix, uint64_t count, const uint64_t *start,
                            const uint64_t *end) {
  if (!ix || (count && (!start || !end))) return fail(FMI_EINVAL, "NULL argument");
  if (ix->on_device) return fail(FMI_ESTATE, "set the karyotype before fmi_index_to_device");
  std::vector<uint64_t> off(count + 1, 0);
  for (uint64_t c = 0; c < count; ++c) {
    if (end[c] + 1 < start[c]) return fail(FMI_EINVAL, "contig end < start - 1");
    off[c + 1] = off[c] + (end[c] - start[c] + 1);
  }
  if (off[count] != ix->n) return fail(FMI_EINVAL, "contig lengths do not add up to the text length");
  ix->kstart.assign(start, start + count);
  ix->kend.assign(end, end + count);
  ix->koffset = off;
  ix->knames.assign(count, std::string());
  return FMI_OK;
}

int fmi_index_set_contig_names(fmi_index *ix, uint64_t count, const char *const *names) {
  if (!ix || (count && !names)) return fail(FMI_EINVAL, "NULL argument");
  if (count != ix->kstart.size()) return fail(FMI_EINVAL, "count differs from the karyotype's");
  std::vector<std::string> v(count);
  for (uint64_t c = 0; c < count; ++c) {
    if (!names[c]) return fail(FMI_EINVAL, "NULL name");
    v[c] = names[c];
    if (v[c].size() > 4096) return fail(FMI_ERANGE, "contig name longer than 4096 bytes");
  }
  ix->knames = std::move(v);
  return FMI_OK;
}

uint64_t fmi_index_contig_count(const fmi_index *ix) { return ix ? ix->kstart.size() : 0; }

const char *fmi_index_contig_name(const fmi_index *ix, uint64_t c) {
  if (!ix || c >= ix->kstart.size()) return nullptr;
  return c < ix->knames.size() ? ix->knames[c].c_str() : "";
}

int fmi_device_count(int *count) {
  if (!count) return fail(FMI_EINVAL, "NULL argument");
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) { *count = 0; return hip_fail(e, "hipGetDeviceCount"); }
  return FMI_OK;
}

int fmi_device_alloc(void **d_ptr, size_t bytes) {
  if (!d_ptr) return fail(FMI_EINVAL, "NULL argument");
  *d_ptr = nullptr;
  hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
  if (e != hipSuccess) return fail(FMI_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return FMI_OK;
}

int fmi_device_free(void *d_ptr) {
  HIP_TRY(hipFree(d_ptr), "hipFree");
  return FMI_OK;
}

int fmi_memcpy_h2d(void *d_dst, const void *src, size_t bytes, fmi_stream s) {
  if (!bytes) return FMI_OK;
  if (!d_dst || !src) return fail(FMI_EINVAL, "NULL argument");
  HIP_TRY(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(s)),
          "hipMemcpyAsync H2D");
  return FMI_OK;
}

int fmi_memcpy_d2h(void *dst, const void *d_src, size_t bytes, fmi_stream s) {
  if (!bytes) return FMI_OK;
  if (!dst || !d_src) return fail(FMI_EINVAL, "NULL argument");
  HIP_TRY(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(s)),
          "hipMemcpyAsync D2H");
  return FMI_OK;
}

int fmi_memset_d(void *d_dst, int value, size_t bytes, fmi_stream s) {
  if (!bytes) return FMI_OK;
  if (!d_dst) return fail(FMI_EINVAL, "NULL argument");
  hipStream_t st = static_cast<hipStream_t>(s);
  // no memset nodes in a graph: a fill kernel while the stream captures
  if (capturing(st)) {
    HIP_TRY(fmi::launch_fill(d_dst, (uint8_t)value, bytes, st), "fill launch");
    return FMI_OK;
  }
  HIP_TRY(hipMemsetAsync(d_dst, value, bytes, st), "hipMemsetAsync");
  return FMI_OK;
}

int fmi_device_name(char *name, size_t size) {
  if (!name || size == 0) return fail(FMI_EINVAL, "NULL argument");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  std::snprintf(name, size, "%s:%s", prop.name, prop.gcnArchName);
  return FMI_OK;
}

int fmi_device_cu_count(int *cus) {
  if (!cus) return fail(FMI_EINVAL, "NULL argument");
  return cu_count(cus);
}

int fmi_device_mem_info(uint64_t *free_bytes, uint64_t *total_bytes) {
  if (!free_bytes || !total_bytes) return fail(FMI_EINVAL, "NULL argument");
  size_t f = 0, t = 0;
  HIP_TRY(hipMemGetInfo(&f, &t), "hipMemGetInfo");
  *free_bytes = f;
  *total_bytes = t;
  return FMI_OK;
}

int fmi_stream_create(fmi_stream *s) {
  if (!s) return fail(FMI_EINVAL, "NULL argument");
  hipStream_t st = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
  *s = st;
  return FMI_OK;
}

int fmi_stream_destroy(fmi_stream s) {
  if (!s) return fail(FMI_EINVAL, "the default stream cannot be destroyed");
  hipStream_t st = static_cast<hipStream_t>(s);
  int rc = fmi::release_stream_state(st);
  if (rc) return rc;
  HIP_TRY(hipStreamDestroy(st), "hipStreamDestroy");
  return FMI_OK;
}

int fmi_stream_sync(fmi_stream s) {
  HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(s)), "hipStreamSynchronize");
  return FMI_OK;
}

int fmi_graph_begin(fmi_stream s) {
  if (!s) return fail(FMI_EINVAL, "capture needs a stream of its own (fmi_stream_create)");
  HIP_TRY(hipStreamBeginCapture(static_cast<hipStream_t>(s), hipStreamCaptureModeRelaxed),
          "hipStreamBeginCapture");
  return FMI_OK;
}

int fmi_graph_end(fmi_stream s, fmi_graph *g) {
  if (!s || !g) return fail(FMI_EINVAL, "NULL argument");
  *g = nullptr;
  hipStream_t st = static_cast<hipStream_t>(s);
  std::string key;
  int rc = stream_key(st, &key);
  if (rc) return rc;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  unsigned long long id = 0;
  HIP_TRY(hipStreamGetCaptureInfo(st, &cs, &id), "hipStreamGetCaptureInfo");
  std::vector<std::string> keys{key};
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto m = g_capture_members.find(id);
    if (m != g_capture_members.end()) {
      keys.insert(keys.end(), m->second.begin(), m->second.end());
      g_capture_members.erase(m);
    }
  }
  hipGraph_t graph = nullptr;
  HIP_TRY(hipStreamEndCapture(st, &graph), "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  {
    // the streams' counters and scratch are now baked into the graph
    std::lock_guard<std::mutex> lk(g_mu);
    for (const std::string &k : keys) g_streams[k].graphs++;
    g_graph_stream[exec] = keys;
  }
  *g = exec;
  return FMI_OK;
}

int fmi_stream_wait(fmi_stream waiter, fmi_stream signaler) {
  if (!waiter || !signaler || waiter == signaler)
    return fail(FMI_EINVAL, "two different streams of fmi_stream_create");
  hipStream_t w = static_cast<hipStream_t>(waiter), sg = static_cast<hipStream_t>(signaler);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  unsigned long long id = 0;
  HIP_TRY(hipStreamGetCaptureInfo(sg, &cs, &id), "hipStreamGetCaptureInfo");
  const bool joins = cs == hipStreamCaptureStatusActive && !capturing(w);
  hipEvent_t ev = nullptr;
  HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreateWithFlags");
  hipError_t e = hipEventRecord(ev, sg);
  if (e == hipSuccess) e = hipStreamWaitEvent(w, ev, 0);
  (void)hipEventDestroy(ev);
  if (e != hipSuccess) return hip_fail(e, "fmi_stream_wait");
  if (joins) {
    // the waiter now records into the signaler's capture
    std::string key;
    int rc = stream_key(w, &key);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    g_capture_members[id].push_back(key);
  }
  return FMI_OK;
}

int fmi_graph_launch(fmi_graph g, fmi_stream s) {
  if (!g) return fail(FMI_EINVAL, "NULL graph");
  HIP_TRY(hipGraphLaunch(static_cast<hipGraphExec_t>(g), static_cast<hipStream_t>(s)),
          "hipGraphLaunch");
  return FMI_OK;
}

int fmi_graph_destroy(fmi_graph g) {
  if (!g) return FMI_OK;
  HIP_TRY(hipGraphExecDestroy(static_cast<hipGraphExec_t>(g)), "hipGraphExecDestroy");
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_graph_stream.find(g);
  if (it != g_graph_stream.end()) {
    for (const std::string &k : it->second) {
      auto st = g_streams.find(k);
      if (st != g_streams.end() && st->second.graphs) st->second.graphs--;
    }
    g_graph_stream.erase(it);
  }
  return FMI_OK;
}

int fmi_last_kernel_ms(float *ms) {
  if (!ms) return fail(FMI_EINVAL, "NULL argument");
  if (!g_timed) return fail(FMI_ESTATE, "no kernel launched on this thread");
  HIP_TRY(hipEventSynchronize(g_ev1), "hipEventSynchronize");
  HIP_TRY(hipEventElapsedTime(ms, g_ev0, g_ev1), "hipEventElapsedTime");
  return FMI_OK;
}

int fmi_timing_begin(void) {
  if (g_region) return fail(FMI_ESTATE, "timing region already open");
  g_region = true;
  g_used = 0;
  return FMI_OK;
}

int fmi_timing_end(uint32_t *launches, double *sum_ms, float *min_ms, float *max_ms) {
  if (!g_region) return fail(FMI_ESTATE, "no timing region open");
  if (!launches || !sum_ms || !min_ms || !max_ms) return fail(FMI_EINVAL, "NULL argument");
  g_region = false;
  double sum = 0;
  float mn = 0, mx = 0;
  for (size_t i = 0; i < g_used; ++i) {
    float ms = 0;
    HIP_TRY(hipEventSynchronize(g_pool[i].second), "hipEventSynchronize");
    HIP_TRY(hipEventElapsedTime(&ms, g_pool[i].first, g_pool[i].second), "hipEventElapsedTime");
    sum += ms;
    if (i == 0 || ms < mn) mn = ms;
    if (i == 0 || ms > mx) mx = ms;
  }
  *launches = (uint32_t)g_used;
  *sum_ms = sum;
  *min_ms = mn;
  *max_ms = mx;
  return FMI_OK;
}

int fmi_exact_batch(const fmi_index *ix, int view, int direction, const uint8_t *d_reads,
                    const uint64_t *d_offsets, uint32_t nreads, fmi_hit *d_out, fmi_hit *d_anchor,
                    fmi_stream s) {
  int rc = need_device(ix);
  if (rc) return rc;
  uint32_t dsel, comp;
  rc = view_params(view, direction, &dsel, &comp);
  if (rc) return rc;
  if (nreads == 0) return FMI_OK;
  if (!d_reads || !d_offsets || !d_out) return fail(FMI_EINVAL, "NULL buffer");
  hipStream_t st = static_cast<hipStream_t>(s);
  rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_exact(ix->dev, dsel, comp, (uint32_t)direction, d_reads, d_offsets,
                                   nreads, d_out, d_anchor, st);
  return end_timed(st, e, "exact launch");
}

int fmi_exact_count_work(const fmi_index *ix, int view, int direction, const uint8_t *d_reads,
                         const uint64_t *d_offsets, uint32_t nreads, fmi_hit *d_out,
                         fmi_hit *d_anchor, fmi_stream s, uint64_t *work) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (!work) return fail(FMI_EINVAL, "NULL argument");
  for (int i = 0; i < 5; ++i) work[i] = 0;
  uint32_t dsel, comp;
  rc = view_params(view, direction, &dsel, &comp);
  if (rc) return rc;
  if (nreads == 0) return FMI_OK;
  if (!d_reads || !d_offsets || !d_out) return fail(FMI_EINVAL, "NULL buffer");
  hipStream_t st = static_cast<hipStream_t>(s);
  if (capturing(st)) return fail(FMI_ESTATE, "a counting launch cannot be captured");
  Workspace w;
  rc = workspace_for(st, &w);
  if (rc) return rc;
  HIP_TRY(fmi::launch_zero(reinterpret_cast<uint32_t *>(w.stats), 2 * 5, st), "zero stats");
  HIP_TRY(fmi::launch_exact(ix->dev, dsel, comp, (uint32_t)direction, d_reads, d_offsets, nreads,
                            d_out, d_anchor, st, w.stats),
          "exact count launch");
  unsigned long long st5[5] = {0, 0, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(st5, w.stats, sizeof st5, hipMemcpyDeviceToHost, st), "copy stats");
  HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
  for (int i = 0; i < 5; ++i) work[i] = st5[i];
  return FMI_OK;
}

int fmi_exact_vector_batch(const fmi_index *ix, int view, int direction, const uint8_t *d_reads,
                           const uint64_t *d_offsets, uint32_t nreads, int64_t *d_vec_k,
                           int64_t *d_vec_l, int16_t *d_last_good, fmi_stream s) {
  int rc = need_device(ix);
  if (rc) return rc;
  uint32_t dsel, comp;
  rc = view_params(view, direction, &dsel, &comp);
  if (rc) return rc;
  if (nreads == 0) return FMI_OK;
  if (!d_reads || !d_offsets || !d_vec_k || !d_vec_l || !d_last_good) return fail(FMI_EINVAL, "NULL buffer");
  hipStream_t st = static_cast<hipStream_t>(s);
  rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_vector(ix->dev, dsel, comp, (uint32_t)direction, d_reads, d_offsets,
                                    nreads, d_vec_k, d_vec_l, d_last_good, st);
  return end_timed(st, e, "vector launch");
}

int fmi_search1_batch(const fmi_index *ix, int mode, int strand_mask, const uint8_t *d_reads,
                      const uint64_t *d_offsets, uint32_t nreads, uint32_t cap, fmi_hit *d_hits,
                      uint32_t *d_counts, fmi_stream s) {
  return search1_common(ix, mode, strand_mask, d_reads, d_offsets, nreads, cap, d_hits, d_counts,
                        s, true);
}

int fmi_search1_batch_noalloc(const fmi_index *ix, int mode, int strand_mask,
                              const uint8_t *d_reads, const uint64_t *d_offsets, uint32_t nreads,
                              uint32_t cap, fmi_hit *d_hits, uint32_t *d_counts, fmi_stream s) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (mode != FMI_MODE_EDIT && mode != FMI_MODE_MISMATCH) return fail(FMI_EINVAL, "mode must be EDIT or MISMATCH");
  if (strand_mask < 1 || strand_mask > 3) return fail(FMI_EINVAL, "strand_mask must be 1, 2 or 3");
  if (nreads == 0) return FMI_OK;
  if (!d_reads || !d_offsets || !d_counts) return fail(FMI_EINVAL, "NULL buffer");
  if (cap > 0 && !d_hits) return fail(FMI_EINVAL, "d_hits is NULL with cap > 0");
  if (nreads > (1u << 30)) return fail(FMI_ERANGE, "more than 2^30 reads in one batch");
  // the launch needs the queue counter and 5 work counters: 104 bytes
  static_assert(kNoAllocScratchBytes >= 64 + 5 * sizeof(unsigned long long),
                "scratch holds the queue and work counters");
  void *base = nullptr;
  rc = noalloc_scratch(&base);
  if (rc) return rc;
  uint32_t *queue = static_cast<uint32_t *>(base);
  unsigned long long *stats =
      reinterpret_cast<unsigned long long *>(static_cast<char *>(base) + 64);
  hipStream_t st = static_cast<hipStream_t>(s);
  Tuning t;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    t = g_tuning;
  }
  int cus = 0;
  rc = cu_count(&cus);
  if (rc) return rc;
  const uint32_t grid = (uint32_t)cus * (t.wg_per_cu ? t.wg_per_cu : kDefaultWgPerCu);
  // the counter keeps the last call's value: clear it with a kernel (no
  // memset), on the call's stream, before the search reads it
  HIP_TRY(fmi::launch_zero(queue, 1, st), "zero queue");
  const bool own = t.loads == FMI_LOADS_OWN ||
                   (t.loads == FMI_LOADS_AUTO && fmi_index_device_bytes(ix) > kOwnLoadsBytes);
  HIP_TRY(fmi::launch_search1(ix->dev, mode == FMI_MODE_MISMATCH, true, false,
                              (uint32_t)strand_mask, d_reads, d_offsets, nreads, cap, d_hits,
                              d_counts, queue, stats, grid, t.chunk, t.refill_min, own, st),
          "search1 launch");
  return FMI_OK;
}

int fmi_search1_vector_batch(const fmi_index *ix, int mode, int strand_mask,
                             const uint8_t *d_reads, const uint64_t *d_offsets,
                             const int16_t *d_bounds, uint32_t nreads, uint32_t cap,
                             fmi_hit *d_hits, uint32_t *d_counts, fmi_stream s) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (mode != FMI_MODE_EDIT && mode != FMI_MODE_MISMATCH) return fail(FMI_EINVAL, "mode must be EDIT or MISMATCH");
  if (strand_mask < 1 || strand_mask > 3) return fail(FMI_EINVAL, "strand_mask must be 1, 2 or 3");
  if (nreads == 0) return FMI_OK;
  if (!d_reads || !d_offsets || !d_counts) return fail(FMI_EINVAL, "NULL buffer");
  if (cap > 0 && !d_hits) return fail(FMI_EINVAL, "d_hits is NULL with cap > 0");
  if (nreads > (1u << 30)) return fail(FMI_ERANGE, "more than 2^30 reads in one batch");
  hipStream_t st = static_cast<hipStream_t>(s);
  // per-lane vector scratch: sized by the lane count, no read of the offsets
  const uint32_t lanes = fmi::vhelper_lanes(nreads);
  void *scr = nullptr;
  rc = scratch_for(st, fmi::vhelper_scratch_bytes(lanes), &scr);
  if (rc) return rc;
  rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_vhelper(ix->dev, mode == FMI_MODE_MISMATCH, (uint32_t)strand_mask,
                                     d_reads, d_offsets, d_bounds, nreads, cap, d_hits, d_counts,
                                     scr, lanes, st);
  return end_timed(st, e, "search1 vector launch");
}

namespace {

int search_k_common(const fmi_index *ix, int strand_mask, const uint8_t *d_reads,
                    const uint64_t *d_offsets, const int16_t *d_fragsize, uint32_t nreads,
                    uint32_t list_cap, uint32_t anchor_cap, fmi_hitk *d_final, uint32_t *d_nfinal,
                    fmi_hitk *d_anchors, uint32_t *d_nanchors, uint8_t *d_full, fmi_stream s,
                    uint64_t *work) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (strand_mask < 1 || strand_mask > 3) return fail(FMI_EINVAL, "strand_mask must be 1, 2 or 3");
  if (nreads == 0) return FMI_OK;
  if (!d_reads || !d_offsets || !d_fragsize || !d_final || !d_nfinal || !d_full)
    return fail(FMI_EINVAL, "NULL buffer");
  if (d_anchors && !d_nanchors) return fail(FMI_EINVAL, "d_nanchors is NULL");
  if (list_cap == 0 || list_cap > 65536) return fail(FMI_EINVAL, "list_cap must be 1..65536");
  if (d_anchors && (anchor_cap == 0 || anchor_cap > 4096))
    return fail(FMI_EINVAL, "anchor_cap must be 1..4096");
  if (nreads > (1u << 30)) return fail(FMI_ERANGE, "more than 2^30 reads in one batch");
  hipStream_t st = static_cast<hipStream_t>(s);
  Workspace w;
  rc = workspace_for(st, &w);
  if (rc) return rc;
  // lanes: enough for the tasks, at most what the device holds resident,
  // their lists and D within 12 GiB
  int cus = 0;
  rc = cu_count(&cus);
  if (rc) return rc;
  const uint64_t per_lane = fmi::search_k_lane_bytes(list_cap);
  uint64_t lanes = std::min<uint64_t>((2ull * nreads + 255) & ~255ull,
                                      (uint64_t)fmi::search_k_lanes_per_cu() * (uint64_t)cus);
  while (lanes > 256 && lanes * per_lane > (12ull << 30)) lanes -= 256;
  void *scr = nullptr;
  // the front kernel: all its lanes resident
  const uint32_t front_lanes =
      nreads <= fmi::kStagedMaxTasks ? fmi::search_k_front_lanes_per_cu() * (uint32_t)cus : 0u;
  rc = scratch_for(st, fmi::search_k_scratch_bytes((uint32_t)lanes, list_cap) +
                          (front_lanes ? fmi::search_k_front_bytes(nreads) : 0ull), &scr);
  if (rc) return rc;
  if (work) {
    // the counting launch: synchronous, its counters in the stream's stats
    HIP_TRY(fmi::launch_zero(reinterpret_cast<uint32_t *>(w.stats), 2 * 6, st), "zero stats");
    HIP_TRY(fmi::launch_search_k(ix->dev, d_reads, d_offsets, d_fragsize, nreads,
                                 (uint32_t)strand_mask, list_cap, anchor_cap, d_final, d_nfinal,
                                 d_anchors, d_nanchors, d_full, scr, (uint32_t)lanes, front_lanes,
                                 w.queue, st, w.stats),
            "k-error count launch");
    unsigned long long st6[6] = {0, 0, 0, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(st6, w.stats, sizeof st6, hipMemcpyDeviceToHost, st), "copy stats");
    HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (int i = 0; i < 6; ++i) work[i] = st6[i];
    return FMI_OK;
  }
  rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_search_k(ix->dev, d_reads, d_offsets, d_fragsize, nreads,
                                      (uint32_t)strand_mask, list_cap, anchor_cap, d_final,
                                      d_nfinal, d_anchors, d_nanchors, d_full, scr,
                                      (uint32_t)lanes, front_lanes, w.queue, st);
  return end_timed(st, e, "k-error search launch");
}

}  // namespace

int fmi_search_k_batch(const fmi_index *ix, int strand_mask, const uint8_t *d_reads,
                       const uint64_t *d_offsets, const int16_t *d_fragsize, uint32_t nreads,
                       uint32_t list_cap, uint32_t anchor_cap, fmi_hitk *d_final,
                       uint32_t *d_nfinal, fmi_hitk *d_anchors, uint32_t *d_nanchors,
                       uint8_t *d_full, fmi_stream s) {
  return search_k_common(ix, strand_mask, d_reads, d_offsets, d_fragsize, nreads, list_cap,
                         anchor_cap, d_final, d_nfinal, d_anchors, d_nanchors, d_full, s, nullptr);
}

int fmi_search_k_count_work(const fmi_index *ix, int strand_mask, const uint8_t *d_reads,
                            const uint64_t *d_offsets, const int16_t *d_fragsize, uint32_t nreads,
                            uint32_t list_cap, uint32_t anchor_cap, fmi_hitk *d_final,
                            uint32_t *d_nfinal, fmi_hitk *d_anchors, uint32_t *d_nanchors,
                            uint8_t *d_full, fmi_stream s, uint64_t *work) {
  if (!work) return fail(FMI_EINVAL, "NULL argument");
  for (int i = 0; i < 6; ++i) work[i] = 0;
  if (capturing(static_cast<hipStream_t>(s))) return fail(FMI_ESTATE, "a counting launch cannot be captured");
  return search_k_common(ix, strand_mask, d_reads, d_offsets, d_fragsize, nreads, list_cap,
                         anchor_cap, d_final, d_nfinal, d_anchors, d_nanchors, d_full, s, work);
}


int fmi_search1_batch_plain(const fmi_index *ix, int mode, int strand_mask,
                            const uint8_t *d_reads, const uint64_t *d_offsets, uint32_t nreads,
                            uint32_t cap, fmi_hit *d_hits, uint32_t *d_counts, fmi_stream s) {
  return search1_common(ix, mode, strand_mask, d_reads, d_offsets, nreads, cap, d_hits, d_counts,
                        s, false);
}

int fmi_set_tuning(uint32_t wg_per_cu, uint32_t chunk, uint32_t refill_min) {
  if (wg_per_cu > 64 || chunk > 4096 || refill_min > 64) return fail(FMI_EINVAL, "tuning out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  g_tuning.wg_per_cu = wg_per_cu;
  g_tuning.chunk = chunk ? chunk : kDefaultChunk;
  g_tuning.refill_min = refill_min ? refill_min : kDefaultRefill;
  return FMI_OK;
}

int fmi_get_tuning(uint32_t *wg_per_cu, uint32_t *chunk, uint32_t *refill_min) {
  if (!wg_per_cu || !chunk || !refill_min) return fail(FMI_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(g_mu);
  *wg_per_cu = g_tuning.wg_per_cu ? g_tuning.wg_per_cu : kDefaultWgPerCu;
  *chunk = g_tuning.chunk;
  *refill_min = g_tuning.refill_min;
  return FMI_OK;
}

int fmi_set_load_mode(uint32_t mode) {
  if (mode > FMI_LOADS_OWN) return fail(FMI_EINVAL, "load mode must be FMI_LOADS_AUTO, _ALL or _OWN");
  std::lock_guard<std::mutex> lk(g_mu);
  g_tuning.loads = mode;
  return FMI_OK;
}

int fmi_get_load_mode(uint32_t *mode) {
  if (!mode) return fail(FMI_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(g_mu);
  *mode = g_tuning.loads;
  return FMI_OK;
}

int fmi_set_staging(int enable) {
  if (enable != 0 && enable != 1) return fail(FMI_EINVAL, "staging must be 0 or 1");
  std::lock_guard<std::mutex> lk(g_mu);
  g_tuning.staging = enable != 0;
  return FMI_OK;
}

int fmi_get_staging(int *enable) {
  if (!enable) return fail(FMI_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(g_mu);
  *enable = g_tuning.staging ? 1 : 0;
  return FMI_OK;
}

int fmi_locate_batch(const fmi_index *ix, int forward, const int64_t *d_k, const int64_t *d_l,
                     uint32_t nintervals, uint32_t cap, uint64_t *d_pos, uint32_t *d_nrows,
                     fmi_stream s) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (nintervals == 0 || cap == 0) return FMI_OK;
  if (!d_k || !d_l || !d_pos || !d_nrows) return fail(FMI_EINVAL, "NULL buffer");
  if ((uint64_t)nintervals * cap > (1ull << 36)) return fail(FMI_ERANGE, "nintervals * cap too large");
  hipStream_t st = static_cast<hipStream_t>(s);
  rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_locate(ix->dev, forward ? 1u : 0u, d_k, d_l, 1, nintervals, cap, d_pos,
                                    d_nrows, st);
  return end_timed(st, e, "locate launch");
}

int fmi_locate_hits(const fmi_index *ix, int forward, const fmi_hit *d_hits, uint32_t count,
                    uint32_t cap, uint64_t *d_pos, uint32_t *d_nrows, fmi_stream s) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (count == 0 || cap == 0) return FMI_OK;
  if (!d_hits || !d_pos || !d_nrows) return fail(FMI_EINVAL, "NULL buffer");
  if ((uint64_t)count * cap > (1ull << 36)) return fail(FMI_ERANGE, "count * cap too large");
  static_assert(sizeof(fmi_hit) % sizeof(int64_t) == 0 && offsetof(fmi_hit, k) == 0 &&
                    offsetof(fmi_hit, l) == 8,
                "fmi_hit layout");
  hipStream_t st = static_cast<hipStream_t>(s);
  rc = begin_timed(st);
  if (rc) return rc;
  const int64_t *base = reinterpret_cast<const int64_t *>(d_hits);
  hipError_t e = fmi::launch_locate(ix->dev, forward ? 1u : 0u, base, base + 1,
                                    (uint32_t)(sizeof(fmi_hit) / sizeof(int64_t)), count, cap,
                                    d_pos, d_nrows, st);
  return end_timed(st, e, "locate launch");
}

int fmi_inverse_locate_batch(const fmi_index *ix, int forward, const uint64_t *d_text_pos,
                             uint64_t count, uint64_t *d_rows, fmi_stream s) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (count == 0) return FMI_OK;
  if (!d_text_pos || !d_rows) return fail(FMI_EINVAL, "NULL buffer");
  if (count > (1ull << 36)) return fail(FMI_ERANGE, "count too large");
  hipStream_t st = static_cast<hipStream_t>(s);
  rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_isa(ix->dev, forward ? 1u : 0u, d_text_pos, count, d_rows, st);
  return end_timed(st, e, "inverse locate launch");
}

int fmi_encode_bases(const char *src, uint64_t n, uint8_t *dst) {
  if (n == 0) return FMI_OK;
  if (!src || !dst) return fail(FMI_EINVAL, "NULL buffer");
  static const struct Table {
    uint8_t t[256];
    Table() {
      for (int i = 0; i < 256; ++i) t[i] = 4;
      t['A'] = t['a'] = 0;
      t['C'] = t['c'] = 1;
      t['G'] = t['g'] = 2;
      t['T'] = t['t'] = 3;
    }
  } table;
  for (uint64_t i = 0; i < n; ++i) dst[i] = table.t[(uint8_t)src[i]];
  return FMI_OK;
}

int fmi_fastq_encode_batch(const char *d_seq, const int32_t *d_data_indices, uint32_t num_reads,
                           uint64_t nbytes, uint8_t *d_codes, uint64_t *d_offsets, fmi_stream s) {
  if (!d_data_indices || !d_offsets || (nbytes && (!d_seq || !d_codes)))
    return fail(FMI_EINVAL, "NULL buffer");
  if (nbytes < num_reads) return fail(FMI_EINVAL, "fewer bytes than reads (each read ends in a NUL)");
  if (nbytes > (1ull << 40)) return fail(FMI_ERANGE, "nbytes too large");
  hipStream_t st = static_cast<hipStream_t>(s);
  int rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_fastq_encode(d_seq, d_data_indices, num_reads, nbytes, 0, 0, d_codes,
                                          d_offsets, st);
  return end_timed(st, e, "fastq encode launch");
}

int fmi_pack_codes(const uint8_t *d_codes, uint64_t nbases, uint32_t *d_packed, uint16_t *d_nmask,
                   fmi_stream s) {
  if (nbases == 0) return FMI_OK;
  if (!d_codes || !d_packed) return fail(FMI_EINVAL, "NULL buffer");
  hipStream_t st = static_cast<hipStream_t>(s);
  int rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_pack_codes(d_codes, nbases, d_packed, d_nmask, st);
  return end_timed(st, e, "pack launch");
}

namespace {

int resolve_common(const fmi_index *ix, const fmi_hit *d_hits, const uint32_t *d_counts,
                   const uint64_t *d_offsets, uint32_t nreads, uint32_t cap, uint32_t row_cap,
                   int32_t *d_chrom, uint64_t *d_coord, uint32_t record_len, fmi_stream s) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (nreads == 0 || cap == 0 || row_cap == 0) return FMI_OK;
  if (!d_hits || !d_counts || !d_offsets || !d_chrom || !d_coord) return fail(FMI_EINVAL, "NULL buffer");
  if ((uint64_t)2 * nreads * cap * row_cap > (1ull << 36)) return fail(FMI_ERANGE, "output too large");
  hipStream_t st = static_cast<hipStream_t>(s);
  rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_resolve(ix->dev, d_hits, d_counts, d_offsets, nreads, cap, row_cap,
                                     ix->d_koffset, ix->d_kstart, (uint32_t)ix->kstart.size(),
                                     record_len, d_chrom, d_coord, st);
  return end_timed(st, e, "resolve launch");
}

}  // namespace

int fmi_resolve_hits(const fmi_index *ix, const fmi_hit *d_hits, const uint32_t *d_counts,
                     const uint64_t *d_offsets, uint32_t nreads, uint32_t cap, uint32_t row_cap,
                     int32_t *d_chrom, uint64_t *d_coord, fmi_stream s) {
  return resolve_common(ix, d_hits, d_counts, d_offsets, nreads, cap, row_cap, d_chrom, d_coord,
                        0u, s);
}

int fmi_resolve_hits_text(const fmi_index *ix, const fmi_hit *d_hits, const uint32_t *d_counts,
                          const uint64_t *d_offsets, uint32_t nreads, uint32_t cap,
                          uint32_t row_cap, int32_t *d_chrom, uint64_t *d_coord, fmi_stream s) {
  return resolve_common(ix, d_hits, d_counts, d_offsets, nreads, cap, row_cap, d_chrom, d_coord,
                        1u, s);
}

int fmi_seed_map_batch(const fmi_index *ix, const uint8_t *d_reads, const uint64_t *d_offsets,
                       uint32_t nreads, uint32_t seed_len, uint32_t stride,
                       uint32_t filter_seed_mappings, fmi_seed_info *d_seed_info,
                       uint64_t seed_info_cap, fmi_seed_region *d_regions, uint64_t region_cap,
                       uint64_t *nseeds, uint64_t *nrows, fmi_stream s) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (!nseeds || !nrows) return fail(FMI_EINVAL, "NULL argument");
  *nseeds = 0;
  *nrows = 0;
  if (seed_len == 0 || seed_len > FMI_MAX_READ_LEN || stride == 0)
    return fail(FMI_EINVAL, "seed_len must be 1..1000 and stride >= 1");
  if (nreads == 0) return FMI_OK;
  if (!d_reads || !d_offsets) return fail(FMI_EINVAL, "NULL buffer");
  if (region_cap && !d_regions) return fail(FMI_EINVAL, "d_regions is NULL with region_cap > 0");
  if (seed_info_cap && !d_seed_info) return fail(FMI_EINVAL, "d_seed_info is NULL with a capacity");
  std::string err;
  rc = fmi::seed_map(ix, d_reads, d_offsets, nreads, seed_len, stride, filter_seed_mappings,
                     seed_info_cap ? d_seed_info : nullptr, seed_info_cap, d_regions, region_cap,
                     nseeds, nrows, s, &err);
  return rc ? fail(rc, err) : FMI_OK;
}

int fmi_map_inexact_seeds_batch(const fmi_index *ix, const uint8_t *d_reads,
                                const uint64_t *d_offsets, uint32_t nreads, uint32_t seed_size,
                                uint32_t min_seed_size, uint32_t rec_cap, uint32_t region_cap,
                                uint64_t windows_cap, fmi_inexact_region *d_regions,
                                uint32_t *d_nregions, uint8_t *d_status, uint64_t *nwindows,
                                fmi_stream s) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (!nwindows) return fail(FMI_EINVAL, "NULL argument");
  *nwindows = 0;
  if (seed_size == 0 || seed_size > FMI_MAX_READ_LEN || min_seed_size == 0)
    return fail(FMI_EINVAL, "seed_size must be 1..1000 and min_seed_size >= 1");
  if (rec_cap == 0 || rec_cap > 20000 || region_cap == 0 || region_cap > 65536)
    return fail(FMI_EINVAL, "rec_cap must be 1..20000 and region_cap 1..65536");
  if (nreads == 0) return FMI_OK;
  if (!d_reads || !d_offsets || !d_regions || !d_nregions || !d_status)
    return fail(FMI_EINVAL, "NULL buffer");
  std::string err;
  rc = fmi::map_inexact_seeds(ix, d_reads, d_offsets, nreads, seed_size, min_seed_size, rec_cap,
                              region_cap, windows_cap, d_regions, d_nregions, d_status, nwindows,
                              s, &err);
  return rc ? fail(rc, err) : FMI_OK;
}

int fmi_map_reads_batch(const fmi_index *ix, const uint8_t *d_reads, const uint64_t *d_offsets,
                        uint32_t nreads, uint32_t filter_read_mappings, uint32_t cap,
                        fmi_read_mapping *d_out, uint32_t *d_counts, uint8_t *d_status,
                        fmi_stream s) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (filter_read_mappings > 4096) return fail(FMI_EINVAL, "filter_read_mappings must be 0..4096");
  if (nreads == 0) return FMI_OK;
  if (!d_reads || !d_offsets || !d_counts || !d_status) return fail(FMI_EINVAL, "NULL buffer");
  if (cap > 0 && !d_out) return fail(FMI_EINVAL, "d_out is NULL with cap > 0");
  if (nreads > (1u << 30)) return fail(FMI_ERANGE, "more than 2^30 reads in one batch");
  if (ix->kstart.empty()) return fail(FMI_ESTATE, "the index has no karyotype");
  hipStream_t st = static_cast<hipStream_t>(s);
  const uint32_t lanes = fmi::map_reads_lanes(nreads, filter_read_mappings);
  void *scr = nullptr;
  rc = scratch_for(st, fmi::map_reads_scratch_bytes(lanes, filter_read_mappings), &scr);
  if (rc) return rc;
  rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_map_reads(ix->dev, d_reads, d_offsets, nreads, ix->d_koffset,
                                       ix->d_kstart, (uint32_t)ix->kstart.size(),
                                       filter_read_mappings, cap, d_out, d_counts, d_status, scr,
                                       lanes, st);
  return end_timed(st, e, "map reads launch");
}

int fmi_format_results(const fmi_index *ix, uint64_t read_index, int strand, const char *mapping,
                       uint32_t nW, const fmi_hit *hits, uint32_t nhits, const int32_t *chrom,
                       const uint64_t *coord, uint32_t row_cap, char *buf, size_t buf_size,
                       size_t *len, uint64_t *rows_missing) {
  if (!ix || !len || !rows_missing || (nW && !mapping)) return fail(FMI_EINVAL, "NULL argument");
  if (nhits && (!hits || !chrom || !coord || row_cap == 0))
    return fail(FMI_EINVAL, "NULL records or rows");
  if (buf_size && !buf) return fail(FMI_EINVAL, "buf is NULL with buf_size > 0");
  std::string text;
  int rc = fmi::format_results(ix, read_index, strand, mapping ? mapping : "", nW, hits, nhits,
                               chrom, coord, row_cap, &text, rows_missing);
  if (rc) return fail(rc, "format_results");
  *len = text.size();
  if (buf_size) {
    const size_t n = text.size() < buf_size ? text.size() : buf_size - 1;
    std::memcpy(buf, text.data(), n);
    buf[n] = 0;
  }
  return FMI_OK;
}

int fmi_gather_bench(const void *d_buf, uint64_t bytes, uint32_t line_bytes, uint64_t nlines,
                     uint64_t *d_sink, float *ms, uint64_t *lines_done, fmi_stream s) {
  if (!d_buf || !d_sink || !ms || !lines_done) return fail(FMI_EINVAL, "NULL argument");
  if (line_bytes != 16 && line_bytes != 32 && line_bytes != 64 && line_bytes != 128 &&
      line_bytes != 256)
    return fail(FMI_EINVAL, "line_bytes must be 16, 32, 64, 128 or 256");
  if (bytes < (64ull << 20)) return fail(FMI_EINVAL, "buffer must be >= 64 MiB");
  int cus = 0;
  int rc = cu_count(&cus);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(s);
  rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_gather(d_buf, bytes, line_bytes, nlines,
                                    reinterpret_cast<unsigned long long *>(d_sink),
                                    (uint32_t)cus * 16u, st);
  rc = end_timed(st, e, "gather launch");
  if (rc) return rc;
  *lines_done = fmi::gather_loads_done(nlines, (uint32_t)cus * 16u);
  return fmi_last_kernel_ms(ms);
}

int fmi_gather_probe(const void *d_buf, uint64_t bytes, uint32_t lines_in_flight,
                     uint32_t wg_per_cu, int cooperative, uint64_t nlines, uint64_t *d_sink,
                     float *ms, uint64_t *lines_done, fmi_stream s) {
  if (!d_buf || !d_sink || !ms || !lines_done) return fail(FMI_EINVAL, "NULL argument");
  const uint32_t f = lines_in_flight;
  if (f != 1 && f != 2 && f != 4 && f != 8 && f != 16)
    return fail(FMI_EINVAL, "lines_in_flight must be 1, 2, 4, 8 or 16");
  if (wg_per_cu < 1 || wg_per_cu > 64) return fail(FMI_EINVAL, "wg_per_cu must be 1..64");
  if (bytes < (64ull << 20)) return fail(FMI_EINVAL, "buffer must be >= 64 MiB");
  if (nlines == 0 || nlines > (1ull << 34)) return fail(FMI_EINVAL, "nlines must be 1..2^34");
  int cus = 0;
  int rc = cu_count(&cus);
  if (rc) return rc;
  const uint32_t grid = (uint32_t)cus * wg_per_cu;
  hipStream_t st = static_cast<hipStream_t>(s);
  rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_gather_probe(d_buf, bytes, f, cooperative != 0, nlines,
                                          reinterpret_cast<unsigned long long *>(d_sink), grid, st);
  rc = end_timed(st, e, "gather probe launch");
  if (rc) return rc;
  *lines_done = fmi::gather_probe_lines(nlines, grid, f, cooperative != 0);
  return fmi_last_kernel_ms(ms);
}

int fmi_plain_create(const fmi_index *ix, fmi_plain **out) {
  if (!out) return fail(FMI_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = need_device(ix);
  if (rc) return rc;
  std::string err;
  fmi::RefPlain *rp = nullptr;
  rc = fmi::refplain_create(ix, &rp, &err);
  if (rc) return fail(rc, err);
  *out = reinterpret_cast<fmi_plain *>(rp);
  return FMI_OK;
}

int fmi_plain_search1(fmi_plain *p, int mode, const uint8_t *reads, const uint64_t *offsets,
                      uint32_t nreads, uint32_t cap, fmi_hit *hits, uint32_t *counts,
                      double *ms_kernel, double *ms_total) {
  if (!p || !ms_kernel || !ms_total) return fail(FMI_EINVAL, "NULL argument");
  if (mode != FMI_MODE_EDIT && mode != FMI_MODE_MISMATCH) return fail(FMI_EINVAL, "bad mode");
  *ms_kernel = *ms_total = 0;
  if (nreads == 0) return FMI_OK;
  if (!reads || !offsets || !counts || (cap && !hits)) return fail(FMI_EINVAL, "NULL buffer");
  if (nreads > (1u << 26)) return fail(FMI_ERANGE, "more than 2^26 reads in one batch");
  std::string err;
  int rc = fmi::refplain_search1(reinterpret_cast<fmi::RefPlain *>(p), mode == FMI_MODE_MISMATCH,
                                 reads, offsets, nreads, cap, hits, counts, ms_kernel, ms_total,
                                 &err);
  return rc ? fail(rc, err) : FMI_OK;
}

uint64_t fmi_plain_words(const fmi_plain *p) {
  return fmi::refplain_words(reinterpret_cast<const fmi::RefPlain *>(p));
}

int fmi_plain_get_layout(const fmi_plain *p, int forward, uint32_t *desp, uint64_t *count) {
  if (!p || !desp || !count) return fail(FMI_EINVAL, "NULL argument");
  std::string err;
  int rc = fmi::refplain_layout(reinterpret_cast<const fmi::RefPlain *>(p), forward, desp, count,
                                &err);
  return rc ? fail(rc, err) : FMI_OK;
}

uint64_t fmi_plain_device_bytes(const fmi_plain *p) {
  return fmi::refplain_device_bytes(reinterpret_cast<const fmi::RefPlain *>(p));
}

int fmi_plain_free(fmi_plain *p) {
  fmi::refplain_free(reinterpret_cast<fmi::RefPlain *>(p));
  return FMI_OK;
}

int fmi_copy_bench(const void *d_src, void *d_dst, uint64_t bytes, float *ms, fmi_stream s) {
  if (!d_src || !d_dst || !ms) return fail(FMI_EINVAL, "NULL argument");
  if (bytes == 0 || bytes % 16) return fail(FMI_EINVAL, "bytes must be a positive multiple of 16");
  if ((reinterpret_cast<uintptr_t>(d_src) | reinterpret_cast<uintptr_t>(d_dst)) & 15u)
    return fail(FMI_EINVAL, "buffers must be 16-byte aligned");
  int cus = 0;
  int rc = cu_count(&cus);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(s);
  rc = begin_timed(st);
  if (rc) return rc;
  hipError_t e = fmi::launch_copy(d_src, d_dst, bytes, (uint32_t)cus * 8u, st);
  rc = end_timed(st, e, "copy launch");
  if (rc) return rc;
  return fmi_last_kernel_ms(ms);
}

int fmi_search1_count_work(const fmi_index *ix, int mode, int strand_mask,
                           const uint8_t *d_reads, const uint64_t *d_offsets, uint32_t nreads,
                           uint32_t cap, uint64_t *work) {
  int rc = need_device(ix);
  if (rc) return rc;
  if (!work) return fail(FMI_EINVAL, "NULL argument");
  if (mode != FMI_MODE_EDIT && mode != FMI_MODE_MISMATCH) return fail(FMI_EINVAL, "bad mode");
  if (strand_mask < 1 || strand_mask > 3) return fail(FMI_EINVAL, "bad strand_mask");
  for (int i = 0; i < 5; ++i) work[i] = 0;
  if (nreads == 0) return FMI_OK;
  if (!d_reads || !d_offsets) return fail(FMI_EINVAL, "NULL buffer");
  hipStream_t st = nullptr;
  Workspace w;
  rc = workspace_for(st, &w);
  if (rc) return rc;
  uint32_t *counts = nullptr;
  HIP_TRY(hipMalloc(&counts, (size_t)2 * nreads * 4), "hipMalloc(counts)");
  int cus = 0;
  rc = cu_count(&cus);
  if (rc) { (void)hipFree(counts); return rc; }
  HIP_TRY(fmi::launch_zero(w.queue, 1, st), "zero queue");
  HIP_TRY(fmi::launch_zero(reinterpret_cast<uint32_t *>(w.stats), 10, st), "zero stats");
  hipError_t e = fmi::launch_search1(ix->dev, mode == FMI_MODE_MISMATCH, true, true,
                                     (uint32_t)strand_mask, d_reads, d_offsets, nreads, cap, nullptr,
                                     counts, w.queue, w.stats, (uint32_t)cus * kDefaultWgPerCu,
                                     kDefaultChunk, kDefaultRefill, false, st);
  if (e != hipSuccess) { (void)hipFree(counts); return hip_fail(e, "count launch"); }
  unsigned long long st5[5] = {0, 0, 0, 0, 0};
  e = hipMemcpy(st5, w.stats, 40, hipMemcpyDeviceToHost);
  (void)hipFree(counts);
  if (e != hipSuccess) return hip_fail(e, "copy stats");
  for (int i = 0; i < 5; ++i) work[i] = st5[i];
  return FMI_OK;
}

int fmi_search1_count_blocks(const fmi_index *ix, int mode, int strand_mask,
                             const uint8_t *d_reads, const uint64_t *d_offsets, uint32_t nreads,
                             uint64_t *blocks, uint64_t *lf_steps) {
  if (!blocks || !lf_steps) return fail(FMI_EINVAL, "NULL argument");
  uint64_t work[5];
  int rc = fmi_search1_count_work(ix, mode, strand_mask, d_reads, d_offsets, nreads, 0, work);
  *blocks = work[0];
  *lf_steps = work[1];
  return rc;
}

}  // extern "C"

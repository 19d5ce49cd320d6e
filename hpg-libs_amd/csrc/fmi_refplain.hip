// fmi_refplain.hip -- the comparison point: the reference's GPU design
// (gpu/gpu.cu:118-203 glue, gpu/gpu.cu:333-451 kernels) re-created on this
// device and extended to the same 1-error search the engine runs.
//
// What makes it "plain", as in the reference:
//   * the reference's occurrence layout (FM_COMP_64, search/csafm.h:51-81):
//     per symbol b a uint32 checkpoint array desp[b][m] and a uint64 bitmap
//     array count[b][m], reached through device arrays of 4 pointers each
//     (comp_matrix's SA_TYPE **desp), so Occ(b, m) is two dependent loads
//     from two arrays per bound, four per LF step (get_O, csafm.h:166-181);
//   * reads padded to a fixed 1000-byte stride (MAXLINE) with a length
//     array, one thread per read, one launch per strand;
//   * synchronous copies: reads in, then every (read, strand) list of cap
//     records out, with the device idle during both.
// The results are the engine's, record for record (the search logic is the
// same restatement of BWSearch1CPU, search/search.c:1711-1933).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fmi.h"
#include "fmi_device.h"
#include "fmi_index.h"
#include "fmi_internal.h"

namespace fmi {

namespace {

constexpr uint32_t kMaxLine = 1000;   // MAXLINE, search/search.h:67

// One view of one direction in the reference layout.
struct RefView {
  const uint32_t *const *desp;   // device array of 4 device pointers
  const uint64_t *const *count;
  uint32_t C[4], C1[4];
  uint32_t n;
};

__device__ __forceinline__ uint32_t ref_occ(const RefView &V, uint32_t b, uint32_t m) {
  // get_O (csafm.h:166-181): bit (r+1)%64 of word (r+1)/64 holds row r
  const uint32_t w = m >> 6, sh = 63u - (m & 63u);
  const uint32_t *d = V.desp[b];
  const uint64_t *c = V.count[b];
  return d[w] + (uint32_t)__popcll(c[w] << sh);
}

// BWiteration (runtime.h:38-42); a code > 3 (N) gives (n + 1, n).
__device__ __forceinline__ void ref_lf(const RefView &V, uint32_t k, uint32_t l, uint32_t b,
                                       uint32_t &ko, uint32_t &lo) {
  if (b > 3u) { ko = V.n + 1u; lo = V.n; return; }
  ko = V.C1[b] + ref_occ(V, b, k);
  lo = V.C[b] + ref_occ(V, b, l + 1u);
}

struct Rec {
  uint32_t k, l;
  int32_t start, pos, end;
  uint32_t dir;
};

// BWExactSearchBackward / Forward (search/search.h:9-65), final interval.
__device__ void ref_exact(const RefView &V, const uint8_t *W, uint32_t dir, Rec &r) {
  uint32_t k2 = r.k, l2 = r.l;
  int32_t i = r.pos;
  if (dir == 0) {
    for (; i >= r.start; --i) {
      ref_lf(V, k2, l2, W[i], k2, l2);
      if (k2 > l2) break;
    }
  } else {
    for (; i <= r.end; ++i) {
      ref_lf(V, k2, l2, W[i], k2, l2);
      if (k2 > l2) break;
    }
  }
  r.k = k2; r.l = l2; r.pos = i;
}

struct Out {
  fmi_hit *dst;
  uint32_t cap, cnt, strand;
  __device__ __forceinline__ void add(const Rec &r, uint32_t nmm, uint32_t kind, uint32_t base,
                                      int32_t err_pos) {
    if (cnt < cap)
      store_hit(dst + cnt, (int64_t)r.k, (int64_t)r.l, (int16_t)r.start, (int16_t)r.pos,
                (int16_t)r.end, (int16_t)err_pos, (uint8_t)r.dir, (uint8_t)nmm, (uint8_t)kind,
                (uint8_t)base, (uint8_t)strand);
    ++cnt;
  }
};

// BWSimpleSearch1Backward (dir 0, search.c:1711-1793) / Forward (dir 1,
// search.c:1795-1874) over [start, end] from the surviving interval (k, l);
// returns the exact extension (or the dead interval) in k, l.
__device__ void ref_simple1(const RefView &V, const uint8_t *W, uint32_t dir, int32_t start,
                            int32_t end, uint32_t &kio, uint32_t &lio, bool mm_only, Out &O) {
  Rec r{0, 0, start, 0, end, dir};
  uint32_t kn = kio, ln = lio, k, l, ka, la;
  int64_t size_next = (int64_t)ln - (int64_t)kn, size;
  const int32_t step = dir ? 1 : -1;
  for (int32_t i = dir ? start : end; dir ? (i <= end) : (i >= start); i += step) {
    k = kn; l = ln;
    if (k > l) { kio = k; lio = l; return; }
    ref_lf(V, k, l, W[i], kn, ln);
    size = size_next;
    size_next = (int64_t)ln - (int64_t)kn;
    if (size == size_next) continue;
    if (!mm_only) {
      r.k = k; r.l = l; r.pos = i + step;
      ref_exact(V, W, dir, r);
      if (r.k <= r.l) O.add(r, 1, FMI_DELETION, 0xFFu, i);
    }
    for (uint32_t b = 0; b < 4u; ++b) {
      ref_lf(V, k, l, b, ka, la);
      if (ka > la) continue;
      if (!mm_only) {
        r.k = ka; r.l = la; r.pos = i;
        ref_exact(V, W, dir, r);
        if (r.k <= r.l) O.add(r, 1, FMI_INSERTION, b, i);
      }
      if (b != W[i]) {
        r.k = ka; r.l = la; r.pos = i + step;
        ref_exact(V, W, dir, r);
        if (r.k <= r.l) O.add(r, 1, FMI_MISMATCH, b, i);
      }
    }
  }
  kio = kn; lio = ln;
}

struct PlainParams {
  RefView vb, vf;                // the strand's pair (search.c:1876 arguments)
  const uint8_t *reads;          // nreads x 1000 bytes
  const uint64_t *nW;            // read lengths
  fmi_hit *hits;
  uint32_t *counts;
  uint32_t nreads, cap, strand, mm_only;
};

// BWSearch1CPU (search.c:1876-1933) for one strand of one read per thread.
__global__ __launch_bounds__(256) void plain_search1_kernel(PlainParams P) {
  const uint32_t rd = blockIdx.x * blockDim.x + threadIdx.x;
  if (rd >= P.nreads) return;
  const uint8_t *W = P.reads + (uint64_t)rd * kMaxLine;
  const int32_t len = (int32_t)P.nW[rd];
  Out O{P.hits + ((uint64_t)rd * 2u + P.strand) * P.cap, P.cap, 0u, P.strand};
  const int32_t end = len - 1;
  int32_t half = len / 2;
  const uint32_t n = P.vb.n;
  Rec r{0, n, half, end, end, 0};
  ref_exact(P.vb, W, 0, r);
  if (r.k <= r.l) {
    uint32_t k = r.k, l = r.l;
    ref_simple1(P.vb, W, 0, 0, half - 1, k, l, P.mm_only != 0, O);
    if (k <= l) {
      Rec e{k, l, 0, -1, end, 0};
      O.add(e, 0, 0, 0, 0);
    }
  }
  half -= 1;
  Rec f{0, n, 0, 0, half, 1};
  ref_exact(P.vf, W, 1, f);
  if (f.k <= f.l) {
    uint32_t k = f.k, l = f.l;
    ref_simple1(P.vf, W, 1, half + 1, end, k, l, P.mm_only != 0, O);
  }
  P.counts[2u * rd + P.strand] = O.cnt;
}

// The reference layout from the engine's blocks: one thread per 64-row word.
// desp[b][w] = Occ(b, 64w - 1) (0 for w = 0); bit (r+1)%64 of count[b][w]
// is set when row r (64w-1 <= r <= 64w+62) holds b (calculate_O,
// search/preprocess.c:74-130).
__global__ __launch_bounds__(256) void plain_layout_kernel(DevDir D, uint32_t n, uint64_t words,
                                                           uint32_t *d0, uint32_t *d1,
                                                           uint32_t *d2, uint32_t *d3,
                                                           uint64_t *c0, uint64_t *c1,
                                                           uint64_t *c2, uint64_t *c3) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= words) return;
  uint32_t occ[4] = {0, 0, 0, 0};
  if (w > 0) {
    const uint32_t m = (uint32_t)(64 * w - 1);
    const uint32_t j = m / kBlockRows, t = m - j * kBlockRows;
    const Blk B = load_block(D.blk, j);
    for (uint32_t b = 0; b < 4u; ++b) occ[b] = occ_blk(B, D.dollar_blk, D.dollar_off, j, b, t);
  }
  uint64_t bits[4] = {0, 0, 0, 0};
  const int64_t r0 = (int64_t)(64 * w) - 1;
  for (int64_t r = r0 < 0 ? 0 : r0; r <= r0 + 63 && r <= (int64_t)n; ++r) {
    if ((uint32_t)r == D.dollar) continue;
    const uint32_t j = (uint32_t)r / kBlockRows, t = (uint32_t)r - j * kBlockRows;
    const Blk B = load_block(D.blk, j);
    const uint32_t s = code_at(B, t);
    const uint64_t bit = 1ull << (uint32_t)((r + 1) & 63);
    bits[0] |= s == 0u ? bit : 0ull;
    bits[1] |= s == 1u ? bit : 0ull;
    bits[2] |= s == 2u ? bit : 0ull;
    bits[3] |= s == 3u ? bit : 0ull;
  }
  d0[w] = occ[0]; d1[w] = occ[1]; d2[w] = occ[2]; d3[w] = occ[3];
  c0[w] = bits[0]; c1[w] = bits[1]; c2[w] = bits[2]; c3[w] = bits[3];
}

}  // namespace

struct RefPlain {
  uint32_t n = 0;
  uint64_t words = 0;
  void *mem = nullptr;                 // desp + count of both directions
  const uint32_t **ptr_desp = nullptr; // device: [dir][view comp][4]
  const uint64_t **ptr_count = nullptr;
  uint32_t C[2][4] = {}, C1[2][4] = {};
  uint64_t device_bytes = 0;
};

#define RP_TRY(expr, what)                                        \
  do {                                                            \
    hipError_t e_ = (expr);                                       \
    if (e_ != hipSuccess) {                                       \
      *err = std::string(what) + ": " + hipGetErrorString(e_);    \
      return FMI_EDEVICE;                                         \
    }                                                             \
  } while (0)

void refplain_free(RefPlain *rp) {
  if (!rp) return;
  (void)hipFree(rp->mem);
  (void)hipFree(rp->ptr_desp);
  (void)hipFree(rp->ptr_count);
  delete rp;
}

int refplain_create(const fmi_index *ix, RefPlain **out, std::string *err) {
  RefPlain *rp = new RefPlain();
  rp->n = (uint32_t)ix->n;
  rp->words = (ix->n + 2 + 63) / 64;
  const uint64_t dbytes = (rp->words * 4 + 255) & ~255ull;
  const uint64_t cbytes = rp->words * 8;
  rp->device_bytes = 2 * 4 * (dbytes + cbytes);
  hipError_t e = hipMalloc(&rp->mem, rp->device_bytes);
  if (e != hipSuccess) {
    delete rp;
    *err = "reference-layout index: hipMalloc failed";
    return FMI_ENOMEM;
  }
  char *p = static_cast<char *>(rp->mem);
  uint32_t *desp[2][4];
  uint64_t *count[2][4];
  for (int d = 0; d < 2; ++d) {
    for (int b = 0; b < 4; ++b) { desp[d][b] = (uint32_t *)p; p += dbytes; }
    for (int b = 0; b < 4; ++b) { count[d][b] = (uint64_t *)p; p += cbytes; }
    const DevDir &D = ix->dev.d[d];
    const uint64_t grid = (rp->words + 255) / 256;
    hipLaunchKernelGGL(plain_layout_kernel, dim3((uint32_t)grid), dim3(256), 0, nullptr, D,
                       rp->n, rp->words, desp[d][0], desp[d][1], desp[d][2], desp[d][3],
                       count[d][0], count[d][1], count[d][2], count[d][3]);
    e = hipGetLastError();
    if (e != hipSuccess) { refplain_free(rp); *err = "layout launch failed"; return FMI_EDEVICE; }
    for (int b = 0; b < 4; ++b) {
      rp->C[d][b] = D.C[b];
      rp->C1[d][b] = D.C[b] + 1u;
    }
  }
  // pointer tables: [dir][comp][b]; the complement view swaps A<->T and
  // C<->G (reverse_strand_O, csafm.c:36-100)
  const uint32_t *hd[16];
  const uint64_t *hc[16];
  for (int d = 0; d < 2; ++d)
    for (int c = 0; c < 2; ++c)
      for (int b = 0; b < 4; ++b) {
        hd[(d * 2 + c) * 4 + b] = desp[d][c ? 3 - b : b];
        hc[(d * 2 + c) * 4 + b] = count[d][c ? 3 - b : b];
      }
  if (hipMalloc(&rp->ptr_desp, sizeof hd) != hipSuccess ||
      hipMalloc(&rp->ptr_count, sizeof hc) != hipSuccess) {
    refplain_free(rp);
    *err = "hipMalloc(pointer tables)";
    return FMI_ENOMEM;
  }
  e = hipMemcpy(rp->ptr_desp, hd, sizeof hd, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(rp->ptr_count, hc, sizeof hc, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) { refplain_free(rp); *err = "reference-layout upload failed"; return FMI_EDEVICE; }
  *out = rp;
  return FMI_OK;
}

static RefView view_of(const RefPlain *rp, int d, int comp) {
  RefView v;
  v.desp = rp->ptr_desp + (d * 2 + comp) * 4;
  v.count = rp->ptr_count + (d * 2 + comp) * 4;
  for (int b = 0; b < 4; ++b) {
    const int s = comp ? 3 - b : b;   // reverse_strand_C (csafm.c:3-34)
    v.C[b] = rp->C[d][s];
    v.C1[b] = rp->C1[d][s];
  }
  v.n = rp->n;
  return v;
}

int refplain_search1(RefPlain *rp, int mm_only, const uint8_t *reads, const uint64_t *offsets,
                     uint32_t nreads, uint32_t cap, fmi_hit *hits, uint32_t *counts,
                     double *ms_kernel, double *ms_total, std::string *err) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  // the reference's batch layout: reads at a fixed MAXLINE stride
  std::vector<uint8_t> padded((size_t)nreads * kMaxLine, 0);
  std::vector<uint64_t> nW(nreads);
  for (uint32_t i = 0; i < nreads; ++i) {
    const uint64_t len = offsets[i + 1] - offsets[i];
    if (len > kMaxLine) { *err = "read longer than 1000 bases"; return FMI_EINVAL; }
    nW[i] = len;
    std::memcpy(padded.data() + (size_t)i * kMaxLine, reads + offsets[i], len);
  }
  uint8_t *d_reads = nullptr;
  uint64_t *d_nW = nullptr;
  fmi_hit *d_hits = nullptr;
  uint32_t *d_counts = nullptr;
  const size_t hit_bytes = (size_t)nreads * 2 * cap * sizeof(fmi_hit);
  auto cleanup = [&]() {
    (void)hipFree(d_reads);
    (void)hipFree(d_nW);
    (void)hipFree(d_hits);
    (void)hipFree(d_counts);
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FMI_OK;
  do {
    hipError_t e;
    if ((e = hipMalloc(&d_reads, padded.size())) != hipSuccess ||
        (e = hipMalloc(&d_nW, nreads * 8ull)) != hipSuccess ||
        (e = hipMalloc(&d_hits, hit_bytes ? hit_bytes : 32)) != hipSuccess ||
        (e = hipMalloc(&d_counts, nreads * 8ull)) != hipSuccess) {
      *err = "plain search: hipMalloc failed";
      rc = FMI_ENOMEM;
      break;
    }
    if ((e = hipMemcpy(d_reads, padded.data(), padded.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(d_nW, nW.data(), nreads * 8ull, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
      *err = std::string("plain search: ") + hipGetErrorString(e);
      rc = FMI_EDEVICE;
      break;
    }
    const uint32_t grid = (nreads + 255) / 256;
    (void)hipEventRecord(e0, nullptr);
    for (uint32_t strand = 0; strand < 2; ++strand) {
      // + strand: (backward, forward); - strand: (forward_rev, backward_rev)
      PlainParams P{strand ? view_of(rp, 1, 1) : view_of(rp, 0, 0),
                    strand ? view_of(rp, 0, 1) : view_of(rp, 1, 0),
                    d_reads, d_nW, d_hits, d_counts, nreads, cap, strand, mm_only ? 1u : 0u};
      hipLaunchKernelGGL(plain_search1_kernel, dim3(grid), dim3(256), 0, nullptr, P);
    }
    (void)hipEventRecord(e1, nullptr);
    if ((e = hipGetLastError()) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess) {
      *err = std::string("plain search launch: ") + hipGetErrorString(e);
      rc = FMI_EDEVICE;
      break;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_kernel = ms;
    if ((e = hipMemcpy(counts, d_counts, nreads * 8ull, hipMemcpyDeviceToHost)) != hipSuccess ||
        (hit_bytes && (e = hipMemcpy(hits, d_hits, hit_bytes, hipMemcpyDeviceToHost)) != hipSuccess)) {
      *err = std::string("plain search copy-out: ") + hipGetErrorString(e);
      rc = FMI_EDEVICE;
      break;
    }
  } while (0);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  cleanup();
  *ms_total = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
  return rc;
}

uint64_t refplain_device_bytes(const RefPlain *rp) { return rp ? rp->device_bytes : 0; }

uint64_t refplain_words(const RefPlain *rp) { return rp ? rp->words : 0; }

int refplain_layout(const RefPlain *rp, int forward, uint32_t *desp, uint64_t *count,
                    std::string *err) {
  const uint32_t *hd[16];
  const uint64_t *hc[16];
  RP_TRY(hipMemcpy(hd, rp->ptr_desp, sizeof hd, hipMemcpyDeviceToHost), "pointer table");
  RP_TRY(hipMemcpy(hc, rp->ptr_count, sizeof hc, hipMemcpyDeviceToHost), "pointer table");
  const int d = forward ? 1 : 0;
  for (int b = 0; b < 4; ++b) {
    RP_TRY(hipMemcpy(desp + (size_t)b * rp->words, hd[d * 8 + b], rp->words * 4,
                     hipMemcpyDeviceToHost), "desp");
    RP_TRY(hipMemcpy(count + (size_t)b * rp->words, hc[d * 8 + b], rp->words * 8,
                     hipMemcpyDeviceToHost), "count");
  }
  return FMI_OK;
}

}  // namespace fmi

// fmi_seed.hip -- exact seed mapping (fmi_seed_map_batch): bwt_map_exact_seed
// (bwt.c:1200-1294) for every seed of a batch of reads.
//
//   1. seeds per read, exclusive scan -> global seed numbering
//   2. one lane per seed: + strand backward search on BACKWARD, - strand
//      forward search on BACKWARD_REV (both over the seed's seed_len symbols),
//      the cumulative filter of bwt.c:1250-1269 (a signed 32-bit running
//      count, as the reference's `int actual_mappings`), rows to resolve
//   3. exclusive scan of the rows of every (seed, strand)
//   4. one lane per row: locate (getScompValue) on BACKWARD, binsearch over
//      the karyotype offsets, the contig-end check key + len <= offset[idx]
//      with len = seq_end - seq_start (one short, bwt.c:1214), region record.
#include <hip/hip_runtime.h>

#include <cstring>   // rocprim's texture_cache_iterator uses memset

#include <rocprim/rocprim.hpp>

#include <cstdint>
#include <string>

#include "../../include/fmi.h"
#include "fmi_device.h"
#include "fmi_index.h"
#include "fmi_internal.h"

namespace fmi {

namespace {

__global__ void seed_count_kernel(const uint64_t *offsets, uint32_t nreads, uint32_t seed_len,
                                  uint32_t stride, uint64_t *cnt) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= nreads;
       r += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t c = 0;
    if (r < nreads) {
      const uint64_t len = offsets[r + 1] - offsets[r];
      if (len >= seed_len) c = (len - seed_len) / stride + 1;
    }
    cnt[r] = c;
  }
}

// Largest r in [0, n) with base[r] <= g (base is non-decreasing, base[0] = 0).
__device__ __forceinline__ uint64_t owner(const uint64_t *base, uint64_t n, uint64_t g) {
  uint64_t lo = 0, hi = n;   // invariant: base[lo] <= g, answer in [lo, hi)
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) / 2;
    if (base[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

struct SeedParams {
  DevIndex ix;
  const uint8_t *reads;
  const uint64_t *offsets;
  const uint64_t *seed_base;   // nreads + 1
  uint64_t nseeds;
  uint32_t nreads, seed_len, stride;
  int32_t filter;
  fmi_seed_info *info;         // may be NULL
  uint64_t info_cap;
  uint64_t *rows;              // 2 * nseeds: rows of + then - per seed
  int64_t *kl;                 // 2 * nseeds: k of + and k of - (first row of each)
};

__global__ __launch_bounds__(256) void seed_search_kernel(SeedParams P) {
  const DevDir &D = P.ix.d[0];
  const uint32_t n = P.ix.n;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < P.nseeds;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = owner(P.seed_base, P.nreads, g);
    const uint64_t s = (g - P.seed_base[r]) * P.stride;
    const uint8_t *W = P.reads + P.offsets[r] + s;
    uint32_t nb;
    // + strand: BWExactSearchBackward over W[0 .. seed_len-1] on BACKWARD
    uint32_t kp = 0, lp = n;
    for (int32_t i = (int32_t)P.seed_len - 1; i >= 0; --i) {
      const uint32_t sym = W[i];
      if (sym > 3u) { kp = n + 1; lp = n; break; }
      lf_one(D, kp, lp, sym, kp, lp, nb);
      if (kp > lp) break;
    }
    // - strand: BWExactSearchForward on BACKWARD_REV (complemented symbols)
    uint32_t km = 0, lm = n;
    for (uint32_t i = 0; i < P.seed_len; ++i) {
      const uint32_t sym = W[i];
      if (sym > 3u) { km = n + 1; lm = n; break; }
      lf_one(D, km, lm, 3u - sym, km, lm, nb);
      if (km > lm) break;
    }
    // cumulative filter with the reference's int arithmetic
    int32_t actual = 0;
    const int64_t cp = (int64_t)lp - (int64_t)kp + 1, cm = (int64_t)lm - (int64_t)km + 1;
    actual = (int32_t)((uint32_t)actual + (uint32_t)(uint64_t)cp);
    const uint64_t rp = (actual > P.filter || kp > lp) ? 0 : (uint64_t)cp;
    actual = (int32_t)((uint32_t)actual + (uint32_t)(uint64_t)cm);
    const uint64_t rm = (actual > P.filter || km > lm) ? 0 : (uint64_t)cm;
    P.rows[2 * g] = rp;
    P.rows[2 * g + 1] = rm;
    P.kl[2 * g] = kp;
    P.kl[2 * g + 1] = km;
    if (P.info && g < P.info_cap) {
      fmi_seed_info inf;
      inf.k_plus = kp; inf.l_plus = lp; inf.k_minus = km; inf.l_minus = lm;
      inf.rows_plus = (uint32_t)rp; inf.rows_minus = (uint32_t)rm;
      P.info[g] = inf;
    }
  }
}

struct RowParams {
  DevIndex ix;
  const uint64_t *offsets;
  const uint64_t *seed_base;   // nreads + 1
  const uint64_t *row_base;    // 2 * nseeds exclusive scan
  const int64_t *kl;
  const uint64_t *koff;        // kcount + 1
  const uint64_t *kstart;
  uint64_t nslots, nrows;      // nslots = 2 * nseeds
  uint32_t nreads, seed_len, stride, kcount;
  fmi_seed_region *out;
};

__global__ __launch_bounds__(256) void seed_rows_kernel(RowParams P) {
  const DevDir &D = P.ix.d[0];
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < P.nrows;
       q += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t slot = owner(P.row_base, P.nslots, q);
    const uint64_t g = slot >> 1;
    const uint32_t strand = (uint32_t)(slot & 1u);
    const uint64_t j = (uint64_t)P.kl[slot] + (q - P.row_base[slot]);
    const uint64_t r = owner(P.seed_base, P.nreads, g);
    const uint64_t seq_start = (g - P.seed_base[r]) * P.stride;
    const uint64_t seq_end = seq_start + P.seed_len - 1;
    const uint64_t seq_len = P.offsets[r + 1] - P.offsets[r];
    const uint64_t len = seq_end - seq_start;
    const uint64_t key = sa_value(D, P.ix.ratio, P.ix.n, (uint32_t)j);
    uint64_t lo = 0, size = P.kcount;          // binsearch, search/io.h:43-64
    while (size > 0) {
      const uint64_t w = size / 2;
      if (P.koff[lo + w] <= key) { lo += w + 1; size -= w + 1; }
      else size = w;
    }
    fmi_seed_region reg;
    reg.read = (uint32_t)r;
    reg.seed_id = (uint16_t)(g - P.seed_base[r]);
    reg.strand = (uint8_t)strand;
    reg.seq_start = (uint16_t)(strand ? seq_len - seq_end - 1 : seq_start);
    reg.seq_end = (uint16_t)(strand ? seq_len - seq_start - 1 : seq_end);
    reg.chrom_idx = (uint32_t)lo;
    if (lo >= 1 && key + len <= P.koff[lo]) {
      reg.start = P.kstart[lo - 1] + (key - P.koff[lo - 1]) + 1;
      reg.end = reg.start + len;
      reg.valid = 1;
    } else {
      reg.start = 0;
      reg.end = 0;
      reg.valid = 0;
    }
    P.out[q] = reg;
  }
}

inline dim3 grid_for(uint64_t m) {
  uint64_t g = (m + 255) / 256;
  if (g > 65536ull * 4) g = 65536ull * 4;
  if (g == 0) g = 1;
  return dim3((uint32_t)g);
}

#define SEED_TRY(expr)                                            \
  do {                                                            \
    hipError_t e_ = (expr);                                       \
    if (e_ != hipSuccess) {                                       \
      *err = std::string(#expr) + ": " + hipGetErrorString(e_);   \
      rc = FMI_EDEVICE;                                           \
      goto done;                                                  \
    }                                                             \
  } while (0)

}  // namespace

int seed_map(const fmi_index *ix, const uint8_t *reads, const uint64_t *offsets, uint32_t nreads,
             uint32_t seed_len, uint32_t stride, uint32_t filter, fmi_seed_info *info,
             uint64_t info_cap, fmi_seed_region *regions, uint64_t region_cap, uint64_t *nseeds_out,
             uint64_t *nrows_out, void *stream, std::string *err) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = FMI_OK;
  uint64_t *cnt = nullptr, *base = nullptr, *rows = nullptr, *rbase = nullptr;
  int64_t *kl = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0, need = 0;
  uint64_t nseeds = 0, nrows = 0, last_rows = 0, last_base = 0;
  *nseeds_out = 0;
  *nrows_out = 0;
  SEED_TRY(hipMalloc(&cnt, (uint64_t)(nreads + 1) * 8));
  SEED_TRY(hipMalloc(&base, (uint64_t)(nreads + 1) * 8));
  seed_count_kernel<<<grid_for(nreads + 1), 256, 0, s>>>(offsets, nreads, seed_len, stride, cnt);
  SEED_TRY(rocprim::exclusive_scan(nullptr, need, cnt, base, (uint64_t)0, (uint64_t)nreads + 1,
                                   rocprim::plus<uint64_t>(), s));
  tmp_bytes = need;
  SEED_TRY(hipMalloc(&tmp, tmp_bytes));
  SEED_TRY(rocprim::exclusive_scan(tmp, need, cnt, base, (uint64_t)0, (uint64_t)nreads + 1,
                                   rocprim::plus<uint64_t>(), s));
  SEED_TRY(hipMemcpyAsync(&nseeds, base + nreads, 8, hipMemcpyDeviceToHost, s));
  SEED_TRY(hipStreamSynchronize(s));
  *nseeds_out = nseeds;
  if (nseeds == 0) goto done;
  SEED_TRY(hipMalloc(&rows, 2 * nseeds * 8));
  SEED_TRY(hipMalloc(&rbase, 2 * nseeds * 8));
  SEED_TRY(hipMalloc(&kl, 2 * nseeds * 8));
  {
    SeedParams P{ix->dev, reads, offsets, base, nseeds, nreads, seed_len, stride,
                 (int32_t)(filter > (uint32_t)INT32_MAX ? INT32_MAX : filter), info, info_cap, rows, kl};
    seed_search_kernel<<<grid_for(nseeds), 256, 0, s>>>(P);
  }
  SEED_TRY(hipGetLastError());
  SEED_TRY(rocprim::exclusive_scan(nullptr, need, rows, rbase, (uint64_t)0, 2 * nseeds,
                                   rocprim::plus<uint64_t>(), s));
  if (need > tmp_bytes) {
    SEED_TRY(hipFree(tmp));
    tmp = nullptr;
    tmp_bytes = need;
    SEED_TRY(hipMalloc(&tmp, tmp_bytes));
  }
  SEED_TRY(rocprim::exclusive_scan(tmp, need, rows, rbase, (uint64_t)0, 2 * nseeds,
                                   rocprim::plus<uint64_t>(), s));
  SEED_TRY(hipMemcpyAsync(&last_base, rbase + 2 * nseeds - 1, 8, hipMemcpyDeviceToHost, s));
  SEED_TRY(hipMemcpyAsync(&last_rows, rows + 2 * nseeds - 1, 8, hipMemcpyDeviceToHost, s));
  SEED_TRY(hipStreamSynchronize(s));
  nrows = last_base + last_rows;
  *nrows_out = nrows;
  if (nrows && region_cap && regions) {
    RowParams R{ix->dev, offsets, base, rbase, kl, ix->d_koffset, ix->d_kstart, 2 * nseeds,
                nrows < region_cap ? nrows : region_cap, nreads, seed_len, stride,
                (uint32_t)ix->kstart.size(), regions};
    seed_rows_kernel<<<grid_for(R.nrows), 256, 0, s>>>(R);
    SEED_TRY(hipGetLastError());
  }
done:
  if (rc == FMI_OK) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) { *err = hipGetErrorString(e); rc = FMI_EDEVICE; }
  }
  (void)hipFree(cnt);
  (void)hipFree(base);
  (void)hipFree(rows);
  (void)hipFree(rbase);
  (void)hipFree(kl);
  (void)hipFree(tmp);
  return rc;
}

}  // namespace fmi

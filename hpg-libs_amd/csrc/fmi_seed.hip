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

#include <algorithm>
#include <cstdint>
#include <string>

#include "../../include/fmi.h"
#include "fmi_device.h"
#include "fmi_index.h"
#include "fmi_internal.h"
#include "fmi_vhelper.h"

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
  // 2 * nseeds: k of + and k of - (the first row of each) in the low 32 bits
  // (rows are below 2^32 on the device), the seed's read in the high 32
  int64_t *kl;
};

// REGS (seed_len <= 32): the seed's bases are packed into registers once
// (vh_regs) and every base below comes from there, not from a byte load.
template <bool REGS>
__global__ __launch_bounds__(256) void seed_search_kernel(SeedParams P) {
  const DevDir &D = P.ix.d[0];
  const uint32_t n = P.ix.n;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < P.nseeds;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = owner(P.seed_base, P.nreads, g);
    const uint64_t s = (g - P.seed_base[r]) * P.stride;
    const uint8_t *W = P.reads + P.offsets[r] + s;
    VhRegs WR{0ull, 0u, 0};
    if (REGS) WR = vh_regs(W, 0, P.seed_len);
    auto base = [&](int32_t i) -> uint32_t { return REGS ? WR.at(i) : (uint32_t)W[i]; };
    // The first k steps of each search from the k-mer table when its k bases
    // are A/C/G/T and their interval is not empty (else step by step, so a
    // dead interval is the one the LF steps give).
    const uint32_t K = P.ix.kmer_k;
    const bool use_k = K && P.seed_len >= K;
    // + strand: BWExactSearchBackward over W[0 .. seed_len-1] on BACKWARD
    uint32_t kp = 0, lp = n;
    int32_t ip = (int32_t)P.seed_len - 1;
    // - strand: BWExactSearchForward on BACKWARD_REV (complemented symbols)
    uint32_t km = 0, lm = n;
    uint32_t im = 0;
    if (use_k) {
      uint32_t cp = 0u, cm = 0u, bad = 0u;
      for (uint32_t j = 0; j < K; ++j) {
        const uint32_t sp = base(ip - (int32_t)j), sm = base((int32_t)j);
        bad |= (sp > 3u || sm > 3u) ? 1u : 0u;
        cp |= (sp & 3u) << (2u * j);
        cm |= (sm & 3u) << (2u * j);
      }
      cm ^= (1u << (2u * K)) - 1u;
      if (!bad) {
        const uint64_t ep = P.ix.kmer[0][cp], em = P.ix.kmer[0][cm];
        if ((uint32_t)ep <= (uint32_t)(ep >> 32)) {
          kp = (uint32_t)ep; lp = (uint32_t)(ep >> 32); ip -= (int32_t)K;
        }
        if ((uint32_t)em <= (uint32_t)(em >> 32)) {
          km = (uint32_t)em; lm = (uint32_t)(em >> 32); im = K;
        }
      }
    }
    for (int32_t i = ip; i >= 0; --i) {
      uint32_t kn, ln;
      if (lf_sym(D, kp, lp, base(i), 0u, kn, ln) > 3u) { kp = n + 1; lp = n; break; }
      kp = kn; lp = ln;
      if (kp > lp) break;
    }
    for (uint32_t i = im; i < P.seed_len; ++i) {
      uint32_t kn, ln;
      if (lf_sym(D, km, lm, base((int32_t)i), 1u, kn, ln) > 3u) { km = n + 1; lm = n; break; }
      km = kn; lm = ln;
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
    P.kl[2 * g] = (int64_t)((uint64_t)kp | (r << 32));
    P.kl[2 * g + 1] = (int64_t)((uint64_t)km | (r << 32));
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
  const int64_t *kl;           // first row | read << 32 per (seed, strand)
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
    const uint64_t klr = (uint64_t)P.kl[slot];
    const uint64_t j = (klr & 0xFFFFFFFFull) + (q - P.row_base[slot]);
    const uint64_t r = klr >> 32;                 // the seed's read, from the search
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

// ---------------------------------------------------------------- inexact seeds
// bwt_map_inexact_seeds_seq (bwt.c:3262-3302) + bwt_map_inexact_seed
// (bwt.c:1401-1644):
//   1. windows per read (len / seed_size, plus the remainder when it holds
//      at least min_seed_size bases), exclusive scan -> window numbering
//   2. one lane per window (persistent grid): the vector helper on the +
//      strand then the - strand (final_pos carries over between them), each
//      record turned into regions as it is produced (the reference makes the
//      records first and then walks them in the same order), straight into
//      the caller's slot; then the slot's dedup and reversal in place.

__global__ void win_count_kernel(const uint64_t *offsets, uint32_t nreads, uint32_t seed_size,
                                 uint32_t min_seed_size, uint64_t *cnt) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= nreads;
       r += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t c = 0;
    if (r < nreads) {
      const uint64_t len = offsets[r + 1] - offsets[r];
      c = len / seed_size + (len % seed_size >= min_seed_size ? 1 : 0);
    }
    cnt[r] = c;
  }
}

struct InexactParams {
  DevIndex ix;
  const uint8_t *reads;
  const uint64_t *offsets;
  const uint64_t *win_base;    // nreads + 1
  uint64_t nwin;
  uint32_t nreads, seed_size, rec_cap, region_cap;
  const uint64_t *koff;        // kcount + 1
  const uint64_t *kstart;
  uint32_t kcount;
  uint2 *vscratch;             // lanes * (seed_size + 3)
  fmi_inexact_region *out;     // 2 * nwin * region_cap
  uint32_t *nreg;
  uint8_t *status;
};

// Regions of one strand's records (bwt.c:1490-1588), in record order.
struct RegionSink {
  const InexactParams *P;
  fmi_inexact_region *out;
  uint32_t m, nrec, type, read, aux_s, aux_e, seq_len, seed_id;
  int32_t start, len;
  int32_t *final_pos;
  uint8_t st;
  __device__ void add(const VhRec &r) {
    if (nrec++ >= P->rec_cap) { st |= 1; return; }   // the reference's list holds 20000
    // a record as the list stores it: no error fields without an error
    const uint32_t error = r.nmm ? r.kind : 0u;
    const int32_t pos = r.nmm ? r.err_pos : 0;
    int64_t len_calc = len;
    if (error == FMI_DELETION) len_calc--;
    else if (error == FMI_INSERTION) len_calc++;
    if (error == FMI_DELETION || pos == 0 || pos == len - 1) {
      *final_pos = type ? pos : len - 1 - pos;
    } else if (error == FMI_INSERTION) {
      if (type) *final_pos = r.dir ? pos : pos + 1;
      else *final_pos = r.dir ? len - pos : len - pos - 1;
    }
    int32_t final_err = error == FMI_MISMATCH ? -1 : (error == FMI_DELETION ? 1 : 2);
    final_err += start;                    // -1 mismatch, SEED_INSERTION 1, SEED_DELETION 2
    const int64_t k = r.k, l = r.l;
    const int64_t l0 = (l - k + 1) < 50 ? l : k + 1;   // bwt.c:1539-1547
    const uint32_t direction = type ? r.dir : !r.dir;
    const uint64_t size_sa = (uint64_t)P->ix.n + 1;
    for (int64_t j = k; j <= l0; ++j) {
      const uint64_t key =
          direction ? size_sa - sa_value(P->ix.d[1], P->ix.ratio, P->ix.n, (uint32_t)j) -
                          (uint64_t)len_calc - 1
                    : sa_value(P->ix.d[0], P->ix.ratio, P->ix.n, (uint32_t)j);
      uint64_t lo = 0, size = P->kcount;    // binsearch, search/io.h:43-64
      while (size > 0) {
        const uint64_t w = size / 2;
        if (P->koff[lo + w] <= key) { lo += w + 1; size -= w + 1; }
        else size = w;
      }
      if (!(lo >= 1 && key + (uint64_t)len_calc <= P->koff[lo])) continue;
      if (m >= P->region_cap) { st |= 2; continue; }
      fmi_inexact_region R;
      R.start = P->kstart[lo - 1] + (key - P->koff[lo - 1]) + 1;
      R.end = R.start + (uint64_t)len - 1;
      R.chrom_idx = (uint32_t)lo;
      R.read = read;
      R.seq_start = (uint16_t)aux_s;
      R.seq_end = (uint16_t)aux_e;
      R.seq_len = (uint16_t)seq_len;
      R.seed_id = (uint16_t)seed_id;
      R.pos_err = *final_pos;
      R.type_err = final_err;
      R.strand = (uint8_t)(1u - type);
      for (int b = 0; b < 7; ++b) R.pad[b] = 0;
      out[m++] = R;
    }
  }
};

// Held to 128 VGPRs for 4 waves per SIMD (157 unbounded: 3 waves): a lane's
// windows are chains of dependent loads, so a fourth wave to hide them pays
// for the few registers spilled (83.3 -> 76.7 ms on 8 M windows of 18 bases
// over 64 Mbp).  REGS (seed_size <= 32): the window's bases are packed into
// registers once (vh_regs) for both strands, so the helper's LF steps and
// comparisons read no read bytes from memory.
template <bool REGS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void inexact_seed_kernel(InexactParams P) {
  const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint2 *vb = P.vscratch + lane * (uint64_t)(P.seed_size + 3);
  for (uint64_t g = lane; g < P.nwin; g += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t rd = owner(P.win_base, P.nreads, g);
    const uint64_t off = P.offsets[rd];
    const int32_t len = (int32_t)(P.offsets[rd + 1] - off);
    const int32_t w = (int32_t)(g - P.win_base[rd]);
    const int32_t num_seeds = len / (int32_t)P.seed_size;
    const int32_t ws = w * (int32_t)P.seed_size;
    const int32_t we = w < num_seeds ? ws + (int32_t)P.seed_size - 1 : len - 1;
    int32_t final_pos = 0;
    VhRegs WR{0ull, 0u, ws};
    if (REGS && we - ws >= 2) WR = vh_regs(P.reads + off, ws, (uint32_t)(we - ws + 1));
    for (uint32_t s = 0; s < 2; ++s) {
      const uint32_t type = 1u - s;          // bwt.c:1460: type 1 (+) first
      const uint64_t slot = 2 * g + s;
      RegionSink S{&P, P.out + slot * P.region_cap, 0u, 0u, type, (uint32_t)rd,
                   type ? (uint32_t)ws : (uint32_t)(len - we - 1),
                   type ? (uint32_t)we : (uint32_t)(len - ws - 1), (uint32_t)len, (uint32_t)w,
                   ws, we - ws + 1, &final_pos, 0};
      if (we - ws >= 2) {
        if (REGS) vh_run(P.ix, WR, ws, we, s, 0u, vb, S);
        else vh_run(P.ix, VhBytes{P.reads + off}, ws, we, s, 0u, vb, S);
      }
      fmi_inexact_region *out = S.out;
      const uint32_t m = S.m;
      // dedup within 10 bp (bwt.c:1593-1615), the mark kept in pad[0]
      for (int32_t a1 = (int32_t)m - 1; a1 >= 1; --a1) {
        if (out[a1].pad[0]) continue;
        const uint64_t s1 = out[a1].start;
        const uint32_t c1 = out[a1].chrom_idx;
        for (int32_t a2 = a1 - 1; a2 >= 0; --a2) {
          const int32_t d = (int32_t)(uint32_t)(s1 - out[a2].start);
          const int64_t dist = d < 0 ? -(int64_t)d : (int64_t)d;
          if (c1 == out[a2].chrom_idx && dist < 10 && !out[a2].pad[0]) out[a2].pad[0] = 1;
        }
      }
      // kept regions, handed over from the last to the first (bwt.c:1616-1626)
      uint32_t kept = 0;
      for (uint32_t i = 0; i < m; ++i) {
        if (out[i].pad[0]) continue;
        const fmi_inexact_region R = out[i];
        out[kept++] = R;
      }
      for (uint32_t i = 0; i < kept / 2; ++i) {
        const fmi_inexact_region x = out[i];
        out[i] = out[kept - 1 - i];
        out[kept - 1 - i] = x;
      }
      P.nreg[slot] = kept;
      P.status[slot] = S.st;
    }
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
    if (seed_len <= 32u) seed_search_kernel<true><<<grid_for(nseeds), 256, 0, s>>>(P);
    else seed_search_kernel<false><<<grid_for(nseeds), 256, 0, s>>>(P);
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

int map_inexact_seeds(const fmi_index *ix, const uint8_t *reads, const uint64_t *offsets,
                      uint32_t nreads, uint32_t seed_size, uint32_t min_seed_size,
                      uint32_t rec_cap, uint32_t region_cap, uint64_t windows_cap,
                      fmi_inexact_region *regions, uint32_t *nregions, uint8_t *status,
                      uint64_t *nwindows, void *stream, std::string *err) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = FMI_OK;
  uint64_t *cnt = nullptr, *base = nullptr;
  uint2 *vscr = nullptr;
  void *tmp = nullptr;
  size_t need = 0;
  uint64_t nwin = 0;
  *nwindows = 0;
  SEED_TRY(hipMalloc(&cnt, (uint64_t)(nreads + 1) * 8));
  SEED_TRY(hipMalloc(&base, (uint64_t)(nreads + 1) * 8));
  win_count_kernel<<<grid_for(nreads + 1), 256, 0, s>>>(offsets, nreads, seed_size,
                                                        min_seed_size, cnt);
  SEED_TRY(hipGetLastError());
  SEED_TRY(rocprim::exclusive_scan(nullptr, need, cnt, base, (uint64_t)0, (uint64_t)nreads + 1,
                                   rocprim::plus<uint64_t>(), s));
  SEED_TRY(hipMalloc(&tmp, need));
  SEED_TRY(rocprim::exclusive_scan(tmp, need, cnt, base, (uint64_t)0, (uint64_t)nreads + 1,
                                   rocprim::plus<uint64_t>(), s));
  SEED_TRY(hipMemcpyAsync(&nwin, base + nreads, 8, hipMemcpyDeviceToHost, s));
  SEED_TRY(hipStreamSynchronize(s));
  *nwindows = nwin;
  if (nwin > windows_cap) {
    *err = "the batch has " + std::to_string(nwin) + " windows, more than windows_cap";
    rc = FMI_ERANGE;
    goto done;
  }
  if (nwin == 0) goto done;
  {
    uint64_t lanes = (nwin + 255) / 256 * 256;
    const uint64_t per = (uint64_t)(seed_size + 3) * sizeof(uint2);
    const uint64_t max_lanes = std::min<uint64_t>(1ull << 19, ((1ull << 30) / per) / 256 * 256);
    if (lanes > max_lanes) lanes = max_lanes;
    SEED_TRY(hipMalloc(&vscr, lanes * per));
    InexactParams P{ix->dev, reads, offsets, base, nwin, nreads, seed_size, rec_cap, region_cap,
                    ix->d_koffset, ix->d_kstart, (uint32_t)ix->kstart.size(), vscr, regions,
                    nregions, status};
    // windows are at most seed_size bases (the remainder window is shorter)
    if (seed_size <= 32u) inexact_seed_kernel<true><<<dim3((uint32_t)(lanes / 256)), 256, 0, s>>>(P);
    else inexact_seed_kernel<false><<<dim3((uint32_t)(lanes / 256)), 256, 0, s>>>(P);
    SEED_TRY(hipGetLastError());
  }
done:
  if (rc == FMI_OK) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) { *err = hipGetErrorString(e); rc = FMI_EDEVICE; }
  }
  (void)hipFree(cnt);
  (void)hipFree(base);
  (void)hipFree(vscr);
  (void)hipFree(tmp);
  return rc;
}

}  // namespace fmi

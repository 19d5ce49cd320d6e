// fmi_mapread.hip -- the read-level 1-error mapping of __bwt_map_inexact_read
// (bwt.c:1890-2236) up to its list of alignments, per read on the device.
//
// Only the error-free records of the read's 1-error search become
// alignments (bwt.c:2038), and the vector-driven search yields at most one
// per strand: the exact interval of the whole read on the strand pair's
// first index (its backward vector's entry 0, search.c:1505-1508) -- here a
// plain backward search on BACKWARD (+ strand) or FORWARD_REV (- strand).
// The running count of l - k over them is held to filter_read_mappings
// before each strand's rows (bwt.c:2145-2156); each SA row inside a contig
// gives an alignment at its 0-based coordinate (bwt.c:2158-2166); each
// strand's alignments are deduplicated within 10 bp on the reference's int
// positions and handed over from last to first with the secondary flag
// (bwt.c:2190-2233).
//
// Lanes stride over the reads; a lane keeps the current strand's alignments
// in its scratch slice (at most filter + 1 of them: a strand's rows number
// l - k + 1 <= filter + 1 once the count has passed) for the dedup.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/fmi.h"
#include "fmi_device.h"
#include "fmi_internal.h"
#include "fmi_kernels.h"

namespace fmi {

namespace {

struct MapTmp {
  uint64_t pos;
  uint32_t chrom;
  uint32_t del;
};

struct MapParams {
  DevIndex ix;
  const uint8_t *reads;
  const uint64_t *offsets;
  const uint64_t *koff;       // kcount + 1 contig offsets
  const uint64_t *kstart;     // kcount contig starts
  uint32_t kcount, nreads, filter, cap;
  fmi_read_mapping *out;
  uint32_t *counts;
  uint8_t *status;
  MapTmp *scratch;            // lanes * (filter + 1)
};

// BWExactSearchBackward of the whole read from [0, n] on one view.
__device__ void exact_interval(const DevDir &D, uint32_t comp, uint32_t n, const uint8_t *W,
                               int32_t len, uint32_t &k, uint32_t &l) {
  k = 0u;
  l = n;
  for (int32_t i = len - 1; i >= 0; --i) {
    uint32_t a, b;
    if (lf_read(D, k, l, W + i, comp, a, b) > 3u) { k = n + 1u; l = n; return; }
    k = a;
    l = b;
    if (k > l) return;
  }
}

__global__ __launch_bounds__(256) void map_reads_kernel(MapParams P) {
  const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lanes = gridDim.x * blockDim.x;
  MapTmp *tmp = P.scratch + (uint64_t)lane * (P.filter + 1u);
  const uint32_t n = P.ix.n;
  for (uint32_t rd = lane; rd < P.nreads; rd += lanes) {
    const uint64_t off = P.offsets[rd];
    const uint64_t len64 = P.offsets[rd + 1] - off;
    const uint8_t *W = P.reads + off;
    fmi_read_mapping *out = P.out + (uint64_t)rd * P.cap;
    if (len64 < 5 || len64 > kMaxReadLen) {
      P.counts[rd] = 0;
      P.status[rd] = len64 < 5 ? 2u : 4u;
      continue;
    }
    const int32_t len = (int32_t)len64;
    uint64_t tot = 0;
    uint32_t nout = 0;
    uint8_t st = 0;
    for (int type = 1; type >= 0; --type) {
      // + strand: BACKWARD; - strand: FORWARD_REV (bwt.c:1996-2001)
      const DevDir &D = type ? P.ix.d[0] : P.ix.d[1];
      uint32_t k, l;
      exact_interval(D, type ? 0u : 1u, n, W, len, k, l);
      uint32_t m = 0;
      if (k <= l) {
        tot += (uint64_t)(l - k);
        if (tot > P.filter) { st |= 1u; break; }
        for (uint32_t j = k; j <= l; ++j) {
          // direction: the record's dir is 0; + keeps it, - flips it
          const uint64_t sa = sa_value(D, P.ix.ratio, n, j);
          const uint64_t key = type ? sa : (uint64_t)n + 1u - sa - (uint64_t)len - 1u;
          uint32_t lo = 0, cnt = P.kcount;   // upper_bound of key in koff[0 .. kcount)
          while (cnt > 0) {
            const uint32_t w = cnt / 2;
            if (P.koff[lo + w] <= key) { lo += w + 1; cnt -= w + 1; }
            else cnt = w;
          }
          if (lo == 0 || key + (uint64_t)len > P.koff[lo]) continue;
          tmp[m].pos = P.kstart[lo - 1] + (key - P.koff[lo - 1]);
          tmp[m].chrom = lo - 1;
          tmp[m].del = 0u;
          ++m;
        }
      }
      // within 10 bp on the int positions: a later alignment goes when an
      // earlier one (all carry one CIGAR operation) sits within 10
      for (int32_t a1 = (int32_t)m - 1; a1 >= 1; --a1) {
        const int32_t p1 = (int32_t)(uint32_t)tmp[a1].pos;
        for (int32_t a2 = a1 - 1; a2 >= 0; --a2) {
          const int32_t d = (int32_t)((uint32_t)p1 - (uint32_t)tmp[a2].pos);
          const int64_t ad = d < 0 ? -(int64_t)d : (int64_t)d;
          if (tmp[a1].chrom == tmp[a2].chrom && ad < 10) { tmp[a1].del = 1u; break; }
        }
      }
      const bool created_primary = nout == 0;
      bool primary_delete = false;
      for (int32_t a = (int32_t)m - 1; a >= 0; --a) {
        if (tmp[a].del) {
          primary_delete |= created_primary;
          continue;
        }
        if (nout < P.cap) {
          fmi_read_mapping &o = out[nout];
          o.position = tmp[a].pos;
          o.chromosome = tmp[a].chrom;
          o.read = rd;
          o.strand = (uint8_t)(type ? 0 : 1);
          o.secondary = (uint8_t)(nout > 0 ? 1 : 0);
          o.cigar_len = (uint16_t)len;
          o.pad = 0u;
        }
        ++nout;
      }
      if (primary_delete && nout > 0 && P.cap > 0) out[0].secondary = 0;
    }
    P.counts[rd] = nout;
    P.status[rd] = st;
  }
}

}  // namespace

uint32_t map_reads_lanes(uint32_t nreads, uint32_t filter) {
  const uint64_t per = (uint64_t)(filter + 1) * sizeof(MapTmp);
  uint64_t lanes = ((uint64_t)nreads + 255) / 256 * 256;
  uint64_t most = ((1ull << 30) / per) / 256 * 256;
  if (most < 256) most = 256;
  if (lanes > most) lanes = most;
  if (lanes > 64 * 1024) lanes = 64 * 1024;
  return (uint32_t)lanes;
}

uint64_t map_reads_scratch_bytes(uint32_t lanes, uint32_t filter) {
  return (uint64_t)lanes * (filter + 1) * sizeof(MapTmp);
}

hipError_t launch_map_reads(const DevIndex &ix, const uint8_t *reads, const uint64_t *offsets,
                            uint32_t nreads, const uint64_t *koff, const uint64_t *kstart,
                            uint32_t kcount, uint32_t filter, uint32_t cap, fmi_read_mapping *out,
                            uint32_t *counts, uint8_t *status, void *scratch, uint32_t lanes,
                            hipStream_t s) {
  if (nreads == 0 || lanes == 0) return hipSuccess;
  MapParams P{ix,     reads,  offsets, koff,   kstart, kcount, nreads, filter, cap,
              out,    counts, status,  static_cast<MapTmp *>(scratch)};
  hipLaunchKernelGGL(map_reads_kernel, dim3(lanes / 256), dim3(256), 0, s, P);
  return hipGetLastError();
}

}  // namespace fmi

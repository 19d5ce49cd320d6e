// fmi_vhelper.h -- BWSearch1VectorHelper (search/search.c:1491-1709) for one
// window of one read on one strand, shared by the vector-driven 1-error
// search (fmi_search1_vector_batch) and the inexact seed mapping
// (fmi_map_inexact_seeds_batch).
//
// The helper walks the backward vector from the left end of the window and
// the forward vector from the right end, the reverse of the order the
// vectors are built in, so the lane first writes the parts it reads
// (entries 0..n/2 of the backward vector and n/2-1..n of the forward one,
// n + 3 intervals) to a scratch slice.  One record state is carried from
// record to record, as the reference reuses one result: each step assigns
// only the fields the reference assigns, so the error fields a record
// inherits (a mismatch after the first block's insertion keeps position b)
// are the same.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/fmi.h"
#include "fmi_device.h"
#include "fmi_internal.h"

namespace fmi {

struct VhView {
  const DevDir *D;
  uint32_t comp, n;
  __device__ __forceinline__ void lf(uint32_t k, uint32_t l, uint32_t b, uint32_t &ko,
                                     uint32_t &lo) const {
    if (b > 3u) { ko = n + 1u; lo = n; return; }   // N never occurs
    uint32_t nb;
    lf_one(*D, k, l, comp ? 3u - b : b, ko, lo, nb);
  }
  // lf for b = 0..3 on one interval from one read of its block(s):
  // ko[b], lo[b] as lf(k, l, b) gives them, for any k, l.
  __device__ __forceinline__ void lf4(uint32_t k, uint32_t l, uint32_t (&ko)[4],
                                      uint32_t (&lo)[4]) const {
    uint32_t pk[4], pl[4];
    lf_four(*D, k, l, pk, pl);
#pragma unroll
    for (uint32_t b = 0; b < 4u; ++b) {
      const uint32_t s = comp ? 3u - b : b;
      ko[b] = pick4(pk, s);
      lo[b] = pick4(pl, s);
    }
  }
  // The same step for a read base (0-3, or > 3 for an N) with the block(s)
  // read in one batch (lf_sym computes what lf_one does, for any k, l).
  __device__ __forceinline__ void lfs(uint32_t k, uint32_t l, uint32_t sym, uint32_t &ko,
                                      uint32_t &lo) const {
    uint32_t a, b;
    if (lf_sym(*D, k, l, sym, comp, a, b) > 3u) { ko = n + 1u; lo = n; return; }
    ko = a;
    lo = b;
  }
};

// Where the helper reads the window's bases: the read's bytes in memory
// (VhBytes), or a window of at most 32 bases held in registers as 2-bit codes
// and N flags (VhRegs, from vh_regs).  at(i) is W[i]: 0-3, or > 3 for an N.
struct VhBytes {
  const uint8_t *W;
  __device__ __forceinline__ uint32_t at(int32_t i) const { return W[i]; }
};

struct VhRegs {
  uint64_t code;                   // base s + j at bits 2j
  uint32_t nfl;                    // bit j: base s + j is an N
  int32_t s;
  __device__ __forceinline__ uint32_t at(int32_t i) const {
    const uint32_t j = (uint32_t)(i - s);
    return ((nfl >> j) & 1u) ? 4u : (uint32_t)(code >> (2u * j)) & 3u;
  }
};

// W[s .. s + len - 1] (1 <= len <= 32) packed into registers, read from the
// 16-byte chunks that hold it (none past the chunk of its last byte).
__device__ __forceinline__ VhRegs vh_regs(const uint8_t *W, int32_t s, uint32_t len) {
  const uint8_t *a = W + s;
  const uint32_t mis = (uint32_t)(reinterpret_cast<uintptr_t>(a) & 15u);
  const uint4 *c = reinterpret_cast<const uint4 *>(a - mis);
  const uint32_t last = (mis + len - 1u) >> 4;            // 0 .. 2
  uint32_t k0, k1 = 0u, k2 = 0u, f0, f1 = 0u, f2 = 0u;
  pack16(c[0], k0, f0);
  if (last >= 1u) pack16(c[1], k1, f1);
  if (last >= 2u) pack16(c[2], k2, f2);
  const uint64_t lo = ((uint64_t)k1 << 32) | k0;
  const uint32_t sh = 2u * mis;
  VhRegs R;
  R.code = sh ? (lo >> sh) | ((uint64_t)k2 << (64u - sh)) : lo;
  R.nfl = (uint32_t)(((uint64_t)f0 | ((uint64_t)f1 << 16) | ((uint64_t)f2 << 32)) >> mis);
  R.s = s;
  return R;
}

struct VhRec {
  uint32_t k, l;
  int32_t pos, err_pos;
  uint32_t dir, nmm, kind, base;
};

struct VhSink {
  fmi_hit *out;
  uint32_t cap, cnt, mm_only, strand;
  int16_t start, end;
  __device__ __forceinline__ void add(const VhRec &r) {
    if (mm_only && r.nmm && r.kind != FMI_MISMATCH) return;
    if (cnt < cap)
      store_hit(out + cnt, (int64_t)r.k, (int64_t)r.l, start, (int16_t)r.pos, end,
                r.nmm ? (int16_t)r.err_pos : (int16_t)0, (uint8_t)r.dir, (uint8_t)r.nmm,
                r.nmm ? (uint8_t)r.kind : (uint8_t)0, r.nmm ? (uint8_t)r.base : (uint8_t)0,
                (uint8_t)strand);
    ++cnt;
  }
};

// BWExactSearchBackward (dir 0) / Forward (dir 1) on the record state.
template <class Src>
__device__ __forceinline__ void vh_exact(const VhView &V, const Src &W, int32_t start,
                                         int32_t end, uint32_t dir, VhRec &r) {
  uint32_t k2 = r.k, l2 = r.l;
  int32_t i = r.pos;
  if (dir == 0) {
    for (; i >= start; --i) {
      V.lfs(k2, l2, W.at(i), k2, l2);
      if (k2 > l2) break;
    }
  } else {
    for (; i <= end; ++i) {
      V.lfs(k2, l2, W.at(i), k2, l2);
      if (k2 > l2) break;
    }
  }
  r.k = k2; r.l = l2; r.pos = i;
}

template <class Src, class Sink>
__device__ inline void vh_replay(const DevIndex &ix, const Src &W, int32_t start, int32_t end,
                                 uint32_t strand, uint32_t mm_only, const uint2 *vb, Sink &S);

// The helper over W[start..end] (end - start >= 2) on strand 0 (backward,
// forward) or 1 (forward_rev, backward_rev); vb: n + 3 intervals of scratch.
// Every record goes to S.add(r) in the reference's order.  W: VhBytes or
// VhRegs (the bases at start..end).
template <class Src, class Sink>
__device__ inline void vh_run(const DevIndex &ix, const Src &W, int32_t start, int32_t end,
                              uint32_t strand, uint32_t mm_only, uint2 *vb, Sink &S) {
  const uint32_t n_sa = ix.n;
  // + strand: backward / forward; - strand: forward_rev / backward_rev
  const VhView B{strand ? &ix.d[1] : &ix.d[0], strand, n_sa};
  const VhView F{strand ? &ix.d[0] : &ix.d[1], strand, n_sa};
  const int32_t n = end - start;
  int32_t half = n / 2;
  uint2 *vf = vb + (half + 1);     // forward vector entries half-1 .. n

  // BWExactSearchVectorBackward over W[start..end]: entry j = W[start+j..end]
  {
    uint32_t k = 0, l = n_sa;
    bool dead = false;
    for (int32_t j = n; j >= 0; --j) {
      if (!dead) {
        B.lfs(k, l, W.at(start + j), k, l);
        dead = k > l;
      }
      if (j <= half) vb[j] = make_uint2(k, l);
    }
  }
  // BWExactSearchVectorForward: entry j = W[start..start+j]
  {
    uint32_t k = 0, l = n_sa;
    bool dead = false;
    for (int32_t j = 0; j <= n; ++j) {
      if (!dead) {
        F.lfs(k, l, W.at(start + j), k, l);
        dead = k > l;
      }
      if (j >= half - 1) vf[j - (half - 1)] = make_uint2(k, l);
    }
  }
  vh_replay(ix, W, start, end, strand, mm_only, vb, S);
}

// The helper's record sequence from the vectors in vb (entries 0..n/2 of the
// backward vector, then entries n/2-1..n of the forward vector).
template <class Src, class Sink>
__device__ inline void vh_replay(const DevIndex &ix, const Src &W, int32_t start, int32_t end,
                                 uint32_t strand, uint32_t mm_only, const uint2 *vb, Sink &S) {
  const uint32_t n_sa = ix.n;
  const VhView B{strand ? &ix.d[1] : &ix.d[0], strand, n_sa};
  const VhView F{strand ? &ix.d[0] : &ix.d[1], strand, n_sa};
  const int32_t n = end - start;
  int32_t half = n / 2;
  const uint2 *vf = vb + (half + 1);

  VhRec r{0, 0, -1, 0, 0, 0, 0, 0};
  uint2 v = vb[0];
  if (v.x <= v.y) { r.k = v.x; r.l = v.y; r.pos = -1; S.add(r); }
  r.nmm = 1; r.kind = FMI_MATCH; r.base = 0xFFu; r.err_pos = start;

  uint32_t k, l, ka, la;
  int64_t size = (int64_t)v.y - (int64_t)v.x, size_last;
  // an error at W[start]
  size_last = size;
  v = vb[1];
  k = v.x; l = v.y;
  size = (int64_t)l - (int64_t)k;
  if (size != size_last) {
    r.k = k; r.l = l; r.pos = -1;
    r.kind = FMI_DELETION; r.base = 0xFFu; r.err_pos = start;
    S.add(r);
    const uint32_t bw = W.at(start);
    uint32_t kb4[4], lb4[4];
    B.lf4(k, l, kb4, lb4);
    for (uint32_t b = 0; b < 4u; ++b) {
      ka = pick4(kb4, b); la = pick4(lb4, b);
      if (ka > la) continue;
      if (b != bw) {
        r.k = ka; r.l = la; r.pos = -1;
        r.kind = FMI_MISMATCH; r.base = b;
        S.add(r);
      }
      B.lf(ka, la, bw, ka, la);
      if (ka <= la) {
        // search.c:1556 stores kind 3, base INSERTION (3), position b
        r.k = ka; r.l = la; r.pos = -1;
        r.kind = 3u; r.base = FMI_INSERTION; r.err_pos = (int32_t)b;
        S.add(r);
      }
    }
  }
  for (int32_t i = start + 2, j = 2; j <= half; ++i, ++j) {
    size_last = size;
    v = vb[j];
    k = v.x; l = v.y;
    size = (int64_t)l - (int64_t)k;
    if (size == size_last) continue;
    r.k = k; r.l = l; r.pos = i - 2;
    r.kind = FMI_DELETION; r.base = 0xFFu; r.err_pos = i - 1;
    vh_exact(B, W, start, end, 0, r);
    if (r.k <= r.l) S.add(r);
    uint32_t kb4[4], lb4[4];
    B.lf4(k, l, kb4, lb4);
    for (uint32_t b = 0; b < 4u; ++b) {
      ka = pick4(kb4, b); la = pick4(lb4, b);
      if (ka > la) continue;
      if (!mm_only) {
        r.k = ka; r.l = la; r.pos = i - 1;
        r.kind = FMI_INSERTION; r.base = b;
        vh_exact(B, W, start, end, 0, r);
        if (r.k <= r.l) S.add(r);
      } else {
        r.kind = FMI_INSERTION; r.base = b;
      }
      if (b != W.at(i - 1)) {
        r.k = ka; r.l = la; r.pos = i - 2;
        r.kind = FMI_MISMATCH;
        vh_exact(B, W, start, end, 0, r);
        if (r.k <= r.l) S.add(r);
      }
    }
  }

  half -= 1;
  v = vf[n - half];                // entry n
  size = (int64_t)v.y - (int64_t)v.x;
  r.dir = 1;
  // an error at W[end]
  size_last = size;
  v = vf[n - 1 - half];            // entry n - 1
  k = v.x; l = v.y;
  size = (int64_t)l - (int64_t)k;
  if (size != size_last) {
    r.k = k; r.l = l; r.pos = -1;
    r.kind = FMI_DELETION; r.base = 0xFFu; r.err_pos = end;
    S.add(r);
    const uint32_t bw = W.at(end);
    uint32_t kb4[4], lb4[4];
    F.lf4(k, l, kb4, lb4);
    for (uint32_t b = 0; b < 4u; ++b) {
      ka = pick4(kb4, b); la = pick4(lb4, b);
      if (ka > la) continue;
      if (b != bw) {
        r.k = ka; r.l = la; r.pos = -1;
        r.kind = FMI_MISMATCH; r.base = b;
        S.add(r);
      }
      F.lf(ka, la, bw, ka, la);
      if (ka <= la) {
        r.k = ka; r.l = la; r.pos = -1;
        r.kind = FMI_INSERTION; r.base = b;
        S.add(r);
      }
    }
  }
  for (int32_t i = end - 2, j = n - 2; j >= half; --i, --j) {
    size_last = size;
    v = vf[j - half];
    k = v.x; l = v.y;
    size = (int64_t)l - (int64_t)k;
    if (size == size_last) continue;
    r.k = k; r.l = l; r.pos = i + 2;
    r.kind = FMI_DELETION; r.base = 0xFFu; r.err_pos = i + 1;
    vh_exact(F, W, start, end, 1, r);
    if (r.k <= r.l) S.add(r);
    uint32_t kb4[4], lb4[4];
    F.lf4(k, l, kb4, lb4);
    for (uint32_t b = 0; b < 4u; ++b) {
      ka = pick4(kb4, b); la = pick4(lb4, b);
      if (ka > la) continue;
      if (!mm_only) {
        r.k = ka; r.l = la; r.pos = i + 1;
        r.kind = FMI_INSERTION; r.base = b;
        vh_exact(F, W, start, end, 1, r);
        if (r.k <= r.l) S.add(r);
      } else {
        r.kind = FMI_INSERTION; r.base = b;
      }
      if (b != W.at(i + 1)) {
        r.k = ka; r.l = la; r.pos = i + 2;
        r.kind = FMI_MISMATCH;
        vh_exact(F, W, start, end, 1, r);
        if (r.k <= r.l) S.add(r);
      }
    }
  }
}

// vh_run storing at most cap records at out; returns the count of all of them.
__device__ inline uint32_t vh_task(const DevIndex &ix, const uint8_t *W, int32_t start,
                                   int32_t end, uint32_t strand, uint32_t mm_only, uint2 *vb,
                                   fmi_hit *out, uint32_t cap) {
  VhSink S{out, cap, 0u, mm_only, strand, (int16_t)start, (int16_t)end};
  vh_run(ix, VhBytes{W}, start, end, strand, mm_only, vb, S);
  return S.cnt;
}

}  // namespace fmi

// fmi_device.h -- device-side building blocks shared by the HIP kernels:
// occurrence-block loads, ranks, the LF step, the locate walk, record stores.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/fmi.h"
#include "fmi_internal.h"

namespace fmi {

// --------------------------------------------------------------------------
// occurrence block access
// --------------------------------------------------------------------------

struct Blk {
  uint32_t c[4];
  uint64_t lo[3], hi[3];   // bit-planes: bit t%64 of word t/64 = code bit of row t
};

__device__ __forceinline__ Blk load_block(const OccBlock *base, uint32_t j) {
  const uint4 *q = reinterpret_cast<const uint4 *>(base + j);
  const uint4 a = q[0], b = q[1], c = q[2], d = q[3];
  Blk B;
  B.c[0] = a.x; B.c[1] = a.y; B.c[2] = a.z; B.c[3] = a.w;
  B.lo[0] = ((uint64_t)b.y << 32) | b.x;
  B.lo[1] = ((uint64_t)b.w << 32) | b.z;
  B.lo[2] = ((uint64_t)c.y << 32) | c.x;
  B.hi[0] = ((uint64_t)c.w << 32) | c.z;
  B.hi[1] = ((uint64_t)d.y << 32) | d.x;
  B.hi[2] = ((uint64_t)d.w << 32) | d.z;
  return B;
}

// The block held by four 16-byte loads of it.
__device__ __forceinline__ Blk blk_of4(const uint4 &a, const uint4 &b, const uint4 &c,
                                       const uint4 &d) {
  Blk B;
  B.c[0] = a.x; B.c[1] = a.y; B.c[2] = a.z; B.c[3] = a.w;
  B.lo[0] = ((uint64_t)b.y << 32) | b.x;
  B.lo[1] = ((uint64_t)b.w << 32) | b.z;
  B.lo[2] = ((uint64_t)c.y << 32) | c.x;
  B.hi[0] = ((uint64_t)c.w << 32) | c.z;
  B.hi[1] = ((uint64_t)d.y << 32) | d.x;
  B.hi[2] = ((uint64_t)d.w << 32) | d.z;
  return B;
}

// Word i (0..3) of a 16-byte load.
__device__ __forceinline__ uint32_t pick4u(const uint4 &v, uint32_t i) {
  const uint32_t lo = (i & 1u) ? v.y : v.x, hi = (i & 1u) ? v.w : v.z;
  return (i & 2u) ? hi : lo;
}

// Masks selecting the first t (0..192) rows of the block in each plane word.
__device__ __forceinline__ void row_masks(uint32_t t, uint64_t &m0, uint64_t &m1, uint64_t &m2) {
  const uint32_t r = t & 63u, q = t >> 6;
  const uint64_t pm = r ? (~0ull >> (64u - r)) : 0ull;
  m0 = q >= 1u ? ~0ull : pm;
  m1 = q >= 2u ? ~0ull : (q == 1u ? pm : 0ull);
  m2 = q >= 3u ? ~0ull : (q == 2u ? pm : 0ull);
}

// Rows among the first t of the block whose code is s (the '$' row holds 0).
__device__ __forceinline__ uint32_t rank1(const Blk &B, uint32_t s, uint32_t t) {
  uint64_t m0, m1, m2;
  row_masks(t, m0, m1, m2);
  const uint64_t xl = (s & 1u) ? 0ull : ~0ull, xh = (s & 2u) ? 0ull : ~0ull;
  return (uint32_t)(__popcll((B.lo[0] ^ xl) & (B.hi[0] ^ xh) & m0) +
                    __popcll((B.lo[1] ^ xl) & (B.hi[1] ^ xh) & m1) +
                    __popcll((B.lo[2] ^ xl) & (B.hi[2] ^ xh) & m2));
}

// The same for all four codes at once: three popcounts per plane word.
__device__ __forceinline__ void rank4(const Blk &B, uint32_t t, uint32_t (&r)[4]) {
  uint64_t m0, m1, m2;
  row_masks(t, m0, m1, m2);
  const uint64_t a0 = B.lo[0] & m0, a1 = B.lo[1] & m1, a2 = B.lo[2] & m2;
  const uint64_t b0 = B.hi[0] & m0, b1 = B.hi[1] & m1, b2 = B.hi[2] & m2;
  const uint32_t nl = (uint32_t)(__popcll(a0) + __popcll(a1) + __popcll(a2));
  const uint32_t nh = (uint32_t)(__popcll(b0) + __popcll(b1) + __popcll(b2));
  const uint32_t nb = (uint32_t)(__popcll(a0 & b0) + __popcll(a1 & b1) + __popcll(a2 & b2));
  r[3] = nb;
  r[2] = nh - nb;
  r[1] = nl - nb;
  r[0] = t - nl - nh + nb;
}

// a[i] for a runtime i in 0..3 without dynamic register indexing.
__device__ __forceinline__ uint32_t pick4(const uint32_t (&a)[4], uint32_t i) {
  const uint32_t lo = (i & 1u) ? a[1] : a[0];
  const uint32_t hi = (i & 1u) ? a[3] : a[2];
  return (i & 2u) ? hi : lo;
}

__device__ __forceinline__ uint64_t pick3(const uint64_t (&a)[3], uint32_t i) {
  return i == 0u ? a[0] : (i == 1u ? a[1] : a[2]);
}

// Code stored at row t of the block.
__device__ __forceinline__ uint32_t code_at(const Blk &B, uint32_t t) {
  const uint32_t w = t >> 6, b = t & 63u;
  return (uint32_t)((pick3(B.lo, w) >> b) & 1ull) | ((uint32_t)((pick3(B.hi, w) >> b) & 1ull) << 1);
}

// Occ(s, m) for m in block j (t = m - 192*j): checkpoint + in-block rank,
// minus the '$' row (stored as code 0) when it lies before m in the block.
__device__ __forceinline__ uint32_t occ_blk(const Blk &B, uint32_t dollar_blk,
                                            uint32_t dollar_off, uint32_t j, uint32_t s,
                                            uint32_t t) {
  uint32_t v = pick4(B.c, s) + rank1(B, s, t);
  if (s == 0 && j == dollar_blk && dollar_off < t) v -= 1;
  return v;
}

// Symbol of row k (block j, offset t), 4 for the '$' row.
__device__ __forceinline__ uint32_t sym_at(const Blk &B, uint32_t dollar_blk, uint32_t dollar_off,
                                           uint32_t j, uint32_t t) {
  return (j == dollar_blk && t == dollar_off) ? 4u : code_at(B, t);
}

// LF step of one physical symbol s on [k, l] (BWiteration, runtime.h:38-42).
// A size-1 interval needs only row k's block: Occ(s, k+1) = Occ(s, k) +
// [BWT[k] == s].
__device__ __forceinline__ void lf_one(const DevDir &D, uint32_t k, uint32_t l, uint32_t s,
                                       uint32_t &ko, uint32_t &lo, uint32_t &nblk) {
  const uint32_t m2 = l + 1;
  const uint32_t j1 = k / kBlockRows, j2 = m2 / kBlockRows;
  const uint32_t t1 = k - j1 * kBlockRows;
  const Blk B1 = load_block(D.blk, j1);
  const uint32_t c = pick4(D.C, s);
  ko = c + 1u + occ_blk(B1, D.dollar_blk, D.dollar_off, j1, s, t1);
  if (k == l) {
    nblk = 1u;
    lo = ko - 1u + (sym_at(B1, D.dollar_blk, D.dollar_off, j1, t1) == s ? 1u : 0u);
    return;
  }
  Blk B2;
  if (j2 != j1) B2 = load_block(D.blk, j2);
  else B2 = B1;
  nblk = (j2 != j1) ? 2u : 1u;
  lo = c + occ_blk(B2, D.dollar_blk, D.dollar_off, j2, s, m2 - j2 * kBlockRows);
}

// The LF step of all four physical symbols on [k, l] at once (one or two
// block loads): ko[s], lo[s] as lf_one gives them for s.
__device__ __forceinline__ void lf_four(const DevDir &D, uint32_t k, uint32_t l,
                                        uint32_t (&ko)[4], uint32_t (&lo)[4]) {
  const uint32_t m2 = l + 1;
  const uint32_t j1 = k / kBlockRows, j2 = m2 / kBlockRows;
  const uint32_t t1 = k - j1 * kBlockRows, t2 = m2 - j2 * kBlockRows;
  const Blk B1 = load_block(D.blk, j1);
  uint32_t r1[4];
  rank4(B1, t1, r1);
  r1[0] -= (j1 == D.dollar_blk && D.dollar_off < t1) ? 1u : 0u;
#pragma unroll
  for (uint32_t s = 0; s < 4u; ++s) ko[s] = pick4(D.C, s) + 1u + B1.c[s] + r1[s];
  if (k == l) {
    const uint32_t srow = sym_at(B1, D.dollar_blk, D.dollar_off, j1, t1);
#pragma unroll
    for (uint32_t s = 0; s < 4u; ++s) lo[s] = ko[s] - 1u + (srow == s ? 1u : 0u);
    return;
  }
  const Blk B2 = j2 != j1 ? load_block(D.blk, j2) : B1;
  uint32_t r2[4];
  rank4(B2, t2, r2);
  r2[0] -= (j2 == D.dollar_blk && D.dollar_off < t2) ? 1u : 0u;
#pragma unroll
  for (uint32_t s = 0; s < 4u; ++s) lo[s] = pick4(D.C, s) + B2.c[s] + r2[s];
}

// A value the compiler cannot see (no instruction is emitted): the
// register is taken as written here.
__device__ __forceinline__ uint4 opaque4() {
  uint32_t a, b, c, d;
  asm volatile("" : "=v"(a), "=v"(b), "=v"(c), "=v"(d));
  return make_uint4(a, b, c, d);
}

// The LF step of base sym (a read byte: 0-3, or > 3 for an N) on [k, l]:
// the block(s) are read in one batch.  Returns sym; for sym <= 3 ko, lo are
// the step of the physical symbol (comp ? 3 - sym : sym) exactly as lf_one
// gives them (for any k, l); for an N the caller empties the interval.
__device__ __forceinline__ uint32_t lf_sym(const DevDir &D, uint32_t k, uint32_t l,
                                           uint32_t sym, uint32_t comp, uint32_t &ko,
                                           uint32_t &lo) {
  const uint32_t m2 = l + 1;
  const uint32_t j1 = k / kBlockRows, j2 = m2 / kBlockRows;
  const uint32_t t1 = k - j1 * kBlockRows, t2 = m2 - j2 * kBlockRows;
  const bool two = k != l && j2 != j1;
  const uint4 *q1 = reinterpret_cast<const uint4 *>(D.blk + j1);
  const uint4 a0 = q1[0], a1 = q1[1], a2 = q1[2], a3 = q1[3];
  uint4 b0 = opaque4(), b1 = opaque4(), b2 = opaque4(), b3 = opaque4();
  if (two) {
    const uint4 *q2 = reinterpret_cast<const uint4 *>(D.blk + j2);
    b0 = q2[0]; b1 = q2[1]; b2 = q2[2]; b3 = q2[3];
  }
  if (sym > 3u) return sym;
  const Blk B1 = blk_of4(a0, a1, a2, a3);
  const uint32_t s = comp ? 3u - sym : sym;
  const uint32_t c = pick4(D.C, s);
  ko = c + 1u + occ_blk(B1, D.dollar_blk, D.dollar_off, j1, s, t1);
  if (k == l) {
    lo = ko - 1u + (sym_at(B1, D.dollar_blk, D.dollar_off, j1, t1) == s ? 1u : 0u);
  } else {
    const Blk B2 = two ? blk_of4(b0, b1, b2, b3) : B1;
    lo = c + occ_blk(B2, D.dollar_blk, D.dollar_off, j2, s, t2);
  }
  return sym;
}

// The LF step of the read base at wp on [k, l] (lf_sym of *wp): the base's
// load is issued with the block loads, so a step waits for memory once.
__device__ __forceinline__ uint32_t lf_read(const DevDir &D, uint32_t k, uint32_t l,
                                            const uint8_t *wp, uint32_t comp, uint32_t &ko,
                                            uint32_t &lo) {
  return lf_sym(D, k, l, *wp, comp, ko, lo);
}

__device__ __forceinline__ const DevDir &dir_of(const DevIndex &ix, uint32_t d) {
  return d ? ix.d[1] : ix.d[0];
}

// --------------------------------------------------------------------------
// records and locate
// --------------------------------------------------------------------------

__device__ __forceinline__ void store_hit(fmi_hit *dst, int64_t k, int64_t l, int16_t start,
                                          int16_t pos, int16_t end, int16_t err_pos, uint8_t dir,
                                          uint8_t nmm, uint8_t kind, uint8_t base, uint8_t strand) {
  uint4 a, b;
  a.x = (uint32_t)(uint64_t)k; a.y = (uint32_t)((uint64_t)k >> 32);
  a.z = (uint32_t)(uint64_t)l; a.w = (uint32_t)((uint64_t)l >> 32);
  b.x = (uint32_t)(uint16_t)start | ((uint32_t)(uint16_t)pos << 16);
  b.y = (uint32_t)(uint16_t)end | ((uint32_t)(uint16_t)err_pos << 16);
  b.z = (uint32_t)dir | ((uint32_t)nmm << 8) | ((uint32_t)kind << 16) | ((uint32_t)base << 24);
  b.w = (uint32_t)strand;
  uint4 *d = reinterpret_cast<uint4 *>(dst);
  d[0] = a;
  d[1] = b;
}

__device__ __forceinline__ uint32_t sa_value(const DevDir &D, uint32_t ratio, uint32_t n,
                                             uint32_t row) {
  if (D.sa) return D.sa[row];      // expanded index: the full suffix array
  uint32_t i = row, steps = 0;
  while (i % ratio) {
    if (i == D.dollar) {
      i = 0;
    } else {
      const uint32_t j = i / kBlockRows, t = i - j * kBlockRows;
      const Blk B = load_block(D.blk, j);
      const uint32_t b = code_at(B, t);
      i = pick4(D.C, b) + occ_blk(B, D.dollar_blk, D.dollar_off, j, b, t) + 1u;
    }
    ++steps;
  }
  const uint64_t v = (uint64_t)D.S[i / ratio] + steps;
  return (uint32_t)(v % ((uint64_t)n + 1));
}

// --------------------------------------------------------------------------
// text mode (expanded index)
// --------------------------------------------------------------------------

// Number of leading read bases W[p], W[p+d], ... (at most lim <= 16, lim <=
// q) that equal the text bases preceding suffix q of direction dsel in LF
// order (X[q-1], X[q-2], ... for the BWT of X; X[n-q], X[n-q+1], ... for the
// BWT of reverse X), the read's bases complemented when comp is set.  The 16
// bases are compared at once as 2-bit groups; an N base never matches.
__device__ __forceinline__ uint32_t text_match(const DevIndex &ix, uint32_t dsel, uint32_t comp,
                                               uint32_t q, const uint8_t *W, int32_t p, int32_t d,
                                               uint32_t lim) {
  if (lim == 0u) return 0u;
  uint32_t r2 = 0u, bad = 0u;
#pragma unroll
  for (uint32_t i = 0; i < 16u; ++i) {
    const uint32_t sym = W[p + (int32_t)(i < lim ? i : 0u) * d];
    const uint32_t ps = (comp ? (sym ^ 3u) : sym) & 3u;
    r2 |= ps << (2u * i);
    bad |= (sym > 3u ? 1u : 0u) << (2u * i);
  }
  const uint32_t nwords = (ix.n + 15u) >> 4;
  uint32_t t2;
  if (dsel) {                        // X[x0], X[x0+1], ... with x0 = n - q
    const uint32_t x0 = ix.n - q, w = x0 >> 4;
    const uint64_t lo = ix.text[w];
    const uint64_t hi = w + 1u < nwords ? (uint64_t)ix.text[w + 1u] : 0ull;
    t2 = (uint32_t)(((hi << 32) | lo) >> ((x0 & 15u) << 1));
  } else {                           // X[x0], X[x0-1], ... with x0 = q - 1
    const int64_t b = (int64_t)q - 16;   // X[b .. b+15] = X[x0-15 .. x0]
    uint32_t win;
    if (b >= 0) {
      const uint32_t w = (uint32_t)b >> 4;
      const uint64_t lo = ix.text[w];
      const uint64_t hi = w + 1u < nwords ? (uint64_t)ix.text[w + 1u] : 0ull;
      win = (uint32_t)(((hi << 32) | lo) >> (((uint32_t)b & 15u) << 1));
    } else {
      win = (uint32_t)((uint64_t)ix.text[0] << (2u * (uint32_t)(-b)));
    }
    // reverse the 16 groups: bit reversal, then swap the bits of each pair
    const uint32_t v = __builtin_bitreverse32(win);
    t2 = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
  }
  const uint32_t x = r2 ^ t2;
  uint32_t miss = ((x | (x >> 1)) & 0x55555555u) | bad;
  if (lim < 16u) miss |= 1u << (2u * lim);
  return miss ? (uint32_t)__builtin_ctz(miss) >> 1 : 16u;
}

// getRcompValue (search/csafm.h:237-272): start from the ISA sample of the
// next sampled text position at or after m (or of position 0, walking
// around the '$' row, when there is none) and step LF back to m.
__device__ __forceinline__ uint32_t isa_value(const DevDir &D, uint32_t ratio, uint32_t n,
                                              uint32_t m) {
  if (D.isa) return D.isa[m];      // expanded index: the full inverse
  const uint64_t siz = (uint64_t)n + 1;
  uint64_t i = (ratio - (m % ratio)) % ratio;
  const uint64_t k = (uint64_t)m + i;
  uint32_t j;
  if (k < siz) {
    j = D.R[k / ratio];
  } else {
    j = D.R[0];
    i = siz - m;
  }
  while (i) {
    if (j == D.dollar) {
      j = 0;
    } else {
      const uint32_t b = j / kBlockRows, t = j - b * kBlockRows;
      const Blk B = load_block(D.blk, b);
      const uint32_t c = code_at(B, t);
      j = pick4(D.C, c) + occ_blk(B, D.dollar_blk, D.dollar_off, b, c, t) + 1u;
    }
    --i;
  }
  return j;
}

// --------------------------------------------------------------------------
// reads staged in LDS: a lane's slot holds its read's 2-bit codes (words 0-15,
// 16 bases each) and N flags (words 16-23, 32 bases each), word j at
// slot[j * 256]
// --------------------------------------------------------------------------

// 16 read bytes as 2-bit codes (byte i at bits 2i) and N flags (bit i: the
// byte is > 3).
__device__ __forceinline__ void pack16(const uint4 &c, uint32_t &code, uint32_t &flag) {
  const uint32_t v[4] = {c.x, c.y, c.z, c.w};
  code = 0u;
  flag = 0u;
#pragma unroll
  for (uint32_t j = 0; j < 4u; ++j) {
    uint32_t t = v[j] & 0x03030303u;
    t = (t | (t >> 6)) & 0x000F000Fu;
    t = (t | (t >> 12)) & 0xFFu;
    uint32_t z = (v[j] >> 2) & 0x3F3F3F3Fu;
    z = z | (z >> 3);
    z = (z | (z >> 1) | (z >> 2)) & 0x01010101u;
    z = (z | (z >> 7)) & 0x00030003u;
    z = (z | (z >> 14)) & 0xFu;
    code |= t << (8u * j);
    flag |= z << (4u * j);
  }
}

// The 32 read bases from W[p] in feed order (rw, 2-bit codes) and their N
// flags (rb), from the lane's slot (words j at slot[j * 256]); the window
// starts at p (d > 0) or p - 31 (then reversed); words outside the slot are
// clamped, which changes only bases outside the read, never compared.
__device__ __forceinline__ void lds_window(const uint32_t *slot, int32_t p, int32_t d,
                                           uint64_t &rw, uint32_t &rb) {
  const uint32_t u = (uint32_t)((d > 0 ? p : p - 31) + 32);
  const int32_t cw = (int32_t)(u >> 4) - 2, fw = (int32_t)(u >> 5) - 1;
  const uint32_t sh = (u & 15u) << 1, fsh = u & 31u;
  const uint32_t i0 = (uint32_t)min(max(cw, 0), 15), i1 = (uint32_t)min(max(cw + 1, 0), 15);
  const uint32_t i2 = (uint32_t)min(max(cw + 2, 0), 15);
  const uint32_t j0 = 16u + (uint32_t)min(max(fw, 0), 7), j1 = 16u + (uint32_t)min(max(fw + 1, 0), 7);
  const uint32_t w0 = slot[i0 * 256u], w1 = slot[i1 * 256u], w2 = slot[i2 * 256u];
  const uint32_t g0 = slot[j0 * 256u], g1 = slot[j1 * 256u];
  const uint64_t lo = ((uint64_t)w1 << 32) | w0;
  const uint64_t win = sh ? (lo >> sh) | ((uint64_t)w2 << (64u - sh)) : lo;
  const uint32_t fl = fsh ? (g0 >> fsh) | (g1 << (32u - fsh)) : g0;
  if (d > 0) {
    rw = win;
    rb = fl;
  } else {
    const uint64_t v = __builtin_bitreverse64(win);
    rw = ((v >> 1) & 0x5555555555555555ull) | ((v & 0x5555555555555555ull) << 1);
    rb = __builtin_bitreverse32(fl);
  }
}


// Stage the read W of len <= 256 bases into the lane's slot from the 16-byte
// chunks that hold its bytes (none past the chunk of its last byte).  Only
// the words that hold its bases are written; the bases after its end in the
// last words are those of the chunk's other bytes, and no step compares them.
__device__ __forceinline__ void stage_read(const uint8_t *W, uint32_t len, uint32_t *slot) {
  if (len == 0u) return;
  const uint32_t mis = (uint32_t)(reinterpret_cast<uintptr_t>(W) & 15u);
  const uint4 *c = reinterpret_cast<const uint4 *>(W - mis);
  const uint32_t last = (mis + len - 1u) >> 4;            // the last byte's chunk
  const uint32_t nw = (len + 15u) >> 4, sh = 2u * mis;
  uint32_t pk, pf, hf = 0u;
  pack16(c[0], pk, pf);
#pragma unroll
  for (uint32_t w = 0; w < 16u; ++w) {
    if (w < nw) {
      uint32_t nk = 0u, nf = 0u;
      if (w + 1u <= last) pack16(c[w + 1u], nk, nf);
      slot[w * 256u] = sh ? (pk >> sh) | (nk << (32u - sh)) : pk;
      const uint32_t f16 = ((pf >> mis) | (nf << (16u - mis))) & 0xFFFFu;
      if (w & 1u) slot[(16u + (w >> 1)) * 256u] = hf | (f16 << 16);
      else hf = f16;
      pk = nk;
      pf = nf;
    }
  }
  if (nw & 1u) slot[(16u + (nw >> 1)) * 256u] = hf;
}

// The read base at p from the lane's slot: its code 0-3, or 4 for an N.
__device__ __forceinline__ uint32_t lds_base(const uint32_t *slot, int32_t p) {
  const uint32_t u = (uint32_t)p;
  const uint32_t n = (slot[(16u + (u >> 5)) * 256u] >> (u & 31u)) & 1u;
  return n ? 4u : (slot[(u >> 4) * 256u] >> ((u & 15u) << 1)) & 3u;
}

// The 32 text bases that precede suffix q of direction dsel in LF order (X[q-1],
// X[q-2], ... for the BWT of X; X[n-q], X[n-q+1], ... for the BWT of reverse
// X) as 2-bit groups in that order; groups before X[0] or past X[n-1] are 0.
__device__ __forceinline__ uint64_t text_window32(const DevIndex &ix, uint32_t dsel, uint32_t q) {
  const uint32_t nwords = (ix.n + 15u) >> 4;
  const int64_t b = dsel ? (int64_t)(ix.n - q) : (int64_t)q - 32;   // first base in address order
  uint64_t win;
  if (b >= 0) {
    const uint32_t w = (uint32_t)b >> 4, sh = ((uint32_t)b & 15u) << 1;
    const uint64_t t0 = w < nwords ? ix.text[w] : 0u;
    const uint64_t t1 = w + 1u < nwords ? ix.text[w + 1u] : 0u;
    const uint64_t t2 = w + 2u < nwords ? ix.text[w + 2u] : 0u;
    const uint64_t lo = (t1 << 32) | t0;
    win = sh ? (lo >> sh) | (t2 << (64u - sh)) : lo;
  } else {                           // dsel 0 and q < 32: X[0 .. q-1] at the top
    const uint64_t t0 = ix.text[0];
    const uint64_t t1 = 1u < nwords ? ix.text[1] : 0u;
    win = ((t1 << 32) | t0) << (2u * (uint32_t)(-b));
  }
  if (dsel) return win;
  const uint64_t v = __builtin_bitreverse64(win);
  return ((v >> 1) & 0x5555555555555555ull) | ((v & 0x5555555555555555ull) << 1);
}

// text_match for a read staged in the lane's slot, up to 32 bases at once:
// the number of leading read bases W[p], W[p+d], ... (at most lim <= 32,
// lim <= q) that equal the text bases preceding suffix q in LF order, the
// read's bases complemented when comp is set; an N never matches.
__device__ __forceinline__ uint32_t text_match32(const DevIndex &ix, uint32_t dsel, uint32_t comp,
                                                 uint32_t q, const uint32_t *slot, int32_t p,
                                                 int32_t d, uint32_t lim) {
  if (lim == 0u) return 0u;
  uint64_t rw;
  uint32_t rb;
  lds_window(slot, p, d, rw, rb);
  if (comp) rw = ~rw;
  const uint64_t x = rw ^ text_window32(ix, dsel, q);
  const uint64_t miss = (x | (x >> 1)) & 0x5555555555555555ull;
  const uint32_t m = miss ? (uint32_t)__builtin_ctzll(miss) >> 1 : 32u;
  const uint32_t nb = rb ? (uint32_t)__builtin_ctz(rb) : 32u;
  return min(min(m, nb), lim);
}

}  // namespace fmi

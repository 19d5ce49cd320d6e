// fmi_kernels.hip -- HIP kernels of the FM-index read-search engine (gfx950).
//
// All arithmetic is exact integer.  Intervals are held as uint32 on the
// device (rows are 0..n with n + 3 < 2^32) and widened to int64 on output.
//
// The unit of work is one LF step: one 64-byte occurrence block (two when k
// and l+1 fall in different blocks) gives Occ of all four symbols at both
// bounds, so a 4-way branch costs the same fetch as a single step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "../../include/fmi.h"
#include "fmi_internal.h"
#include "fmi_device.h"
#include "fmi_kernels.h"
#include "fmi_vhelper.h"

namespace fmi {

// --------------------------------------------------------------------------
// exact search: BWExactSearchBackward / Forward (search/search.h:9-65)
// --------------------------------------------------------------------------

struct ExactParams {
  DevIndex ix;
  uint32_t dsel, comp, direction, nreads;
  const uint8_t *reads;
  const uint64_t *offsets;
  fmi_hit *out;
  fmi_hit *anchor;
  unsigned long long *work;   // counting launch: blocks, LF queries, SA / ISA, text, table
};

// With a k-mer table the first k steps are one read of it when all k bases
// are A/C/G/T and their interval is not empty, else the first k - 4 steps
// from the smaller table the same way (otherwise the steps are taken one by
// one, so the position and interval where the search dies are the
// reference's).  With an expanded index, a search whose interval has
// narrowed to one row goes on against the text at that row's suffix, 16
// bases at a time (32 for a staged read); where it dies, one LF step from
// the last live row gives the reference's empty interval exactly.
//
// STAGED: the read (at most kStagedMaxLen bases) is in the thread's LDS slot
// (stage_read): the k-mer codes, each LF step's base and the text runs' read
// bases come from there instead of byte loads of W.
template <bool COUNT, bool STAGED>
__device__ __forceinline__ void exact_one(const ExactParams &P, uint32_t r, const uint8_t *W,
                                          int32_t len, const uint32_t *slot) {
  uint32_t wb = 0, wq = 0, ws = 0, wt = 0, wk = 0;   // counting launch
  const DevDir &D = dir_of(P.ix, P.dsel);
  const uint32_t n = P.ix.n;
  const int32_t d = P.direction ? 1 : -1;
  uint32_t k = 0, l = 0, k2 = 0, l2 = n, nb;
  int32_t i = P.direction ? 0 : len - 1;
  // the first K bases from the k-table, else the first K - 4 from the
  // smaller one, else one LF step at a time
  const uint32_t tks[2] = {P.ix.kmer_k, P.ix.kmer2_k};
  uint64_t kw = 0;                   // staged: the first 32 bases in feed order
  uint32_t kn = 0;                   // and their N flags
  if (STAGED && tks[0] && len > 0) lds_window(slot, i, d, kw, kn);
#pragma unroll
  for (uint32_t tt = 0; tt < 2u; ++tt) {
    const uint32_t K = tks[tt];
    if (!K || len < (int32_t)K || k2 != 0u || l2 != n) continue;
    uint32_t code = 0u, bad = 0u;
    if (STAGED) {
      code = (uint32_t)kw & ((1u << (2u * K)) - 1u);
      bad = kn & ((1u << K) - 1u);
    } else {
      for (uint32_t j = 0; j < K; ++j) {
        const uint32_t s = W[i + (int32_t)j * d];
        bad |= s > 3u ? 1u : 0u;
        code |= (s & 3u) << (2u * j);
      }
    }
    if (P.comp) code ^= (1u << (2u * K)) - 1u;
    if (!bad) {
      const uint64_t e = (tt ? P.ix.kmer2 : P.ix.kmer)[P.dsel][code];
      if (COUNT) wk++;
      if ((uint32_t)e <= (uint32_t)(e >> 32)) {
        k = k2 = (uint32_t)e;
        l = l2 = (uint32_t)(e >> 32);
        i += (int32_t)K * d;
      }
    }
  }
  for (; P.direction ? (i <= len - 1) : (i >= 0); i += d) {
    if (k2 == l2 && P.ix.text != nullptr) {
      if ((STAGED ? lds_base(slot, i) : (uint32_t)W[i]) > 3u) { k2 = n + 1; l2 = n; break; }
      // text mode from the single row k2 on
      uint32_t q = D.sa[k2];
      if (COUNT) ws++;
      uint32_t rem = (uint32_t)(d < 0 ? i + 1 : len - i);
      bool died = false;
      while (rem > 0u) {
        const uint32_t lim = min(min(rem, STAGED ? 32u : 16u), q);
        const uint32_t m = STAGED ? text_match32(P.ix, P.dsel, P.comp, q, slot, i, d, lim)
                                  : text_match(P.ix, P.dsel, P.comp, q, W, i, d, lim);
        if (COUNT) wt++;
        q -= m;
        rem -= m;
        i += (int32_t)m * d;
        // a mismatch or an N (m < lim), or the text's start (q == 0)
        if (rem > 0u && (m < lim || q == 0u)) { died = true; break; }
      }
      k = l = D.isa[q];
      if (COUNT) ws++;
      if (!died) { k2 = l2 = k; break; }
      const uint32_t s2 = STAGED ? lds_base(slot, i) : (uint32_t)W[i];
      if (s2 > 3u) { k2 = n + 1; l2 = n; }
      else {
        lf_one(D, k, l, P.comp ? 3u - s2 : s2, k2, l2, nb);
        if (COUNT) { wb += nb; wq++; }
      }
      break;
    }
    uint32_t kn2, ln2;
    if (COUNT) {
      wb += (k2 != l2 && (l2 + 1u) / kBlockRows != k2 / kBlockRows) ? 2u : 1u;
      wq++;
    }
    const uint32_t sym = STAGED ? lf_sym(D, k2, l2, lds_base(slot, i), P.comp, kn2, ln2)
                                : lf_read(D, k2, l2, W + i, P.comp, kn2, ln2);
    if (sym > 3u) { k2 = n + 1; l2 = n; break; }
    k2 = kn2; l2 = ln2;
    if (k2 > l2) break;
    k = k2; l = l2;
  }
  if (P.direction == 0) {
    if (P.anchor)
      store_hit(P.anchor + r, k, l, (int16_t)(i + 1), (int16_t)i, (int16_t)(len - 1), 0, 0, 0, 0, 0, 0);
  } else {
    if (P.anchor)
      store_hit(P.anchor + r, k, l, 0, (int16_t)i, (int16_t)(i - 1), 0, 1, 0, 0, 0, 0);
  }
  store_hit(P.out + r, k2, l2, 0, (int16_t)i, (int16_t)(len - 1), 0, (uint8_t)P.direction, 0, 0, 0, 0);
  if (COUNT) {
    atomicAdd(&P.work[0], (unsigned long long)wb);
    atomicAdd(&P.work[1], (unsigned long long)wq);
    atomicAdd(&P.work[2], (unsigned long long)ws);
    atomicAdd(&P.work[3], (unsigned long long)wt);
    atomicAdd(&P.work[4], (unsigned long long)wk);
  }
}

// One thread per read; a read of at most kStagedMaxLen bases is staged in
// the thread's LDS slot first (24 KB per workgroup of 256).
template <bool COUNT>
__global__ __launch_bounds__(256) void exact_kernel(ExactParams P) {
  __shared__ uint32_t slots[kReadRecordWords * 256u];
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P.nreads) return;
  const uint64_t off = P.offsets[r];
  const uint8_t *W = P.reads + off;
  const int32_t len = (int32_t)(P.offsets[r + 1] - off);
  uint32_t *slot = slots + threadIdx.x;
  if (len <= (int32_t)kStagedMaxLen) {
    stage_read(W, (uint32_t)len, slot);
    exact_one<COUNT, true>(P, r, W, len, slot);
  } else {
    exact_one<COUNT, false>(P, r, W, len, slot);
  }
}

// BWExactSearchVectorBackward / Forward (search/search.c:1407-1489).
struct VectorParams {
  DevIndex ix;
  uint32_t dsel, comp, direction, nreads;
  const uint8_t *reads;
  const uint64_t *offsets;
  int64_t *vk, *vl;
  int16_t *last_good;
};

__global__ __launch_bounds__(256) void vector_kernel(VectorParams P) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P.nreads) return;
  const DevDir &D = dir_of(P.ix, P.dsel);
  const uint64_t off = P.offsets[r];
  const uint8_t *W = P.reads + off;
  const int32_t len = (int32_t)(P.offsets[r + 1] - off);
  const uint32_t n = P.ix.n;
  int64_t *vk = P.vk + off, *vl = P.vl + off;
  int32_t good = 0;
  uint32_t k2 = 0, l2 = n;
  if (P.direction == 0) {
    int32_t i = len - 1, j = len - 1;
    for (; i >= 0; --i, --j) {
      uint32_t kn, ln;
      if (lf_read(D, k2, l2, W + i, P.comp, kn, ln) > 3u) { k2 = n + 1; l2 = n; }
      else { k2 = kn; l2 = ln; }
      vk[j] = k2; vl[j] = l2;
      if (k2 > l2) { --i; --j; break; }
      ++good;
    }
    for (; i >= 0; --i, --j) { vk[j] = k2; vl[j] = l2; }
  } else {
    int32_t i = 0, j = 0;
    for (; i <= len - 1; ++i, ++j) {
      uint32_t kn, ln;
      if (lf_read(D, k2, l2, W + i, P.comp, kn, ln) > 3u) { k2 = n + 1; l2 = n; }
      else { k2 = kn; l2 = ln; }
      vk[j] = k2; vl[j] = l2;
      if (k2 > l2) { ++i; ++j; break; }
      ++good;
    }
    for (; i <= len - 1; ++i, ++j) { vk[j] = k2; vl[j] = l2; }
  }
  P.last_good[r] = (int16_t)good;
}

// --------------------------------------------------------------------------
// 1-error search: BWSearch1CPU (search/search.c:1876-1933) with
// BWSimpleSearch1Backward / Forward (search/search.c:1711-1874)
// --------------------------------------------------------------------------
//
// One lane runs one (read, strand) task as a state machine whose every
// iteration is exactly one LF query, so all lanes of a wavefront issue their
// block loads together:
//   EXACT  exact extension over the anchor half (pass 0: [len/2, len-1]
//          right to left on the first index; pass 1: [0, len/2 - 1] left to
//          right on the second index)
//   SCAN   one node of the other half: the block gives the four children;
//          if the size of the interval does not change, every occurrence
//          extends by W[i] and no branch can succeed (search.c:1747)
//   CONT   exact continuation of one branch (deletion, insertion b,
//          mismatch b, in the reference's order) to the end of the read
// Records are written in the order the reference adds them.

enum : uint32_t { PH_IDLE = 0, PH_EXACT = 1, PH_SCAN = 2, PH_CONT = 3, PH_NEXT_STRAND = 4 };

struct S1Params {
  DevIndex ix;
  const uint8_t *reads;
  const uint64_t *offsets;
  fmi_hit *hits;
  uint32_t *counts;
  uint32_t *queue;                 // work counter (persistent launch)
  unsigned long long *stats;       // count mode: [0] blocks, [1] LF queries,
                                   // [2] SA / ISA reads, [3] text steps,
                                   // [4] k-mer table lookups
  uint32_t ntasks, strand_mask, cap, chunk;
  uint32_t refill_min;             // idle lanes before a wave takes new tasks
  const uint4 *rec;                // staged launch: the read records (kRecWords words each)
  const uint32_t *list;            // when set: the tasks are list[0 .. *list_n)
  const uint32_t *list_n;
};

// The staged launch keeps every lane's read in LDS for the whole task: a
// streaming kernel first packs each read into a 96-byte record, words 0..15
// the 2-bit codes (base i at bits 2 (i % 16) of word i / 16), 16..23 the N
// flags (base i at bit i % 32 of word 16 + i / 32), and a lane copies its
// read's record into its LDS slot (the same layout) when it takes the task,
// so that no step reads the read from memory again.  The reads longer than
// kStagedMaxLen are listed and searched by the unstaged kernel afterwards.
constexpr uint32_t kSlotWords = 24;   // 16 code words + 8 flag words per lane
constexpr uint32_t kRecWords = kSlotWords;

struct S1Lane {
  uint32_t phase, pass, strand, read, nout, comp, bi;
  const uint8_t *W;
  int32_t len, i, cp;
  uint32_t k, l, kn, ln, qk, ql, wsym;
  uint32_t ck[4], cl[4];           // children of the SCAN node, physical symbol order
  uint32_t dsel;                   // the pass's direction: 0 BWT of X, 1 of reverse X
  // Text mode (expanded index): a single-row interval is held as the text
  // position q of its suffix in the pass's direction; kTqNone = rows in k, l.
  uint32_t tq, tqn, cq;            // node, child of W[i], continuation
  uint32_t kc[2];                  // k-mer table index of each pass's first k bases
                                   // (+ strand; kKReady unless unknown or with an N)
  uint32_t nlook, ntext, nkmer;    // SA / ISA reads, text steps, table lookups (count mode)
};

constexpr uint32_t kTqNone = 0xFFFFFFFFu;
constexpr uint32_t kTextRowTag = 0x80000000u;   // k's high word: k, l hold q

// 16 read bases from byte wp of the 32 bytes c0, c1 (towards higher bytes
// when up, else lower) as 2-bit groups in that order (w, physical symbols)
// and their N flags (nb, bit i: a byte > 3).
__device__ __forceinline__ void window16(const uint4 &c0, const uint4 &c1, uint32_t wp, bool up,
                                         uint32_t &w, uint32_t &nb) {
  const uint32_t v[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
  uint64_t codes = 0ull;
  uint32_t nfl = 0u;
#pragma unroll
  for (uint32_t j = 0; j < 8u; ++j) {
    uint32_t t = v[j] & 0x03030303u;              // 4 codes -> 8 bits
    t = (t | (t >> 6)) & 0x000F000Fu;
    t = (t | (t >> 12)) & 0xFFu;
    uint32_t z = (v[j] >> 2) & 0x3F3F3F3Fu;       // a byte > 3 -> its bit 0
    z = z | (z >> 3);
    z = (z | (z >> 1) | (z >> 2)) & 0x01010101u;
    z = (z | (z >> 7)) & 0x00030003u;
    z = (z | (z >> 14)) & 0xFu;
    codes |= (uint64_t)t << (8u * j);
    nfl |= z << (4u * j);
  }
  if (up) {
    w = (uint32_t)(codes >> (2u * wp));
    nb = (nfl >> wp) & 0xFFFFu;
  } else {
    const uint32_t s0 = wp - 15u;                   // bytes s0 .. wp
    const uint32_t r = __builtin_bitreverse32((uint32_t)(codes >> (2u * s0)));
    w = ((r >> 1) & 0x55555555u) | ((r & 0x55555555u) << 1);
    nb = __builtin_bitreverse32((nfl >> s0) & 0xFFFFu) >> 16;
  }
}

// k-mer table index of a pass's first k bases, with this bit set when it is
// known (a table, at least k bases, no N among them).
constexpr uint32_t kKmerReady = 0x80000000u;

// The k-mer table index of both passes' first k bases of the read at
// reads + off (pass 1: W[0], W[1], ...; pass 0: W[len-1], W[len-2], ...),
// from two aligned 16-byte chunks at each end of the read, each holding one
// of those bases; 0 where unknown.
__device__ __forceinline__ void kmer_codes(const uint8_t *reads, uint64_t off, uint32_t len,
                                           uint32_t K, uint32_t &kc0, uint32_t &kc1) {
  kc0 = kc1 = 0u;
  if (!K || len < K) return;
  const uint8_t *a = reads + off, *z = reads + off + (len - 1u);
  const uint32_t ma = (uint32_t)(reinterpret_cast<uintptr_t>(a) & 15u);
  const uint32_t mz = (uint32_t)(reinterpret_cast<uintptr_t>(z) & 15u);
  const uint8_t *a0 = a - ma, *z1 = z - mz;
  const uint8_t *a1 = ma + K - 1u >= 16u ? a0 + 16 : a0;
  const uint8_t *z0 = K - 1u > mz ? z1 - 16 : z1;
  const uint4 ca0 = *reinterpret_cast<const uint4 *>(a0);
  const uint4 ca1 = *reinterpret_cast<const uint4 *>(a1);
  const uint4 cz0 = *reinterpret_cast<const uint4 *>(z0);
  const uint4 cz1 = *reinterpret_cast<const uint4 *>(z1);
  const uint32_t kmask = (1u << (2u * K)) - 1u, nmask = (1u << K) - 1u;
  uint32_t w, nb;
  window16(ca0, ca1, ma, true, w, nb);
  kc1 = (nb & nmask) ? 0u : ((w & kmask) | kKmerReady);
  window16(cz0, cz1, 16u + mz, false, w, nb);
  kc0 = (nb & nmask) ? 0u : ((w & kmask) | kKmerReady);
}

struct RecParams {
  const uint8_t *reads;
  const uint64_t *offsets;
  uint32_t ntasks;
  uint32_t *rec;                   // kRecWords words per task
  uint32_t *list, *list_n;         // the reads longer than kStagedMaxLen
};

// Eight lanes per read, each packing 32 bases (from the three aligned
// 16-byte chunks that hold them, none past the chunk of the read's last
// byte); the first lane of a longer read lists it instead.
__global__ __launch_bounds__(256) void read_records_kernel(RecParams P) {
  const uint64_t gid = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  const uint64_t t64 = gid >> 3;
  const uint32_t j = (uint32_t)(gid & 7u);
  if (t64 >= P.ntasks) return;
  const uint32_t t = (uint32_t)t64;
  const uint64_t off = P.offsets[t];
  const uint32_t len = (uint32_t)(P.offsets[t + 1] - off);
  uint32_t *rec = P.rec + (size_t)t * kRecWords;
  if (len > kStagedMaxLen) {
    if (j == 0u) P.list[atomicAdd(P.list_n, 1u)] = t;
    return;
  }
  if (32u * j >= len) return;
  const uint8_t *a = P.reads + off + 32u * j;
  const uint32_t mis = (uint32_t)(reinterpret_cast<uintptr_t>(a) & 15u);
  const uint8_t *c = a - mis;
  const uint8_t *z = P.reads + off + (len - 1u);
  const uint8_t *zc = z - (reinterpret_cast<uintptr_t>(z) & 15u);   // chunk of the last byte
  const uint4 c0 = *reinterpret_cast<const uint4 *>(c);
  const uint4 c1 = c + 16 <= zc ? *reinterpret_cast<const uint4 *>(c + 16) : make_uint4(0, 0, 0, 0);
  const uint4 c2 = c + 32 <= zc ? *reinterpret_cast<const uint4 *>(c + 32) : make_uint4(0, 0, 0, 0);
  uint32_t k0, k1, k2, f0, f1, f2;
  pack16(c0, k0, f0);
  pack16(c1, k1, f1);
  pack16(c2, k2, f2);
  const uint64_t lo = ((uint64_t)k1 << 32) | k0;
  const uint32_t sh = 2u * mis;
  const uint64_t codes = sh ? (lo >> sh) | ((uint64_t)k2 << (64u - sh)) : lo;
  const uint64_t fl = (uint64_t)f0 | ((uint64_t)f1 << 16) | ((uint64_t)f2 << 32);
  *reinterpret_cast<uint2 *>(rec + 2 * j) = make_uint2((uint32_t)codes, (uint32_t)(codes >> 32));
  rec[16 + j] = (uint32_t)(fl >> mis);
}

template <bool MM_ONLY, bool COUNT, bool TEXT, bool OWN, bool LDS = false>
struct S1 {
  const S1Params &P;
  S1Lane &L;
  uint32_t *slot;                  // staged launch: this lane's LDS words (stride 256)

  __device__ __forceinline__ int32_t step() const { return L.pass ? 1 : -1; }
  __device__ __forceinline__ bool in_read(int32_t p) const { return p >= 0 && p <= L.len - 1; }
  __device__ __forceinline__ bool in_exact(int32_t p) const {
    return L.pass ? (p <= L.len / 2 - 1) : (p >= L.len / 2);
  }

  // C[s] of the pass's direction (s may differ per lane).
  __device__ __forceinline__ uint32_t c_of(uint32_t s) const {
    const uint32_t a0 = L.dsel ? P.ix.d[1].C[0] : P.ix.d[0].C[0];
    const uint32_t a1 = L.dsel ? P.ix.d[1].C[1] : P.ix.d[0].C[1];
    const uint32_t a2 = L.dsel ? P.ix.d[1].C[2] : P.ix.d[0].C[2];
    const uint32_t a3 = L.dsel ? P.ix.d[1].C[3] : P.ix.d[0].C[3];
    const uint32_t lo = (s & 1u) ? a1 : a0, hi = (s & 1u) ? a3 : a2;
    return (s & 2u) ? hi : lo;
  }

  // Physical symbol of the text base that precedes suffix q (q >= 1) in the
  // pass's direction: X[q-1] for the BWT of X, X[n-q] for the BWT of
  // reverse X.
  __device__ __forceinline__ uint32_t text_sym(uint32_t q) const {
    const uint32_t x = L.dsel ? P.ix.n - q : q - 1u;
    return (P.ix.text[x >> 4] >> ((x & 15u) << 1)) & 3u;
  }

  __device__ __forceinline__ void emit(uint32_t k, uint32_t l, uint32_t q, uint8_t nmm,
                                       uint8_t kind, uint8_t base, int16_t err_pos) {
    if (L.nout < P.cap) {
      // A single-row text-mode interval is stored as its suffix's text
      // position, tagged in k's high word; text_rows_kernel turns it into
      // the row (ISA[q]) after the search, so no lane waits for that read
      // here (count mode counts it).
      uint64_t kk = k, ll = l;
      if (q != kTqNone) {
        if (COUNT) L.nlook++;
        kk = ((uint64_t)kTextRowTag << 32) | q;
        ll = q;
      }
      int16_t start, end, pos;
      if (nmm == 0) {               // the exact record of pass 0
        start = 0; end = (int16_t)(L.len - 1); pos = -1;
      } else if (L.pass == 0) {     // bound_result(start, half-1), pos = start-1
        start = 0; end = (int16_t)(L.len / 2 - 1); pos = -1;
      } else {                      // bound_result(half+1, end), pos = end+1
        start = (int16_t)(L.len / 2); end = (int16_t)(L.len - 1); pos = (int16_t)L.len;
      }
      if (!COUNT)
        store_hit(P.hits + (size_t)(2 * L.read + L.strand) * P.cap + L.nout, (int64_t)kk,
                  (int64_t)ll, start, pos,
                  end, err_pos, (uint8_t)L.pass, nmm, kind,
                  base, (uint8_t)L.strand);
    }
    L.nout++;
  }

  // A task is one read: its + strand, then its - strand when both are asked
  // for, so that every lane carries one true-strand and one wrong-strand
  // search and lanes finish at similar times.
  // The - strand is started from the main loop (PH_NEXT_STRAND), not from
  // here: calling start_strand would close a call cycle that the compiler
  // can only lower as recursion on a scratch stack.
  __device__ __forceinline__ void finish_task() {
    P.counts[2 * L.read + L.strand] = L.nout;
    L.phase = (P.strand_mask == 3u && L.strand == 0u) ? PH_NEXT_STRAND : PH_IDLE;
  }

  __device__ __forceinline__ void start_strand(uint32_t strand) {
    L.strand = strand;
    L.nout = 0;
    enter_pass(0);
  }

  __device__ __forceinline__ void end_pass() {
    if (L.pass == 0) enter_pass(1);
    else finish_task();
  }

  __device__ __forceinline__ void enter_pass(uint32_t pass) {
    L.pass = pass;
    // strand 0: (backward, forward); strand 1: (forward_rev, backward_rev)
    const bool second = (pass ^ L.strand) != 0u;
    L.dsel = second ? 1u : 0u;
    L.tq = L.tqn = L.cq = kTqNone;
    L.comp = L.strand;
    // the pass's k-mer table index, complemented on the - strand (0: none
    // known; the first iteration reads the bases)
    const uint32_t kc = pass ? L.kc[1] : L.kc[0];
    const uint32_t kmask = (1u << (2u * P.ix.kmer_k)) - 1u;
    L.bi = (kc & kKReady) && L.comp ? kc ^ kmask : kc;
    L.k = 0;
    L.l = P.ix.n;
    L.i = pass ? 0 : L.len - 1;
    if (in_read(L.i) && in_exact(L.i)) L.phase = PH_EXACT;
    else enter_scan();
  }

  __device__ __forceinline__ void enter_scan() {
    L.kn = L.k;
    L.ln = L.l;
    L.tqn = L.tq;
    L.i = L.pass ? L.len / 2 : L.len / 2 - 1;
    scan_next();
  }

  __device__ __forceinline__ void scan_next() {
    if (!in_read(L.i)) {
      if (L.pass == 0 && L.kn <= L.ln) emit(L.kn, L.ln, L.tqn, 0, 0, 0, 0);
      end_pass();
      return;
    }
    if (L.kn > L.ln) { end_pass(); return; }
    L.k = L.kn;
    L.l = L.ln;
    L.tq = L.tqn;
    L.phase = PH_SCAN;
  }

  // The node's next branch from L.bi on (0: deletion; 1 + 2b: insertion of
  // b; 2 + 2b: mismatch to b), in that order: the branches that exist form
  // a 9-bit mask and the loop visits only those.  A branch whose
  // continuation is already past the end of the read succeeds at once (its
  // record is stored and the next branch is taken); otherwise the lane
  // continues it (PH_CONT).  With no branch left the scan moves on.  In text
  // mode the node's children are its suffix extended by one text base.
  __device__ __forceinline__ void branch_next() {
    uint32_t ok = MM_ONLY ? 0u : 1u;
#pragma unroll
    for (uint32_t b = 0; b < 4u; ++b) {
      const uint32_t ps = L.comp ? 3u - b : b;
      const bool nonempty = pick4(L.ck, ps) <= pick4(L.cl, ps);
      if (nonempty && !MM_ONLY) ok |= 1u << (1u + 2u * b);
      if (nonempty && b != L.wsym) ok |= 1u << (2u + 2u * b);
    }
    ok &= ~0u << min(L.bi, 31u);
    while (ok) {
      const uint32_t bi = (uint32_t)__builtin_ctz(ok);
      ok &= ok - 1u;
      L.bi = bi;
      if (bi == 0u) {
        L.qk = L.k; L.ql = L.l; L.cq = L.tq; L.cp = L.i + step();
      } else {
        const uint32_t b = (bi - 1u) >> 1;
        const uint32_t ps = L.comp ? 3u - b : b;
        L.qk = pick4(L.ck, ps);
        L.ql = pick4(L.cl, ps);
        L.cq = L.tq == kTqNone ? kTqNone : L.tq - 1u;
        L.cp = (bi & 1u) ? L.i : L.i + step();
      }
      if (in_read(L.cp)) {
        L.phase = PH_CONT;
        return;
      }
      emit_branch();
    }
    L.bi = 9u;
    L.i += step();
    scan_next();
  }

  __device__ __forceinline__ void emit_branch() {
    const uint32_t bi = L.bi;
    if (bi == 0) emit(L.qk, L.ql, L.cq, 1, FMI_DELETION, 0xFF, (int16_t)L.i);
    else emit(L.qk, L.ql, L.cq, 1, (bi & 1u) ? FMI_INSERTION : FMI_MISMATCH,
              (uint8_t)((bi - 1) >> 1), (int16_t)L.i);
  }

  // What a lane needs to start a task: the read's offset and length and the
  // k-mer table index of both passes' first k bases (kmer_codes).
  __device__ __forceinline__ void task_info(uint32_t task, uint64_t &off, uint32_t &len,
                                            uint32_t &kc0, uint32_t &kc1) const {
    off = P.offsets[task];
    len = (uint32_t)(P.offsets[task + 1] - off);
    kmer_codes(P.reads, off, len, P.ix.kmer_k, kc0, kc1);
  }

  __device__ __forceinline__ void start_with(uint32_t task, uint64_t off, uint32_t len,
                                             uint32_t kc0, uint32_t kc1) {
    L.read = task;
    L.W = P.reads + off;
    L.len = (int32_t)len;
    L.kc[0] = kc0;
    L.kc[1] = kc1;
    const uint32_t strand = (P.strand_mask == 2u) ? 1u : 0u;
    if (P.strand_mask != 3u) P.counts[2 * L.read + (strand ^ 1u)] = 0;
    start_strand(strand);
  }

  // Staged launch: the read's codes and N flags, copied from its record into
  // the lane's LDS slot (only the 16-byte chunks that hold its bases; the
  // rest of the slot keeps an earlier read's words, which no step reads);
  // then the k-mer table index of both passes' first k bases is taken from
  // the slot (as kmer_codes takes it from the read) and the task starts as
  // start_with starts it.
  __device__ __forceinline__ void stage_with(uint32_t task, uint32_t len) {
    const uint4 *r = P.rec + (size_t)task * (kRecWords / 4u);
    const uint32_t nc = (len + 63u) >> 6, nf = (len + 127u) >> 7;
    uint4 c0 = opaque4(), c1 = opaque4(), c2 = opaque4(), c3 = opaque4();
    uint4 f0 = opaque4(), f1 = opaque4();
    if (nc > 0u) c0 = r[0];
    if (nc > 1u) c1 = r[1];
    if (nc > 2u) c2 = r[2];
    if (nc > 3u) c3 = r[3];
    if (nf > 0u) f0 = r[4];
    if (nf > 1u) f1 = r[5];
    if (nc > 0u) { slot[0 * 256] = c0.x; slot[1 * 256] = c0.y; slot[2 * 256] = c0.z; slot[3 * 256] = c0.w; }
    if (nc > 1u) { slot[4 * 256] = c1.x; slot[5 * 256] = c1.y; slot[6 * 256] = c1.z; slot[7 * 256] = c1.w; }
    if (nc > 2u) { slot[8 * 256] = c2.x; slot[9 * 256] = c2.y; slot[10 * 256] = c2.z; slot[11 * 256] = c2.w; }
    if (nc > 3u) { slot[12 * 256] = c3.x; slot[13 * 256] = c3.y; slot[14 * 256] = c3.z; slot[15 * 256] = c3.w; }
    if (nf > 0u) { slot[16 * 256] = f0.x; slot[17 * 256] = f0.y; slot[18 * 256] = f0.z; slot[19 * 256] = f0.w; }
    if (nf > 1u) { slot[20 * 256] = f1.x; slot[21 * 256] = f1.y; slot[22 * 256] = f1.z; slot[23 * 256] = f1.w; }
    uint32_t kc0 = 0u, kc1 = 0u;
    const uint32_t K = P.ix.kmer_k;
    if (K && len >= K) {
      const uint32_t kmask = (1u << (2u * K)) - 1u, nmask = (1u << K) - 1u;
      uint64_t w;
      uint32_t nb;
      slot_window(0, 1, w, nb);                     // W[0], W[1], ...
      kc1 = (nb & nmask) ? 0u : (((uint32_t)w & kmask) | kKmerReady);
      slot_window((int32_t)len - 1, -1, w, nb);     // W[len-1], W[len-2], ...
      kc0 = (nb & nmask) ? 0u : (((uint32_t)w & kmask) | kKmerReady);
    }
    start_with(task, 0u, len, kc0, kc1);
  }

  // Staged launch: the 32 read bases from W[p] in feed order (rw, 2-bit
  // physical symbols) and their N flags (rb, bit i), from the lane's LDS
  // slot.  The window starts at base s = p (d > 0) or p - 31 (then it is
  // reversed); word indexes outside the slot are clamped, which changes only
  // bases outside the read, and no step compares those.
  __device__ __forceinline__ void slot_window(int32_t p, int32_t d, uint64_t &rw,
                                              uint32_t &rb) const {
    const uint32_t u = (uint32_t)((d > 0 ? p : p - 31) + 32);   // s + 32: 1 .. 287
    const int32_t cw = (int32_t)(u >> 4) - 2, fw = (int32_t)(u >> 5) - 1;
    const uint32_t sh = (u & 15u) << 1, fs = u & 31u;
    const uint32_t i0 = (uint32_t)min(max(cw, 0), 15), i1 = (uint32_t)min(max(cw + 1, 0), 15);
    const uint32_t i2 = (uint32_t)min(cw + 2, 15);
    const uint32_t j0 = 16u + (uint32_t)max(fw, 0), j1 = 16u + (uint32_t)min(fw + 1, 7);
    const uint32_t w0 = slot[i0 * 256u], w1 = slot[i1 * 256u], w2 = slot[i2 * 256u];
    const uint32_t f0 = slot[j0 * 256u], f1 = slot[j1 * 256u];
    const uint64_t lo = ((uint64_t)w1 << 32) | w0;
    const uint64_t win = sh ? (lo >> sh) | ((uint64_t)w2 << (64u - sh)) : lo;
    const uint32_t fl = fs ? (f0 >> fs) | (f1 << (32u - fs)) : f0;
    if (d > 0) {
      rw = win;
      rb = fl;
    } else {
      const uint64_t v = __builtin_bitreverse64(win);
      rw = ((v >> 1) & 0x5555555555555555ull) | ((v & 0x5555555555555555ull) << 1);
      rb = __builtin_bitreverse32(fl);
    }
  }

  __device__ __forceinline__ void start(uint32_t task) {
    uint64_t off;
    uint32_t len, kc0, kc1;
    task_info(task, off, len, kc0, kc1);
    start_with(task, off, len, kc0, kc1);
  }

  // The 32 text bases that precede suffix q in the pass's direction, in
  // feed order as 2-bit groups (physical symbols), from the three text words
  // t0, t1, t2 from word w on that advance() read for q (0 past the last
  // word): w = (n - q) / 16 for the BWT of reverse X, (q - 32) / 16 (0 when
  // q < 32) for the BWT of X.
  __device__ __forceinline__ uint64_t text_window(uint32_t q, uint32_t t0, uint32_t t1,
                                                  uint32_t t2) const {
    const uint64_t lo = ((uint64_t)t1 << 32) | t0;
    if (L.dsel) {                    // X[x0], X[x0+1], ... with x0 = n - q
      const uint32_t sh = ((P.ix.n - q) & 15u) << 1;
      return sh ? (lo >> sh) | ((uint64_t)t2 << (64u - sh)) : lo;
    }
    const int64_t b = (int64_t)q - 32;   // X[b .. b+31] = X[q-32 .. q-1]
    uint64_t win;
    if (b >= 0) {
      const uint32_t sh = ((uint32_t)b & 15u) << 1;
      win = sh ? (lo >> sh) | ((uint64_t)t2 << (64u - sh)) : lo;
    } else {                         // q < 32: X[0 .. q-1] end the window
      win = lo << min(2u * (uint32_t)(-b), 62u);
    }
    // reverse the 32 groups: bit reversal, then swap the bits of each pair
    const uint64_t v = __builtin_bitreverse64(win);
    return ((v >> 1) & 0x5555555555555555ull) | ((v & 0x5555555555555555ull) << 1);
  }

  // First mismatch (0..lim) of the read window r (feed order, complemented
  // on the - strand) against the text window t; bad marks N bases (bit 2i).
  __device__ __forceinline__ static uint32_t run_length(uint64_t t, uint64_t r, uint64_t bad,
                                                        uint32_t lim) {
    if (lim == 0u) return 0u;
    const uint64_t x = r ^ t;
    uint64_t miss = ((x | (x >> 1)) & 0x5555555555555555ull) | bad;
    if (lim < 32u) miss |= 1ull << (2u * lim);
    return miss ? (uint32_t)__builtin_ctzll(miss) >> 1 : 32u;
  }

  // The transition of one text step on a text-mode interval (suffix q): m of
  // lim read bases matched the text that precedes the suffix (no occurrence
  // block); tsym is the text base before q (4 at q = 0).  A single-row
  // interval has one non-empty child, the one of the text base, so a SCAN
  // node branches exactly where the read's base differs from the text; the
  // matching positions before it are passed over at once.
  __device__ __forceinline__ void text_apply(bool cont, uint32_t q, uint32_t lim, uint32_t m,
                                             uint32_t r0, uint32_t tsym) {
    const int32_t p = cont ? L.cp : L.i;
    const int32_t d = step();
    if (COUNT) L.ntext++;
    if (L.phase == PH_SCAN) {
      if (m > 0u) {
        L.kn = 0; L.ln = 0; L.tqn = q - m;
        L.i += (int32_t)m * d;
        scan_next();
        return;
      }
#pragma unroll
      for (uint32_t s = 0; s < 4; ++s) {
        L.ck[s] = s == tsym ? 0u : 1u;
        L.cl[s] = 0u;
      }
      L.wsym = r0;
      L.kn = P.ix.n + 1; L.ln = P.ix.n; L.tqn = kTqNone;
      L.bi = 0;
      branch_next();
      return;
    }
    uint32_t rem;
    if (!cont) rem = (uint32_t)(L.pass ? L.len / 2 - p : p - L.len / 2 + 1);
    else rem = (uint32_t)(d < 0 ? p + 1 : L.len - p);
    const bool died = m < lim || (m == q && m < rem);
    if (!cont) {                     // PH_EXACT
      if (died) { end_pass(); return; }
      L.tq = q - m;
      L.k = 0; L.l = 0;
      L.i += (int32_t)m * d;
      if (in_read(L.i) && in_exact(L.i)) return;
      enter_scan();
      return;
    }
    if (died) {                      // continuation died
      L.bi++;
      branch_next();
      return;
    }
    L.cq = q - m;
    L.cp += (int32_t)m * d;
    if (in_read(L.cp)) return;
    emit_branch();
    L.bi++;
    branch_next();
  }

  // What one iteration of a lane reads.
  enum : uint32_t { F_TEXT = 0, F_KPREP = 1, F_KLOOK = 2, F_SA = 3, F_LF = 4 };

  // One iteration of the lane's state machine: one LF query, one text step,
  // the SA read that enters text mode, or one k-mer table lookup (in two
  // iterations: the read bases, then the entry), and the transition it
  // drives.  Every lane first works out the addresses of what it reads and
  // all lanes issue their loads before the first wait, so a wavefront whose
  // lanes are in different kinds of step waits for one memory round trip per
  // iteration, not one per kind; the transitions, ALU work, follow.
  //
  // The k-mer lookup: (0, n) in EXACT is the phase's start (every LF step
  // gives k >= 1); its first kmer_k steps are one table entry, the interval
  // of those read bases (complemented on the - strand, as the LF steps take
  // them); an N among them empties the interval as its LF step would.  The
  // prepared entry index is kept in bi (unused in EXACT) with kKReady set.
  static constexpr uint32_t kKReady = kKmerReady;

  __device__ __forceinline__ void advance(unsigned long long &nblk_acc, unsigned long long &nq_acc) {
    const bool cont = L.phase == PH_CONT;
    const uint32_t tq = cont ? L.cq : L.tq;
    const uint32_t ak = cont ? L.qk : L.k;
    const uint32_t al = cont ? L.ql : L.l;
    const int32_t p = cont ? L.cp : L.i;
    const int32_t d = step();
    const uint32_t K = P.ix.kmer_k;
    const uint32_t *text = P.ix.text;

    // ---- what to read (ALU only)
    uint32_t kind, rlim = 1u, lim = 0u;
    const uint32_t *a0 = nullptr;    // the word the kind reads (not F_LF, F_KPREP)
    bool two = false, v1 = false, v2 = false, next = false;
    uint32_t j1 = 0u, j2 = 0u;
    if (TEXT && tq != kTqNone) {
      kind = F_TEXT;
      uint32_t rem;
      if (L.phase == PH_EXACT) rem = (uint32_t)(L.pass ? L.len / 2 - p : p - L.len / 2 + 1);
      else rem = (uint32_t)(d < 0 ? p + 1 : L.len - p);
      lim = min(min(rem, 32u), tq);
      rlim = max(lim, 1u);
      const uint32_t nwords = (P.ix.n + 15u) >> 4;
      uint32_t w;
      if (L.dsel) {
        w = min((P.ix.n - tq) >> 4, nwords - 1u);   // q = 0 (lim = 0) is clamped
      } else {
        const int64_t b = (int64_t)tq - 32;
        w = b >= 0 ? (uint32_t)b >> 4 : 0u;
      }
      v1 = w + 1u < nwords;
      v2 = w + 2u < nwords;
      next = (w & 3u) >= 2u && w + 4u - (w & 3u) < nwords;   // words in the next 16 bytes
      a0 = text + w;
    } else if (L.phase == PH_EXACT && K && ak == 0u && al == P.ix.n &&
               (L.pass ? L.len / 2 - L.i : L.i - L.len / 2 + 1) >= (int32_t)K) {
      if (L.bi & kKReady) {
        kind = F_KLOOK;
        const uint64_t *tab = L.dsel ? P.ix.kmer[1] : P.ix.kmer[0];
        a0 = reinterpret_cast<const uint32_t *>(tab + (L.bi & ~kKReady));
      } else {
        kind = F_KPREP;
        rlim = K;
      }
    } else if (TEXT && ak == al) {
      kind = F_SA;                   // a single row: continue on the text from its suffix
      a0 = (L.dsel ? P.ix.d[1].sa : P.ix.d[0].sa) + ak;
    } else {
      kind = F_LF;
      j1 = ak / kBlockRows;
      j2 = (al + 1u) / kBlockRows;
      // A size-1 interval needs only row k's block (Occ(s, k+1) = Occ(s, k) +
      // [BWT[k] == s]); otherwise the block of l+1 too when it differs.
      two = ak != al && j2 != j1;
    }
    // The read bases W[p], W[p+d], ... (rlim <= 32 of them) lie in the
    // aligned 16-byte chunk of W[p] and the one or two after it in the step's
    // direction (x1, x2); a lane reads x1 / x2 only when they hold one of
    // its bases.  On the compact index no step compares more than a k-mer
    // (at most 15 bases), so x2 is never needed there.
    // (The staged launch reads them from the lane's LDS slot instead.)
    const uint8_t *pa = LDS ? nullptr : L.W + p;
    const uint32_t mis = LDS ? 0u : (uint32_t)(reinterpret_cast<uintptr_t>(pa) & 15u);
    const uint8_t *pal = LDS ? reinterpret_cast<const uint8_t *>(P.rec + (size_t)L.read * (kRecWords / 4u))
                             : pa - mis;
    const uint8_t *x1 = d > 0 ? pal + 16 : pal - 16;
    const uint8_t *x2 = d > 0 ? pal + 32 : pal - 32;
    const bool n1 = !LDS && (d > 0 ? mis + rlim - 1u >= 16u : rlim - 1u > mis);
    const bool n2 = !LDS && TEXT && (d > 0 ? mis + rlim - 1u >= 32u : rlim - 1u > mis + 16u);
    // First line: the occupancy block of row k (four 16-byte loads), or the
    // aligned 16 bytes that hold the word(s) the kind reads (and the next 16
    // bytes when text words continue there).  Second line (the lanes whose
    // LF query needs it): the block of row l+1.
    const uint8_t *f0, *g0 = nullptr;
    if (kind == F_LF) {
      const OccBlock *blk = L.dsel ? P.ix.d[1].blk : P.ix.d[0].blk;
      f0 = reinterpret_cast<const uint8_t *>(blk + j1);
      g0 = reinterpret_cast<const uint8_t *>(blk + j2);
    } else {
      const uint8_t *w0 = kind == F_KPREP ? pal : reinterpret_cast<const uint8_t *>(a0);
      f0 = w0 - (reinterpret_cast<uintptr_t>(w0) & 15u);
    }
    const bool lf = kind == F_LF;

    // ---- the loads: one batch, no value used before the last is issued.
    // The loads only some lanes make come last, each group in the branch of
    // its lanes; a lane that skips one holds an opaque value there (no
    // constant for the compiler to fold into the branch, where it would
    // wait for the load; no instruction either), and those bytes or words
    // are never compared.
    //
    // Two forms of the batch, chosen at launch (results are the same): with
    // OWN every lane issues only the loads its step needs; otherwise every
    // lane issues the first seven, pointing those it does not need at the
    // chunk of W[p] or the first 16 bytes it reads (more L1 requests, fewer
    // instructions and branches — faster while the index's lines come from
    // the caches, slower once its accesses spread over tens of GB).
    uint4 rp = opaque4();
    if (!LDS) rp = *reinterpret_cast<const uint4 *>(pal);
    const uint4 fx0 = *reinterpret_cast<const uint4 *>(f0);
    uint4 fx1 = opaque4(), fx2 = opaque4(), fx3 = opaque4(), rx1 = opaque4(), rx2 = opaque4();
    uint4 gx0 = opaque4(), gx1 = opaque4(), gx2 = opaque4(), gx3 = opaque4();
    if (OWN) {
      if (lf || next) fx1 = *reinterpret_cast<const uint4 *>(f0 + 16);
      if (lf) {
        fx2 = *reinterpret_cast<const uint4 *>(f0 + 32);
        fx3 = *reinterpret_cast<const uint4 *>(f0 + 48);
      }
      if (n1) rx1 = *reinterpret_cast<const uint4 *>(x1);
      if (n2) rx2 = *reinterpret_cast<const uint4 *>(x2);
    } else {
      fx1 = *reinterpret_cast<const uint4 *>(lf || next ? f0 + 16 : f0);
      fx2 = *reinterpret_cast<const uint4 *>(lf ? f0 + 32 : f0);
      fx3 = *reinterpret_cast<const uint4 *>(lf ? f0 + 48 : f0);
      if (!LDS) {
        rx1 = *reinterpret_cast<const uint4 *>(n1 ? x1 : pal);
        if (TEXT) rx2 = *reinterpret_cast<const uint4 *>(n2 ? x2 : pal);
      }
    }
    if (two) {
      gx0 = *reinterpret_cast<const uint4 *>(g0);
      gx1 = *reinterpret_cast<const uint4 *>(g0 + 16);
      gx2 = *reinterpret_cast<const uint4 *>(g0 + 32);
      gx3 = *reinterpret_cast<const uint4 *>(g0 + 48);
    }
    // the 48 (32 on the compact index) bytes in address order; W[p] is byte wp
    const uint4 rc0 = d > 0 ? rp : (TEXT ? rx2 : rx1);
    const uint4 rc1 = d > 0 ? rx1 : (TEXT ? rx1 : rp);
    const uint4 rc2 = d > 0 ? rx2 : rp;
    const uint32_t wp = d > 0 ? mis : (TEXT ? 32u : 16u) + mis;

    // ---- the read window: 32 bases from W[p] in feed order as 2-bit
    // groups (rw, physical symbols) and their N flags (rb, bit i)
    uint64_t rw;
    uint32_t rb;
    if (LDS) {
      slot_window(p, d, rw, rb);
    } else if (!TEXT) {
      uint32_t w32;
      window16(rc0, rc1, wp, d > 0, w32, rb);
      rw = w32;
    } else {
    uint64_t clo = 0ull, nfl = 0ull;
    uint32_t chi = 0u;
    {
      const uint32_t v[12] = {rc0.x, rc0.y, rc0.z, rc0.w, rc1.x, rc1.y,
                              rc1.z, rc1.w, rc2.x, rc2.y, rc2.z, rc2.w};
      uint32_t any = 0u;                            // some byte > 3 (an N)
#pragma unroll
      for (uint32_t j = 0; j < 12u; ++j) {
        uint32_t t = v[j] & 0x03030303u;            // 4 codes -> 8 bits
        t = (t | (t >> 6)) & 0x000F000Fu;
        t = (t | (t >> 12)) & 0xFFu;
        if (j < 8u) clo |= (uint64_t)t << (8u * j);
        else chi |= t << (8u * (j - 8u));
        any |= v[j] & 0xFCFCFCFCu;
      }
      if (any) {                                    // the N flags, only where needed
#pragma unroll
        for (uint32_t j = 0; j < 12u; ++j) {
          uint32_t z = (v[j] >> 2) & 0x3F3F3F3Fu;   // a byte > 3 -> its bit 0
          z = z | (z >> 3);
          z = (z | (z >> 1) | (z >> 2)) & 0x01010101u;
          z = (z | (z >> 7)) & 0x00030003u;
          z = (z | (z >> 14)) & 0xFu;
          nfl |= (uint64_t)z << (4u * j);
        }
      }
    }
    if (d > 0) {                     // bytes wp .. wp+31
      const uint32_t sh = 2u * wp;
      rw = sh ? (clo >> sh) | ((uint64_t)chi << (64u - sh)) : clo;
      rb = (uint32_t)(nfl >> wp);
    } else {                         // W[p-31 .. p] = bytes wp-31 .. wp
      const uint32_t s0 = wp - 31u, sh = 2u * s0;    // s0 >= 1
      const uint64_t win = (clo >> sh) | ((uint64_t)chi << (64u - sh));
      const uint64_t v = __builtin_bitreverse64(win);
      rw = ((v >> 1) & 0x5555555555555555ull) | ((v & 0x5555555555555555ull) << 1);
      rb = __builtin_bitreverse32((uint32_t)(nfl >> s0));
    }
    }
    const uint32_t r0 = (rb & 1u) ? 4u : (uint32_t)(rw & 3u);   // W[p] (4 for any N)

    // Every load of the batch is waited for here, also those whose value
    // this lane's transition does not use: a load left outstanding at the end
    // of the iteration would make the next iteration wait, before it issues
    // its batch, for everything issued since, the records stored included.
    asm volatile("" ::"v"(rp.x), "v"(rx1.x), "v"(rx2.x), "v"(fx0.x), "v"(fx1.x), "v"(fx2.x),
                 "v"(fx3.x), "v"(gx0.x), "v"(gx1.x), "v"(gx2.x), "v"(gx3.x));

    // ---- the transitions
    if (kind == F_TEXT) {
      const uint32_t sel = (uint32_t)(reinterpret_cast<uintptr_t>(a0) & 15u) >> 2;
      const uint32_t t0 = pick4u(fx0, sel);
      const uint32_t t1 = !v1 ? 0u : (sel == 3u ? fx1.x : pick4u(fx0, sel + 1u));
      const uint32_t t2w = !v2 ? 0u : (sel >= 2u ? pick4u(fx1, sel - 2u) : pick4u(fx0, sel + 2u));
      const uint64_t tw = text_window(tq, t0, t1, t2w);
      uint64_t spread = rb;          // N flag of base i -> bit 2i
      spread = (spread | (spread << 16)) & 0x0000FFFF0000FFFFull;
      spread = (spread | (spread << 8)) & 0x00FF00FF00FF00FFull;
      spread = (spread | (spread << 4)) & 0x0F0F0F0F0F0F0F0Full;
      spread = (spread | (spread << 2)) & 0x3333333333333333ull;
      spread = (spread | (spread << 1)) & 0x5555555555555555ull;
      const uint32_t m = run_length(tw, L.comp ? ~rw : rw, spread, lim);
      text_apply(cont, tq, lim, m, r0, tq >= 1u ? (uint32_t)(tw & 3u) : 4u);
      return;
    }
    if (kind == F_SA) {
      if (COUNT) L.nlook++;
      const uint32_t q = pick4u(fx0, (uint32_t)(reinterpret_cast<uintptr_t>(a0) & 15u) >> 2);
      if (cont) L.cq = q;
      else L.tq = q;
      return;
    }
    uint32_t rk, rl;
    int32_t adv = 1;
    if (kind == F_KPREP) {
      const uint32_t kmask = (1u << (2u * K)) - 1u;
      const uint32_t code = (uint32_t)(L.comp ? ~rw : rw) & kmask;
      if ((rb & ((1u << K) - 1u)) == 0u) { L.bi = code | kKReady; return; }
      if (COUNT) L.nkmer++;
      end_pass();                    // an N: the interval is empty
      return;
    }
    if (kind == F_KLOOK) {
      const bool hi = (reinterpret_cast<uintptr_t>(a0) & 8u) != 0u;
      rk = hi ? fx0.z : fx0.x;
      rl = hi ? fx0.w : fx0.y;
      adv = (int32_t)K;
      if (COUNT) L.nkmer++;
      L.bi = 0u;
    } else {
    const uint32_t sym = r0;
    const uint32_t m2 = al + 1;
    const uint32_t t1 = ak - j1 * kBlockRows, t2 = m2 - j2 * kBlockRows;
    const bool single = ak == al;
    const uint32_t dollar_blk = L.dsel ? P.ix.d[1].dollar_blk : P.ix.d[0].dollar_blk;
    const uint32_t dollar_off = L.dsel ? P.ix.d[1].dollar_off : P.ix.d[0].dollar_off;
    const Blk B1 = blk_of4(fx0, fx1, fx2, fx3);
    const Blk B2 = two ? blk_of4(gx0, gx1, gx2, gx3) : B1;
    if (COUNT) { nblk_acc += two ? 2u : 1u; nq_acc += 1u; }
    const uint32_t dfix1 = (j1 == dollar_blk && dollar_off < t1) ? 1u : 0u;
    const uint32_t dfix2 = (j2 == dollar_blk && dollar_off < t2) ? 1u : 0u;
    const uint32_t srow = single ? sym_at(B1, dollar_blk, dollar_off, j1, t1) : 4u;

    if (L.phase == PH_SCAN) {
      uint32_t r1[4], r2[4];
      rank4(B1, t1, r1);
      r1[0] -= dfix1;
#pragma unroll
      for (uint32_t s = 0; s < 4; ++s) L.ck[s] = c_of(s) + 1u + pick4(B1.c, s) + r1[s];
      if (single) {
#pragma unroll
        for (uint32_t s = 0; s < 4; ++s) L.cl[s] = L.ck[s] - 1u + (srow == s ? 1u : 0u);
      } else {
        rank4(B2, t2, r2);
        r2[0] -= dfix2;
#pragma unroll
        for (uint32_t s = 0; s < 4; ++s) L.cl[s] = c_of(s) + pick4(B2.c, s) + r2[s];
      }
      L.wsym = sym;
      const int64_t size_node = (int64_t)L.l - (int64_t)L.k;
      if (sym < 4u) {
        const uint32_t ps = L.comp ? 3u - sym : sym;
        L.kn = pick4(L.ck, ps);
        L.ln = pick4(L.cl, ps);
      } else {
        L.kn = P.ix.n + 1;
        L.ln = P.ix.n;
      }
      L.tqn = kTqNone;
      if (size_node == (int64_t)L.ln - (int64_t)L.kn) {
        L.i += step();
        scan_next();
      } else {
        L.bi = 0;
        branch_next();
      }
      return;
    }

    if (sym < 4u) {
      const uint32_t s = L.comp ? 3u - sym : sym;
      const uint32_t c = c_of(s);
      rk = c + 1u + pick4(B1.c, s) + rank1(B1, s, t1) - (s == 0u ? dfix1 : 0u);
      if (single) rl = rk - 1u + (srow == s ? 1u : 0u);
      else rl = c + pick4(B2.c, s) + rank1(B2, s, t2) - (s == 0u ? dfix2 : 0u);
    } else {
      rk = P.ix.n + 1;
      rl = P.ix.n;
    }
    }
    if (!cont) {                     // PH_EXACT
      if (rk > rl) { end_pass(); return; }
      L.k = rk; L.l = rl;
      L.i += adv * step();
      if (in_read(L.i) && in_exact(L.i)) return;
      enter_scan();
      return;
    }
    if (rk > rl) {                   // continuation died
      L.bi++;
      branch_next();
      return;
    }
    L.qk = rk; L.ql = rl;
    L.cp += step();
    if (in_read(L.cp)) return;
    emit_branch();
    L.bi++;
    branch_next();
  }
};

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

template <bool PERSISTENT, bool MM_ONLY, bool COUNT, bool TEXT, bool OWN, bool LDS = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void search1_kernel(S1Params P) {
  // staged launch: word j of thread t's slot at slots[j * 256 + t], so the
  // lanes of a wavefront reading the same word hit different banks
  __shared__ uint32_t slots[LDS ? kSlotWords * 256u : 1u];
  S1Lane L;
  L.phase = PH_IDLE;
  L.nlook = 0;
  L.ntext = 0;
  L.nkmer = 0;
  S1<MM_ONLY, COUNT, TEXT, OWN, LDS> m{P, L, slots + threadIdx.x};
  unsigned long long nblk = 0, nq = 0;
  const uint32_t lane = lane_id();

  if (!PERSISTENT) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < P.ntasks) m.start(t);
    while (L.phase != PH_IDLE) {
      if (L.phase == PH_NEXT_STRAND) m.start_strand(1u);
      else m.advance(nblk, nq);
    }
  } else {
    // Wavefront-level work queue: the wave owns tasks [wq_next, wq_end) and
    // hands them to idle lanes by rank among the idle lanes (ballot +
    // popcount); a new chunk is taken with one atomic when it runs dry.
    // Lanes are refilled only once refill_min of them are idle (or the whole
    // wave is), so the task set-up runs for many lanes at a time.
    //
    // The next (up to) 64 tasks of the chunk are read ahead, one per lane
    // (their offsets and k-mer indexes, task_info, or their record's header
    // in the staged launch): lane i holds task wbase + i, and a lane that
    // takes a task gets it from its holder with register shuffles, so a
    // refill waits for no memory except once per 64 tasks (and, staged, for
    // the records of the tasks it starts).  With a task list the queue runs
    // over list positions.
    const uint32_t ntasks = P.list ? *P.list_n : P.ntasks;
    if (ntasks == 0u) return;   // an empty list: no wave touches the work counter
    uint32_t wq_next = 0, wq_end = 0, wbase = 0, wend = 0;
    uint64_t h_off = 0;
    uint32_t h_len = 0, h_kc0 = 0, h_kc1 = 0;
    bool more = true;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    while (true) {
      uint64_t idle = __ballot(L.phase == PH_IDLE);
      const uint32_t nidle0 = (uint32_t)__popcll(idle);
      if (nidle0 < P.refill_min && nidle0 < 64u) idle = 0ull;
      while (idle != 0ull && more) {
        if (wq_next >= wq_end) {
          uint32_t base = 0;
          if (lane == 0) base = atomicAdd(P.queue, P.chunk);
          base = __shfl(base, 0);
          if (base >= ntasks) { more = false; break; }
          wq_next = base;
          wq_end = min(base + P.chunk, ntasks);
        }
        if (wq_next >= wend) {       // read the next 64 tasks ahead
          wbase = wq_next;
          wend = min(wbase + 64u, wq_end);
          if (wbase + lane < wend) {
            const uint32_t h_task = P.list ? P.list[wbase + lane] : wbase + lane;
            if (LDS) {
              const uint32_t len = (uint32_t)(P.offsets[h_task + 1] - P.offsets[h_task]);
              h_len = len > kStagedMaxLen ? 0xFFFFFFFFu : len;   // longer: the unstaged kernel's
            } else {
              m.task_info(h_task, h_off, h_len, h_kc0, h_kc1);
            }
          }
        }
        const uint32_t nidle = (uint32_t)__popcll(idle);
        const uint32_t rank = (uint32_t)__popcll(idle & lt_mask);
        const uint32_t avail = wend - wq_next;
        const int src = (int)((wq_next - wbase + rank) & 63u);   // the holder of task wq_next + rank
        uint32_t task = wq_next + rank;
        if (P.list && rank < avail) task = P.list[task];
        const uint32_t len = __shfl(h_len, src);
        if (LDS) {
          if (L.phase == PH_IDLE && rank < avail && len != 0xFFFFFFFFu) m.stage_with(task, len);
        } else {
          const uint32_t kc0 = __shfl(h_kc0, src);
          const uint32_t kc1 = __shfl(h_kc1, src);
          const uint32_t off_lo = __shfl((uint32_t)h_off, src);
          const uint32_t off_hi = __shfl((uint32_t)(h_off >> 32), src);
          if (L.phase == PH_IDLE && rank < avail)
            m.start_with(task, ((uint64_t)off_hi << 32) | off_lo, len, kc0, kc1);
        }
        wq_next += min(avail, nidle);
        idle = __ballot(L.phase == PH_IDLE);
      }
      if (__ballot(L.phase != PH_IDLE) == 0ull) break;
      if (L.phase == PH_NEXT_STRAND) m.start_strand(1u);
      else if (L.phase != PH_IDLE) m.advance(nblk, nq);
    }
  }
  if (COUNT) {
    atomicAdd(&P.stats[0], nblk);
    atomicAdd(&P.stats[1], nq);
    atomicAdd(&P.stats[2], (unsigned long long)L.nlook);
    atomicAdd(&P.stats[3], (unsigned long long)L.ntext);
    atomicAdd(&P.stats[4], (unsigned long long)L.nkmer);
  }
}

// The rows of the records search1_kernel stored as text positions (k's high
// word kTextRowTag, l = q): k = l = ISA[q] in the record's direction (dir ^
// strand: the pass's direction).  One thread per (read, strand) slot.
__global__ __launch_bounds__(256) void text_rows_kernel(DevIndex ix, fmi_hit *hits,
                                                        const uint32_t *counts, uint32_t nslots,
                                                        uint32_t cap) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  const uint32_t c = min(counts[t], cap);
  uint4 *h = reinterpret_cast<uint4 *>(hits + (size_t)t * cap);
  for (uint32_t j = 0; j < c; ++j) {
    const uint4 a = h[2 * j];
    if (a.y != kTextRowTag) continue;
    const uint4 b = h[2 * j + 1];
    const uint32_t dsel = (b.z & 0xFFu) ^ (b.w & 0xFFu);
    const uint32_t row = (dsel ? ix.d[1].isa : ix.d[0].isa)[a.x];
    h[2 * j] = make_uint4(row, 0u, row, 0u);
  }
}

__global__ void zero_u32_kernel(uint32_t *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Byte fill (fmi_memset_d while a stream captures: no memset graph nodes).
__global__ void fill_u8_kernel(uint8_t *p, uint8_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// --------------------------------------------------------------------------
// mapping-layer 1-error search: BWSearch1VectorHelper (search/search.c:
// 1491-1709) over the interval vectors of bwt.c:1940-1950 (or, for a seed
// window, bwt.c:1447-1457)
// --------------------------------------------------------------------------
//
// One lane per (read, strand) task at a time runs the helper (vh_task,
// fmi_vhelper.h) over the read or its window; the lanes stride over the
// tasks, and each owns a vector scratch slice of kMaxReadLen + 3 intervals,
// so the scratch is sized by the lane count alone (no read of the offsets
// on the host).

struct VhParams {
  DevIndex ix;
  const uint8_t *reads;
  const uint64_t *offsets;
  const int16_t *bounds;    // NULL or [start, end] per read, relative to the read
  fmi_hit *hits;
  uint32_t *counts;
  uint2 *scratch;           // lanes * (kMaxReadLen + 3) intervals
  uint32_t nreads, cap, strand_mask, mm_only;
};

__global__ __launch_bounds__(256) void vhelper_kernel(VhParams P) {
  const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lanes = gridDim.x * blockDim.x;
  uint2 *vb = P.scratch + (uint64_t)lane * (kMaxReadLen + 3u);
  for (uint32_t t = lane; t < 2u * P.nreads; t += lanes) {
    const uint32_t rd = t >> 1, strand = t & 1u;
    const uint64_t off = P.offsets[rd];
    const uint64_t len64 = P.offsets[rd + 1] - off;
    const int32_t len = len64 > kMaxReadLen ? -1 : (int32_t)len64;
    const uint8_t *W = P.reads + off;
    int32_t start = 0, end = len - 1;
    if (P.bounds) { start = P.bounds[2 * rd]; end = P.bounds[2 * rd + 1]; }
    if (!(P.strand_mask & (1u << strand)) || len < 0 || start < 0 || end >= len ||
        end - start < 2) {
      P.counts[t] = 0;
      continue;
    }
    P.counts[t] = vh_task(P.ix, W, start, end, strand, P.mm_only, vb,
                          P.hits + (uint64_t)t * P.cap, P.cap);
  }
}

// --------------------------------------------------------------------------
// k-mer interval table
// --------------------------------------------------------------------------
// The interval of every string of k symbols fed to LF from the whole index
// (symbol j at bits 2j..2j+1 of the entry's index), so an exact search can
// take its first k steps in one read.  Entry = k | l << 32; an empty
// interval is stored as (1, 0).

__global__ __launch_bounds__(256) void kmer_table_kernel(DevDir D, uint32_t n, uint32_t k,
                                                         uint64_t count, uint64_t *out) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < count;
       c += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0u, hi = n, nb;
    for (uint32_t j = 0; j < k; ++j) {
      lf_one(D, lo, hi, (uint32_t)(c >> (2u * j)) & 3u, lo, hi, nb);
      if (lo > hi) break;
    }
    out[c] = lo <= hi ? ((uint64_t)hi << 32 | lo) : 1ull;
  }
}

// --------------------------------------------------------------------------
// locate: getScompValue (search/csafm.h:208-235)
// --------------------------------------------------------------------------
// Each LF step of the walk needs only the row's own block: the symbol b at
// row i and Occ(b, i+1) = Occ(b, i) + 1.

struct LocateParams {
  DevIndex ix;
  uint32_t dsel, nintervals, cap, stride;   // stride: int64 words between intervals
  const int64_t *k, *l;
  uint64_t *pos;
  uint32_t *nrows;
};

__global__ __launch_bounds__(256) void locate_kernel(LocateParams P) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t iv = g / P.cap;
  const uint32_t j = (uint32_t)(g - iv * P.cap);
  if (iv >= P.nintervals) return;
  const int64_t k = P.k[iv * P.stride], l = P.l[iv * P.stride];
  // Rows outside 0..n are not in the index: such an interval yields no rows.
  const bool valid = k >= 0 && l <= (int64_t)P.ix.n;
  const int64_t cnt = (valid && k <= l) ? (l - k + 1) : 0;
  const uint32_t rows = cnt < (int64_t)P.cap ? (uint32_t)cnt : P.cap;
  if (j == 0) P.nrows[iv] = rows;
  if (j >= rows) return;
  const DevDir &D = dir_of(P.ix, P.dsel);
  P.pos[iv * P.cap + j] = sa_value(D, P.ix.ratio, P.ix.n, (uint32_t)(k + j));
}

// --------------------------------------------------------------------------
// inverse locate: getRcompValue (search/csafm.h:237-272)
// --------------------------------------------------------------------------

struct IsaParams {
  DevIndex ix;
  uint32_t dsel;
  uint64_t count;
  const uint64_t *tpos;
  uint64_t *rows;
};

__global__ __launch_bounds__(256) void isa_kernel(IsaParams P) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= P.count) return;
  const uint64_t m = P.tpos[g];
  // Text positions beyond n are not in the index.
  if (m > (uint64_t)P.ix.n) {
    P.rows[g] = ~0ull;
    return;
  }
  P.rows[g] = isa_value(dir_of(P.ix, P.dsel), P.ix.ratio, P.ix.n, (uint32_t)m);
}

// --------------------------------------------------------------------------
// full SA / ISA and 2-bit text from the samples (fmi_index_expand)
// --------------------------------------------------------------------------
// Walk w starts at the ISA sample of text position w * ratio and steps LF
// back over the ratio - 1 positions below it (position 0 alone for w = 0);
// one more walk starts at row 0 (the suffix "$" at position n) and covers
// the positions above the last sample.  Every position, and so every row, is
// written exactly once.

__global__ __launch_bounds__(256) void expand_walk_kernel(DevDir D, uint32_t n, uint32_t ratio,
                                                          uint32_t nsamp, uint32_t *sa,
                                                          uint32_t *isa) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w > nsamp) return;
  uint32_t pos, row, cnt;
  if (w < nsamp) {
    pos = w * ratio;
    row = D.R[w];
    cnt = w == 0 ? 1u : ratio;
  } else {
    pos = n;
    row = 0;
    cnt = n - (nsamp - 1) * ratio;
  }
  for (uint32_t t = 0; t < cnt; ++t) {
    isa[pos] = row;
    sa[row] = pos;
    if (t + 1 == cnt) break;
    // row of suffix pos - 1: its BWT symbol is X[pos - 1] (pos >= 1 here)
    const uint32_t j = row / kBlockRows, o = row - j * kBlockRows;
    const Blk B = load_block(D.blk, j);
    const uint32_t c = code_at(B, o);
    row = pick4(D.C, c) + occ_blk(B, D.dollar_blk, D.dollar_off, j, c, o) + 1u;
    --pos;
  }
}

// X[j] is the BWT symbol of the row of suffix j + 1.
__global__ __launch_bounds__(256) void pack_text_kernel(DevDir D, uint32_t n, const uint32_t *isa,
                                                        uint32_t *text, uint32_t nwords) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwords) return;
  uint32_t word = 0;
  for (uint32_t t = 0; t < 16u; ++t) {
    const uint32_t x = w * 16u + t;
    if (x >= n) break;
    const uint32_t row = isa[x + 1];
    const uint32_t j = row / kBlockRows;
    word |= code_at(load_block(D.blk, j), row - j * kBlockRows) << (2u * t);
  }
  text[w] = word;
}

// --------------------------------------------------------------------------
// position resolution (bwt.c:2158-2166, binsearch search/io.h:43-64)
// --------------------------------------------------------------------------

struct ResolveParams {
  DevIndex ix;
  const fmi_hit *hits;
  const uint32_t *counts;
  const uint64_t *offsets;
  const uint64_t *koff;     // kcount + 1 entries
  const uint64_t *kstart;
  uint32_t nslots, cap, row_cap;
  uint32_t kcount;
  uint32_t record_len;      // 0: the read's length (bwt.c:2031-2036); 1: end - start + 1
  int32_t *chrom;
  uint64_t *coord;
};

__global__ __launch_bounds__(256) void resolve_kernel(ResolveParams P) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t per_slot = (uint64_t)P.cap * P.row_cap;
  const uint64_t slot = g / per_slot;
  if (slot >= P.nslots) return;
  const uint32_t rem = (uint32_t)(g - slot * per_slot);
  const uint32_t h = rem / P.row_cap, j = rem - h * P.row_cap;
  const uint32_t stored = min(P.counts[slot], P.cap);
  const uint64_t o = slot * per_slot + (uint64_t)h * P.row_cap + j;
  if (h >= stored) return;
  const fmi_hit hit = P.hits[slot * P.cap + h];
  const bool valid = hit.k >= 0 && hit.l <= (int64_t)P.ix.n;
  const int64_t nrows = (valid && hit.k <= hit.l) ? hit.l - hit.k + 1 : 0;
  if ((int64_t)j >= nrows) { P.chrom[o] = -1; P.coord[o] = 0; return; }
  const uint32_t read = (uint32_t)(slot >> 1), strand = (uint32_t)(slot & 1u);
  // bwt.c:2031-2036 measure the read; manage_single_result measures the
  // record's own window (enW = end - start + 1, search/results.h:146-152)
  const uint64_t len = P.record_len ? (uint64_t)(int64_t)(hit.end - hit.start + 1)
                                    : P.offsets[read + 1] - P.offsets[read];
  uint64_t len_calc = len;
  if (hit.num_mismatches) {
    if (hit.err_kind == FMI_DELETION) len_calc--;
    else if (hit.err_kind == FMI_INSERTION) len_calc++;
  }
  const uint32_t direction = strand ? (hit.dir ? 0u : 1u) : hit.dir;
  const uint32_t row = (uint32_t)(hit.k + j);
  uint64_t key;
  if (direction) key = (uint64_t)P.ix.n + 1 - sa_value(P.ix.d[1], P.ix.ratio, P.ix.n, row) - len_calc - 1;
  else key = sa_value(P.ix.d[0], P.ix.ratio, P.ix.n, row);
  // upper_bound of key in koff[0 .. kcount)
  uint64_t lo = 0, size = P.kcount;
  while (size > 0) {
    const uint64_t w = size / 2;
    if (P.koff[lo + w] <= key) { lo += w + 1; size -= w + 1; }
    else size = w;
  }
  if (lo >= 1 && key + len_calc <= P.koff[lo]) {
    P.chrom[o] = (int32_t)(lo - 1);
    P.coord[o] = P.kstart[lo - 1] + key - P.koff[lo - 1];
  } else {
    P.chrom[o] = -1;
    P.coord[o] = 0;
  }
}

// --------------------------------------------------------------------------
// record compaction: the stored records of every (read, strand) slot,
// slot after slot, in one dense array (exclusive scan of min(count, cap))
// --------------------------------------------------------------------------
// Three launches: per-1024-slot sums; one workgroup scans the block sums;
// each block rescans its slots and copies its records.

constexpr uint32_t kScanBlock = 1024;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *lds, uint32_t *total) {
  const uint32_t t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (uint32_t d = 1; d < blockDim.x; d <<= 1) {
    const uint32_t x = t >= d ? lds[t - d] : 0u;
    __syncthreads();
    lds[t] += x;
    __syncthreads();
  }
  const uint32_t incl = lds[t];
  *total = lds[blockDim.x - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kScanBlock) void compact_sums_kernel(const uint32_t *counts,
                                                                 uint32_t nslots, uint32_t cap,
                                                                 uint32_t *block_sums) {
  __shared__ uint32_t lds[kScanBlock];
  const uint32_t s = blockIdx.x * kScanBlock + threadIdx.x;
  const uint32_t v = s < nslots ? min(counts[s], cap) : 0u;
  uint32_t total;
  block_exclusive_scan(v, lds, &total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void compact_scan_kernel(uint32_t *block_sums,
                                                                 uint32_t nblocks,
                                                                 unsigned long long *total_out) {
  __shared__ uint32_t lds[kScanBlock];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nblocks; base += kScanBlock) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < nblocks ? block_sums[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, lds, &total);
    if (i < nblocks) block_sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(kScanBlock) void compact_copy_kernel(const uint32_t *counts,
                                                                 const fmi_hit *hits,
                                                                 uint32_t nslots, uint32_t cap,
                                                                 const uint32_t *block_offsets,
                                                                 fmi_hit *out) {
  __shared__ uint32_t lds[kScanBlock];
  const uint32_t s = blockIdx.x * kScanBlock + threadIdx.x;
  const uint32_t v = s < nslots ? min(counts[s], cap) : 0u;
  uint32_t total;
  const uint32_t ex = block_exclusive_scan(v, lds, &total);
  if (s >= nslots) return;
  const uint64_t dst = (uint64_t)block_offsets[blockIdx.x] + ex;
  const uint4 *src = reinterpret_cast<const uint4 *>(hits + (uint64_t)s * cap);
  uint4 *d = reinterpret_cast<uint4 *>(out + dst);
  for (uint32_t j = 0; j < 2u * v; ++j) d[j] = src[j];
}

// --------------------------------------------------------------------------
// random-gather roofline probe
// --------------------------------------------------------------------------

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33; x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

// Each lane issues the loads of four random lines (VEC 16-byte pieces each)
// before consuming any, so 4*VEC loads are in flight per lane; line indexes
// come from a 64-bit mix over a power-of-two line count (a mask).
template <int VEC>
__global__ __launch_bounds__(256) void gather_kernel(const uint4 *buf, uint64_t line_mask,
                                                     uint64_t per_thread,
                                                     unsigned long long *sink) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t st = mix64(tid * 0x9E3779B97F4A7C15ull + 1);
  uint32_t acc = 0;
  for (uint64_t it = 0; it < per_thread; it += 4) {
    uint4 x[4][VEC];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      st = mix64(st + u + 1);
      const uint4 *p = buf + (st & line_mask) * VEC;
#pragma unroll
      for (int v = 0; v < VEC; ++v) x[u][v] = p[v];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < VEC; ++v) acc ^= x[u][v].x ^ x[u][v].y ^ x[u][v].z ^ x[u][v].w;
  }
  if (acc == 0x9E3779B9u) atomicAdd(sink, 1ull);   // keeps the loads live
}

// 64-byte lines at random, ILP lines in flight per lane (COOP = false: each
// lane loads its own lines as four 16-byte pieces) or per quad of lanes
// (COOP = true: the four lanes of a quad load one 16-byte piece each of the
// quad's line, so one instruction fetches 16 whole lines).  per_thread
// iterations of ILP loads; lines loaded = threads * per_thread (COOP: / 4).
template <int ILP, bool COOP>
__global__ __launch_bounds__(256) void gather64_kernel(const uint4 *buf, uint64_t line_mask,
                                                       uint64_t per_thread,
                                                       unsigned long long *sink) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t piece = COOP ? (threadIdx.x & 3u) : 0u;
  uint64_t st = mix64((COOP ? tid >> 2 : tid) * 0x9E3779B97F4A7C15ull + 1);
  uint32_t acc = 0;
  for (uint64_t it = 0; it < per_thread; it += ILP) {
    if (COOP) {
      uint4 x[ILP];
#pragma unroll
      for (int u = 0; u < ILP; ++u) {
        st = mix64(st + u + 1);
        x[u] = buf[(st & line_mask) * 4 + piece];
      }
#pragma unroll
      for (int u = 0; u < ILP; ++u) acc ^= x[u].x ^ x[u].y ^ x[u].z ^ x[u].w;
    } else {
      uint4 x[ILP][4];
#pragma unroll
      for (int u = 0; u < ILP; ++u) {
        st = mix64(st + u + 1);
        const uint4 *p = buf + (st & line_mask) * 4;
#pragma unroll
        for (int v = 0; v < 4; ++v) x[u][v] = p[v];
      }
#pragma unroll
      for (int u = 0; u < ILP; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc ^= x[u][v].x ^ x[u][v].y ^ x[u][v].z ^ x[u][v].w;
    }
  }
  if (acc == 0x9E3779B9u) atomicAdd(sink, 1ull);   // keeps the loads live
}

// Streaming copy of n16 16-byte words, grid-stride: a kernel of known HBM
// traffic (n16*16 bytes read and written) that calibrates the profiler's
// fetch/write counters and measures the streaming bandwidth.
__global__ __launch_bounds__(256) void copy_kernel(const uint4 *__restrict__ src,
                                                   uint4 *__restrict__ dst, uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    dst[i] = src[i];
}

// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------

hipError_t launch_exact(const DevIndex &ix, uint32_t dsel, uint32_t comp, uint32_t direction,
                        const uint8_t *reads, const uint64_t *offsets, uint32_t nreads,
                        fmi_hit *out, fmi_hit *anchor, hipStream_t s, unsigned long long *work) {
  ExactParams P{ix, dsel, comp, direction, nreads, reads, offsets, out, anchor, work};
  const uint32_t grid = (nreads + 255) / 256;
  if (work) hipLaunchKernelGGL(exact_kernel<true>, dim3(grid), dim3(256), 0, s, P);
  else hipLaunchKernelGGL(exact_kernel<false>, dim3(grid), dim3(256), 0, s, P);
  return hipGetLastError();
}

hipError_t launch_vector(const DevIndex &ix, uint32_t dsel, uint32_t comp, uint32_t direction,
                         const uint8_t *reads, const uint64_t *offsets, uint32_t nreads,
                         int64_t *vk, int64_t *vl, int16_t *last_good, hipStream_t s) {
  VectorParams P{ix, dsel, comp, direction, nreads, reads, offsets, vk, vl, last_good};
  const uint32_t grid = (nreads + 255) / 256;
  hipLaunchKernelGGL(vector_kernel, dim3(grid), dim3(256), 0, s, P);
  return hipGetLastError();
}

uint64_t search1_staged_scratch_bytes(uint32_t nreads) {
  return ((64ull + 4ull * nreads + 127ull) & ~127ull) + (uint64_t)kRecWords * 4ull * nreads;
}

hipError_t launch_read_records(const uint8_t *reads, const uint64_t *offsets, uint32_t nreads,
                               uint32_t *rec, uint32_t *list, uint32_t *list_n, hipStream_t s) {
  static_assert(kRecWords == kReadRecordWords, "read record size");
  if (nreads == 0) return hipSuccess;
  RecParams R{reads, offsets, nreads, rec, list, list_n};
  hipLaunchKernelGGL(read_records_kernel, dim3((uint32_t)((8ull * nreads + 255) / 256)), dim3(256),
                     0, s, R);
  return hipGetLastError();
}

hipError_t launch_search1(const DevIndex &ix, bool mm_only, bool persistent, bool count,
                          uint32_t strand_mask, const uint8_t *reads, const uint64_t *offsets,
                          uint32_t nreads, uint32_t cap, fmi_hit *hits, uint32_t *counts,
                          uint32_t *queue, unsigned long long *stats, uint32_t grid_persistent,
                          uint32_t chunk, uint32_t refill_min, bool own_loads, hipStream_t s,
                          void *staged) {
  const uint32_t ntasks = nreads;   // one task per read (both strands in turn)
  S1Params P{ix, reads, offsets, hits, counts, queue, stats, ntasks, strand_mask, cap, chunk,
             refill_min, nullptr, nullptr, nullptr};
  if (ntasks == 0) return hipSuccess;
  const bool text = ix.text != nullptr;   // expanded index: text mode
  if (persistent) {
    if (grid_persistent == 0 || chunk == 0 || refill_min == 0 || refill_min > 64)
      return hipErrorInvalidValue;
    const dim3 g(grid_persistent), b(256);
    if (staged && !count && ntasks <= kStagedMaxTasks) {
      // scratch: [0] the unstaged kernel's work counter, [1] the list's
      // length; the list from byte 64; the records from the next 128 bytes
      uint32_t *head = static_cast<uint32_t *>(staged);
      uint32_t *list = head + 16;
      uint32_t *rec = reinterpret_cast<uint32_t *>(static_cast<char *>(staged) +
                                                   ((64ull + 4ull * ntasks + 127ull) & ~127ull));
      hipError_t e = launch_zero(head, 2, s);
      if (e != hipSuccess) return e;
      RecParams R{reads, offsets, ntasks, rec, list, head + 1};
      hipLaunchKernelGGL(read_records_kernel, dim3((uint32_t)((8ull * ntasks + 255) / 256)),
                         dim3(256), 0, s, R);
      P.rec = reinterpret_cast<const uint4 *>(rec);
#define FMI_S1_STAGED(MM)                                                                      \
  do {                                                                                         \
    if (!text) hipLaunchKernelGGL((search1_kernel<true, MM, false, false, false, true>), g, b, 0, s, P); \
    else if (own_loads)                                                                        \
      hipLaunchKernelGGL((search1_kernel<true, MM, false, true, true, true>), g, b, 0, s, P);  \
    else hipLaunchKernelGGL((search1_kernel<true, MM, false, true, false, true>), g, b, 0, s, P); \
  } while (0)
      if (mm_only) FMI_S1_STAGED(true);
      else FMI_S1_STAGED(false);
#undef FMI_S1_STAGED
      // the reads longer than a slot, unstaged, from the list
      P.rec = nullptr;
      P.list = list;
      P.list_n = head + 1;
      P.queue = head;
    }
#define FMI_S1_LAUNCH(MM, CNT)                                                                \
  do {                                                                                        \
    if (!text) hipLaunchKernelGGL((search1_kernel<true, MM, CNT, false, false>), g, b, 0, s, P); \
    else if (own_loads)                                                                       \
      hipLaunchKernelGGL((search1_kernel<true, MM, CNT, true, true>), g, b, 0, s, P);         \
    else hipLaunchKernelGGL((search1_kernel<true, MM, CNT, true, false>), g, b, 0, s, P);     \
  } while (0)
    if (mm_only) {
      if (count) FMI_S1_LAUNCH(true, true);
      else FMI_S1_LAUNCH(true, false);
    } else {
      if (count) FMI_S1_LAUNCH(false, true);
      else FMI_S1_LAUNCH(false, false);
    }
#undef FMI_S1_LAUNCH
  } else {
    const dim3 g((ntasks + 255) / 256), b(256);
    if (mm_only) {
      if (text) hipLaunchKernelGGL((search1_kernel<false, true, false, true, true>), g, b, 0, s, P);
      else hipLaunchKernelGGL((search1_kernel<false, true, false, false, false>), g, b, 0, s, P);
    } else {
      if (text) hipLaunchKernelGGL((search1_kernel<false, false, false, true, true>), g, b, 0, s, P);
      else hipLaunchKernelGGL((search1_kernel<false, false, false, false, false>), g, b, 0, s, P);
    }
  }
  if (!count && ix.text != nullptr && cap > 0) {
    const uint64_t nslots = 2ull * ntasks;
    hipLaunchKernelGGL(text_rows_kernel, dim3((uint32_t)((nslots + 255) / 256)), dim3(256), 0, s,
                       ix, hits, counts, (uint32_t)nslots, cap);
  }
  return hipGetLastError();
}

uint64_t compact_scratch_bytes(uint32_t nslots) {
  return ((uint64_t)(nslots + kScanBlock - 1) / kScanBlock) * 4 + 64;
}

hipError_t launch_compact(const uint32_t *counts, const fmi_hit *hits, uint32_t nslots,
                          uint32_t cap, fmi_hit *out, unsigned long long *total, void *scratch,
                          hipStream_t s) {
  const uint32_t nblocks = (nslots + kScanBlock - 1) / kScanBlock;
  if (nblocks == 0 || cap == 0) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t *>(total), 2u);
    return hipGetLastError();
  }
  uint32_t *sums = static_cast<uint32_t *>(scratch);
  hipLaunchKernelGGL(compact_sums_kernel, dim3(nblocks), dim3(kScanBlock), 0, s, counts, nslots, cap, sums);
  hipLaunchKernelGGL(compact_scan_kernel, dim3(1), dim3(kScanBlock), 0, s, sums, nblocks, total);
  hipLaunchKernelGGL(compact_copy_kernel, dim3(nblocks), dim3(kScanBlock), 0, s, counts, hits, nslots,
                     cap, sums, out);
  return hipGetLastError();
}

hipError_t launch_fill(void *p, uint8_t v, uint64_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  const uint64_t grid = std::min<uint64_t>((bytes + 255) / 256, 65536);
  hipLaunchKernelGGL(fill_u8_kernel, dim3((uint32_t)grid), dim3(256), 0, s,
                     static_cast<uint8_t *>(p), v, bytes);
  return hipGetLastError();
}

hipError_t launch_zero(uint32_t *p, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

uint32_t vhelper_lanes(uint32_t nreads) {
  const uint64_t want = (2ull * nreads + 255) / 256 * 256;
  return (uint32_t)std::min<uint64_t>(want, kVhelperMaxLanes);
}

uint64_t vhelper_scratch_bytes(uint32_t lanes) {
  return (uint64_t)lanes * (kMaxReadLen + 3u) * sizeof(uint2);
}

hipError_t launch_vhelper(const DevIndex &ix, bool mm_only, uint32_t strand_mask,
                          const uint8_t *reads, const uint64_t *offsets, const int16_t *bounds,
                          uint32_t nreads, uint32_t cap, fmi_hit *hits, uint32_t *counts,
                          void *scratch, uint32_t lanes, hipStream_t s) {
  VhParams P{ix, reads, offsets, bounds, hits, counts, static_cast<uint2 *>(scratch), nreads, cap,
             strand_mask, mm_only ? 1u : 0u};
  if (nreads == 0 || lanes == 0) return hipSuccess;
  hipLaunchKernelGGL(vhelper_kernel, dim3(lanes / 256), dim3(256), 0, s, P);
  return hipGetLastError();
}

hipError_t launch_locate(const DevIndex &ix, uint32_t dsel, const int64_t *k, const int64_t *l,
                         uint32_t stride, uint32_t nintervals, uint32_t cap, uint64_t *pos,
                         uint32_t *nrows, hipStream_t s) {
  LocateParams P{ix, dsel, nintervals, cap, stride, k, l, pos, nrows};
  const uint64_t total = (uint64_t)nintervals * cap;
  const uint64_t grid = (total + 255) / 256;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(locate_kernel, dim3((uint32_t)grid), dim3(256), 0, s, P);
  return hipGetLastError();
}

hipError_t launch_expand(const DevDir &D, uint32_t n, uint32_t ratio, uint32_t *sa, uint32_t *isa,
                         hipStream_t s) {
  const uint32_t nsamp = n / ratio + 1;
  const uint32_t grid = (nsamp + 1 + 255) / 256;
  hipLaunchKernelGGL(expand_walk_kernel, dim3(grid), dim3(256), 0, s, D, n, ratio, nsamp, sa, isa);
  return hipGetLastError();
}

hipError_t launch_kmer_table(const DevDir &D, uint32_t n, uint32_t k, uint64_t *out, hipStream_t s) {
  const uint64_t count = 1ull << (2 * k);
  const uint64_t grid = std::min<uint64_t>((count + 255) / 256, 65536ull);
  hipLaunchKernelGGL(kmer_table_kernel, dim3((uint32_t)grid), dim3(256), 0, s, D, n, k, count, out);
  return hipGetLastError();
}

hipError_t launch_pack_text(const DevDir &D, uint32_t n, const uint32_t *isa, uint32_t *text,
                            hipStream_t s) {
  const uint32_t nwords = (n + 15) / 16;
  if (nwords == 0) return hipSuccess;
  hipLaunchKernelGGL(pack_text_kernel, dim3((nwords + 255) / 256), dim3(256), 0, s, D, n, isa, text,
                     nwords);
  return hipGetLastError();
}

hipError_t launch_isa(const DevIndex &ix, uint32_t dsel, const uint64_t *tpos, uint64_t count,
                      uint64_t *rows, hipStream_t s) {
  IsaParams P{ix, dsel, count, tpos, rows};
  const uint64_t grid = (count + 255) / 256;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(isa_kernel, dim3((uint32_t)grid), dim3(256), 0, s, P);
  return hipGetLastError();
}

hipError_t launch_resolve(const DevIndex &ix, const fmi_hit *hits, const uint32_t *counts,
                          const uint64_t *offsets, uint32_t nreads, uint32_t cap,
                          uint32_t row_cap, const uint64_t *koff, const uint64_t *kstart,
                          uint32_t kcount, uint32_t record_len, int32_t *chrom, uint64_t *coord,
                          hipStream_t s) {
  ResolveParams P{ix, hits, counts, offsets, koff, kstart, 2u * nreads, cap, row_cap, kcount,
                  record_len, chrom, coord};
  const uint64_t total = (uint64_t)2 * nreads * cap * row_cap;
  const uint64_t grid = (total + 255) / 256;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(resolve_kernel, dim3((uint32_t)grid), dim3(256), 0, s, P);
  return hipGetLastError();
}

hipError_t launch_gather(const void *buf, uint64_t bytes, uint32_t line_bytes, uint64_t nloads,
                         unsigned long long *sink, uint32_t grid, hipStream_t s) {
  uint64_t nlines = 1;
  while (nlines * 2 <= bytes / line_bytes) nlines *= 2;   // power of two <= bytes/line
  const uint64_t threads = (uint64_t)grid * 256;
  uint64_t per_thread = (nloads + threads - 1) / threads;
  per_thread = (per_thread + 3) & ~3ull;
  const uint4 *b = (const uint4 *)buf;
  switch (line_bytes) {
    case 16: hipLaunchKernelGGL(gather_kernel<1>, dim3(grid), dim3(256), 0, s, b, nlines - 1, per_thread, sink); break;
    case 32: hipLaunchKernelGGL(gather_kernel<2>, dim3(grid), dim3(256), 0, s, b, nlines - 1, per_thread, sink); break;
    case 64: hipLaunchKernelGGL(gather_kernel<4>, dim3(grid), dim3(256), 0, s, b, nlines - 1, per_thread, sink); break;
    case 128: hipLaunchKernelGGL(gather_kernel<8>, dim3(grid), dim3(256), 0, s, b, nlines - 1, per_thread, sink); break;
    case 256: hipLaunchKernelGGL(gather_kernel<16>, dim3(grid), dim3(256), 0, s, b, nlines - 1, per_thread, sink); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Iterations per lane of gather64_kernel for at least nlines lines.
static uint64_t probe_per_thread(uint64_t nlines, uint32_t grid, uint32_t ilp, bool coop) {
  const uint64_t units = (uint64_t)grid * 256 / (coop ? 4 : 1);
  uint64_t per = (nlines + units - 1) / units;
  return (per + ilp - 1) / ilp * ilp;
}

uint64_t gather_probe_lines(uint64_t nlines, uint32_t grid, uint32_t ilp, bool coop) {
  return probe_per_thread(nlines, grid, ilp, coop) * ((uint64_t)grid * 256 / (coop ? 4 : 1));
}

hipError_t launch_gather_probe(const void *buf, uint64_t bytes, uint32_t ilp, bool coop,
                               uint64_t nlines, unsigned long long *sink, uint32_t grid,
                               hipStream_t s) {
  uint64_t n = 1;
  while (n * 2 <= bytes / 64) n *= 2;
  const uint64_t per = probe_per_thread(nlines, grid, ilp, coop);
  const uint4 *b = (const uint4 *)buf;
#define FMI_PROBE(I, C) \
  hipLaunchKernelGGL((gather64_kernel<I, C>), dim3(grid), dim3(256), 0, s, b, n - 1, per, sink)
  switch (ilp * 2 + (coop ? 1 : 0)) {
    case 2: FMI_PROBE(1, false); break;
    case 3: FMI_PROBE(1, true); break;
    case 4: FMI_PROBE(2, false); break;
    case 5: FMI_PROBE(2, true); break;
    case 8: FMI_PROBE(4, false); break;
    case 9: FMI_PROBE(4, true); break;
    case 16: FMI_PROBE(8, false); break;
    case 17: FMI_PROBE(8, true); break;
    case 32: FMI_PROBE(16, false); break;
    case 33: FMI_PROBE(16, true); break;
    default: return hipErrorInvalidValue;
  }
#undef FMI_PROBE
  return hipGetLastError();
}

hipError_t launch_copy(const void *src, void *dst, uint64_t bytes, uint32_t grid, hipStream_t s) {
  const uint64_t n16 = bytes / 16;
  if (n16 == 0 || grid == 0) return hipSuccess;
  hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst,
                     n16);
  return hipGetLastError();
}

uint64_t gather_loads_done(uint64_t nloads, uint32_t grid) {
  const uint64_t threads = (uint64_t)grid * 256;
  uint64_t per_thread = (nloads + threads - 1) / threads;
  per_thread = (per_thread + 3) & ~3ull;
  return per_thread * threads;
}

}  // namespace fmi

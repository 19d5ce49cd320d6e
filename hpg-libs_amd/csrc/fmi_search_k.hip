// fmi_search_k.hip -- the k-error search BWSearchCPU (search/search.c:
// 1171-1405) on the GPU: the read is cut into len / fragsize fragments and
// searched with up to fragments - 1 errors from every fragment in turn,
// breadth-first over bounded result lists (partial exact extension with the
// D lower bounds, branch steps, changes of direction through SA / ISA, final
// extension), exactly as the reference orders, truncates and updates them.
//
// One lane runs one (read, strand) task at a time and takes the next from a
// device counter.  Its four working lists and D live in a scratch slice of
// the lane; the final list and the anchors are written straight to the
// caller's output slots of the task.  The control flow is the reference's,
// so lanes diverge freely: this path is about identical results for the
// k = 2 reads of a mixed batch, not about occupancy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>

#include "../../include/fmi.h"
#include "fmi_device.h"
#include "fmi_internal.h"
#include "fmi_kernels.h"

namespace fmi {

namespace {

constexpr uint32_t kFrontPasses = 3;
constexpr uint32_t kFrontLen = 256;
// A lane's D: FMI_MAX_READ_LEN bytes, or the front passes' D side by side
// (kFrontPasses x kFrontLen), 16-byte aligned.
constexpr uint32_t kDStride = 1024;
static_assert(kDStride >= FMI_MAX_READ_LEN && kDStride >= kFrontPasses * kFrontLen, "D stride");
struct KList {
  fmi_hitk *list;
  uint32_t n, cap;
};

struct KView {
  const DevDir *d;
  uint32_t comp, dsel;               // dsel: 0 BWT of X, 1 BWT of reverse X
};

struct KCtx {
  const DevIndex *ix;
  // the counting launch's work counters (KWORK), else null: what the launch
  // reads at random, in the line model of the 1-error search's count
  uint32_t *wk;
  KView B, F;                        // the search's `backward` / `forward`
  const uint8_t *W;
  int32_t nW;
  uint8_t *D;
  bool text;                         // expanded index: single rows step on the text
  // the exact searches and D of the first passes, computed beforehand: pass
  // b's D at fd + b * kFrontLen (expanded into the lane's scratch when it
  // takes the task), its r and anchor at fr[2 * b * fstep]
  const uint8_t *fd;
  const uint4 *fr;
  uint32_t fstep;
  // the read's codes and N flags in the lane's LDS slot (words j at
  // slot[j * 256]), or null: read bytes from W
  const uint32_t *slot;
};

enum : uint32_t { KW_BLOCKS = 0, KW_LF = 1, KW_SAISA = 2, KW_TEXT = 3, KW_TABLE = 4, KW_LIST = 5,
                  KW_N = 6 };
#define KWORK(c, i, v)                          \
  do {                                          \
    if ((c).wk) (c).wk[(i)] += (uint32_t)(v);   \
  } while (0)

// The 32 text bases before suffix q in LF order (BWT of X: X[q-1], X[q-2],
// ...; of reverse X: X[n-q], X[n-q+1], ...) from the words t0, t1, t2 that
// start at word w (see front_step).
__device__ __forceinline__ uint64_t ftext_window(uint32_t dsel, uint32_t n, uint32_t q, uint32_t t0,
                                                 uint32_t t1, uint32_t t2) {
  const uint64_t lo = ((uint64_t)t1 << 32) | t0;
  if (dsel) {
    const uint32_t sh = ((n - q) & 15u) << 1;
    return sh ? (lo >> sh) | ((uint64_t)t2 << (64u - sh)) : lo;
  }
  const int64_t b = (int64_t)q - 32;
  uint64_t win;
  if (b >= 0) {
    const uint32_t sh = ((uint32_t)b & 15u) << 1;
    win = sh ? (lo >> sh) | ((uint64_t)t2 << (64u - sh)) : lo;
  } else {
    win = lo << min(2u * (uint32_t)(-b), 62u);
  }
  const uint64_t v = __builtin_bitreverse64(win);
  return ((v >> 1) & 0x5555555555555555ull) | ((v & 0x5555555555555555ull) << 1);
}

// On an expanded index a one-row interval is stepped on the text: the LF
// steps of W[i], W[i + step], ... (count positions) from the row whose
// suffix is q keep one row while the read's bases equal the text bases
// before q (in the view's LF order).  Returns the positions matched, the
// suffix after them in q_end, and whether a step inside count fails (a base
// differs or is N, or the text start is reached: the '$' row has no
// successor), where klf would give an empty interval.
__device__ __forceinline__ int32_t k_run(const KCtx &c, const KView &v, uint32_t q, int32_t i,
                                         int32_t step, int32_t count, uint32_t &q_end,
                                         bool &dies) {
  int32_t done = 0;
  dies = false;
  if (c.slot) {
    // 32 bases per compare, the read's from the LDS slot
    const uint32_t n = c.ix->n, nwords = (n + 15u) >> 4;
    while (done < count) {
      uint32_t lim = (uint32_t)(count - done) < 32u ? (uint32_t)(count - done) : 32u;
      if (lim > q) lim = q;
      if (lim == 0u) { dies = true; break; }
      uint64_t rw;
      uint32_t rb;
      lds_window(c.slot, i + done * step, step, rw, rb);
      uint32_t w;
      if (v.dsel) {
        w = min((n - q) >> 4, nwords - 1u);
      } else {
        const int64_t b = (int64_t)q - 32;
        w = b >= 0 ? (uint32_t)b >> 4 : 0u;
      }
      const uint32_t t0 = c.ix->text[w];
      const uint32_t t1 = w + 1u < nwords ? c.ix->text[w + 1u] : 0u;
      const uint32_t t2 = w + 2u < nwords ? c.ix->text[w + 2u] : 0u;
      const uint64_t tw = ftext_window(v.dsel, n, q, t0, t1, t2);
      uint64_t spread = rb;                           // N flag of base j -> bit 2j
      spread = (spread | (spread << 16)) & 0x0000FFFF0000FFFFull;
      spread = (spread | (spread << 8)) & 0x00FF00FF00FF00FFull;
      spread = (spread | (spread << 4)) & 0x0F0F0F0F0F0F0F0Full;
      spread = (spread | (spread << 2)) & 0x3333333333333333ull;
      spread = (spread | (spread << 1)) & 0x5555555555555555ull;
      const uint64_t x = (v.comp ? ~rw : rw) ^ tw;
      uint64_t miss = ((x | (x >> 1)) & 0x5555555555555555ull) | spread;
      if (lim < 32u) miss |= 1ull << (2u * lim);
      const uint32_t m = miss ? (uint32_t)__builtin_ctzll(miss) >> 1 : 32u;
      KWORK(c, KW_TEXT, 1);
      done += (int32_t)m;
      q -= m;
      if (m < lim) { dies = true; break; }
    }
    q_end = q;
    return done;
  }
  while (done < count) {
    uint32_t lim = (uint32_t)(count - done) < 16u ? (uint32_t)(count - done) : 16u;
    if (lim > q) lim = q;
    if (lim == 0u) { dies = true; break; }
    const uint32_t m = text_match(*c.ix, v.dsel, v.comp, q, c.W, i + done * step, step, lim);
    KWORK(c, KW_TEXT, 1);
    done += (int32_t)m;
    q -= m;
    if (m < lim) { dies = true; break; }
  }
  q_end = q;
  return done;
}

__device__ __forceinline__ uint32_t k_sa(const KCtx &c, const KView &v, int64_t row) {
  KWORK(c, KW_SAISA, 1);
  return v.d->sa[row];
}
__device__ __forceinline__ uint32_t k_isa(const KCtx &c, const KView &v, uint32_t q) {
  KWORK(c, KW_SAISA, 1);
  return v.d->isa[q];
}
// occurrence blocks an LF query on [k, l] reads (one when l + 1 shares k's
// block or the interval is one row)
__device__ __forceinline__ uint32_t k_blocks(int64_t k, int64_t l) {
  return (k != l && (uint64_t)(l + 1) / kBlockRows != (uint64_t)k / kBlockRows) ? 2u : 1u;
}

// BWiteration (runtime.h:38-42) on a view; N (code > 3) gives (n + 1, n).
// Only live intervals reach it; a dead one is passed through unchanged
// rather than read past the blocks.
__device__ __forceinline__ void klf(const KCtx &c, const KView &v, int64_t k, int64_t l, uint32_t b,
                                    int64_t &ko, int64_t &lo) {
  const uint32_t n = c.ix->n;
  if (b > 3u) { ko = (int64_t)n + 1; lo = n; return; }
  if (k > l) { ko = k; lo = l; return; }
  uint32_t a, e, nb;
  lf_one(*v.d, (uint32_t)k, (uint32_t)l, v.comp ? 3u - b : b, a, e, nb);
  KWORK(c, KW_BLOCKS, nb);
  KWORK(c, KW_LF, 1);
  ko = a;
  lo = e;
}

// The k-mer table's interval of the kmer_k bases W[i], W[i + step], ... on
// view v (in the order its LF steps take them, complemented on the - strand),
// when the index has a table, all of them are A/C/G/T and the interval is not
// empty; otherwise the caller steps one base at a time, so where and how a
// search dies stays the reference's.
__device__ __forceinline__ bool k_table(const KCtx &c, const KView &v, int32_t i, int32_t step,
                                        int64_t &k, int64_t &l) {
  const uint32_t K = c.ix->kmer_k;
  uint32_t code = 0u, bad = 0u;
  if (c.slot) {
    uint64_t rw;
    uint32_t rb;
    lds_window(c.slot, i, step, rw, rb);
    code = (uint32_t)rw & ((1u << (2u * K)) - 1u);
    bad = rb & ((1u << K) - 1u);
  } else {
    for (uint32_t j = 0; j < K; ++j) {
      const uint32_t s = c.W[i + (int32_t)j * step];
      bad |= s > 3u ? 1u : 0u;
      code |= (s & 3u) << (2u * j);
    }
  }
  if (bad) return false;
  if (v.comp) code ^= (1u << (2u * K)) - 1u;
  const uint64_t e = c.ix->kmer[v.dsel][code];
  KWORK(c, KW_TABLE, 1);
  if ((uint32_t)e > (uint32_t)(e >> 32)) return false;
  k = (uint32_t)e;
  l = (uint32_t)(e >> 32);
  return true;
}

// klf of the read base at wp: on a live interval the base and the block(s)
// are read in one batch (lf_read); a dead one reads only the base (an N still
// gives (n + 1, n), as klf does).
__device__ __forceinline__ void klfw(const KCtx &c, const KView &v, int64_t k, int64_t l,
                                     const uint8_t *wp, int64_t &ko, int64_t &lo) {
  const uint32_t n = c.ix->n;
  if (k > l) {
    if (*wp > 3u) { ko = (int64_t)n + 1; lo = n; }
    else { ko = k; lo = l; }
    return;
  }
  uint32_t a, e;
  KWORK(c, KW_BLOCKS, k_blocks(k, l));
  KWORK(c, KW_LF, 1);
  if (lf_read(*v.d, (uint32_t)k, (uint32_t)l, wp, v.comp, a, e) > 3u) {
    ko = (int64_t)n + 1;
    lo = n;
    return;
  }
  ko = a;
  lo = e;
}

// add_result (results.h:93-103): keep, or return true from the caller.
#define KADD(r, L)                                                   \
  do {                                                               \
    KWORK(c, KW_LIST, 1);                                            \
    if ((L).n < (L).cap) (L).list[(L).n++] = (r);                    \
    else return true;                                                \
  } while (0)

__device__ __forceinline__ void k_add_mm(fmi_hitk &r, uint8_t kind, uint8_t base, int16_t pos) {
  if (r.num_mismatches < FMI_KMAX) {
    r.err_kind[r.num_mismatches] = kind;
    r.err_base[r.num_mismatches] = base;
    r.err_pos[r.num_mismatches] = pos;
  }
  r.num_mismatches++;
}
__device__ __forceinline__ void k_mod2(fmi_hitk &r, uint8_t kind, uint8_t base) {
  if (r.num_mismatches - 1 >= FMI_KMAX) return;
  r.err_kind[r.num_mismatches - 1] = kind;
  r.err_base[r.num_mismatches - 1] = base;
}
__device__ __forceinline__ void k_mod1(fmi_hitk &r, uint8_t kind) {
  if (r.num_mismatches - 1 >= FMI_KMAX) return;
  r.err_kind[r.num_mismatches - 1] = kind;
}
__device__ __forceinline__ void k_change(fmi_hitk &r, int64_t k, int64_t l, int32_t pos) {
  r.k = k;
  r.l = l;
  r.pos = (int16_t)pos;
}

__device__ __forceinline__ void k_clear(fmi_hitk &r) {
  r.k = 0; r.l = 0; r.start = 0; r.pos = 0; r.end = 0; r.dir = 0; r.num_mismatches = 0;
}

// BWExactSearchBackward / Forward (search.h:9-65) with the anchor.  In a
// text run the empty interval where the search dies is not formed: the
// callers only keep a live result, and the anchor holds the last live one.
__device__ __forceinline__ void k_exact(const KCtx &c, const KView &v, int dir, fmi_hitk &r, fmi_hitk *anchor) {
  int64_t k = 0, l = 0, k2 = r.k, l2 = r.l;
  int32_t i, i0 = r.pos;
  // from the whole index, the first kmer_k steps can be one table read
  const int32_t K = (int32_t)c.ix->kmer_k;
  const bool whole = K && k2 == 0 && l2 == (int64_t)c.ix->n;
  int64_t kt, lt;
  if (dir == 0) {
    if (whole && r.pos - r.start + 1 >= K && k_table(c, v, r.pos, -1, kt, lt)) {
      k = k2 = kt; l = l2 = lt; i0 = r.pos - K;
    }
    for (i = i0; i >= r.start; --i) {
      if (c.text && k2 == l2) {
        uint32_t q2;
        bool dies;
        const int32_t done = k_run(c, v, k_sa(c, v, k2), i, -1, i - r.start + 1, q2, dies);
        if (done) { k = l = k2 = l2 = k_isa(c, v, q2); }
        i -= done;
        if (dies) { k2 = (int64_t)c.ix->n + 1; l2 = c.ix->n; }
        break;
      }
      klfw(c, v, k2, l2, c.W + i, k2, l2);
      if (k2 > l2) break;
      k = k2; l = l2;
    }
    if (anchor) {
      k_clear(*anchor);
      anchor->dir = 0; anchor->start = (int16_t)(i + 1); anchor->end = r.pos;
      anchor->k = k; anchor->l = l; anchor->pos = (int16_t)i;
    }
  } else {
    if (whole && r.end - r.pos + 1 >= K && k_table(c, v, r.pos, 1, kt, lt)) {
      k = k2 = kt; l = l2 = lt; i0 = r.pos + K;
    }
    for (i = i0; i <= r.end; ++i) {
      if (c.text && k2 == l2) {
        uint32_t q2;
        bool dies;
        const int32_t done = k_run(c, v, k_sa(c, v, k2), i, 1, r.end - i + 1, q2, dies);
        if (done) { k = l = k2 = l2 = k_isa(c, v, q2); }
        i += done;
        if (dies) { k2 = (int64_t)c.ix->n + 1; l2 = c.ix->n; }
        break;
      }
      klfw(c, v, k2, l2, c.W + i, k2, l2);
      if (k2 > l2) break;
      k = k2; l = l2;
    }
    if (anchor) {
      k_clear(*anchor);
      anchor->dir = 1; anchor->start = r.pos; anchor->end = (int16_t)(i - 1);
      anchor->k = k; anchor->l = l; anchor->pos = (int16_t)i;
    }
  }
  r.k = k2; r.l = l2; r.pos = (int16_t)i;
}

// The exact extension of the final helpers: klf over W[i], W[i + step], ...
// (count positions) until the interval empties; a one-row interval runs on
// the text and its row is read back only when it survives.
__device__ __forceinline__ void k_extend(const KCtx &c, const KView &v, int64_t &k, int64_t &l,
                                         int32_t i, int32_t step, int32_t count) {
  for (; count > 0; i += step, --count) {
    if (c.text && k == l) {
      uint32_t q2;
      bool dies;
      const int32_t done = k_run(c, v, k_sa(c, v, k), i, step, count, q2, dies);
      if (dies) { k = (int64_t)c.ix->n + 1; l = c.ix->n; }
      else if (done) { k = l = k_isa(c, v, q2); }
      return;
    }
    klfw(c, v, k, l, c.W + i, k, l);
    if (k > l) return;
  }
}

// BWExactFinalResultBackward / Forward (search.h:66-158): the block filter,
// then an exact extension to the end; only a surviving record is changed.
__device__ __forceinline__ bool k_final1(const KCtx &c, const KView &v, int dir, fmi_hitk &r, KList &fin,
                         int32_t bs, int32_t lb) {
  const int32_t start = r.start, pos = r.pos, end = r.end;
  int64_t k = r.k, l = r.l;
  const int32_t cb = pos / bs;
  if (dir == 0) {
    if ((cb < lb) || (pos == start - 1)) {
    } else if (cb > lb) {
      return false;
    } else if (!((pos + 1) % bs)) {
      return false;
    }
    k_extend(c, v, k, l, pos, -1, pos - start + 1);
    if (k <= l) { k_change(r, k, l, start - 1); KADD(r, fin); }
  } else {
    if ((cb > lb) || (pos == end + 1)) {
    } else if (cb < lb) {
      return false;
    } else if (!(pos % bs)) {
      return false;
    }
    k_extend(c, v, k, l, pos, 1, end - pos + 1);
    if (k <= l) { k_change(r, k, l, end + 1); KADD(r, fin); }
  }
  return false;
}

// BWExactFinalResultsBackward / Forward (search.c:3-114)
__device__ __forceinline__ bool k_finals(const KCtx &c, const KView &v, int dir, KList &prev, KList &out,
                         int32_t bs, int32_t lb) {
  for (uint32_t ii = 0; ii < prev.n; ++ii) {
    fmi_hitk &r = prev.list[ii];
    KWORK(c, KW_LIST, 1);
    const int32_t start = r.start, pos = r.pos, end = r.end;
    int64_t k = r.k, l = r.l;
    const int32_t cb = pos / bs;
    if (dir == 0) {
      if ((cb < lb) || (pos == start - 1)) {
      } else if (cb > lb) {
        continue;
      } else if (!((pos + 1) % bs)) {
        continue;
      }
      k_extend(c, v, k, l, pos, -1, pos - start + 1);
      if (k <= l) { k_change(r, k, l, start - 1); KADD(r, out); }
    } else {
      if ((cb > lb) || (pos == end + 1)) {
      } else if (cb < lb) {
        continue;
      } else if (!(pos % bs)) {
        continue;
      }
      k_extend(c, v, k, l, pos, 1, end - pos + 1);
      if (k <= l) { k_change(r, k, l, end + 1); KADD(r, out); }
    }
  }
  prev.n = 0;
  return false;
}

// One record's branch step at its pos (BWBranchPartialResults*, search.c:
// 648-1000, and BWBranchFinalResults*, search.c:116-469): deletion, then
// insertion and mismatch of each base, as the previous error allows; the
// record is updated field by field as the reference updates it.
template <bool FINAL>
__device__ __forceinline__ bool k_branch(const KCtx &c, const KView &v, int dir, fmi_hitk &r, KList &out,
                         int32_t bs, int32_t lb) {
  const int32_t pos = r.pos;
  const int32_t step = dir ? 1 : -1;
  int32_t last_err_pos;
  uint32_t last_err_kind, last_err_base;
  const int32_t rnm = (int32_t)r.num_mismatches - 1;
  if (rnm > -1 && rnm < FMI_KMAX) {
    last_err_pos = r.err_pos[rnm]; last_err_kind = r.err_kind[rnm]; last_err_base = r.err_base[rnm];
  } else {
    last_err_pos = -10; last_err_kind = 0; last_err_base = 0xFFu;
  }
  const int64_t k = r.k, l = r.l;
  int64_t ka, la;
  bool no_previous = true;
  const uint32_t wp = c.W[pos];
  // the children of (k, l) for b = 0..3, from one read of the block(s): the
  // four klf calls of the loops below, which the record stores in between
  // would otherwise make four dependent reads
  int64_t kb[4], lb4[4];
  if (k > l) {
#pragma unroll
    for (uint32_t b = 0; b < 4u; ++b) { kb[b] = k; lb4[b] = l; }
  } else {
    uint32_t ko[4], lo[4];
    lf_four(*v.d, (uint32_t)k, (uint32_t)l, ko, lo);
    KWORK(c, KW_BLOCKS, k_blocks(k, l));
    KWORK(c, KW_LF, 1);
#pragma unroll
    for (uint32_t b = 0; b < 4u; ++b) {
      const uint32_t s = v.comp ? 3u - b : b;
      kb[b] = pick4(ko, s);
      lb4[b] = pick4(lo, s);
    }
  }
#define KEMIT()                                                          \
  do {                                                                   \
    if (FINAL) {                                                         \
      if (k_final1(c, v, dir, r, out, bs, lb)) return true;              \
    } else {                                                             \
      KADD(r, out);                                                      \
    }                                                                    \
  } while (0)
  k_add_mm(r, FMI_DELETION, 0xFF, (int16_t)pos);
  if (last_err_pos == pos - step) {
    if (last_err_kind == FMI_MISMATCH) {
      if (wp != last_err_base) { k_change(r, k, l, pos + step); KEMIT(); }
      const uint32_t wl = c.W[last_err_pos];
      for (uint32_t b = 0; b < 4u; ++b) {
        ka = kb[b]; la = lb4[b];
        if (ka > la) continue;
        if (b != wp) {
          if (b != wl) { k_change(r, ka, la, pos); k_mod2(r, FMI_INSERTION, (uint8_t)b); KEMIT(); }
          k_change(r, ka, la, pos + step); k_mod2(r, FMI_MISMATCH, (uint8_t)b); KEMIT();
        }
      }
      no_previous = false;
    } else if (last_err_kind == FMI_DELETION) {
      k_change(r, k, l, pos + step); KEMIT();
      const uint32_t wl = c.W[last_err_pos];
      for (uint32_t b = 0; b < 4u; ++b) {
        ka = kb[b]; la = lb4[b];
        if (ka > la) continue;
        if (b != wp && b != wl) {
          k_change(r, ka, la, pos + step); k_mod2(r, FMI_MISMATCH, (uint8_t)b); KEMIT();
        }
      }
      no_previous = false;
    }
  } else if (last_err_pos == pos) {
    for (uint32_t b = 0; b < 4u; ++b) {
      ka = kb[b]; la = lb4[b];
      if (ka > la) continue;
      if (b != wp) {
        k_change(r, ka, la, pos); k_mod2(r, FMI_INSERTION, (uint8_t)b); KEMIT();
        if (wp != last_err_base) { r.pos = (int16_t)(pos + step); k_mod1(r, FMI_MISMATCH); KEMIT(); }
      }
    }
    no_previous = false;
  }
  if (no_previous) {
    k_change(r, k, l, pos + step); KEMIT();
    for (uint32_t b = 0; b < 4u; ++b) {
      ka = kb[b]; la = lb4[b];
      if (ka > la) continue;
      if (b != wp) {
        k_change(r, ka, la, pos); k_mod2(r, FMI_INSERTION, (uint8_t)b); KEMIT();
        r.pos = (int16_t)(pos + step); k_mod1(r, FMI_MISMATCH); KEMIT();
      }
    }
  }
#undef KEMIT
  return false;
}

template <bool FINAL>
__device__ __forceinline__ bool k_branches(const KCtx &c, const KView &v, int dir, KList &prev, KList &out,
                           int32_t bs, int32_t lb) {
  for (uint32_t ii = 0; ii < prev.n; ++ii) {
    fmi_hitk &r = prev.list[ii];
    KWORK(c, KW_LIST, 1);
    if (r.dir != dir) continue;
    if (dir == 0 ? (r.pos < r.start) : (r.pos > r.end)) {
      if (FINAL && dir == 0) {
        if (k_final1(c, v, 0, r, out, bs, lb)) return true;
      } else {
        KADD(r, out);
      }
      continue;
    }
    if (k_branch<FINAL>(c, v, dir, r, out, bs, lb)) return true;
  }
  prev.n = 0;
  return false;
}

// BWExactPartialResultsBackward / Forward (search.c:471-646)
__device__ __forceinline__ bool k_partials(const KCtx &c, uint8_t max_errors, const KView &v, int dir, KList &prev,
                           KList &next, KList &next_i, int32_t bs, int32_t lb) {
  for (uint32_t ii = 0; ii < prev.n; ++ii) {
    fmi_hitk &r = prev.list[ii];
    KWORK(c, KW_LIST, 1);
    const int32_t start = r.start, pos = r.pos, end = r.end;
    int64_t kn = r.k, ln = r.l, k = 0, l = 0;
    int64_t res, res_next = ln - kn;
    const int32_t cb = pos / bs;
    const uint8_t eb = (uint8_t)(max_errors - r.num_mismatches);
    int32_t lbp;
    bool complete;
    if (dir == 0) {
      if ((cb < lb) || (pos == start - 1)) { lbp = start; complete = true; }
      else if (cb > lb) { lbp = ((pos + 1) % bs) ? (cb - 1) * bs : cb * bs; complete = false; }
      else if ((pos + 1) % bs) { lbp = start; complete = true; }
      else { lbp = cb * bs; complete = false; }
      for (int32_t i = pos; i >= lbp; --i) {
        k = kn; l = ln;
        if (k > l) break;
        if (c.text && k == l) {
          // one row: no size change (no record) until the step that empties it
          uint32_t q2;
          bool dies;
          const int32_t done = k_run(c, v, k_sa(c, v, k), i, -1, i - lbp + 1, q2, dies);
          i -= done;
          if (!dies) {
            if (done) kn = ln = k_isa(c, v, q2);
            break;
          }
          res_next = -1;                     // the step at i empties the row
          kn = (int64_t)c.ix->n + 1; ln = c.ix->n;
          if (eb < c.D[i]) break;
          if (done) k = l = k_isa(c, v, q2);
          k_change(r, k, l, i);
          KADD(r, next);
          break;
        }
        klfw(c, v, k, l, c.W + i, kn, ln);
        res = res_next; res_next = ln - kn;
        if (res == res_next || eb < c.D[i]) continue;
        k_change(r, k, l, i);
        KADD(r, next);
      }
      if (complete && kn <= ln) { k_change(r, kn, ln, start - 1); KADD(r, next_i); }
    } else {
      if ((cb > lb) || (pos == end + 1)) { lbp = end; complete = true; }
      else if (cb < lb) { lbp = (pos % bs) ? (cb + 2) * bs - 1 : (cb + 1) * bs - 1; complete = false; }
      else if (pos % bs) { lbp = end; complete = true; }
      else { lbp = (cb + 1) * bs - 1; complete = false; }
      for (int32_t i = pos; i <= lbp; ++i) {
        k = kn; l = ln;
        if (k > l) break;
        if (c.text && k == l) {
          uint32_t q2;
          bool dies;
          const int32_t done = k_run(c, v, k_sa(c, v, k), i, 1, lbp - i + 1, q2, dies);
          i += done;
          if (!dies) {
            if (done) kn = ln = k_isa(c, v, q2);
            break;
          }
          res_next = -1;
          kn = (int64_t)c.ix->n + 1; ln = c.ix->n;
          if (eb < c.D[i]) break;
          if (done) k = l = k_isa(c, v, q2);
          k_change(r, k, l, i);
          KADD(r, next);
          break;
        }
        klfw(c, v, k, l, c.W + i, kn, ln);
        res = res_next; res_next = ln - kn;
        if (res == res_next || eb < c.D[i]) continue;
        k_change(r, k, l, i);
        KADD(r, next);
      }
      if (complete && kn <= ln) { k_change(r, kn, ln, end + 1); KADD(r, next_i); }
    }
  }
  prev.n = 0;
  return false;
}

// change_direction (search.h:160-198): min / max ISA on the other index of
// every row's mirrored suffix; a position outside the text is left out.
__device__ __forceinline__ void k_change_dir(const KCtx &c, const KView &b, const KView &f, fmi_hitk &r) {
  const int64_t siz = (int64_t)c.ix->n + 1;
  int64_t ki = siz - 1, li = 0;
  int32_t err_offset = 0;
  for (uint32_t rr = 0; rr < r.num_mismatches && rr < FMI_KMAX; ++rr) {
    if (r.err_kind[rr] == FMI_DELETION) err_offset--;
    else if (r.err_kind[rr] == FMI_INSERTION) err_offset++;
  }
  for (int64_t i = r.k; i <= r.l; ++i) {
    const int64_t aux = siz - (int64_t)sa_value(*b.d, c.ix->ratio, c.ix->n, (uint32_t)i) -
                        (r.end - r.start + 2) - err_offset;
    if (aux < 0 || aux >= siz) continue;
    const int64_t aux2 = isa_value(*f.d, c.ix->ratio, c.ix->n, (uint32_t)aux);
    if (aux2 < ki) ki = aux2;
    if (aux2 > li) li = aux2;
  }
  r.k = ki;
  r.l = li;
}

// BWChangeDirectionBackward / Forward (search.c:1002-1043)
__device__ __forceinline__ void k_turn(const KCtx &c, const KView &b, const KView &f, KList &L, bool to_fwd,
                       int32_t bound) {
  for (uint32_t ii = L.n; ii > 0; --ii) {
    fmi_hitk &r = L.list[ii - 1];
    KWORK(c, KW_LIST, 1);
    if (to_fwd) {
      if (r.dir == 1) break;
      if (r.pos == r.start - 1) {
        k_change_dir(c, b, f, r);
        r.pos = (int16_t)(r.end + 1); r.end = (int16_t)bound; r.dir = 1;
      }
    } else {
      if (r.dir == 0) break;
      if (r.pos == r.end + 1) {
        k_change_dir(c, b, f, r);
        r.pos = (int16_t)(r.start - 1); r.start = (int16_t)bound; r.dir = 0;
      }
    }
  }
}

// calculateDBackward / Forward (search.c:1045-1169), with the reference's
// 8-bit counter and int8 `last`.
// One walk of calculateD*: klf over W[i], W[i + step], ... (count
// positions), restarting from the whole index where the interval empties
// and counting the restarts in z; D[i] = z.  One-row stretches run on the
// text; the interval left at the end is not used by the callers.
__device__ __forceinline__ void k_dwalk(const KCtx &c, const KView &v, int64_t k, int64_t l,
                                        int32_t i, int32_t step, int32_t count, uint8_t &z) {
  const int64_t nmax = c.ix->n;
  const int32_t K = (int32_t)c.ix->kmer_k;
  uint8_t *D = c.D;
  while (count > 0) {
    int64_t kt, lt;
    if (K && k == 0 && l == nmax && count >= K && k_table(c, v, i, step, kt, lt)) {
      // K steps from the whole index, none of them empty
      for (int32_t t = 0; t < K; ++t) D[i + t * step] = z;
      i += K * step;
      count -= K;
      k = kt; l = lt;
      continue;
    }
    if (c.text && k == l) {
      uint32_t q2;
      bool dies;
      const int32_t done = k_run(c, v, k_sa(c, v, k), i, step, count, q2, dies);
      for (int32_t t = 0; t < done; ++t) D[i + t * step] = z;
      i += done * step;
      count -= done;
      if (!dies) return;
      k = 0; l = nmax; z++;                // the step at i empties the row
      D[i] = z;
      i += step;
      --count;
      continue;
    }
    klfw(c, v, k, l, c.W + i, k, l);
    if (k > l) { k = 0; l = nmax; z++; }
    D[i] = z;
    i += step;
    --count;
  }
}

__device__ __forceinline__ void k_calc_d(const KCtx &c, const fmi_hitk &r, bool backward) {
  const int64_t nmax = c.ix->n;
  uint8_t z = 0;
  int8_t last;
  uint8_t *D = c.D;
  const int32_t nW = c.nW;
  if (backward) {
    k_dwalk(c, c.B, r.k, r.l, r.pos, -1, r.pos + 1, z);
    if (r.end == (int16_t)(nW - 1)) {
      last = (int8_t)D[0];
    } else {
      k_dwalk(c, c.F, 0, nmax, r.end + 1, 1, nW - (r.end + 1), z);
      last = (int8_t)D[nW - 1];
      for (int32_t i = r.end + 1; i < nW; ++i) D[i] = (uint8_t)(last - D[i]);
    }
    for (int32_t i = r.pos; i >= 0; --i) D[i] = (uint8_t)(last - D[i]);
  } else {
    k_dwalk(c, c.F, r.k, r.l, r.pos, 1, nW - r.pos, z);
    if (r.start == 0) {
      last = (int8_t)D[nW - 1];
    } else {
      k_dwalk(c, c.B, 0, nmax, r.start - 1, -1, r.start, z);
      last = (int8_t)D[0];
      for (int32_t i = r.start - 1; i >= 0; --i) D[i] = (uint8_t)(last - D[i]);
    }
    for (int32_t i = r.pos; i < nW; ++i) D[i] = (uint8_t)(last - D[i]);
  }
}

// One pass of BWSearchCPU (search.c:1171-1405): the exact search of
// fragment i, then up to fragments - 1 rounds of partial extension and
// branching and the final step.  A forward pass (the middle fragments
// i = half - 1 .. 1 and fragment 0, search.c:1214-1296) searches on
// `forward` left to right and, for a middle fragment, turns its finished
// partials to `backward`; a backward pass (fragments half .. fragments - 2 and
// the last, search.c:1298-1401) mirrors it.  Each pass starts from empty
// working lists; the final list and the anchors carry over.
__device__ __forceinline__ bool k_pass(const KCtx &c, KList &prev, KList &next, KList &prev_i,
                                       KList &next_i, KList &fin, KList &anc, int32_t fs,
                                       int32_t fragments, int32_t half, bool fwd, int32_t i) {
  const int32_t nW = c.nW;
  const int64_t lsa = c.ix->n;
  const uint8_t maxe = (uint8_t)(fragments - 1);
  const bool two = fwd ? i > 0 : i < fragments - 1;   // a middle fragment
  const KView &X = fwd ? c.F : c.B;
  const KView &Y = fwd ? c.B : c.F;
  const int dx = fwd ? 1 : 0, dy = 1 - dx;
  const int32_t lb = fwd ? half - 1 : 0;
  const int32_t bound = fwd ? 0 : nW - 1;
  fmi_hitk r, ra;
  bool flow = true;
  int32_t err_count = fragments - 1;
  prev.n = 0; prev_i.n = 0; next.n = 0; next_i.n = 0;
  k_clear(r);
  r.dir = (uint8_t)dx;
  r.start = (int16_t)(fs * i);
  r.end = (int16_t)(fwd || i < fragments - 1 ? fs * (i + 1) - 1 : nW - 1);
  k_change(r, 0, lsa, fwd ? r.start : r.end);
  KCtx cf = c;
  const int32_t pb = fwd ? half - 1 - i : i;          // the pass's place in the order
  const bool front = c.fd != nullptr && pb < (int32_t)kFrontPasses;
  if (front) {
    const uint4 *fr = c.fr + 2ull * (uint64_t)pb * c.fstep;
    const uint4 a = fr[0];
    const int16_t pos = (int16_t)fr[1].x;
    k_clear(ra);
    ra.dir = (uint8_t)dx;
    ra.k = a.z; ra.l = a.w; ra.pos = pos;
    if (dx == 0) { ra.start = (int16_t)(pos + 1); ra.end = r.pos; }
    else { ra.start = r.pos; ra.end = (int16_t)(pos - 1); }
    r.k = a.x; r.l = a.y; r.pos = pos;
    cf.D = const_cast<uint8_t *>(c.fd) + (uint32_t)pb * kFrontLen;
  } else {
    k_exact(c, X, dx, r, &ra);
  }
  if (anc.list) KADD(ra, anc);
  if (fwd) r.end = (int16_t)(nW - 1);
  else r.start = 0;
  if (r.k > r.l) return false;
  if (!front) k_calc_d(c, r, !fwd);
  KADD(r, prev);
  KList &out1 = two ? prev_i : fin;          // where X's complete partials go
  while (err_count > 0) {
    if (k_partials(cf, maxe, X, dx, prev, next, out1, fs, lb)) { flow = false; break; }
    if (two) {
      k_turn(cf, X, Y, prev_i, !fwd, bound);
      if (k_partials(cf, maxe, Y, dy, prev_i, next_i, fin, fs, lb)) { flow = false; break; }
    }
    if (err_count == 1) break;
    if (k_branches<false>(cf, X, dx, next, prev, 0, 0)) { flow = false; break; }
    if (two && k_branches<false>(cf, Y, dy, next_i, prev_i, 0, 0)) { flow = false; break; }
    err_count--;
  }
  if (flow) {
    k_branches<true>(cf, X, dx, next, out1, fs, lb);
    if (two) {
      k_turn(cf, X, Y, prev_i, !fwd, bound);
      k_finals(cf, Y, dy, prev_i, fin, fs, lb);
      k_branches<true>(cf, Y, dy, next_i, fin, fs, lb);
    }
  }
  return false;
}

// BWSearchCPU (search.c:1171-1405) for one strand of one read.
__device__ __forceinline__ bool k_search(const KCtx &c, KList &prev, KList &next, KList &prev_i, KList &next_i,
                         KList &fin, KList &anc, int32_t fs, int type) {
  fmi_hitk r, ra;
  const int32_t nW = c.nW;
  const int32_t fragments = nW / fs;
  int32_t half = fragments / 2;
  if (fragments % 2) half++;
  const int64_t lsa = c.ix->n;
  anc.n = 0;

  if (fragments <= 1) {
    k_clear(r);
    if (type == 0) {
      r.dir = 1; k_change(r, 0, lsa, 0); r.start = 0; r.end = (int16_t)(nW - 1);
      k_exact(c, c.F, 1, r, &ra);
    } else {
      r.dir = 0; k_change(r, 0, lsa, nW - 1); r.start = 0; r.end = (int16_t)(nW - 1);
      k_exact(c, c.B, 0, r, &ra);
    }
    if (anc.list) KADD(ra, anc);
    if (r.k <= r.l) KADD(r, fin);
    return false;
  }

  // passes in the reference's order: fragments half-1 .. 1 forward, then
  // fragment 0, then fragments half .. fragments-2 backward, then the last
  for (int32_t b = 0; b < fragments; ++b) {
    const bool fwd = b < half;
    const int32_t i = b < half - 1 ? half - 1 - b : (b == half - 1 ? 0 : b);
    if (k_pass(c, prev, next, prev_i, next_i, fin, anc, fs, fragments, half, fwd, i)) return true;
  }
  return false;
}

#undef KADD
#undef KWORK

struct KParams {
  DevIndex ix;
  const uint8_t *reads;
  const uint64_t *offsets;
  const int16_t *fragsize;
  fmi_hitk *fin, *anchors;
  uint32_t *nfin, *nanchors;
  uint8_t *full;
  fmi_hitk *scratch;                 // per lane: 4 lists of list_cap records
  uint8_t *dscratch;                 // per lane: kDStride bytes of D
  uint32_t *counter;
  unsigned long long *work;          // counting launch: KW_N counters
  uint4 *front_r;                    // per (pass, task): the exact search's r and anchor
  uint4 *front_m;                    // per (pass, task): the walks' restart marks
  const uint4 *front_rec;            // the read records (kReadRecordWords words per read)
  uint32_t *tasks, *ntasks_dev;      // when set: the main kernel's tasks are tasks[0 .. *ntasks_dev)
  uint32_t ntasks, list_cap, anchor_cap, strand_mask;
};

// ---- The first passes' exact searches and D, as a flat state machine.
//
// One unit is (pass b < kFrontPasses, task) of a read of at most kFrontLen
// bases: k_exact on the pass's fragment (r and its anchor), then, when r is
// live, the D walks of k_calc_d.  Each lane runs one unit at a time and takes
// the next at once when it ends; every iteration is one memory access of
// the lane (an LF query's block(s), an SA or ISA entry, 32 text bases, a
// k-mer table entry), issued by all lanes together, then the lane's
// transition.  The read is staged in the lane's LDS slot from its record.
// D is not written position by position: a walk marks where it restarts
// from the whole index (a bit per position in the slot), the unit stores the
// marks, and the next launch writes D[i], the number of restarts after
// position i in walk order (last - D[i] of calculateD*), from them, one
// thread per 16 positions.

enum : uint32_t { FK_TEXT = 0, FK_SA = 1, FK_ISA = 2, FK_TABLE = 3, FK_LF = 4, FK_UNIT = 5,
                  FK_RECORD = 6, FK_MSA = 7, FK_MTEXT = 8 };
enum : uint32_t { FP_EXACT = 0, FP_WALK1 = 1, FP_WALK2 = 2, FP_IDLE = 3, FP_UNIT = 4,
                  FP_RECORD = 5 };
constexpr uint32_t kFQNone = 0xFFFFFFFFu;
constexpr uint32_t kFSlotWords = 32;             // 24 record words, 8 restart-mask words
constexpr uint32_t F_TAB2FAIL = 1u;              // the (k - 4)-table cannot be used here
constexpr uint32_t F_TABFAIL = 2u;               // the k-table cannot be used here
constexpr uint32_t F_MOVED = 4u;                 // exact: the text run matched a base
constexpr uint32_t F_ISA = 8u;                   // exact: the row after the run is to be read
constexpr uint32_t F_RUNDIED = 16u;              // exact: ... and the run died

struct FLane {
  uint32_t u, rd, ph, flags, dsel, strand, fwd;
  int32_t nW, i, step, cnt, f0, f1, p1;
  uint32_t k, l, q, ka, la;
  uint32_t r, q0, q1, q2, q3;      // walk: the rows of a multi-row text run (r > 1)
};

// A walk on an interval of 2..kMRows rows goes on against the text at each
// row's suffix at once: the interval stays non-empty while one row's text
// matches, and empties at the first base that no row's text continues with,
// which is all a walk needs (D counts where intervals empty).
constexpr uint32_t kMRows = 4;

// A new unit's read offsets and fragment size are in (the FK_UNIT access):
// its pass, fragment and view; false when the unit has no pass to run.
__device__ __forceinline__ bool front_unit(const KParams &P, FLane &L, uint64_t off0,
                                           uint64_t off1, int32_t fs) {
  const uint32_t b = L.u / P.ntasks, t = L.u - b * P.ntasks;
  const uint32_t strand = t & 1u;
  if (!(P.strand_mask & (1u << strand))) return false;
  const int32_t len = (int32_t)(off1 - off0);
  if (fs <= 0 || len > (int32_t)kFrontLen || len / fs - 1 > FMI_KMAX) return false;
  const int32_t fragments = len / fs;
  if (fragments <= 1 || (int32_t)b >= fragments) return false;
  int32_t half = fragments / 2;
  if (fragments % 2) half++;
  const bool fwd = (int32_t)b < half;
  const int32_t fi = (int32_t)b < half - 1 ? half - 1 - (int32_t)b
                                           : ((int32_t)b == half - 1 ? 0 : (int32_t)b);
  L.nW = len;
  L.strand = strand;
  L.fwd = fwd ? 1u : 0u;
  L.f0 = fs * fi;
  L.f1 = fwd || fi < fragments - 1 ? fs * (fi + 1) - 1 : len - 1;
  return true;
}

// The read's record is in (the FK_RECORD access): its codes and N flags into
// the slot, the restart marks cleared, and k_exact on X (forward: `forward`
// left to right; backward: `backward` right to left) set to start from the
// whole index.
__device__ __forceinline__ void front_stage(const KParams &P, uint32_t *slot, FLane &L,
                                            const uint4 &c0, const uint4 &c1, const uint4 &c2,
                                            const uint4 &c3, const uint4 &g0, const uint4 &g1) {
  const uint32_t nc = ((uint32_t)L.nW + 63u) >> 6, nf = ((uint32_t)L.nW + 127u) >> 7;
  if (nc > 0u) { slot[0 * 256] = c0.x; slot[1 * 256] = c0.y; slot[2 * 256] = c0.z; slot[3 * 256] = c0.w; }
  if (nc > 1u) { slot[4 * 256] = c1.x; slot[5 * 256] = c1.y; slot[6 * 256] = c1.z; slot[7 * 256] = c1.w; }
  if (nc > 2u) { slot[8 * 256] = c2.x; slot[9 * 256] = c2.y; slot[10 * 256] = c2.z; slot[11 * 256] = c2.w; }
  if (nc > 3u) { slot[12 * 256] = c3.x; slot[13 * 256] = c3.y; slot[14 * 256] = c3.z; slot[15 * 256] = c3.w; }
  if (nf > 0u) { slot[16 * 256] = g0.x; slot[17 * 256] = g0.y; slot[18 * 256] = g0.z; slot[19 * 256] = g0.w; }
  if (nf > 1u) { slot[20 * 256] = g1.x; slot[21 * 256] = g1.y; slot[22 * 256] = g1.z; slot[23 * 256] = g1.w; }
#pragma unroll
  for (uint32_t j = 24u; j < kFSlotWords; ++j) slot[j * 256u] = 0u;
  L.dsel = L.fwd ? 1u - L.strand : L.strand;
  L.step = L.fwd ? 1 : -1;
  L.i = L.fwd ? L.f0 : L.f1;
  L.cnt = L.f1 - L.f0 + 1;
  L.k = 0u;
  L.l = P.ix.n;
  L.ka = 0u;
  L.la = 0u;
  L.q = kFQNone;
  L.r = 0u;
  L.flags = 0u;
  L.ph = FP_EXACT;
}

// D of a finished unit from its restart marks (8 words: bit x set where a
// walk restarted from the whole index at position x), 16 positions from
// 16c, the marks' word holding them being mw.  The walk from p1 runs to the read's end in the pass's direction (A);
// the other walk, if any, covers B = [0, f0) (forward pass) or (f1, nW)
// (backward pass), and its restarts come after all of A's in walk order.
// D[x] is the number of restarts after x in walk order: forward, x in A:
// restarts above x plus B's, x in B: restarts below x; backward, x in A:
// restarts below x plus B's, x in B: restarts above x.  The fragment's own
// positions, which no walk visits, get a value no step reads.
// pw: the restarts below word mw's first position; tot: all of them; cntB:
// those of the second walk.
__device__ __forceinline__ uint4 front_d_chunk(uint32_t pw, uint32_t mw, uint32_t tot,
                                               uint32_t cntB, bool fwd, int32_t p1, int32_t c) {
  const uint32_t hoff = 16u * ((uint32_t)c & 1u);
  uint32_t v[4] = {0u, 0u, 0u, 0u};
  if (((mw >> hoff) & 0xFFFFu) == 0u) {
    // no restart among the 16: one value below the walks' boundary p1, one above
    const uint32_t b0 = pw + (uint32_t)__builtin_popcount(mw & ((1u << hoff) - 1u));
    const uint32_t dlo = fwd ? b0 : b0 + cntB;
    const uint32_t dhi = fwd ? tot - b0 + cntB : tot - b0;
    const int32_t sp = min(max((fwd ? p1 : p1 + 1) - 16 * c, 0), 16);
    const uint32_t rlo = (dlo & 0xFFu) * 0x01010101u, rhi = (dhi & 0xFFu) * 0x01010101u;
#pragma unroll
    for (int32_t qd = 0; qd < 4; ++qd) {
      const int32_t e = min(max(sp - 4 * qd, 0), 4);
      const uint32_t mh = e >= 4 ? 0u : (0xFFFFFFFFu << (8u * (uint32_t)e));
      v[qd] = (rlo & ~mh) | (rhi & mh);
    }
    return make_uint4(v[0], v[1], v[2], v[3]);
  }
#pragma unroll
  for (int32_t j = 0; j < 16; ++j) {
    const int32_t x = 16 * c + j;
    const uint32_t b = hoff + (uint32_t)j;
    const uint32_t bx = pw + (uint32_t)__builtin_popcount(mw & ((1u << b) - 1u));
    const uint32_t bx1 = pw + (uint32_t)__builtin_popcount(mw & ((2u << b) - 1u));
    uint32_t d;
    if (fwd) d = x >= p1 ? tot - bx1 + cntB : bx;
    else d = x <= p1 ? bx + cntB : tot - bx1;
    v[j >> 2] |= (d & 0xFFu) << (8u * (uint32_t)(j & 3));
  }
  return make_uint4(v[0], v[1], v[2], v[3]);
}

// One iteration of a lane's unit: choose the access (ALU and LDS only),
// issue it together with the other lanes' (no value used before the last
// load is issued), then the transition.
template <bool COUNT>
__device__ __forceinline__ void front_step(const KParams &P, uint32_t *slot, FLane &L,
                                           uint32_t *wk) {
  const uint32_t n = P.ix.n;
  const uint32_t K = P.ix.kmer_k;
  const DevDir &Dd = L.dsel ? P.ix.d[1] : P.ix.d[0];
  const bool textmode = P.ix.text != nullptr && Dd.sa != nullptr;
  const bool exact = L.ph == FP_EXACT;
  uint64_t rw;
  uint32_t rb;
  lds_window(slot, L.i, L.step, rw, rb);
  const uint64_t rwc = L.strand ? ~rw : rw;          // complemented on the - strand

  // ---- what to read
  uint32_t kind;
  uint32_t lim = 0u, code = 0u, j1 = 0u, j2 = 0u, tk = 0u;
  bool two = false, nextw = false;
  const uint32_t *a0 = nullptr;
  if (L.ph == FP_UNIT) {
    kind = FK_UNIT;
  } else if (L.ph == FP_RECORD) {
    kind = FK_RECORD;
  } else if (L.q != kFQNone) {
    if (exact && (L.flags & F_ISA)) {
      kind = FK_ISA;                                  // the row after the text run
      a0 = Dd.isa + L.q;
    } else {
      kind = FK_TEXT;
      lim = min(min((uint32_t)L.cnt, 32u), L.q);
      const uint32_t nwords = (n + 15u) >> 4;
      uint32_t w;
      if (L.dsel) {
        w = min((n - L.q) >> 4, nwords - 1u);
      } else {
        const int64_t b = (int64_t)L.q - 32;
        w = b >= 0 ? (uint32_t)b >> 4 : 0u;
      }
      nextw = (w & 3u) >= 2u && w + 4u - (w & 3u) < nwords;
      a0 = P.ix.text + w;
    }
  } else if (L.r > 1u) {
    kind = FK_MTEXT;
  } else if (textmode && L.k == L.l) {
    kind = FK_SA;
    a0 = Dd.sa + L.k;
  } else {
    kind = FK_LF;
    // From the whole index (the exact search's start, a walk's start or
    // restart) the first k bases are one table read when they are A/C/G/T
    // and their entry is not empty, else the first k - 4 from the smaller
    // table, else LF queries one base at a time.
    const uint32_t K2 = P.ix.kmer2_k;
    if (K && L.k == 0u && L.l == n) {
      if (!(L.flags & F_TABFAIL) && (uint32_t)L.cnt >= K && (rb & ((1u << K) - 1u)) == 0u) {
        kind = FK_TABLE;
        tk = K;
        code = (uint32_t)rwc & ((1u << (2u * K)) - 1u);
        a0 = reinterpret_cast<const uint32_t *>((L.dsel ? P.ix.kmer[1] : P.ix.kmer[0]) + code);
      } else if (K2 && !(L.flags & F_TAB2FAIL) && (uint32_t)L.cnt >= K2 &&
                 (rb & ((1u << K2) - 1u)) == 0u) {
        kind = FK_TABLE;
        tk = K2;
        code = (uint32_t)rwc & ((1u << (2u * K2)) - 1u);
        a0 = reinterpret_cast<const uint32_t *>((L.dsel ? P.ix.kmer2[1] : P.ix.kmer2[0]) + code);
      }
    }
    if (kind == FK_LF && !exact && textmode && L.l > L.k && L.l - L.k < kMRows) {
      kind = FK_MSA;                                  // the rows' suffixes, then their text
      a0 = Dd.sa + L.k;
    }
    if (kind == FK_LF) {
      j1 = L.k / kBlockRows;
      j2 = (L.l + 1u) / kBlockRows;
      two = L.k != L.l && j2 != j1;
    }
  }
  // multi-row text: the words of each row's 32 text bases (as FK_TEXT's)
  const uint32_t nwords = (n + 15u) >> 4;
  uint32_t mw0 = 0u, mw1 = 0u, mw2 = 0u, mw3 = 0u;
  if (kind == FK_MTEXT) {
    auto wordof = [&](uint32_t q) -> uint32_t {
      if (L.dsel) return min((n - q) >> 4, nwords - 1u);
      const int64_t b = (int64_t)q - 32;
      return b >= 0 ? (uint32_t)b >> 4 : 0u;
    };
    mw0 = wordof(L.q0);
    mw1 = wordof(L.q1);
    mw2 = L.r > 2u ? wordof(L.q2) : 0u;
    mw3 = L.r > 3u ? wordof(L.q3) : 0u;
  }
  const bool lf = kind == FK_LF;
  const bool unit = kind == FK_UNIT, record = kind == FK_RECORD;
  const uint8_t *f0;
  if (lf) {
    f0 = reinterpret_cast<const uint8_t *>(Dd.blk + j1);
  } else if (unit) {
    f0 = reinterpret_cast<const uint8_t *>(P.offsets + L.rd);
  } else if (record) {
    f0 = reinterpret_cast<const uint8_t *>(P.front_rec + (size_t)L.rd * (kReadRecordWords / 4u));
  } else if (kind == FK_MTEXT) {
    f0 = reinterpret_cast<const uint8_t *>(P.ix.text);
  } else {
    const uint8_t *w0 = reinterpret_cast<const uint8_t *>(a0);
    f0 = w0 - (reinterpret_cast<uintptr_t>(w0) & 15u);
  }

  // ---- the loads
  uint4 fx0 = opaque4();
  uint4 fx1 = opaque4(), fx2 = opaque4(), fx3 = opaque4();
  uint4 gx0 = opaque4(), gx1 = opaque4(), gx2 = opaque4(), gx3 = opaque4();
  uint64_t o0 = 0ull, o1 = 0ull;
  int32_t fsv = 0;
  if (unit) {                                         // the unit's read offsets and fragment size
    const uint64_t *op = reinterpret_cast<const uint64_t *>(f0);
    o0 = op[0];
    o1 = op[1];
    fsv = P.fragsize[L.rd];
  } else if (kind != FK_MTEXT) {
    fx0 = *reinterpret_cast<const uint4 *>(f0);
  }
  if (lf || nextw || record) fx1 = *reinterpret_cast<const uint4 *>(f0 + 16);
  if (lf || record) {
    fx2 = *reinterpret_cast<const uint4 *>(f0 + 32);
    fx3 = *reinterpret_cast<const uint4 *>(f0 + 48);
  }
  if (record) {
    gx0 = *reinterpret_cast<const uint4 *>(f0 + 64);
    gx1 = *reinterpret_cast<const uint4 *>(f0 + 80);
  }
  const bool mtext = kind == FK_MTEXT;
  if (kind == FK_MSA && (L.k & 3u) + (L.l - L.k + 1u) > 4u)
    fx1 = *reinterpret_cast<const uint4 *>(f0 + 16);
  if (mtext) {
    auto nxt = [&](uint32_t w) { return (w & 3u) >= 2u && w + 4u - (w & 3u) < nwords; };
    const uint4 *tx = reinterpret_cast<const uint4 *>(P.ix.text);
    fx0 = tx[mw0 >> 2];
    fx1 = tx[mw1 >> 2];
    if (L.r > 2u) fx2 = tx[mw2 >> 2];
    if (L.r > 3u) fx3 = tx[mw3 >> 2];
    if (nxt(mw0)) gx0 = tx[(mw0 >> 2) + 1u];
    if (nxt(mw1)) gx1 = tx[(mw1 >> 2) + 1u];
    if (L.r > 2u && nxt(mw2)) gx2 = tx[(mw2 >> 2) + 1u];
    if (L.r > 3u && nxt(mw3)) gx3 = tx[(mw3 >> 2) + 1u];
  }
  if (two) {
    const uint8_t *g0 = reinterpret_cast<const uint8_t *>(Dd.blk + j2);
    gx0 = *reinterpret_cast<const uint4 *>(g0);
    gx1 = *reinterpret_cast<const uint4 *>(g0 + 16);
    gx2 = *reinterpret_cast<const uint4 *>(g0 + 32);
    gx3 = *reinterpret_cast<const uint4 *>(g0 + 48);
  }
  asm volatile("" ::"v"(fx0.x), "v"(fx1.x), "v"(fx2.x), "v"(fx3.x), "v"(gx0.x), "v"(gx1.x),
               "v"(gx2.x), "v"(gx3.x), "v"((uint32_t)o0), "v"((uint32_t)o1), "v"(fsv));
  if (unit) {
    if (front_unit(P, L, o0, o1, fsv)) L.ph = FP_RECORD;
    else L.ph = FP_IDLE;
    return;
  }
  if (record) {
    front_stage(P, slot, L, fx0, fx1, fx2, fx3, gx0, gx1);
    return;
  }

  // ---- the transitions
  const uint32_t sel = (uint32_t)(reinterpret_cast<uintptr_t>(a0) & 15u) >> 2;
  bool restart = false;      // walk: the step at i emptied the interval
  bool ended = false;        // the exact search or the walk is over
  if (kind == FK_MSA) {
    if (COUNT) wk[KW_SAISA] += L.l - L.k + 1u;
    const uint32_t c0 = L.k & 3u;
    auto sa_at = [&](uint32_t j) -> uint32_t {   // SA[k + j], from the two chunks
      const uint32_t x = c0 + j;
      return x < 4u ? pick4u(fx0, x) : pick4u(fx1, x - 4u);
    };
    L.r = L.l - L.k + 1u;
    L.q0 = sa_at(0u);
    L.q1 = sa_at(1u);
    L.q2 = L.r > 2u ? sa_at(2u) : 0u;
    L.q3 = L.r > 3u ? sa_at(3u) : 0u;
    return;
  }
  if (kind == FK_MTEXT) {
    if (COUNT) wk[KW_TEXT] += L.r;
    uint64_t spread = rb;                             // N flag of base j -> bit 2j
    spread = (spread | (spread << 16)) & 0x0000FFFF0000FFFFull;
    spread = (spread | (spread << 8)) & 0x00FF00FF00FF00FFull;
    spread = (spread | (spread << 4)) & 0x0F0F0F0F0F0F0F0Full;
    spread = (spread | (spread << 2)) & 0x3333333333333333ull;
    spread = (spread | (spread << 1)) & 0x5555555555555555ull;
    const uint32_t lc = min((uint32_t)L.cnt, 32u);   // every row compares the same bases
    auto run = [&](uint32_t q, uint32_t w, const uint4 &a, const uint4 &b) -> uint32_t {
      const uint32_t sl = w & 3u;
      const uint32_t t0 = pick4u(a, sl);
      const uint32_t t1 = w + 1u >= nwords ? 0u : (sl == 3u ? b.x : pick4u(a, sl + 1u));
      const uint32_t t2 = w + 2u >= nwords ? 0u : (sl >= 2u ? pick4u(b, sl - 2u) : pick4u(a, sl + 2u));
      const uint32_t lim = min(lc, q);
      if (lim == 0u) return 0u;
      const uint64_t x = rwc ^ ftext_window(L.dsel, n, q, t0, t1, t2);
      uint64_t miss = ((x | (x >> 1)) & 0x5555555555555555ull) | spread;
      if (lim < 32u) miss |= 1ull << (2u * lim);
      return miss ? (uint32_t)__builtin_ctzll(miss) >> 1 : 32u;
    };
    const uint32_t m0 = run(L.q0, mw0, fx0, gx0), m1 = run(L.q1, mw1, fx1, gx1);
    const uint32_t m2 = L.r > 2u ? run(L.q2, mw2, fx2, gx2) : 0u;
    const uint32_t m3 = L.r > 3u ? run(L.q3, mw3, fx3, gx3) : 0u;
    const bool s0 = m0 == lc, s1 = m1 == lc, s2 = L.r > 2u && m2 == lc, s3 = L.r > 3u && m3 == lc;
    const uint32_t ns = (s0 ? 1u : 0u) + (s1 ? 1u : 0u) + (s2 ? 1u : 0u) + (s3 ? 1u : 0u);
    if (ns == 0u) {
      // the interval empties at the first base no row continues with
      const uint32_t mm = max(max(m0, m1), max(m2, m3));
      L.i += (int32_t)mm * L.step;
      L.cnt -= (int32_t)mm;
      L.r = 0u;
      restart = true;
    } else {
      // the survivors, in order, lc bases further on
      uint32_t nq0 = 0u, nq1 = 0u, nq2 = 0u, nq3 = 0u, x = 0u;
      auto put = [&](bool keep, uint32_t v) {
        if (!keep) return;
        if (x == 0u) nq0 = v;
        else if (x == 1u) nq1 = v;
        else if (x == 2u) nq2 = v;
        else nq3 = v;
        x++;
      };
      put(s0, L.q0 - lc);
      put(s1, L.q1 - lc);
      put(s2, L.q2 - lc);
      put(s3, L.q3 - lc);
      L.i += (int32_t)lc * L.step;
      L.cnt -= (int32_t)lc;
      if (ns == 1u) {
        L.r = 0u;
        L.q = nq0;                                    // one row: single-row text from here
        L.k = L.l = 0u;
      } else {
        L.r = ns;
        L.q0 = nq0; L.q1 = nq1; L.q2 = nq2; L.q3 = nq3;
      }
      if (L.cnt == 0) {
        L.r = 0u;
        L.q = kFQNone;
        ended = true;
      }
    }
  }
  if (kind == FK_SA) {
    if (COUNT) wk[KW_SAISA]++;
    L.q = pick4u(fx0, sel);
    L.flags &= ~(F_MOVED | F_ISA | F_RUNDIED);
    return;
  }
  if (kind == FK_MTEXT) {
    // done above
  } else if (kind == FK_ISA) {
    if (COUNT) wk[KW_SAISA]++;
    const uint32_t row = pick4u(fx0, sel);
    L.ka = L.la = row;
    L.k = L.l = row;
    if (L.flags & F_RUNDIED) { L.k = n + 1u; L.l = n; }
    L.q = kFQNone;
    ended = true;
  } else if (kind == FK_TEXT) {
    if (COUNT) wk[KW_TEXT]++;
    const uint32_t nwords = (n + 15u) >> 4;
    const uint32_t w = (uint32_t)(a0 - P.ix.text);
    const uint32_t t0 = pick4u(fx0, sel);
    const uint32_t t1 = w + 1u >= nwords ? 0u : (sel == 3u ? fx1.x : pick4u(fx0, sel + 1u));
    const uint32_t t2 = w + 2u >= nwords ? 0u : (sel >= 2u ? pick4u(fx1, sel - 2u) : pick4u(fx0, sel + 2u));
    const uint64_t tw = ftext_window(L.dsel, n, L.q, t0, t1, t2);
    uint64_t spread = rb;                             // N flag of base j -> bit 2j
    spread = (spread | (spread << 16)) & 0x0000FFFF0000FFFFull;
    spread = (spread | (spread << 8)) & 0x00FF00FF00FF00FFull;
    spread = (spread | (spread << 4)) & 0x0F0F0F0F0F0F0F0Full;
    spread = (spread | (spread << 2)) & 0x3333333333333333ull;
    spread = (spread | (spread << 1)) & 0x5555555555555555ull;
    uint32_t m = 0u;
    if (lim > 0u) {
      const uint64_t x = rwc ^ tw;
      uint64_t miss = ((x | (x >> 1)) & 0x5555555555555555ull) | spread;
      if (lim < 32u) miss |= 1ull << (2u * lim);
      m = miss ? (uint32_t)__builtin_ctzll(miss) >> 1 : 32u;
    }
    L.i += (int32_t)m * L.step;
    L.cnt -= (int32_t)m;
    L.q -= m;
    if (m > 0u) L.flags |= F_MOVED;
    const bool dies = m < lim || lim == 0u;
    if (!dies && L.cnt > 0) return;                  // the run goes on
    if (exact) {
      if (L.flags & F_MOVED) {                        // the row after the matched bases
        L.flags |= F_ISA | (dies ? F_RUNDIED : 0u);
        return;
      }
      L.k = n + 1u;                                   // died on its first base
      L.l = n;
      L.q = kFQNone;
      ended = true;
    } else {
      L.q = kFQNone;
      if (!dies) {
        ended = true;
      } else {
        restart = true;
      }
    }
  } else if (kind == FK_TABLE) {
    if (COUNT) wk[KW_TABLE]++;
    const bool hi = (reinterpret_cast<uintptr_t>(a0) & 8u) != 0u;
    const uint32_t ek = hi ? fx0.z : fx0.x, el = hi ? fx0.w : fx0.y;
    if (ek <= el) {
      L.k = ek;
      L.l = el;
      if (exact) { L.ka = ek; L.la = el; }
      L.i += (int32_t)tk * L.step;
      L.cnt -= (int32_t)tk;
      L.flags &= ~(F_TABFAIL | F_TAB2FAIL);
      if (L.cnt == 0) ended = true;
    } else {
      // the string is absent: the smaller table, or one base at a time
      L.flags |= tk == K ? F_TABFAIL : F_TAB2FAIL;
      return;
    }
  } else {
    // LF of W[i] on [k, l] (an N gives the empty (n + 1, n))
    const uint32_t sym = (rb & 1u) ? 4u : (uint32_t)(rw & 3u);
    if (COUNT) { wk[KW_LF]++; wk[KW_BLOCKS] += two ? 2u : 1u; }
    uint32_t kn, ln;
    if (sym > 3u) {
      kn = n + 1u;
      ln = n;
    } else {
      const uint32_t s = L.strand ? 3u - sym : sym;
      const uint32_t c = pick4(Dd.C, s);
      const uint32_t t1 = L.k - j1 * kBlockRows, t2 = L.l + 1u - j2 * kBlockRows;
      const Blk B1 = blk_of4(fx0, fx1, fx2, fx3);
      kn = c + 1u + occ_blk(B1, Dd.dollar_blk, Dd.dollar_off, j1, s, t1);
      if (L.k == L.l) {
        ln = kn - 1u + (sym_at(B1, Dd.dollar_blk, Dd.dollar_off, j1, t1) == s ? 1u : 0u);
      } else {
        const Blk B2 = two ? blk_of4(gx0, gx1, gx2, gx3) : B1;
        ln = c + occ_blk(B2, Dd.dollar_blk, Dd.dollar_off, j2, s, t2);
      }
    }
    L.flags &= ~(F_TABFAIL | F_TAB2FAIL);
    if (exact) {
      L.k = kn;
      L.l = ln;
      if (kn > ln) {
        ended = true;                                 // r dies here; the anchor keeps (ka, la)
      } else {
        L.ka = kn;
        L.la = ln;
        L.i += L.step;
        L.cnt--;
        if (L.cnt == 0) ended = true;
      }
    } else if (kn > ln) {
      restart = true;
    } else {
      L.k = kn;
      L.l = ln;
      L.i += L.step;
      L.cnt--;
      if (L.cnt == 0) ended = true;
    }
  }
  if (restart) {                                      // walk: D counts one more from i on
    slot[(24u + ((uint32_t)L.i >> 5)) * 256u] |= 1u << ((uint32_t)L.i & 31u);
    L.r = 0u;
    L.q = kFQNone;
    L.k = 0u;
    L.l = n;
    L.i += L.step;
    L.cnt--;
    L.flags = 0u;
    if (L.cnt == 0) ended = true;
  }
  if (!ended) return;
  if (exact) {
    // r = (k, l) at position i, the anchor the last live interval
    P.front_r[2ull * L.u] = make_uint4(L.k, L.l, L.ka, L.la);
    P.front_r[2ull * L.u + 1ull] = make_uint4((uint32_t)L.i, 0u, 0u, 0u);   // bit 17 of .z: D follows
    if (L.k > L.l) { L.ph = FP_IDLE; return; }
    // walk 1 (k_calc_d): on the same view from r.pos to the read's end
    L.p1 = L.i;
    L.cnt = L.fwd ? L.nW - L.i : L.i + 1;
    L.q = kFQNone;
    L.r = 0u;
    L.flags = 0u;
    L.ph = FP_WALK1;
    return;
  }
  if (L.ph == FP_WALK1 && (L.fwd ? L.f0 > 0 : L.f1 < L.nW - 1)) {
    // walk 2: on the other view from the fragment's other side, whole index
    L.dsel ^= 1u;
    L.step = -L.step;
    L.i = L.fwd ? L.f0 - 1 : L.f1 + 1;
    L.cnt = L.fwd ? L.f0 : L.nW - 1 - L.f1;
    L.k = 0u;
    L.l = n;
    L.q = kFQNone;
    L.r = 0u;
    L.flags = 0u;
    L.ph = FP_WALK2;
    return;
  }
  // the marks and what D needs besides them; D is written by the next launch
  uint4 *fm = P.front_m + 2ull * L.u;
  fm[0] = make_uint4(slot[24u * 256u], slot[25u * 256u], slot[26u * 256u], slot[27u * 256u]);
  fm[1] = make_uint4(slot[28u * 256u], slot[29u * 256u], slot[30u * 256u], slot[31u * 256u]);
  P.front_r[2ull * L.u + 1ull] = make_uint4((uint32_t)L.p1, (uint32_t)L.f0 | ((uint32_t)L.f1 << 16),
                                            (uint32_t)L.nW | (L.fwd << 16) | (1u << 17), 0u);
  L.ph = FP_IDLE;
}

template <bool COUNT>
__global__ __launch_bounds__(256) void search_k_front_kernel(KParams P) {
  __shared__ uint32_t fsl[kFSlotWords * 256u];
  uint32_t *slot = fsl + threadIdx.x;
  const uint32_t nunits = P.ntasks * kFrontPasses;
  const uint32_t G = gridDim.x * 256u;
  uint32_t unext = blockIdx.x * 256u + threadIdx.x;   // units unext, unext + G, ...
  uint32_t wk[KW_N] = {0, 0, 0, 0, 0, 0};
  FLane L;
  L.ph = FP_IDLE;
  while (true) {
    if (L.ph == FP_IDLE) {
      if (unext >= nunits) break;
      L.u = unext;
      L.rd = (unext % P.ntasks) >> 1;
      unext += G;
      L.ph = FP_UNIT;
    }
    front_step<COUNT>(P, slot, L, wk);
  }
  if (COUNT)
    for (uint32_t i = 0; i < KW_N; ++i) atomicAdd(&P.work[i], (unsigned long long)wk[i]);
}

// D of unit u (if its r is live) into dst[0 .. nW), from its marks: done by
// the main kernel's lane when it takes the task, so that the D the pass reads
// is freshly written into the lane's own scratch.
__device__ __forceinline__ void front_expand_d(const KParams &P, uint64_t u, uint8_t *dst) {
  const uint4 e = P.front_r[2ull * u + 1ull];
  if (!(e.z & (1u << 17))) return;
  const int32_t nW = (int32_t)(e.z & 0xFFFFu);
  const uint4 a = P.front_m[2ull * u], b = P.front_m[2ull * u + 1ull];
  const uint32_t m[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  const bool fwd = ((e.z >> 16) & 1u) != 0u;
  const int32_t p1 = (int32_t)e.x, f0 = (int32_t)(e.y & 0xFFFFu), f1 = (int32_t)(e.y >> 16);
  // the second walk lies below f0 (forward pass) or from f1 + 1 on
  const int32_t xb = fwd ? f0 : f1 + 1;
  uint32_t tot = 0u, below = 0u;
#pragma unroll
  for (uint32_t j = 0; j < 8u; ++j) {
    const uint32_t c = (uint32_t)__builtin_popcount(m[j]);
    tot += c;
    if ((int32_t)(32u * j + 32u) <= xb) below += c;
    else if ((int32_t)(32u * j) < xb)
      below += (uint32_t)__builtin_popcount(m[j] & ((1u << ((uint32_t)xb & 31u)) - 1u));
  }
  const uint32_t cntB = fwd ? below : tot - below;
  uint32_t pre = 0u;
#pragma unroll
  for (uint32_t w = 0; w < 8u; ++w) {
    if ((int32_t)(32u * w) < nW)
      *reinterpret_cast<uint4 *>(dst + 32u * w) = front_d_chunk(pre, m[w], tot, cntB, fwd, p1, (int32_t)(2u * w));
    if ((int32_t)(32u * w + 16u) < nW)
      *reinterpret_cast<uint4 *>(dst + 32u * w + 16u) =
          front_d_chunk(pre, m[w], tot, cntB, fwd, p1, (int32_t)(2u * w + 1u));
    pre += (uint32_t)__builtin_popcount(m[w]);
  }
}

// After the front: a task whose every pass the front ran and whose exact
// fragment search died in each of them (most tasks of the strand a read was
// not drawn from) has no partial to extend: its result is the passes'
// anchors (add_result, results.h:93-103, up to anchor_cap; a full anchor
// list ends the search as BWSearchCPU's return 1) and an empty final list,
// written here; every other task goes to the main kernel's list.
static_assert(sizeof(fmi_hitk) == 64 && offsetof(fmi_hitk, start) == 16 &&
                  offsetof(fmi_hitk, pos) == 18 && offsetof(fmi_hitk, end) == 20 &&
                  offsetof(fmi_hitk, dir) == 22 && offsetof(fmi_hitk, num_mismatches) == 23,
              "fmi_hitk layout");
__global__ __launch_bounds__(256) void search_k_triage_kernel(KParams P) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= P.ntasks) return;
  const uint32_t rd = t >> 1, strand = t & 1u;
  const uint64_t off = P.offsets[rd];
  const int32_t len = (int32_t)(P.offsets[rd + 1] - off);
  const int32_t fs = P.fragsize[rd];
  bool trivial = false;
  int32_t fragments = 0;
  if ((P.strand_mask & (1u << strand)) && fs > 0 && len <= (int32_t)kFrontLen &&
      len / fs - 1 <= FMI_KMAX) {
    fragments = len / fs;
    if (fragments >= 2 && fragments <= (int32_t)kFrontPasses) {
      trivial = true;
      for (int32_t b = 0; b < fragments; ++b) {
        const uint4 a = P.front_r[2ull * ((uint64_t)b * P.ntasks + t)];
        if (a.x <= a.y) trivial = false;              // r live: partials to extend
      }
    }
  }
  if (!trivial) {
    P.tasks[atomicAdd(P.ntasks_dev, 1u)] = t;
    return;
  }
  int32_t half = fragments / 2;
  if (fragments % 2) half++;
  uint32_t na = 0u;
  bool full = false;
  for (int32_t b = 0; b < fragments; ++b) {
    if (!P.anchors) break;
    if (na >= P.anchor_cap) { full = true; break; }
    const bool fwd = b < half;
    const int32_t i = b < half - 1 ? half - 1 - b : (b == half - 1 ? 0 : b);
    const int32_t start = fs * i;
    const int32_t end = fwd || i < fragments - 1 ? fs * (i + 1) - 1 : len - 1;
    const uint64_t u = (uint64_t)b * P.ntasks + t;
    const uint4 a = P.front_r[2ull * u];
    const int16_t pos = (int16_t)P.front_r[2ull * u + 1ull].x;
    // the anchor k_exact leaves (k_clear'ed, no errors), as four 16-byte
    // stores: k, l (k, l < 2^32), start, pos, end, dir, num_mismatches = 0
    const int16_t as = fwd ? (int16_t)start : (int16_t)(pos + 1);
    const int16_t ae = fwd ? (int16_t)(pos - 1) : (int16_t)end;
    uint4 *dst = reinterpret_cast<uint4 *>(P.anchors + (uint64_t)t * P.anchor_cap + na);
    dst[0] = make_uint4(a.z, 0u, a.w, 0u);
    dst[1] = make_uint4((uint32_t)(uint16_t)as | ((uint32_t)(uint16_t)pos << 16),
                        (uint32_t)(uint16_t)ae | ((fwd ? 1u : 0u) << 16), 0u, 0u);
    dst[2] = make_uint4(0u, 0u, 0u, 0u);
    dst[3] = make_uint4(0u, 0u, 0u, 0u);
    na++;
  }
  P.nfin[t] = 0;
  if (P.anchors) P.nanchors[t] = na;
  P.full[t] = full ? 1 : 0;
}

template <bool COUNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void search_k_kernel(KParams P) {
  __shared__ uint32_t kslots[kReadRecordWords * 256u];   // the lanes' reads (front tasks)
  const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  fmi_hitk *sl = P.scratch + lane * 4ull * P.list_cap;
  uint32_t wk[KW_N] = {0, 0, 0, 0, 0, 0};
  const uint32_t ntasks = P.tasks ? *P.ntasks_dev : P.ntasks;
  while (true) {
    const uint32_t q = atomicAdd(P.counter, 1u);
    if (q >= ntasks) break;
    const uint32_t t = P.tasks ? P.tasks[q] : q;
    const uint32_t rd = t >> 1, strand = t & 1u;
    const uint64_t off = P.offsets[rd];
    const int32_t len = (int32_t)(P.offsets[rd + 1] - off);
    const int32_t fs = P.fragsize[rd];
    P.nfin[t] = 0;
    if (P.anchors) P.nanchors[t] = 0;
    if (!(P.strand_mask & (1u << strand))) { P.full[t] = 0; continue; }
    if (fs <= 0 || len > (int32_t)FMI_MAX_READ_LEN || len / fs - 1 > FMI_KMAX) {
      P.full[t] = 2;                 // not searched: fragment size out of range
      continue;
    }
    KCtx c;
    c.ix = &P.ix;
    c.wk = COUNT ? wk : nullptr;
    // + strand: (backward, forward); - strand: (forward_rev, backward_rev)
    c.B = KView{strand ? &P.ix.d[1] : &P.ix.d[0], strand, strand ? 1u : 0u};
    c.F = KView{strand ? &P.ix.d[0] : &P.ix.d[1], strand, strand ? 0u : 1u};
    c.text = P.ix.text != nullptr && P.ix.d[0].sa != nullptr && P.ix.d[1].sa != nullptr;
    c.W = P.reads + off;
    c.nW = len;
    c.D = P.dscratch + lane * (uint64_t)kDStride;
    const bool fr = P.front_r != nullptr && len <= (int32_t)kFrontLen;
    c.fd = fr ? c.D : nullptr;
    c.fr = P.front_r + 2ull * t;
    c.fstep = P.ntasks;
    c.slot = nullptr;
    if (fr) {
      // D of the front's passes, pass b at D + b * kFrontLen
      const int32_t nb = min(len / fs, (int32_t)kFrontPasses);
      for (int32_t b = 0; b < nb; ++b)
        front_expand_d(P, (uint64_t)b * P.ntasks + t, c.D + (uint32_t)b * kFrontLen);
      // the read's codes and N flags from its record into the lane's slot
      uint32_t *sw = kslots + threadIdx.x;
      const uint4 *r = P.front_rec + (size_t)rd * (kReadRecordWords / 4u);
      const uint32_t nc = ((uint32_t)len + 63u) >> 6, nf = ((uint32_t)len + 127u) >> 7;
      for (uint32_t j = 0; j < nc; ++j) {
        const uint4 v4 = r[j];
        sw[(4u * j) * 256u] = v4.x; sw[(4u * j + 1u) * 256u] = v4.y;
        sw[(4u * j + 2u) * 256u] = v4.z; sw[(4u * j + 3u) * 256u] = v4.w;
      }
      for (uint32_t j = 0; j < nf; ++j) {
        const uint4 v4 = r[4u + j];
        sw[(16u + 4u * j) * 256u] = v4.x; sw[(17u + 4u * j) * 256u] = v4.y;
        sw[(18u + 4u * j) * 256u] = v4.z; sw[(19u + 4u * j) * 256u] = v4.w;
      }
      c.slot = sw;
    }
    KList prev{sl, 0, P.list_cap}, next{sl + P.list_cap, 0, P.list_cap};
    KList prev_i{sl + 2 * P.list_cap, 0, P.list_cap}, next_i{sl + 3 * P.list_cap, 0, P.list_cap};
    KList fin{P.fin + (uint64_t)t * P.list_cap, 0, P.list_cap};
    KList anc{P.anchors ? P.anchors + (uint64_t)t * P.anchor_cap : nullptr, 0, P.anchor_cap};
    const bool f = k_search(c, prev, next, prev_i, next_i, fin, anc, fs, strand ? 0 : 1);
    P.nfin[t] = fin.n;
    if (P.anchors) P.nanchors[t] = anc.n;
    P.full[t] = f ? 1 : 0;
  }
  if (COUNT)
    for (uint32_t i = 0; i < KW_N; ++i) atomicAdd(&P.work[i], (unsigned long long)wk[i]);
}

}  // namespace

uint64_t search_k_lane_bytes(uint32_t list_cap) {
  return 4ull * list_cap * sizeof(fmi_hitk) + kDStride;
}

uint64_t search_k_scratch_bytes(uint32_t lanes, uint32_t list_cap) {
  return (uint64_t)lanes * search_k_lane_bytes(list_cap);
}

// The front kernel's scratch: r / anchor (32 bytes) and the restart marks
// (32 bytes) per unit, then the read records, the list of longer reads and
// its length.
uint64_t search_k_front_bytes(uint32_t nreads) {
  return (uint64_t)kFrontPasses * 2ull * nreads * 64ull +
         (uint64_t)nreads * (4ull * kReadRecordWords + 4ull + 8ull) + 64ull;
}

uint32_t search_k_front_lanes_per_cu() {
  static const uint32_t lanes = [] {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, search_k_front_kernel<false>, 256,
                                                     0) != hipSuccess ||
        blocks < 1)
      blocks = 1;
    return 256u * (uint32_t)blocks;
  }();
  return lanes;
}

// Three waves per SIMD (168 registers, a few spilled): 14% faster than two
// without spills on the 256 Mbp mixed batch, and faster than four.
uint32_t search_k_lanes_per_cu() {
  static const uint32_t lanes = [] {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, search_k_kernel<false>, 256, 0) !=
            hipSuccess ||
        blocks < 1)
      blocks = 2;
    return 256u * (uint32_t)blocks;
  }();
  return lanes;
}

hipError_t launch_search_k(const DevIndex &ix, const uint8_t *reads, const uint64_t *offsets,
                           const int16_t *fragsize, uint32_t nreads, uint32_t strand_mask,
                           uint32_t list_cap, uint32_t anchor_cap, fmi_hitk *fin, uint32_t *nfin,
                           fmi_hitk *anchors, uint32_t *nanchors, uint8_t *full, void *scratch,
                           uint32_t lanes, uint32_t front_lanes, uint32_t *counter,
                           hipStream_t s, unsigned long long *work) {
  KParams P;
  P.work = work;
  P.ix = ix;
  P.reads = reads;
  P.offsets = offsets;
  P.fragsize = fragsize;
  P.fin = fin;
  P.anchors = anchors;
  P.nfin = nfin;
  P.nanchors = nanchors;
  P.full = full;
  P.scratch = static_cast<fmi_hitk *>(scratch);
  P.dscratch = static_cast<uint8_t *>(scratch) + (uint64_t)lanes * 4ull * list_cap * sizeof(fmi_hitk);
  P.counter = counter;
  P.ntasks = 2u * nreads;
  P.list_cap = list_cap;
  P.anchor_cap = anchor_cap;
  P.strand_mask = strand_mask;
  P.front_r = nullptr;
  P.front_m = nullptr;
  P.front_rec = nullptr;
  P.tasks = nullptr;
  P.ntasks_dev = nullptr;
  hipError_t e = launch_zero(counter, 1, s);
  if (e != hipSuccess) return e;
  if (front_lanes) {
    // the first passes' exact searches and D, from the read records
    uint8_t *fb = static_cast<uint8_t *>(scratch) + search_k_scratch_bytes(lanes, list_cap);
    const uint64_t units = (uint64_t)kFrontPasses * 2ull * nreads;
    P.front_r = reinterpret_cast<uint4 *>(fb);
    P.front_m = reinterpret_cast<uint4 *>(fb + units * 32ull);
    uint32_t *rec = reinterpret_cast<uint32_t *>(fb + units * 64ull);
    uint32_t *list = rec + (uint64_t)kReadRecordWords * nreads;
    uint32_t *list_n = list + nreads;
    P.ntasks_dev = list_n + 1;
    P.tasks = list_n + 4;                 // 2 x nreads entries
    e = launch_zero(list_n, 2, s);
    if (e != hipSuccess) return e;
    e = launch_read_records(reads, offsets, nreads, rec, list, list_n, s);
    if (e != hipSuccess) return e;
    const uint32_t fl = (uint32_t)std::min<uint64_t>((units + 255ull) & ~255ull, front_lanes);
    P.front_rec = reinterpret_cast<const uint4 *>(rec);
    if (work) hipLaunchKernelGGL(search_k_front_kernel<true>, dim3(fl / 256u), dim3(256), 0, s, P);
    else hipLaunchKernelGGL(search_k_front_kernel<false>, dim3(fl / 256u), dim3(256), 0, s, P);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(search_k_triage_kernel, dim3((2u * nreads + 255u) / 256u), dim3(256), 0, s, P);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (work) hipLaunchKernelGGL(search_k_kernel<true>, dim3(lanes / 256), dim3(256), 0, s, P);
  else hipLaunchKernelGGL(search_k_kernel<false>, dim3(lanes / 256), dim3(256), 0, s, P);
  return hipGetLastError();
}

}  // namespace fmi

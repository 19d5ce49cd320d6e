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

#include <cstdint>

#include "../../include/fmi.h"
#include "fmi_device.h"
#include "fmi_internal.h"
#include "fmi_kernels.h"

namespace fmi {

namespace {

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
};

enum : uint32_t { KW_BLOCKS = 0, KW_LF = 1, KW_SAISA = 2, KW_TEXT = 3, KW_TABLE = 4, KW_LIST = 5,
                  KW_N = 6 };
#define KWORK(c, i, v)                          \
  do {                                          \
    if ((c).wk) (c).wk[(i)] += (uint32_t)(v);   \
  } while (0)

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
  for (uint32_t j = 0; j < K; ++j) {
    const uint32_t s = c.W[i + (int32_t)j * step];
    bad |= s > 3u ? 1u : 0u;
    code |= (s & 3u) << (2u * j);
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
  k_exact(c, X, dx, r, &ra);
  if (anc.list) KADD(ra, anc);
  if (fwd) r.end = (int16_t)(nW - 1);
  else r.start = 0;
  if (r.k > r.l) return false;
  k_calc_d(c, r, !fwd);
  KADD(r, prev);
  KList &out1 = two ? prev_i : fin;          // where X's complete partials go
  while (err_count > 0) {
    if (k_partials(c, maxe, X, dx, prev, next, out1, fs, lb)) { flow = false; break; }
    if (two) {
      k_turn(c, X, Y, prev_i, !fwd, bound);
      if (k_partials(c, maxe, Y, dy, prev_i, next_i, fin, fs, lb)) { flow = false; break; }
    }
    if (err_count == 1) break;
    if (k_branches<false>(c, X, dx, next, prev, 0, 0)) { flow = false; break; }
    if (two && k_branches<false>(c, Y, dy, next_i, prev_i, 0, 0)) { flow = false; break; }
    err_count--;
  }
  if (flow) {
    k_branches<true>(c, X, dx, next, out1, fs, lb);
    if (two) {
      k_turn(c, X, Y, prev_i, !fwd, bound);
      k_finals(c, Y, dy, prev_i, fin, fs, lb);
      k_branches<true>(c, Y, dy, next_i, fin, fs, lb);
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
  uint8_t *dscratch;                 // per lane: FMI_MAX_READ_LEN bytes of D
  uint32_t *counter;
  unsigned long long *work;          // counting launch: KW_N counters
  uint32_t ntasks, list_cap, anchor_cap, strand_mask;
};

template <bool COUNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void search_k_kernel(KParams P) {
  const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  fmi_hitk *sl = P.scratch + lane * 4ull * P.list_cap;
  uint32_t wk[KW_N] = {0, 0, 0, 0, 0, 0};
  while (true) {
    const uint32_t t = atomicAdd(P.counter, 1u);
    if (t >= P.ntasks) break;
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
    c.D = P.dscratch + lane * (uint64_t)FMI_MAX_READ_LEN;
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
  return 4ull * list_cap * sizeof(fmi_hitk) + FMI_MAX_READ_LEN;
}

uint64_t search_k_scratch_bytes(uint32_t lanes, uint32_t list_cap) {
  return (uint64_t)lanes * search_k_lane_bytes(list_cap);
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
                           uint32_t lanes, uint32_t *counter, hipStream_t s,
                           unsigned long long *work) {
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
  hipError_t e = launch_zero(counter, 1, s);
  if (e != hipSuccess) return e;
  if (work) hipLaunchKernelGGL(search_k_kernel<true>, dim3(lanes / 256), dim3(256), 0, s, P);
  else hipLaunchKernelGGL(search_k_kernel<false>, dim3(lanes / 256), dim3(256), 0, s, P);
  return hipGetLastError();
}

}  // namespace fmi

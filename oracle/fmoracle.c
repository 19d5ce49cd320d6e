/*
 * fmoracle.c -- CPU restatement of the reference FM-index search path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and
 * the CPU-baseline leg of the benchmark load this library; the HIP engine
 * never links or calls it.
 *
 * Reference paths are relative to c/src/aligners/bwt/ of opencb/hpg-libs.
 * Each function states the reference lines whose behaviour it restates.
 * Symbols are the codes 0..3 of the alphabet "ACGT" (bwt_commons.c:60-91);
 * a read code > 3 (an N) is this build's policy: a fifth symbol, after T,
 * that never occurs in the text, so a LF step on it yields the empty
 * interval (n + 1, n).  The reference leaves that case undefined (its base
 * table only sets ACGT).
 */
#include "fmoracle.h"

#include <stdbool.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* builder                                                                  */
/* ------------------------------------------------------------------------ */

/* Prefix doubling over (rank[i], rank[i+h]) pairs; '$' has rank 0 and is the
 * unique smallest symbol, so ties never need positions beyond the text. */
static const int64_t *g_rank;
static uint64_t g_h, g_n1;
static int cmp_suffix(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  if (g_rank[x] != g_rank[y]) return g_rank[x] < g_rank[y] ? -1 : 1;
  int64_t rx = x + g_h < g_n1 ? g_rank[x + g_h] : -1;
  int64_t ry = y + g_h < g_n1 ? g_rank[y + g_h] : -1;
  return rx < ry ? -1 : (rx > ry ? 1 : 0);
}

int fmo_sa_naive(const uint8_t *text, uint64_t n, uint64_t *sa) {
  uint64_t n1 = n + 1;
  int64_t *rank = (int64_t *)malloc(n1 * sizeof(int64_t));
  int64_t *tmp = (int64_t *)malloc(n1 * sizeof(int64_t));
  if (!rank || !tmp) { free(rank); free(tmp); return -1; }
  for (uint64_t i = 0; i < n; i++) { rank[i] = (int64_t)text[i] + 1; sa[i] = i; }
  rank[n] = 0; sa[n] = n;
  g_rank = rank; g_n1 = n1;
  for (uint64_t h = 1;; h <<= 1) {
    g_h = h;
    qsort(sa, n1, sizeof(uint64_t), cmp_suffix);
    tmp[sa[0]] = 0;
    for (uint64_t i = 1; i < n1; i++)
      tmp[sa[i]] = tmp[sa[i - 1]] + (cmp_suffix(&sa[i - 1], &sa[i]) < 0 ? 1 : 0);
    memcpy(rank, tmp, n1 * sizeof(int64_t));
    if ((uint64_t)rank[sa[n1 - 1]] == n1 - 1) break;
    if (h > n1) break;
  }
  free(rank); free(tmp);
  return 0;
}

/* B.vec convention of dbwt.c:732-763: the n symbols of the BWT with the '$'
 * row left out; `dollar` is the row that holds '$' (the row of suffix 0). */
void fmo_bwt_from_sa(const uint8_t *text, uint64_t n, const uint64_t *sa,
                     uint8_t *bwt, uint64_t *dollar) {
  uint64_t o = 0;
  for (uint64_t i = 0; i <= n; i++) {
    if (sa[i] == 0) { *dollar = i; continue; }
    bwt[o++] = text[sa[i] - 1];
  }
}

static inline uint8_t bwt_row(const uint8_t *bwt, uint64_t dollar, uint64_t i) {
  /* row i of the full BWT, i != dollar (preprocess.c:33-34) */
  return i > dollar ? bwt[i - 1] : bwt[i];
}

static inline uint64_t occ_dir(const fmo_dir *d, unsigned b, uint64_t m) {
  /* get_O, csafm.h:166-181 (FM_COMP_64): checkpoint + popcount of the bits
   * of rows < m within word m/64. */
  uint64_t w = m >> 6;
  unsigned sh = 63u - (unsigned)(m & 63u);
  return (uint64_t)d->desp[b][w] + (uint64_t)__builtin_popcountll(d->count[b][w] << sh);
}

static inline unsigned bwt_sym_dir(const fmo_dir *d, uint64_t m) {
  /* get_B_from_O, csafm.h:183-206: probe bit (m+1)%64 of each symbol's
   * bitmap; 0xFF for the '$' row. */
  m++;
  uint64_t w = m >> 6;
  unsigned bit = (unsigned)(m & 63u);
  for (unsigned b = 0; b < 4; b++)
    if ((d->count[b][w] >> bit) & 1u) return b;
  return 0xFFu;
}

int fmo_dir_from_bwt(fmo_dir *d, const uint8_t *bwt, uint64_t n, uint64_t dollar,
                     uint32_t ratio) {
  memset(d, 0, sizeof(*d));
  d->n = n; d->dollar = dollar; d->ratio = ratio;
  /* calculate_C, preprocess.c:44-72 */
  uint64_t cnt[4] = {0, 0, 0, 0};
  for (uint64_t i = 0; i < n; i++) cnt[bwt[i] & 3]++;
  uint64_t acc = 0;
  for (int b = 0; b < 4; b++) { d->C[b] = (uint32_t)acc; d->C1[b] = (uint32_t)(acc + 1); acc += cnt[b]; }
  /* calculate_O, preprocess.c:74-130 */
  uint64_t siz = n + 2;
  d->words = siz / 64 + (siz % 64 ? 1 : 0);
  for (int b = 0; b < 4; b++) {
    d->desp[b] = (uint32_t *)calloc(d->words, sizeof(uint32_t));
    d->count[b] = (uint64_t *)calloc(d->words, sizeof(uint64_t));
    if (!d->desp[b] || !d->count[b]) return -1;
  }
  uint64_t run[4] = {0, 0, 0, 0};
  for (uint64_t row = 0; row <= n; row++) {
    uint64_t pos = (row + 1) >> 6;
    unsigned bit = (unsigned)((row + 1) & 63u);
    if (bit == 0)
      for (int b = 0; b < 4; b++) d->desp[b][pos] = (uint32_t)run[b];
    if (row == dollar) continue;
    unsigned s = bwt_row(bwt, dollar, row);
    run[s]++;
    d->count[s][pos] |= (uint64_t)1 << bit;
  }
  /* calculate_S_and_R, preprocess.c:7-42: walk LF from row 0 (suffix n)
   * down to suffix 1, sampling S by row and R by text position.  ratio 0
   * leaves the samples out (search only; get_SA/get_ISA then unusable). */
  if (ratio == 0) return 0;
  uint64_t n1 = n + 1;
  d->nsamp = n1 / ratio + (n1 % ratio ? 1 : 0);
  d->S = (uint32_t *)calloc(d->nsamp, sizeof(uint32_t));
  d->R = (uint32_t *)calloc(d->nsamp, sizeof(uint32_t));
  if (!d->S || !d->R) return -1;
  if (dollar % ratio == 0) d->S[dollar / ratio] = 0;
  d->R[0] = (uint32_t)dollar;
  uint64_t i = 0;
  for (uint64_t j = n1 - 1; j > 0; j--) {
    if (i % ratio == 0) d->S[i / ratio] = (uint32_t)j;
    if (j % ratio == 0) d->R[j / ratio] = (uint32_t)i;
    unsigned s = bwt_row(bwt, dollar, i);
    i = d->C[s] + occ_dir(d, s, i + 1);
  }
  return 0;
}

void fmo_dir_free(fmo_dir *d) {
  for (int b = 0; b < 4; b++) { free(d->desp[b]); free(d->count[b]); }
  free(d->S); free(d->R);
  memset(d, 0, sizeof(*d));
}

int fmo_index_from_bwts(fmo_index *ix, const uint8_t *bwt, uint64_t dollar,
                        const uint8_t *bwt_rev, uint64_t dollar_rev, uint64_t n,
                        uint32_t ratio) {
  memset(ix, 0, sizeof(*ix));
  int rc0 = 0, rc1 = 0;
  /* the two directions are independent (bwt.c:772-809 builds them in turn) */
#ifdef _OPENMP
#pragma omp parallel sections num_threads(2)
#endif
  {
#ifdef _OPENMP
#pragma omp section
#endif
    rc0 = fmo_dir_from_bwt(&ix->bwd, bwt, n, dollar, ratio);
#ifdef _OPENMP
#pragma omp section
#endif
    rc1 = fmo_dir_from_bwt(&ix->fwd, bwt_rev, n, dollar_rev, ratio);
  }
  return (rc0 || rc1) ? -1 : 0;
}

int fmo_index_build_naive(fmo_index *ix, const uint8_t *text, uint64_t n, uint32_t ratio) {
  uint64_t *sa = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
  uint8_t *rev = (uint8_t *)malloc(n + 1);
  uint8_t *b1 = (uint8_t *)malloc(n + 1), *b2 = (uint8_t *)malloc(n + 1);
  int rc = -1;
  if (!sa || !rev || !b1 || !b2) goto out;
  uint64_t d1, d2;
  if (fmo_sa_naive(text, n, sa)) goto out;
  fmo_bwt_from_sa(text, n, sa, b1, &d1);
  for (uint64_t i = 0; i < n; i++) rev[i] = text[n - 1 - i];   /* Xi, bwt.c:791-793 */
  if (fmo_sa_naive(rev, n, sa)) goto out;
  fmo_bwt_from_sa(rev, n, sa, b2, &d2);
  rc = fmo_index_from_bwts(ix, b1, d1, b2, d2, n, ratio);
out:
  free(sa); free(rev); free(b1); free(b2);
  return rc;
}

void fmo_index_free(fmo_index *ix) { fmo_dir_free(&ix->bwd); fmo_dir_free(&ix->fwd); }

int fmo_dir_set_samples(fmo_index *ix, int fwd, uint32_t ratio, const uint32_t *S,
                        const uint32_t *R, uint64_t nsamp) {
  fmo_dir *d = fwd ? &ix->fwd : &ix->bwd;
  uint64_t n1 = d->n + 1;
  if (ratio == 0 || nsamp != n1 / ratio + (n1 % ratio ? 1 : 0)) return -1;
  uint32_t *s2 = (uint32_t *)malloc(nsamp * sizeof(uint32_t));
  uint32_t *r2 = (uint32_t *)malloc(nsamp * sizeof(uint32_t));
  if (!s2 || !r2) { free(s2); free(r2); return -1; }
  memcpy(s2, S, nsamp * sizeof(uint32_t));
  memcpy(r2, R, nsamp * sizeof(uint32_t));
  free(d->S); free(d->R);
  d->S = s2; d->R = r2; d->nsamp = nsamp; d->ratio = ratio;
  return 0;
}

uint64_t fmo_dir_words(const fmo_index *ix, int fwd) { return fwd ? ix->fwd.words : ix->bwd.words; }

void fmo_dir_export(const fmo_index *ix, int fwd, uint32_t *C, uint32_t *C1,
                    uint32_t *desp, uint64_t *count, uint32_t *S, uint32_t *R,
                    uint64_t *dollar, uint64_t *nsamp) {
  const fmo_dir *d = fwd ? &ix->fwd : &ix->bwd;
  for (int b = 0; b < 4; b++) {
    if (C) C[b] = d->C[b];
    if (C1) C1[b] = d->C1[b];
    if (desp) memcpy(desp + (size_t)b * d->words, d->desp[b], d->words * sizeof(uint32_t));
    if (count) memcpy(count + (size_t)b * d->words, d->count[b], d->words * sizeof(uint64_t));
  }
  if (S) memcpy(S, d->S, d->nsamp * sizeof(uint32_t));
  if (R) memcpy(R, d->R, d->nsamp * sizeof(uint32_t));
  if (dollar) *dollar = d->dollar;
  if (nsamp) *nsamp = d->nsamp;
}

/* ------------------------------------------------------------------------ */
/* runtime                                                                  */
/* ------------------------------------------------------------------------ */

static inline const fmo_dir *view_dir(const fmo_index *ix, int view) {
  return (view == FMO_VIEW_BACKWARD || view == FMO_VIEW_BACKWARD_REV) ? &ix->bwd : &ix->fwd;
}
static inline int view_comp(int view) { return view >= FMO_VIEW_BACKWARD_REV; }

uint64_t fmo_size_sa(const fmo_index *ix) { return ix->bwd.n + 1; }   /* S.siz, runtime.h:86 */

uint64_t fmo_occ(const fmo_index *ix, int view, unsigned b, uint64_t m) {
  const fmo_dir *d = view_dir(ix, view);
  return occ_dir(d, view_comp(view) ? 3u - b : b, m);
}

/* BWiteration, runtime.h:38-42, on the view's (possibly complement-swapped,
 * csafm.c:3-100) C/C1/O. */
static inline void lf(const fmo_dir *d, int comp, int64_t k, int64_t l, unsigned b,
                      int64_t *ko, int64_t *lo) {
  if (b > 3u) {               /* N: C1 = n + 1, C = n, never occurs */
    *ko = (int64_t)d->n + 1; *lo = (int64_t)d->n; return;
  }
  unsigned s = comp ? 3u - b : b;
  *ko = (int64_t)((uint64_t)d->C1[s] + occ_dir(d, s, (uint64_t)k));
  *lo = (int64_t)((uint64_t)d->C[s] + occ_dir(d, s, (uint64_t)l + 1));
}

void fmo_lf(const fmo_index *ix, int view, int64_t k, int64_t l, unsigned b,
            int64_t *ko, int64_t *lo) {
  lf(view_dir(ix, view), view_comp(view), k, l, b, ko, lo);
}

uint64_t fmo_get_sa(const fmo_index *ix, int fwd, uint64_t m) {
  /* getScompValue, csafm.h:208-235 */
  const fmo_dir *d = fwd ? &ix->fwd : &ix->bwd;
  if (!d->S) return UINT64_MAX;          /* built without samples */
  uint64_t i = m, j = 0;
  while (i % d->ratio) {
    unsigned b = bwt_sym_dir(d, i);
    if (b == 0xFFu) i = 0;
    else i = d->C[b] + occ_dir(d, b, i + 1);
    j++;
  }
  return ((uint64_t)d->S[i / d->ratio] + j) % (d->n + 1);
}

uint64_t fmo_get_isa(const fmo_index *ix, int fwd, uint64_t m) {
  /* getRcompValue, csafm.h:237-272 */
  const fmo_dir *d = fwd ? &ix->fwd : &ix->bwd;
  if (!d->R) return UINT64_MAX;          /* built without samples */
  uint64_t siz = d->n + 1;
  uint64_t i = (d->ratio - (m % d->ratio)) % d->ratio;
  uint64_t k = m + i, j;
  if (k < siz) j = d->R[k / d->ratio];
  else { j = d->R[0]; i = siz - m; }
  while (i) {
    unsigned b = bwt_sym_dir(d, j);
    if (b == 0xFFu) j = 0;
    else j = d->C[b] + occ_dir(d, b, j + 1);
    i--;
  }
  return j;
}

/* ------------------------------------------------------------------------ */
/* search                                                                   */
/* ------------------------------------------------------------------------ */

/* BWExactSearchBackward / BWExactSearchForward, search/search.h:9-65.
 * The last non-empty interval before death is the anchor; both start from
 * (k, l) = (0, 0) when no step succeeds. */
static void exact_dir(const fmo_dir *d, int comp, int dir, const uint8_t *W,
                      fmo_hit *r, fmo_hit *anchor) {
  int64_t k = 0, l = 0, k2 = r->k, l2 = r->l;
  int16_t i;
  if (dir == 0) {
    for (i = r->pos; i >= r->start; i--) {
      lf(d, comp, k2, l2, W[i], &k2, &l2);
      if (k2 > l2) break;
      k = k2; l = l2;
    }
    if (anchor) {
      anchor->dir = 0; anchor->num_mismatches = 0;
      anchor->start = (int16_t)(i + 1); anchor->end = r->pos;
      anchor->k = k; anchor->l = l; anchor->pos = i;
    }
  } else {
    for (i = r->pos; i <= r->end; i++) {
      lf(d, comp, k2, l2, W[i], &k2, &l2);
      if (k2 > l2) break;
      k = k2; l = l2;
    }
    if (anchor) {
      anchor->dir = 1; anchor->num_mismatches = 0;
      anchor->start = r->pos; anchor->end = (int16_t)(i - 1);
      anchor->k = k; anchor->l = l; anchor->pos = i;
    }
  }
  r->k = k2; r->l = l2; r->pos = i;
}

void fmo_exact(const fmo_index *ix, int view, int dir, const uint8_t *W,
               fmo_hit *r, fmo_hit *anchor) {
  exact_dir(view_dir(ix, view), view_comp(view), dir, W, r, anchor);
}

int16_t fmo_exact_vector(const fmo_index *ix, int view, int dir, const uint8_t *W,
                         int16_t start, int16_t end, int64_t k, int64_t l,
                         int64_t *vec_k, int64_t *vec_l) {
  /* BWExactSearchVectorBackward (search.c:1407-1447) and
   * BWExactSearchVectorForward (search.c:1449-1489): the interval after
   * each step; after death the dead interval is repeated to the end. */
  const fmo_dir *d = view_dir(ix, view);
  int comp = view_comp(view);
  int16_t good = 0, i, j;
  if (k > l || start > end) return 0;
  int64_t k2 = k, l2 = l;
  if (dir == 0) {
    for (i = end, j = (int16_t)(end - start); i >= start; i--, j--) {
      lf(d, comp, k2, l2, W[i], &k2, &l2);
      vec_k[j] = k2; vec_l[j] = l2;
      if (k2 > l2) { i--; j--; break; }
      good++;
    }
    for (; i >= start; i--, j--) { vec_k[j] = k2; vec_l[j] = l2; }
  } else {
    for (i = start, j = 0; i <= end; i++, j++) {
      lf(d, comp, k2, l2, W[i], &k2, &l2);
      vec_k[j] = k2; vec_l[j] = l2;
      if (k2 > l2) { i++; j++; break; }
      good++;
    }
    for (; i <= end; i++, j++) { vec_k[j] = k2; vec_l[j] = l2; }
  }
  return good;
}

typedef struct {
  const uint8_t *W;
  fmo_hit *out;
  uint32_t cap, cnt;
  int mm_only;
  uint8_t strand;
} s1_sink;

static inline void sink_add(s1_sink *s, const fmo_hit *h) {
  /* add_result (results.h:93-103) keeps the first max_results records; the
   * reference then stops, which leaves the kept prefix unchanged.  Here the
   * search goes on so that the total count is known. */
  if (s->cnt < s->cap) { s->out[s->cnt] = *h; s->out[s->cnt].strand = s->strand; }
  s->cnt++;
}

/* BWSimpleSearch1Backward (dir 0, search.c:1711-1793) and
 * BWSimpleSearch1Forward (dir 1, search.c:1795-1874).  res carries the
 * surviving exact interval; on exit res->k,l hold the exact extension over
 * the scanned half (or the dead interval) and res->pos = -1. */
static void simple_search1(const fmo_dir *d, int comp, int dir, s1_sink *s, fmo_hit *res) {
  const uint8_t *W = s->W;
  int16_t start, end;
  if (dir == 0) { start = res->start; end = res->pos; }
  else          { start = res->pos;   end = res->end; }
  const int step = dir ? 1 : -1;

  fmo_hit r;
  memset(&r, 0, sizeof(r));
  r.dir = (uint8_t)dir; r.start = start; r.end = end;
  r.num_mismatches = 1; r.err_kind = FMO_MATCH; r.err_base = 0xFF; r.err_pos = start;

  int64_t kn = res->k, ln = res->l, k, l, ka, la;
  int64_t size_next = ln - kn, size;
  int16_t i = dir ? start : end;
  for (; dir ? (i <= end) : (i >= start); i = (int16_t)(i + step)) {
    k = kn; l = ln;
    if (k > l) { res->k = k; res->l = l; res->pos = -1; return; }
    lf(d, comp, k, l, W[i], &kn, &ln);
    size = size_next; size_next = ln - kn;
    if (size == size_next) continue;   /* every occurrence extends by W[i] */

    if (!s->mm_only) {                 /* deletion of W[i] */
      r.k = k; r.l = l; r.pos = (int16_t)(i + step);
      exact_dir(d, comp, dir, W, &r, NULL);
      if (r.k <= r.l) {
        r.err_kind = FMO_DELETION; r.err_base = 0xFF; r.err_pos = i;
        sink_add(s, &r);
      }
    }
    for (unsigned b = 0; b < 4; b++) {
      lf(d, comp, k, l, b, &ka, &la);
      if (ka > la) continue;
      if (!s->mm_only) {               /* insertion of b before W[i] */
        r.k = ka; r.l = la; r.pos = i;
        exact_dir(d, comp, dir, W, &r, NULL);
        if (r.k <= r.l) {
          r.err_kind = FMO_INSERTION; r.err_base = (uint8_t)b; r.err_pos = i;
          sink_add(s, &r);
        }
      }
      if (b != (unsigned)W[i]) {       /* W[i] -> b */
        r.k = ka; r.l = la; r.pos = (int16_t)(i + step);
        exact_dir(d, comp, dir, W, &r, NULL);
        if (r.k <= r.l) {
          r.err_kind = FMO_MISMATCH; r.err_base = (uint8_t)b; r.err_pos = i;
          sink_add(s, &r);
        }
      }
    }
  }
  res->k = kn; res->l = ln; res->pos = -1;
}

uint32_t fmo_search1(const fmo_index *ix, int strand, int mismatch_only,
                     const uint8_t *W, int16_t len, fmo_hit *out, uint32_t cap,
                     fmo_hit *anchors) {
  /* BWSearch1CPU, search.c:1876-1933.  Strand pairing of bwt.c:1693-1701:
   * + strand (backward, forward); - strand (forward_rev, backward_rev). */
  const int vb = strand ? FMO_VIEW_FORWARD_REV : FMO_VIEW_BACKWARD;
  const int vf = strand ? FMO_VIEW_BACKWARD_REV : FMO_VIEW_FORWARD;
  const fmo_dir *db = view_dir(ix, vb), *df = view_dir(ix, vf);
  const int cb = view_comp(vb), cf = view_comp(vf);
  s1_sink s = {W, out, cap, 0, mismatch_only, (uint8_t)strand};

  int16_t start = 0, end = (int16_t)(len - 1);
  int16_t n = (int16_t)(end - start + 1), half = (int16_t)(n / 2);
  int64_t k0 = 0, l0 = (int64_t)fmo_size_sa(ix) - 1;

  fmo_hit r;
  memset(&r, 0, sizeof(r));
  r.dir = 0; r.start = half; r.end = end; r.k = k0; r.l = l0; r.pos = end;
  exact_dir(db, cb, 0, W, &r, anchors ? &anchors[0] : NULL);
  if (r.k <= r.l) {
    r.start = start; r.pos = (int16_t)(half - 1);
    simple_search1(db, cb, 0, &s, &r);
    if (r.k <= r.l) sink_add(&s, &r);   /* the exact match */
  }

  half--;
  memset(&r, 0, sizeof(r));
  r.dir = 1; r.start = start; r.end = half; r.k = k0; r.l = l0; r.pos = start;
  exact_dir(df, cf, 1, W, &r, anchors ? &anchors[1] : NULL);
  if (r.k <= r.l) {
    r.pos = (int16_t)(half + 1); r.end = end;
    simple_search1(df, cf, 1, &s, &r);
  }
  if (anchors) { anchors[0].strand = (uint8_t)strand; anchors[1].strand = (uint8_t)strand; }
  return s.cnt;
}

/* BWSearch1VectorHelper, search.c:1491-1709, driven by the four interval
 * vectors of bwt.c:1940-1950 (or bwt.c:1447-1457 for a seed [start, end]).
 * One result r is reused for every record, so its error fields carry over
 * from one record to the next exactly as the reference's do: only the fields
 * each step assigns change. */
typedef struct {
  int64_t k, l;
  int16_t pos;
  uint8_t dir, nmm, kind, base;
  int16_t err_pos;
} vh_state;

static void vh_add(s1_sink *s, const vh_state *r, int16_t start, int16_t end) {
  if (s->mm_only && r->nmm && r->kind != FMO_MISMATCH) return;
  fmo_hit h;
  memset(&h, 0, sizeof(h));
  h.k = r->k; h.l = r->l; h.start = start; h.end = end; h.pos = r->pos;
  h.dir = r->dir; h.num_mismatches = r->nmm;
  if (r->nmm) { h.err_kind = r->kind; h.err_base = r->base; h.err_pos = r->err_pos; }
  sink_add(s, &h);
}

/* BWExactSearchBackward/Forward on the state (search.h:9-65). */
static void vh_exact(const fmo_dir *d, int comp, int dir, const uint8_t *W, int16_t start,
                     int16_t end, vh_state *r) {
  fmo_hit t;
  memset(&t, 0, sizeof(t));
  t.k = r->k; t.l = r->l; t.pos = r->pos; t.start = start; t.end = end;
  exact_dir(d, comp, dir, W, &t, NULL);
  r->k = t.k; r->l = t.l; r->pos = t.pos;
}

uint32_t fmo_search1_vector(const fmo_index *ix, int strand, int mismatch_only,
                            const uint8_t *W, int16_t start, int16_t end, fmo_hit *out,
                            uint32_t cap) {
  /* + strand: vectors on (backward, forward), bwt.c:1940-1944,1996;
   * - strand: (forward_rev, backward_rev), bwt.c:1946-1950,2000. */
  const int vb = strand ? FMO_VIEW_FORWARD_REV : FMO_VIEW_BACKWARD;
  const int vf = strand ? FMO_VIEW_BACKWARD_REV : FMO_VIEW_FORWARD;
  const fmo_dir *db = view_dir(ix, vb), *df = view_dir(ix, vf);
  const int cb = view_comp(vb), cf = view_comp(vf);
  s1_sink s = {W, out, cap, 0, mismatch_only, (uint8_t)strand};
  if (end - start < 2) return 0;   /* the helper reads outside its vectors below 3 bases */

  const int16_t n = (int16_t)(end - start);
  const int64_t lsa = (int64_t)fmo_size_sa(ix) - 1;
  int64_t *vk = (int64_t *)malloc(4 * (size_t)(n + 1) * sizeof(int64_t));
  int64_t *vl = vk + (n + 1), *vki = vl + (n + 1), *vli = vki + (n + 1);
  fmo_exact_vector(ix, vb, 0, W, start, end, 0, lsa, vk, vl);
  fmo_exact_vector(ix, vf, 1, W, start, end, 0, lsa, vki, vli);

  int16_t half = (int16_t)(n / 2), i, j;
  vh_state r;
  memset(&r, 0, sizeof(r));
  r.dir = 0;
  if (vk[0] <= vl[0]) { r.k = vk[0]; r.l = vl[0]; r.pos = -1; vh_add(&s, &r, start, end); }
  r.nmm = 1; r.kind = FMO_MATCH; r.base = 0xFF; r.err_pos = start;

  int64_t k, l, ka, la, size, size_last;
  /* error at W[start]: the suffix interval W[start+1..end] */
  size = vl[0] - vk[0];
  size_last = size;
  k = vk[1]; l = vl[1];
  size = l - k;
  if (size != size_last) {
    r.k = k; r.l = l; r.pos = -1;
    r.kind = FMO_DELETION; r.base = 0xFF; r.err_pos = start;
    vh_add(&s, &r, start, end);
    for (unsigned b = 0; b < 4; b++) {
      lf(db, cb, k, l, b, &ka, &la);
      if (ka > la) continue;
      const uint8_t bw = W[start];
      if (b != bw) {
        r.k = ka; r.l = la; r.pos = -1;
        r.kind = FMO_MISMATCH; r.base = (uint8_t)b;
        vh_add(&s, &r, start, end);
      }
      lf(db, cb, ka, la, bw, &ka, &la);
      if (ka <= la) {
        /* modify_last_mismatch3(&r, 3, INSERTION, b), search.c:1556: kind 3,
         * base INSERTION, position b */
        r.k = ka; r.l = la; r.pos = -1;
        r.kind = 3; r.base = FMO_INSERTION; r.err_pos = (int16_t)b;
        vh_add(&s, &r, start, end);
      }
    }
  }
  for (i = (int16_t)(start + 2), j = 2; j <= half; i++, j++) {
    size_last = size;
    k = vk[j]; l = vl[j];
    size = l - k;
    if (size == size_last) continue;
    r.k = k; r.l = l; r.pos = (int16_t)(i - 2);
    r.kind = FMO_DELETION; r.base = 0xFF; r.err_pos = (int16_t)(i - 1);
    vh_exact(db, cb, 0, W, start, end, &r);
    if (r.k <= r.l) vh_add(&s, &r, start, end);
    for (unsigned b = 0; b < 4; b++) {
      lf(db, cb, k, l, b, &ka, &la);
      if (ka > la) continue;
      r.k = ka; r.l = la; r.pos = (int16_t)(i - 1);
      r.kind = FMO_INSERTION; r.base = (uint8_t)b;
      vh_exact(db, cb, 0, W, start, end, &r);
      if (r.k <= r.l) vh_add(&s, &r, start, end);
      if (b != W[i - 1]) {
        r.k = ka; r.l = la; r.pos = (int16_t)(i - 2);
        r.kind = FMO_MISMATCH;
        vh_exact(db, cb, 0, W, start, end, &r);
        if (r.k <= r.l) vh_add(&s, &r, start, end);
      }
    }
  }

  half--;
  size = vli[n] - vki[n];
  r.dir = 1;
  /* error at W[end]: the prefix interval W[start..end-1] */
  size_last = size;
  k = vki[n - 1]; l = vli[n - 1];
  size = l - k;
  if (size != size_last) {
    r.k = k; r.l = l; r.pos = -1;
    r.kind = FMO_DELETION; r.base = 0xFF; r.err_pos = end;
    vh_add(&s, &r, start, end);
    for (unsigned b = 0; b < 4; b++) {
      lf(df, cf, k, l, b, &ka, &la);
      if (ka > la) continue;
      const uint8_t bw = W[end];
      if (b != bw) {
        r.k = ka; r.l = la; r.pos = -1;
        r.kind = FMO_MISMATCH; r.base = (uint8_t)b;
        vh_add(&s, &r, start, end);
      }
      lf(df, cf, ka, la, bw, &ka, &la);
      if (ka <= la) {
        r.k = ka; r.l = la; r.pos = -1;
        r.kind = FMO_INSERTION; r.base = (uint8_t)b;
        vh_add(&s, &r, start, end);
      }
    }
  }
  for (i = (int16_t)(end - 2), j = (int16_t)(n - 2); j >= half; i--, j--) {
    size_last = size;
    k = vki[j]; l = vli[j];
    size = l - k;
    if (size == size_last) continue;
    r.k = k; r.l = l; r.pos = (int16_t)(i + 2);
    r.kind = FMO_DELETION; r.base = 0xFF; r.err_pos = (int16_t)(i + 1);
    vh_exact(df, cf, 1, W, start, end, &r);
    if (r.k <= r.l) vh_add(&s, &r, start, end);
    for (unsigned b = 0; b < 4; b++) {
      lf(df, cf, k, l, b, &ka, &la);
      if (ka > la) continue;
      r.k = ka; r.l = la; r.pos = (int16_t)(i + 1);
      r.kind = FMO_INSERTION; r.base = (uint8_t)b;
      vh_exact(df, cf, 1, W, start, end, &r);
      if (r.k <= r.l) vh_add(&s, &r, start, end);
      if (b != W[i + 1]) {
        r.k = ka; r.l = la; r.pos = (int16_t)(i + 2);
        r.kind = FMO_MISMATCH;
        vh_exact(df, cf, 1, W, start, end, &r);
        if (r.k <= r.l) vh_add(&s, &r, start, end);
      }
    }
  }
  free(vk);
  return s.cnt;
}

uint64_t fmo_seed_map(const fmo_index *ix, const uint8_t *W, uint32_t len, uint32_t seed_len,
                      uint32_t stride, int32_t filter, const uint64_t *koff,
                      const uint64_t *kstart, uint64_t kcount, uint32_t read,
                      fmo_seed_region *out, uint64_t cap, int64_t *info) {
  /* bwt_map_exact_seed, bwt.c:1200-1294, for each seed of the tiling */
  uint64_t nrow = 0;
  if (len < seed_len) return 0;
  const uint32_t nseeds = (len - seed_len) / stride + 1;
  const int64_t lsa = (int64_t)fmo_size_sa(ix) - 1;
  for (uint32_t q = 0; q < nseeds; q++) {
    const uint64_t seq_start = (uint64_t)q * stride, seq_end = seq_start + seed_len - 1;
    const uint64_t slen = seq_end - seq_start;                   /* one short, bwt.c:1214 */
    const uint8_t *code = W + seq_start;
    int32_t actual = 0;                                         /* int actual_mappings */
    for (int type = 1; type >= 0; type--) {
      fmo_hit r;
      memset(&r, 0, sizeof(r));
      r.k = 0; r.l = lsa; r.start = 0; r.end = (int16_t)(seed_len - 1);
      if (type == 1) { r.pos = r.end; fmo_exact(ix, FMO_VIEW_BACKWARD, 0, code, &r, NULL); }
      else { r.pos = 0; fmo_exact(ix, FMO_VIEW_BACKWARD_REV, 1, code, &r, NULL); }
      actual = (int32_t)((uint32_t)actual + (uint32_t)(r.l - r.k + 1));
      const int skip = actual > filter;
      int64_t rows = (!skip && r.k <= r.l) ? r.l - r.k + 1 : 0;
      if (info) {
        info[6 * q + (type ? 0 : 2)] = r.k;
        info[6 * q + (type ? 1 : 3)] = r.l;
        info[6 * q + (type ? 4 : 5)] = rows;
      }
      for (int64_t j = 0; j < rows; j++) {
        const uint64_t key = fmo_get_sa(ix, 0, (uint64_t)(r.k + j));
        const uint64_t idx = fmo_binsearch(koff, kcount, key);
        fmo_seed_region g;
        memset(&g, 0, sizeof(g));
        g.read = read; g.seed_id = (uint16_t)q; g.strand = (uint8_t)(!type);
        g.seq_start = (uint16_t)(type ? seq_start : len - seq_end - 1);
        g.seq_end = (uint16_t)(type ? seq_end : len - seq_start - 1);
        g.chrom_idx = (uint32_t)idx;
        if (idx >= 1 && key + slen <= koff[idx]) {
          g.start = kstart[idx - 1] + (key - koff[idx - 1]) + 1;
          g.end = g.start + slen;
          g.valid = 1;
        }
        if (nrow < cap) out[nrow] = g;
        nrow++;
      }
    }
  }
  return nrow;
}

/* bwt_map_inexact_seed (bwt.c:1401-1644) for the window [ss, se] of W:
 * writes both strands' region lists; *final_pos carries across records and
 * strands as the reference's local does (it is only assigned for some
 * error kinds, bwt.c:1508-1525). */
static void inexact_seed(const fmo_index *ix, const uint8_t *W, uint32_t seq_len, uint32_t ss,
                         uint32_t se, uint32_t seed_id, const uint64_t *koff,
                         const uint64_t *kstart, uint64_t kcount, uint32_t read,
                         uint32_t rec_cap, uint32_t region_cap, fmo_hit *recs,
                         fmo_inexact_region *out2, uint32_t *nreg2, uint8_t *status2) {
  const int64_t len = (int64_t)se - (int64_t)ss + 1;
  const uint64_t size_sa = fmo_size_sa(ix);
  const int max_mappings = 50;                                 /* bwt.c:1409 */
  int final_pos = 0, final_err;
  for (int type = 1; type >= 0; type--) {
    const int s = !type;                                        /* + strand first */
    fmo_inexact_region *out = out2 + (uint64_t)s * region_cap;
    uint8_t st = 0;
    const uint32_t cnt = fmo_search1_vector(ix, s, 0, W, (int16_t)ss, (int16_t)se, recs, rec_cap);
    if (cnt > rec_cap) st |= 1;
    const uint32_t nrec = cnt < rec_cap ? cnt : rec_cap;
    const uint32_t aux_s = type ? ss : seq_len - se - 1;
    const uint32_t aux_e = type ? se : seq_len - ss - 1;
    uint32_t m = 0;                                             /* regions in tmp list */
    for (uint32_t ii = 0; ii < nrec; ii++) {
      const fmo_hit *r = &recs[ii];
      const int error = r->num_mismatches ? r->err_kind : 0;
      int64_t len_calc = len;
      if (error == FMO_DELETION) len_calc--;
      else if (error == FMO_INSERTION) len_calc++;
      const int pos = r->err_pos;
      if (error == FMO_DELETION || pos == 0 || pos == len - 1) {
        final_pos = type ? pos : (int)(len - 1 - pos);
      } else if (error == FMO_INSERTION) {
        if (type) final_pos = r->dir ? pos : pos + 1;
        else final_pos = r->dir ? (int)(len - pos) : (int)(len - pos - 1);
      }
      if (error == FMO_MISMATCH) final_err = -1;
      else if (error == FMO_DELETION) final_err = 1;             /* SEED_INSERTION */
      else final_err = 2;                                        /* SEED_DELETION */
      final_err += (int)ss;
      int64_t k0 = r->k, l0 = r->l;
      if (!((r->l - r->k + 1) < max_mappings)) l0 = r->k + 1;    /* LIMIT EXCEEDED */
      const int direction = type ? r->dir : !r->dir;
      for (int64_t j = k0; j <= l0; j++) {
        const uint64_t key = direction
            ? size_sa - fmo_get_sa(ix, 1, (uint64_t)j) - (uint64_t)len_calc - 1
            : fmo_get_sa(ix, 0, (uint64_t)j);
        const uint64_t idx = fmo_binsearch(koff, kcount, key);
        if (!(idx >= 1 && key + (uint64_t)len_calc <= koff[idx])) continue;
        if (m >= region_cap) { st |= 2; m++; continue; }
        fmo_inexact_region g;
        memset(&g, 0, sizeof(g));
        g.start = kstart[idx - 1] + (key - koff[idx - 1]) + 1;
        g.end = g.start + (uint64_t)len - 1;
        g.chrom_idx = (uint32_t)idx; g.read = read;
        g.seq_start = (uint16_t)aux_s; g.seq_end = (uint16_t)aux_e;
        g.seq_len = (uint16_t)seq_len; g.seed_id = (uint16_t)seed_id;
        g.pos_err = final_pos; g.type_err = final_err;
        g.strand = (uint8_t)(!type);
        out[m++] = g;
      }
    }
    if (m > region_cap) m = region_cap;
    /* dedup within 10 bp (bwt.c:1593-1615): an earlier region of the same
     * contig closer than 10 to a later one that is kept is dropped; the
     * distance is abs() of the 32-bit truncated difference */
    uint8_t del[m ? m : 1];
    memset(del, 0, sizeof(del));
    for (int a1 = (int)m - 1; a1 >= 1; a1--) {
      if (del[a1]) continue;
      for (int a2 = a1 - 1; a2 >= 0; a2--) {
        int32_t d = (int32_t)(uint32_t)(out[a1].start - out[a2].start);
        int64_t dist = d < 0 ? -(int64_t)d : d;
        if (out[a1].chrom_idx == out[a2].chrom_idx && dist < 10 && !del[a2]) del[a2] = 1;
      }
    }
    /* the kept regions are handed over from the last to the first
     * (bwt.c:1616-1626) */
    uint32_t kept = 0;
    for (uint32_t i = 0; i < m; i++)
      if (!del[i]) out[kept++] = out[i];
    for (uint32_t i = 0; i < kept / 2; i++) {
      fmo_inexact_region t = out[i];
      out[i] = out[kept - 1 - i];
      out[kept - 1 - i] = t;
    }
    nreg2[s] = kept;
    status2[s] = st;
  }
}

uint32_t fmo_map_inexact_seeds(const fmo_index *ix, const uint8_t *W, uint32_t len,
                               uint32_t seed_size, uint32_t min_seed_size, const uint64_t *koff,
                               const uint64_t *kstart, uint64_t kcount, uint32_t read,
                               uint32_t rec_cap, uint32_t region_cap, fmo_inexact_region *out,
                               uint32_t *nregions, uint8_t *status, uint32_t wcap) {
  /* bwt_map_inexact_seeds_seq, bwt.c:3262-3302 */
  const uint32_t num_seeds = len / seed_size;
  const uint32_t nwin = num_seeds + (len % seed_size >= min_seed_size ? 1 : 0);
  fmo_hit *recs = (fmo_hit *)malloc((size_t)(rec_cap ? rec_cap : 1) * sizeof(fmo_hit));
  for (uint32_t w = 0; w < nwin && w < wcap; w++) {
    const uint32_t ss = w * seed_size;
    const uint32_t se = w < num_seeds ? ss + seed_size - 1 : len - 1;
    inexact_seed(ix, W, len, ss, se, w, koff, kstart, kcount, read, rec_cap, region_cap, recs,
                 out + (uint64_t)2 * w * region_cap, nregions + 2 * w, status + 2 * w);
  }
  free(recs);
  return nwin;
}

/* __bwt_map_inexact_read (bwt.c:1890-2236) for one read, up to the list of
 * alignments it builds: the vector-driven 1-error search of each strand (+
 * then -, records kept in a list of filter capacity, bwt.c:1984-2003), only
 * its error-free records (bwt.c:2026-2038), the running count of l - k over
 * them checked against filter before each record's rows (bwt.c:2145-2156;
 * over the filter: the strand's alignments are dropped, status bit 0, no
 * later strand), one alignment per SA row inside a contig at 0-based
 * start[idx-1] + key - offset[idx-1] (bwt.c:2158-2166), then per strand the
 * within-10-bp dedup on the int positions (bwt.c:2190-2215), the survivors
 * appended from last to first with the secondary flag num_mappings > 0 and
 * the primary restored on the list's first when a primary was deleted
 * (bwt.c:2217-2233).  Reads under 5 bases give no alignment and status bit 1
 * (the reference's one unmapped placeholder, bwt.c:1899-1915).  At most cap
 * alignments are stored; returns their total. */
static void map_read_strand(const fmo_index *ix, const uint8_t *W, uint32_t len, int type,
                            uint32_t filter, const uint64_t *koff, const uint64_t *kstart,
                            uint64_t kcount, uint32_t read, fmo_hit *recs,
                            fmo_read_mapping *tmp, uint8_t *del, uint64_t *tot, int *exceeded,
                            fmo_read_mapping *out, uint32_t cap, uint32_t *nout) {
  /* bwt.c:1990-2003: type 1 is the + strand (backward, forward), type 0 the
   * - strand (forward_rev, backward_rev) */
  const uint32_t nrec = fmo_search1_vector(ix, type ? 0 : 1, 0, W, 0, (int16_t)(len - 1), recs,
                                           filter);
  const uint32_t stored = nrec < filter ? nrec : filter;
  uint32_t n = 0;
  for (uint32_t ii = 0; ii < stored; ii++) {
    const fmo_hit *r = &recs[ii];
    const int error = r->num_mismatches ? r->err_kind : 0;
    if (error != 0) continue;                          /* bwt.c:2038 */
    const uint64_t direction = type ? r->dir : !r->dir;
    const uint64_t len_calc = len;
    const uint64_t k_start = (uint64_t)r->k, l_start = (uint64_t)r->l;
    *tot += l_start - k_start;
    if (*tot > filter) { *exceeded = 1; return; }     /* bwt.c:2147-2156 */
    for (uint64_t j = k_start; j <= l_start; j++) {
      const uint64_t key = direction ? fmo_size_sa(ix) - fmo_get_sa(ix, 1, j) - len_calc - 1
                                     : fmo_get_sa(ix, 0, j);
      const uint64_t idx = fmo_binsearch(koff, kcount, key);
      if (key + len_calc <= koff[idx]) {
        fmo_read_mapping *m = &tmp[n++];
        m->position = kstart[idx - 1] + (key - koff[idx - 1]);
        m->chromosome = (uint32_t)(idx - 1);
        m->read = read;
        m->strand = (uint8_t)!type;
        /* created with (num_mappings > 0), bwt.c:2176 */
        m->secondary = (uint8_t)(*nout > 0);
        m->cigar_len = (uint16_t)len;
        m->pad = 0;
      }
    }
  }
  /* bwt.c:2190-2215: every alignment carries one CIGAR operation ("<len>M"),
   * so a pair within 10 of each other deletes the later one (a1) */
  memset(del, 0, n);
  for (int a1 = (int)n - 1; a1 >= 1; a1--) {
    if (del[a1]) continue;
    for (int a2 = a1 - 1; a2 >= 0; a2--) {
      /* int positions and abs() of their int difference */
      const int32_t p1 = (int32_t)(uint32_t)tmp[a1].position;
      const int32_t p2 = (int32_t)(uint32_t)tmp[a2].position;
      const int32_t d = (int32_t)((uint32_t)p1 - (uint32_t)p2);
      const uint64_t dist = (uint64_t)(int64_t)(d < 0 ? -(int64_t)d : (int64_t)d);
      if (tmp[a1].chromosome == tmp[a2].chromosome && dist < 10 && !del[a2]) {
        /* alig_1 ops (1) < alig_2 ops (1) never holds */
        del[a1] = 1;
      }
    }
  }
  int primary_delete = 0;
  for (int m = (int)n - 1; m >= 0; m--) {
    if (del[m]) {
      if (!tmp[m].secondary) primary_delete = 1;
    } else {
      tmp[m].secondary = (uint8_t)(*nout > 0);
      if (*nout < cap) out[*nout] = tmp[m];
      (*nout)++;
    }
  }
  if (primary_delete && *nout > 0 && cap > 0) out[0].secondary = 0;
}

uint32_t fmo_map_read(const fmo_index *ix, const uint8_t *W, uint32_t len, uint32_t filter,
                      const uint64_t *koff, const uint64_t *kstart, uint64_t kcount,
                      uint32_t read, fmo_read_mapping *out, uint32_t cap, uint8_t *status) {
  *status = 0;
  if (len < 5) { *status = 2; return 0; }
  fmo_hit *recs = (fmo_hit *)malloc((size_t)(filter ? filter : 1) * sizeof(fmo_hit));
  /* each strand's exact record has l - k <= filter rows here */
  fmo_read_mapping *tmp = (fmo_read_mapping *)malloc((size_t)(filter + 1) * sizeof(fmo_read_mapping));
  uint8_t *del = (uint8_t *)malloc((size_t)filter + 1);
  uint64_t tot = 0;
  int exceeded = 0;
  uint32_t nout = 0;
  for (int type = 1; type >= 0 && !exceeded; type--)
    map_read_strand(ix, W, len, type, filter, koff, kstart, kcount, read, recs, tmp, del, &tot,
                    &exceeded, out, cap, &nout);
  if (exceeded) *status |= 1;
  free(recs);
  free(tmp);
  free(del);
  return nout;
}

uint64_t fmo_binsearch(const uint64_t *array, uint64_t size, uint64_t key) {
  /* binsearch, search/io.h:43-64: first index whose value is > key. */
  if (!array) return 0;
  uint64_t lo = 0;
  while (size > 0) {
    uint64_t w = size / 2;
    if (array[lo + w] <= key) { lo += w + 1; size -= w + 1; }
    else size = w;
  }
  return lo;
}

/* ------------------------------------------------------------------------ */
/* batch drivers (CPU baseline)                                             */
/* ------------------------------------------------------------------------ */

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

double fmo_search1_batch(const fmo_index *ix, const uint8_t *reads,
                         const uint64_t *offsets, uint64_t nreads, int mismatch_only,
                         uint32_t cap, fmo_hit *hits, uint32_t *counts, int threads) {
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
  omp_set_num_threads(threads);
  long chunk = (long)(nreads / ((uint64_t)threads * 10));
  if (chunk < 1) chunk = 1;
#endif
  double t0 = now_s();
#ifdef _OPENMP
#pragma omp parallel
  {
    fmo_hit *scratch = hits ? NULL : (fmo_hit *)malloc((size_t)(cap ? cap : 1) * sizeof(fmo_hit));
#pragma omp for schedule(dynamic, chunk)
#else
  {
    fmo_hit *scratch = hits ? NULL : (fmo_hit *)malloc((size_t)(cap ? cap : 1) * sizeof(fmo_hit));
#endif
    for (long i = 0; i < (long)nreads; i++) {
      const uint8_t *W = reads + offsets[i];
      int16_t len = (int16_t)(offsets[i + 1] - offsets[i]);
      for (int s = 0; s < 2; s++) {
        fmo_hit *dst = hits ? hits + ((size_t)i * 2 + s) * cap : scratch;
        counts[2 * i + s] = fmo_search1(ix, s, mismatch_only, W, len, dst, cap, NULL);
      }
    }
    free(scratch);
  }
  return now_s() - t0;
}

double fmo_exact_batch(const fmo_index *ix, const uint8_t *reads,
                       const uint64_t *offsets, uint64_t nreads, int64_t *out_kl,
                       int threads) {
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
  omp_set_num_threads(threads);
  long chunk = (long)(nreads / ((uint64_t)threads * 10));
  if (chunk < 1) chunk = 1;
#endif
  double t0 = now_s();
  int64_t lsa = (int64_t)fmo_size_sa(ix) - 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, chunk)
#endif
  for (long i = 0; i < (long)nreads; i++) {
    const uint8_t *W = reads + offsets[i];
    int16_t len = (int16_t)(offsets[i + 1] - offsets[i]);
    fmo_hit r;
    memset(&r, 0, sizeof(r));
    r.k = 0; r.l = lsa; r.start = 0; r.end = (int16_t)(len - 1); r.pos = (int16_t)(len - 1);
    fmo_exact(ix, FMO_VIEW_BACKWARD, 0, W, &r, NULL);
    out_kl[4 * i + 0] = r.k; out_kl[4 * i + 1] = r.l;
    memset(&r, 0, sizeof(r));
    r.k = 0; r.l = lsa; r.start = 0; r.end = (int16_t)(len - 1); r.pos = 0;
    fmo_exact(ix, FMO_VIEW_BACKWARD_REV, 1, W, &r, NULL);
    out_kl[4 * i + 2] = r.k; out_kl[4 * i + 3] = r.l;
  }
  return now_s() - t0;
}

/* fmo_exact_batch followed by get_SA(j, backward) over the first cap rows
 * of each interval (bwt.c:1271-1277): rows located and their position sum,
 * per read and strand. */
double fmo_exact_locate_batch(const fmo_index *ix, const uint8_t *reads,
                              const uint64_t *offsets, uint64_t nreads, uint32_t cap,
                              uint32_t *nrows, uint64_t *possum, int threads) {
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
  omp_set_num_threads(threads);
  long chunk = (long)(nreads / ((uint64_t)threads * 10));
  if (chunk < 1) chunk = 1;
#endif
  double t0 = now_s();
  int64_t lsa = (int64_t)fmo_size_sa(ix) - 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, chunk)
#endif
  for (long i = 0; i < (long)nreads; i++) {
    const uint8_t *W = reads + offsets[i];
    int16_t len = (int16_t)(offsets[i + 1] - offsets[i]);
    for (int s = 0; s < 2; s++) {
      fmo_hit r;
      memset(&r, 0, sizeof(r));
      r.k = 0; r.l = lsa; r.start = 0; r.end = (int16_t)(len - 1);
      r.pos = s ? 0 : (int16_t)(len - 1);
      fmo_exact(ix, s ? FMO_VIEW_BACKWARD_REV : FMO_VIEW_BACKWARD, s, W, &r, NULL);
      uint32_t m = 0;
      uint64_t sum = 0;
      for (int64_t j = r.k; j <= r.l && m < cap; j++, m++) sum += fmo_get_sa(ix, 0, (uint64_t)j);
      nrows[2 * i + s] = m;
      possum[2 * i + s] = sum;
    }
  }
  return now_s() - t0;
}

double fmo_map_inexact_seeds_batch(const fmo_index *ix, const uint8_t *reads,
                                   const uint64_t *offsets, uint64_t nreads, uint32_t seed_size,
                                   uint32_t min_seed_size, const uint64_t *koff,
                                   const uint64_t *kstart, uint64_t kcount, uint32_t rec_cap,
                                   uint32_t region_cap, uint64_t *nregions_total, int threads) {
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
  omp_set_num_threads(threads);
  long chunk = (long)(nreads / ((uint64_t)threads * 10));
  if (chunk < 1) chunk = 1;
#endif
  double t0 = now_s();
  uint64_t total = 0;
#ifdef _OPENMP
#pragma omp parallel reduction(+ : total)
#endif
  {
    uint32_t wmax = FMO_MAX_LEN / (seed_size ? seed_size : 1) + 1;
    fmo_inexact_region *out =
        (fmo_inexact_region *)malloc((size_t)2 * wmax * region_cap * sizeof(fmo_inexact_region));
    uint32_t *nreg = (uint32_t *)malloc((size_t)2 * wmax * sizeof(uint32_t));
    uint8_t *st = (uint8_t *)malloc((size_t)2 * wmax);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, chunk)
#endif
    for (long i = 0; i < (long)nreads; i++) {
      const uint32_t len = (uint32_t)(offsets[i + 1] - offsets[i]);
      const uint32_t nw = fmo_map_inexact_seeds(ix, reads + offsets[i], len, seed_size,
                                                min_seed_size, koff, kstart, kcount, (uint32_t)i,
                                                rec_cap, region_cap, out, nreg, st, wmax);
      for (uint32_t t = 0; t < 2 * nw && t < 2 * wmax; t++) total += nreg[t];
    }
    free(out); free(nreg); free(st);
  }
  *nregions_total = total;
  return now_s() - t0;
}

/* ------------------------------------------------------------------------ */
/* k-error search: BWSearchCPU, search/search.c:1171-1405, and its helpers   */
/* search/search.c:3-1169, search/search.h:66-198                           */
/* ------------------------------------------------------------------------ */

/* One view of the pair the search runs on (its `backward` / `forward`
 * arguments). */
typedef struct { const fmo_dir *d; int comp; } kview;

static inline void klf(kview v, int64_t k, int64_t l, unsigned b, int64_t *ko, int64_t *lo) {
  lf(v.d, v.comp, k, l, b, ko, lo);
}

/* add_result (results.h:93-103): copy into the list, or return true from
 * the enclosing function when it is full. */
#define KADD(r, L)                                          \
  do {                                                      \
    if ((L)->n < (L)->cap) (L)->list[(L)->n++] = *(r);      \
    else return true;                                       \
  } while (0)

/* add_mismatch / modify_last_mismatch* (results.h:59-83) */
static inline void k_add_mm(fmo_hitk *r, uint8_t kind, uint8_t base, int16_t pos) {
  if (r->num_mismatches < FMO_KMAX) {
    r->err_kind[r->num_mismatches] = kind;
    r->err_base[r->num_mismatches] = base;
    r->err_pos[r->num_mismatches] = pos;
  }
  r->num_mismatches++;
}
static inline void k_mod2(fmo_hitk *r, uint8_t kind, uint8_t base) {
  if (r->num_mismatches - 1 >= FMO_KMAX) return;
  r->err_kind[r->num_mismatches - 1] = kind;
  r->err_base[r->num_mismatches - 1] = base;
}
static inline void k_mod1(fmo_hitk *r, uint8_t kind) {
  if (r->num_mismatches - 1 >= FMO_KMAX) return;
  r->err_kind[r->num_mismatches - 1] = kind;
}
static inline void k_change(fmo_hitk *r, int64_t k, int64_t l, int16_t pos) { r->k = k; r->l = l; r->pos = pos; }

/* BWExactSearchBackward / Forward (search.h:9-65) on a k-record */
static void k_exact(kview v, int dir, const uint8_t *W, fmo_hitk *r, fmo_hitk *anchor) {
  fmo_hit t;
  memset(&t, 0, sizeof(t));
  t.k = r->k; t.l = r->l; t.start = r->start; t.pos = r->pos; t.end = r->end;
  fmo_hit a;
  exact_dir(v.d, v.comp, dir, W, &t, anchor ? &a : NULL);
  r->k = t.k; r->l = t.l; r->pos = t.pos;
  if (anchor) {
    memset(anchor, 0, sizeof(*anchor));
    anchor->dir = a.dir; anchor->k = a.k; anchor->l = a.l;
    anchor->start = a.start; anchor->pos = a.pos; anchor->end = a.end;
  }
}

/* BWExactFinalResultBackward (search.h:66-111): the block filter, then an
 * exact extension to start-1; only a surviving record is changed and kept. */
static bool k_final1_bwd(const uint8_t *W, kview v, fmo_hitk *r, klist *fin, int16_t bs, int16_t lb) {
  int16_t start = r->start, pos = r->pos;
  int64_t k = r->k, l = r->l;
  int16_t cb = (int16_t)(pos / bs);
  if ((cb < lb) || (pos == start - 1)) {
  } else if (cb > lb) {
    return false;
  } else if (!((pos + 1) % bs)) {
    return false;
  }
  for (int16_t i = pos; i >= start; i--) {
    klf(v, k, l, W[i], &k, &l);
    if (k > l) break;
  }
  if (k <= l) {
    k_change(r, k, l, (int16_t)(start - 1));
    KADD(r, fin);
  }
  return false;
}

/* BWExactFinalResultForward (search.h:113-158) */
static bool k_final1_fwd(const uint8_t *W, kview v, fmo_hitk *r, klist *fin, int16_t bs, int16_t lb) {
  int16_t pos = r->pos, end = r->end;
  int64_t k = r->k, l = r->l;
  int16_t cb = (int16_t)(pos / bs);
  if ((cb > lb) || (pos == end + 1)) {
  } else if (cb < lb) {
    return false;
  } else if (!(pos % bs)) {
    return false;
  }
  for (int16_t i = pos; i <= end; i++) {
    klf(v, k, l, W[i], &k, &l);
    if (k > l) break;
  }
  if (k <= l) {
    k_change(r, k, l, (int16_t)(end + 1));
    KADD(r, fin);
  }
  return false;
}

/* BWExactFinalResultsBackward / Forward (search.c:3-114) */
static bool k_finals(const uint8_t *W, kview v, int dir, klist *prev, klist *out, int16_t bs, int16_t lb) {
  for (uint64_t ii = 0; ii < prev->n; ii++) {
    fmo_hitk *r = &prev->list[ii];
    int16_t start = r->start, pos = r->pos, end = r->end;
    int64_t k = r->k, l = r->l;
    int16_t cb = (int16_t)(pos / bs);
    if (dir == 0) {
      if ((cb < lb) || (pos == start - 1)) {
      } else if (cb > lb) {
        continue;
      } else if (!((pos + 1) % bs)) {
        continue;
      }
      for (int16_t i = pos; i >= start; i--) {
        klf(v, k, l, W[i], &k, &l);
        if (k > l) break;
      }
      if (k <= l) { k_change(r, k, l, (int16_t)(start - 1)); KADD(r, out); }
    } else {
      if ((cb > lb) || (pos == end + 1)) {
      } else if (cb < lb) {
        continue;
      } else if (!(pos % bs)) {
        continue;
      }
      for (int16_t i = pos; i <= end; i++) {
        klf(v, k, l, W[i], &k, &l);
        if (k > l) break;
      }
      if (k <= l) { k_change(r, k, l, (int16_t)(end + 1)); KADD(r, out); }
    }
  }
  prev->n = 0;
  return false;
}

/* The branch step of one record at its pos, shared by the partial
 * (BWBranchPartialResults*, search.c:648-1000; each branch record is kept)
 * and final (BWBranchFinalResults*, search.c:116-469; each branch record is
 * extended to the end by k_final1_*) variants.  The record itself is
 * changed as the reference changes it, field by field. */
static bool k_branch(const uint8_t *W, kview v, int dir, int final, fmo_hitk *r, klist *out,
                     int16_t bs, int16_t lb) {
  const int16_t pos = r->pos;
  const int16_t step = dir ? 1 : -1;
  int16_t last_err_pos;
  uint8_t last_err_kind, last_err_base;
  int rnm = (int)r->num_mismatches - 1;
  if (rnm > -1) {
    last_err_pos = r->err_pos[rnm]; last_err_kind = r->err_kind[rnm]; last_err_base = r->err_base[rnm];
  } else {
    last_err_pos = -10; last_err_kind = 0; last_err_base = 0xFF;
  }
  int64_t k = r->k, l = r->l, ka, la;
  int no_previous = 1;
#define KEMIT()                                                              \
  do {                                                                       \
    if (final) {                                                             \
      if (dir ? k_final1_fwd(W, v, r, out, bs, lb) : k_final1_bwd(W, v, r, out, bs, lb)) return true; \
    } else {                                                                 \
      KADD(r, out);                                                          \
    }                                                                        \
  } while (0)
  k_add_mm(r, FMO_DELETION, 0xFF, pos);
  if (last_err_pos == pos - step) {            /* previous MISMATCH or DELETION */
    if (last_err_kind == FMO_MISMATCH) {
      if (W[pos] != last_err_base) { k_change(r, k, l, (int16_t)(pos + step)); KEMIT(); }
      for (unsigned b = 0; b < 4; b++) {
        klf(v, k, l, b, &ka, &la);
        if (ka > la) continue;
        if (b != W[pos]) {
          if (b != W[last_err_pos]) {
            k_change(r, ka, la, pos); k_mod2(r, FMO_INSERTION, (uint8_t)b); KEMIT();
          }
          k_change(r, ka, la, (int16_t)(pos + step)); k_mod2(r, FMO_MISMATCH, (uint8_t)b); KEMIT();
        }
      }
      no_previous = 0;
    } else if (last_err_kind == FMO_DELETION) {
      k_change(r, k, l, (int16_t)(pos + step)); KEMIT();
      for (unsigned b = 0; b < 4; b++) {
        klf(v, k, l, b, &ka, &la);
        if (ka > la) continue;
        if (b != W[pos] && b != W[last_err_pos]) {
          k_change(r, ka, la, (int16_t)(pos + step)); k_mod2(r, FMO_MISMATCH, (uint8_t)b); KEMIT();
        }
      }
      no_previous = 0;
    }
  } else if (last_err_pos == pos) {            /* previous INSERTION */
    for (unsigned b = 0; b < 4; b++) {
      klf(v, k, l, b, &ka, &la);
      if (ka > la) continue;
      if (b != W[pos]) {
        k_change(r, ka, la, pos); k_mod2(r, FMO_INSERTION, (uint8_t)b); KEMIT();
        if (W[pos] != last_err_base) {
          r->pos = (int16_t)(pos + step); k_mod1(r, FMO_MISMATCH); KEMIT();
        }
      }
    }
    no_previous = 0;
  }
  if (no_previous) {                           /* previous MATCH */
    k_change(r, k, l, (int16_t)(pos + step)); KEMIT();
    for (unsigned b = 0; b < 4; b++) {
      klf(v, k, l, b, &ka, &la);
      if (ka > la) continue;
      if (b != W[pos]) {
        k_change(r, ka, la, pos); k_mod2(r, FMO_INSERTION, (uint8_t)b); KEMIT();
        r->pos = (int16_t)(pos + step); k_mod1(r, FMO_MISMATCH); KEMIT();
      }
    }
  }
#undef KEMIT
  return false;
}

static bool k_branches(const uint8_t *W, kview v, int dir, int final, klist *prev, klist *out,
                       int16_t bs, int16_t lb) {
  for (uint64_t ii = 0; ii < prev->n; ii++) {
    fmo_hitk *r = &prev->list[ii];
    if (r->dir != dir) continue;
    if (dir == 0 ? (r->pos < r->start) : (r->pos > r->end)) {
      /* nothing left to branch on: kept (partial) or extended (final) */
      if (final) {
        if (dir == 0) { if (k_final1_bwd(W, v, r, out, bs, lb)) return true; }
        else KADD(r, out);
      } else {
        KADD(r, out);
      }
      continue;
    }
    if (k_branch(W, v, dir, final, r, out, bs, lb)) return true;
  }
  prev->n = 0;
  return false;
}

/* BWExactPartialResultsBackward / Forward (search.c:471-646) */
static bool k_partials(const uint8_t *W, const uint8_t *D, uint8_t max_errors, kview v, int dir,
                       klist *prev, klist *next, klist *next_i, int16_t bs, int16_t lb) {
  for (uint64_t ii = 0; ii < prev->n; ii++) {
    fmo_hitk *r = &prev->list[ii];
    int16_t start = r->start, pos = r->pos, end = r->end;
    uint8_t nm = r->num_mismatches;
    int64_t kn = r->k, ln = r->l, k = 0, l = 0;
    int64_t res, res_next = ln - kn;
    int16_t cb = (int16_t)(pos / bs), lbp;
    uint8_t eb = (uint8_t)(max_errors - nm);
    int complete;
    if (dir == 0) {
      if ((cb < lb) || (pos == start - 1)) { lbp = start; complete = 1; }
      else if (cb > lb) {
        lbp = (int16_t)(((pos + 1) % bs) ? (cb - 1) * bs : cb * bs);
        complete = 0;
      } else if ((pos + 1) % bs) { lbp = start; complete = 1; }
      else { lbp = (int16_t)(cb * bs); complete = 0; }
      for (int16_t i = pos; i >= lbp; i--) {
        k = kn; l = ln;
        if (k > l) break;
        klf(v, k, l, W[i], &kn, &ln);
        res = res_next; res_next = ln - kn;
        if (res == res_next || eb < D[i]) continue;
        k_change(r, k, l, i);
        KADD(r, next);
      }
      if (complete && kn <= ln) { k_change(r, kn, ln, (int16_t)(start - 1)); KADD(r, next_i); }
    } else {
      if ((cb > lb) || (pos == end + 1)) { lbp = end; complete = 1; }
      else if (cb < lb) {
        lbp = (int16_t)((pos % bs) ? (cb + 2) * bs - 1 : (cb + 1) * bs - 1);
        complete = 0;
      } else if (pos % bs) { lbp = end; complete = 1; }
      else { lbp = (int16_t)((cb + 1) * bs - 1); complete = 0; }
      for (int16_t i = pos; i <= lbp; i++) {
        k = kn; l = ln;
        if (k > l) break;
        klf(v, k, l, W[i], &kn, &ln);
        res = res_next; res_next = ln - kn;
        if (res == res_next || eb < D[i]) continue;
        k_change(r, k, l, i);
        KADD(r, next);
      }
      if (complete && kn <= ln) { k_change(r, kn, ln, (int16_t)(end + 1)); KADD(r, next_i); }
    }
  }
  prev->n = 0;
  return false;
}

/* change_direction (search.h:160-198): the interval of the same occurrences
 * on the other index, as the min / max ISA of every row's mirrored suffix. */
static void k_change_dir(const fmo_index *ix, kview b, kview f, fmo_hitk *r) {
  const int64_t siz = (int64_t)b.d->n + 1;
  int64_t ki = siz - 1, li = 0;
  int16_t err_offset = 0;
  for (uint8_t rr = 0; rr < r->num_mismatches && rr < FMO_KMAX; rr++) {
    if (r->err_kind[rr] == FMO_DELETION) err_offset--;
    else if (r->err_kind[rr] == FMO_INSERTION) err_offset++;
  }
  const int fb = b.d == &ix->fwd, ff = f.d == &ix->fwd;
  for (int64_t i = r->k; i <= r->l; i++) {
    int64_t aux = siz - (int64_t)fmo_get_sa(ix, fb, (uint64_t)i) - (r->end - r->start + 2) - err_offset;
    /* a position outside the text has no ISA (getRcompValue would walk
     * ~2^32 steps there); such a row is left out */
    if (aux < 0 || aux >= siz) continue;
    int64_t aux2 = (int64_t)fmo_get_isa(ix, ff, (uint64_t)aux);
    if (aux2 < ki) ki = aux2;
    if (aux2 > li) li = aux2;
  }
  r->k = ki;
  r->l = li;
}

/* BWChangeDirectionBackward / Forward (search.c:1002-1043): from the end of
 * the list back, every record of the list's direction that reached its end
 * turns to the other direction. */
static void k_turn(const fmo_index *ix, kview b, kview f, klist *L, int to_fwd, int16_t bound) {
  for (uint64_t ii = L->n; ii > 0; ii--) {
    fmo_hitk *r = &L->list[ii - 1];
    if (to_fwd) {
      if (r->dir == 1) break;
      if (r->pos == r->start - 1) {
        k_change_dir(ix, b, f, r);
        r->pos = (int16_t)(r->end + 1); r->end = bound; r->dir = 1;
      }
    } else {
      if (r->dir == 0) break;
      if (r->pos == r->end + 1) {
        k_change_dir(ix, b, f, r);
        r->pos = (int16_t)(r->start - 1); r->start = bound; r->dir = 0;
      }
    }
  }
}

/* calculateDBackward / Forward (search.c:1045-1169): lower bounds on the
 * errors needed in each suffix / prefix, with the reference's 8-bit types. */
static void k_calc_d(uint8_t *D, const uint8_t *W, int64_t nW, kview b, kview f, const fmo_hitk *r,
                     int backward) {
  int64_t k, l;
  uint8_t z = 0;
  int8_t last;
  const int64_t lb = (int64_t)b.d->n, lf_ = (int64_t)f.d->n;
  k = r->k; l = r->l;
  if (backward) {
    for (int16_t i = r->pos; i >= 0; i--) {
      klf(b, k, l, W[i], &k, &l);
      if (k > l) { k = 0; l = lb; z++; }
      D[i] = z;
    }
    if (r->end == (int16_t)(nW - 1)) {
      last = (int8_t)D[0];
    } else {
      k = 0; l = lb;
      for (int64_t i = r->end + 1; i < nW; i++) {
        klf(f, k, l, W[i], &k, &l);
        if (k > l) { k = 0; l = lf_; z++; }
        D[i] = z;
      }
      last = (int8_t)D[nW - 1];
      for (int64_t i = r->end + 1; i < nW; i++) D[i] = (uint8_t)(last - D[i]);
    }
    for (int16_t i = r->pos; i >= 0; i--) D[i] = (uint8_t)(last - D[i]);
  } else {
    for (int64_t i = r->pos; i < nW; i++) {
      klf(f, k, l, W[i], &k, &l);
      if (k > l) { k = 0; l = lf_; z++; }
      D[i] = z;
    }
    if (r->start == 0) {
      last = (int8_t)D[nW - 1];
    } else {
      k = 0; l = lb;
      for (int16_t i = (int16_t)(r->start - 1); i >= 0; i--) {
        klf(b, k, l, W[i], &k, &l);
        if (k > l) { k = 0; l = lb; z++; }
        D[i] = z;
      }
      last = (int8_t)D[0];
      for (int16_t i = (int16_t)(r->start - 1); i >= 0; i--) D[i] = (uint8_t)(last - D[i]);
    }
    for (int64_t i = r->pos; i < nW; i++) D[i] = (uint8_t)(last - D[i]);
  }
}

static bool k_search(const fmo_index *ix, const uint8_t *W, int64_t nW, kview B, kview F,
                     klist *prev, klist *next, klist *prev_i, klist *next_i, klist *fin,
                     klist *anc, int16_t fs, int type, uint8_t *D) {
  fmo_hitk r, ra;
  const int16_t fragments = (int16_t)(nW / fs);
  int16_t half = (int16_t)(fragments / 2);
  if (fragments % 2) half++;
  int err_count, flow;
  const int64_t lsa = (int64_t)B.d->n;     /* size_SA - 1 */
  if (anc) anc->n = 0;

  if (fragments <= 1) {
    memset(&r, 0, sizeof(r));
    if (type == 0) {
      r.dir = 1; k_change(&r, 0, lsa, 0); r.start = 0; r.end = (int16_t)(nW - 1);
      k_exact(F, 1, W, &r, anc ? &ra : NULL);
    } else {
      r.dir = 0; k_change(&r, 0, lsa, (int16_t)(nW - 1)); r.start = 0; r.end = (int16_t)(nW - 1);
      k_exact(B, 0, W, &r, anc ? &ra : NULL);
    }
    if (anc) KADD(&ra, anc);
    if (r.k <= r.l) KADD(&r, fin);
    return false;
  }

  /* blocks half-1 .. 1: exact on the forward index, then both ways */
  for (int16_t i = (int16_t)(half - 1); i > 0; i--) {
    flow = 1;
    err_count = fragments - 1;
    prev->n = 0; prev_i->n = 0; next->n = 0; next_i->n = 0;
    memset(&r, 0, sizeof(r));
    r.dir = 1; k_change(&r, 0, lsa, (int16_t)(fs * i));
    r.start = (int16_t)(fs * i); r.end = (int16_t)(fs * (i + 1) - 1);
    k_exact(F, 1, W, &r, anc ? &ra : NULL);
    if (anc) KADD(&ra, anc);
    r.end = (int16_t)(nW - 1);
    if (r.k <= r.l) {
      k_calc_d(D, W, nW, B, F, &r, 0);
      KADD(&r, prev);
      while (err_count > 0) {
        if (k_partials(W, D, (uint8_t)(fragments - 1), F, 1, prev, next, prev_i, fs, (int16_t)(half - 1))) { flow = 0; break; }
        k_turn(ix, F, B, prev_i, 0, 0);
        if (k_partials(W, D, (uint8_t)(fragments - 1), B, 0, prev_i, next_i, fin, fs, (int16_t)(half - 1))) { flow = 0; break; }
        if (err_count == 1) break;
        if (k_branches(W, F, 1, 0, next, prev, 0, 0)) { flow = 0; break; }
        if (k_branches(W, B, 0, 0, next_i, prev_i, 0, 0)) { flow = 0; break; }
        err_count--;
      }
      if (flow) {
        k_branches(W, F, 1, 1, next, prev_i, fs, (int16_t)(half - 1));
        k_turn(ix, F, B, prev_i, 0, 0);
        k_finals(W, B, 0, prev_i, fin, fs, (int16_t)(half - 1));
        k_branches(W, B, 0, 1, next_i, fin, fs, (int16_t)(half - 1));
      }
    }
  }

  /* block 0 */
  flow = 1;
  err_count = fragments - 1;
  prev->n = 0; prev_i->n = 0; next->n = 0; next_i->n = 0;
  memset(&r, 0, sizeof(r));
  r.dir = 1; k_change(&r, 0, lsa, 0); r.start = 0; r.end = (int16_t)(fs - 1);
  k_exact(F, 1, W, &r, anc ? &ra : NULL);
  if (anc) KADD(&ra, anc);
  r.end = (int16_t)(nW - 1);
  if (r.k <= r.l) {
    k_calc_d(D, W, nW, B, F, &r, 0);
    KADD(&r, prev);
    while (err_count > 0) {
      if (k_partials(W, D, (uint8_t)(fragments - 1), F, 1, prev, next, fin, fs, (int16_t)(half - 1))) { flow = 0; break; }
      if (err_count == 1) break;
      if (k_branches(W, F, 1, 0, next, prev, 0, 0)) { flow = 0; break; }
      err_count--;
    }
    if (flow) k_branches(W, F, 1, 1, next, fin, fs, (int16_t)(half - 1));
  }

  /* blocks half .. fragments-2: exact on the backward index, then both ways */
  for (int16_t i = half; i < fragments - 1; i++) {
    flow = 1;
    err_count = fragments - 1;
    prev->n = 0; prev_i->n = 0; next->n = 0; next_i->n = 0;
    memset(&r, 0, sizeof(r));
    r.dir = 0; k_change(&r, 0, lsa, (int16_t)(fs * (i + 1) - 1));
    r.start = (int16_t)(fs * i); r.end = (int16_t)(fs * (i + 1) - 1);
    k_exact(B, 0, W, &r, anc ? &ra : NULL);
    if (anc) KADD(&ra, anc);
    r.start = 0;
    if (r.k <= r.l) {
      k_calc_d(D, W, nW, B, F, &r, 1);
      KADD(&r, prev);
      while (err_count > 0) {
        if (k_partials(W, D, (uint8_t)(fragments - 1), B, 0, prev, next, prev_i, fs, 0)) { flow = 0; break; }
        k_turn(ix, B, F, prev_i, 1, (int16_t)(nW - 1));
        if (k_partials(W, D, (uint8_t)(fragments - 1), F, 1, prev_i, next_i, fin, fs, 0)) { flow = 0; break; }
        if (err_count == 1) break;
        if (k_branches(W, B, 0, 0, next, prev, 0, 0)) { flow = 0; break; }
        if (k_branches(W, F, 1, 0, next_i, prev_i, 0, 0)) { flow = 0; break; }
        err_count--;
      }
      if (flow) {
        k_branches(W, B, 0, 1, next, prev_i, fs, 0);
        k_turn(ix, B, F, prev_i, 1, (int16_t)(nW - 1));
        k_finals(W, F, 1, prev_i, fin, fs, 0);
        k_branches(W, F, 1, 1, next_i, fin, fs, 0);
      }
    }
  }

  /* block fragments-1 (the last block takes the remainder) */
  flow = 1;
  err_count = fragments - 1;
  prev->n = 0; prev_i->n = 0; next->n = 0; next_i->n = 0;
  memset(&r, 0, sizeof(r));
  r.dir = 0; k_change(&r, 0, lsa, (int16_t)(nW - 1));
  r.start = (int16_t)(fs * (fragments - 1)); r.end = (int16_t)(nW - 1);
  k_exact(B, 0, W, &r, anc ? &ra : NULL);
  if (anc) KADD(&ra, anc);
  r.start = 0;
  if (r.k <= r.l) {
    k_calc_d(D, W, nW, B, F, &r, 1);
    KADD(&r, prev);
    while (err_count > 0) {
      if (k_partials(W, D, (uint8_t)(fragments - 1), B, 0, prev, next, fin, fs, 0)) { flow = 0; break; }
      if (err_count == 1) break;
      if (k_branches(W, B, 0, 0, next, prev, 0, 0)) { flow = 0; break; }
      err_count--;
    }
    if (flow) k_branches(W, B, 0, 1, next, fin, fs, 0);
  }
  return false;
}

int fmo_search_k(const fmo_index *ix, int strand, const uint8_t *W, int16_t len, int16_t fragsize,
                 uint32_t list_cap, uint32_t anchor_cap, fmo_hitk *final_out, uint32_t *nfinal,
                 fmo_hitk *anchors_out, uint32_t *nanchors) {
  /* bwt_map_inexact_read_2 (bwt.c:2430-2470): + strand (backward, forward,
   * type 1); - strand (forward_rev, backward_rev, type 0) */
  kview B, F;
  B.d = strand ? &ix->fwd : &ix->bwd; B.comp = strand;
  F.d = strand ? &ix->bwd : &ix->fwd; F.comp = strand;
  klist L[5], A;
  fmo_hitk *mem = (fmo_hitk *)calloc((size_t)5 * (list_cap ? list_cap : 1), sizeof(fmo_hitk));
  uint8_t *D = (uint8_t *)calloc((size_t)(len > 0 ? len : 1), 1);
  for (int i = 0; i < 5; i++) { L[i].list = mem + (size_t)i * (list_cap ? list_cap : 1); L[i].n = 0; L[i].cap = list_cap; }
  A.list = anchors_out; A.n = 0; A.cap = anchor_cap;
  bool full = k_search(ix, W, len, B, F, &L[0], &L[1], &L[2], &L[3], &L[4], anchors_out ? &A : NULL,
                       fragsize, strand ? 0 : 1, D);
  for (uint64_t i = 0; i < L[4].n; i++) final_out[i] = L[4].list[i];
  *nfinal = (uint32_t)L[4].n;
  if (nanchors) *nanchors = (uint32_t)A.n;
  free(mem);
  free(D);
  return full ? 1 : 0;
}

/* fmo_search_k on both strands of every read (fragsize[i] per read), OpenMP
 * over reads as fmo_search1_batch; task t = 2 * read + strand owns
 * final_out[t * list_cap ..] / nfinal[t], anchors_out[t * anchor_cap ..] /
 * nanchors[t] and full[t]. */
double fmo_search_k_batch(const fmo_index *ix, const uint8_t *reads, const uint64_t *offsets,
                          uint64_t nreads, const int16_t *fragsize, uint32_t list_cap,
                          uint32_t anchor_cap, fmo_hitk *final_out, uint32_t *nfinal,
                          fmo_hitk *anchors_out, uint32_t *nanchors, uint8_t *full, int threads) {
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
  omp_set_num_threads(threads);
  long chunk = (long)(nreads / ((uint64_t)threads * 10));
  if (chunk < 1) chunk = 1;
#endif
  double t0 = now_s();
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, chunk)
#endif
  for (long i = 0; i < (long)nreads; i++) {
    const uint8_t *W = reads + offsets[i];
    int16_t len = (int16_t)(offsets[i + 1] - offsets[i]);
    for (int s = 0; s < 2; s++) {
      const size_t t = (size_t)i * 2 + s;
      full[t] = (uint8_t)fmo_search_k(ix, s, W, len, fragsize[i], list_cap, anchor_cap,
                                      final_out + t * list_cap, &nfinal[t],
                                      anchors_out ? anchors_out + t * anchor_cap : NULL,
                                      &nanchors[t]);
    }
  }
  return now_s() - t0;
}

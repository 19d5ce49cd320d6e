/*
 * ref_harness.c -- thin driver around the reference's own search core.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) from the
 * reference sources where they lie (c/src/aligners/bwt/{search,dbwt}/ and
 * bwt_commons.c of opencb/hpg-libs) into oracle/_ref/libfmref_ref.so.  It is
 * used to pin the oracle restatement, to write the golden fixtures, and as
 * the reference CPU baseline.  Nothing in the product links it.
 *
 * The index is built exactly as bwt_generate_index_files does
 * (bwt.c:744-819): direct_bwt -> calculate_C -> calculate_O ->
 * calculate_S_and_R, then the same on the reversed text, and the complement
 * views of load_bwt_index (runtime.h:97-124).  The harness keeps the arrays
 * in memory instead of writing and reading the index directory; only the BWT
 * goes through a file because direct_bwt writes one.
 */
#include "search/search.h"
#include "search/preprocess.h"
#include "search/io.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>

typedef struct {
  int64_t k, l;
  int16_t start, pos, end, err_pos;
  uint8_t dir, num_mismatches, err_kind, err_base;
  uint8_t strand, pad[3];
} ref_hit;

typedef struct {
  bwt_index backward, forward, backward_rev, forward_rev;
  bwt_config_t cfg;
  ref_vector B, Bi;
  uint64_t n;
  int wrapped;   /* arrays owned by the caller (ref_wrap) */
} ref_ctx;

static bwt_index *view_of(ref_ctx *c, int view) {
  switch (view) {
    case 0: return &c->backward;
    case 1: return &c->forward;
    case 2: return &c->backward_rev;
    default: return &c->forward_rev;
  }
}

void *ref_build(const uint8_t *text, uint64_t n, uint32_t ratio, const char *workdir) {
  ref_ctx *c = (ref_ctx *)calloc(1, sizeof(ref_ctx));
  c->n = n;
  bwt_init_replace_table(&c->cfg, "ACGT");
  c->cfg.duplicate_strand = false;

  ref_vector X;
  X.vector = (uint8_t *)malloc(n + 1);
  memcpy(X.vector, text, n);
  X.vector[n] = 0;
  X.n = n;
  X.dollar = 0;
  calculate_and_save_B(&X, workdir, "B");
  read_ref_vector(&c->B, workdir, "B");

  calculate_C(&c->backward.C, &c->backward.C1, &c->B, 4);
  calculate_O(&c->backward.O, &c->B, 4);
  calculate_S_and_R(&c->backward.S, &c->backward.R, &c->B, &c->backward.C, &c->backward.O, ratio);

  ref_vector Xi;
  Xi.vector = (uint8_t *)malloc(n + 1);
  memcpy(Xi.vector, text, n);
  revstring(Xi.vector, n);
  Xi.vector[n] = 0;
  Xi.n = n;
  Xi.dollar = 0;
  calculate_and_save_B(&Xi, workdir, "Bi");
  read_ref_vector(&c->Bi, workdir, "Bi");

  calculate_C(&c->forward.C, &c->forward.C1, &c->Bi, 4);
  calculate_O(&c->forward.O, &c->Bi, 4);
  calculate_S_and_R(&c->forward.S, &c->forward.R, &c->Bi, &c->backward.C, &c->forward.O, ratio);

  reverse_strand_C(&c->backward_rev.C, &c->backward.C, &c->backward_rev.C1, &c->backward.C1, c->cfg);
  reverse_strand_O(&c->backward_rev.O, &c->backward.O, c->cfg);
  c->backward_rev.S = c->backward.S;
  c->backward_rev.R = c->backward.R;
  reverse_strand_C(&c->forward_rev.C, &c->forward.C, &c->forward_rev.C1, &c->forward.C1, c->cfg);
  reverse_strand_O(&c->forward_rev.O, &c->forward.O, c->cfg);
  c->forward_rev.S = c->forward.S;
  c->forward_rev.R = c->forward.R;

  /* direct_bwt releases its input text itself (dbwt.c:526). */
  return c;
}

/* Writes the index directory bwt_generate_index_files (bwt.c:744-819)
 * would write, with the reference's own writers: config.txt, C, C1, O, S,
 * R, Oi, Si, Ri and the karyotype file `index` (contigs c of the text:
 * names chr1.., start[c], end[c]).  save_exome_file leaves its stream open
 * (search/io.c:83-103), so the streams are flushed here. */
int ref_save_dir(void *h, const char *dir, uint64_t count, const uint64_t *start,
                 const uint64_t *end) {
  ref_ctx *c = (ref_ctx *)h;
  save_config("ACGT", false, dir);
  save_vector(&c->backward.C, dir, "C");
  save_vector(&c->backward.C1, dir, "C1");
  save_comp_matrix(&c->backward.O, dir, "O");
  save_comp_vector(&c->backward.S, dir, "S");
  save_comp_vector(&c->backward.R, dir, "R");
  save_comp_matrix(&c->forward.O, dir, "Oi");
  save_comp_vector(&c->forward.S, dir, "Si");
  save_comp_vector(&c->forward.R, dir, "Ri");
  exome *ex = exome_new();
  if (count > (uint64_t)ex->max_chromosomes) { exome_free(ex); return -1; }
  for (uint64_t i = 0; i < count; i++) {
    snprintf(ex->chromosome + i * IDMAX, IDMAX, "chr%llu", (unsigned long long)(i + 1));
    ex->start[i] = start[i];
    ex->end[i] = end[i];
  }
  ex->size = count;
  save_exome_file(ex, false, dir);
  fflush(NULL);
  exome_free(ex);
  return 0;
}

/* Wraps index arrays the caller owns (the reference layout: C, C1, per-symbol
 * desp/count, S, R) into the reference's bwt_index structs, without running
 * the reference builder.  Used for large texts, where dbwt and the serial
 * S/R walk would take minutes; the arrays must equal what the reference
 * builder writes.  d = 0 backward, 1 forward. */
typedef struct {
  const uint32_t *C, *C1;
  uint32_t *desp[4];
  uint64_t *count[4];
  uint64_t words;
  uint32_t *S, *R;
  uint64_t nsamp;
} ref_dir_arrays;

static void wrap_dir(bwt_index *ix, const ref_dir_arrays *a, uint64_t n, uint32_t ratio) {
  ix->C.n = 4; ix->C1.n = 4;
  ix->C.vector = (SA_TYPE *)malloc(4 * sizeof(SA_TYPE));
  ix->C1.vector = (SA_TYPE *)malloc(4 * sizeof(SA_TYPE));
  for (int b = 0; b < 4; b++) { ix->C.vector[b] = a->C[b]; ix->C1.vector[b] = a->C1[b]; }
  ix->O.siz = (SA_TYPE)(n + 2);
  ix->O.n_desp = 4; ix->O.m_desp = (SA_TYPE)a->words;
  ix->O.n_count = 4; ix->O.m_count = (SA_TYPE)a->words;
  ix->O.desp = (SA_TYPE **)malloc(4 * sizeof(SA_TYPE *));
  ix->O.count = (FM_COMP_TYPE **)malloc(4 * sizeof(FM_COMP_TYPE *));
  for (int b = 0; b < 4; b++) { ix->O.desp[b] = a->desp[b]; ix->O.count[b] = a->count[b]; }
  ix->S.vector = a->S; ix->S.siz = (SA_TYPE)(n + 1); ix->S.n = (SA_TYPE)a->nsamp; ix->S.ratio = ratio;
  ix->R.vector = a->R; ix->R.siz = (SA_TYPE)(n + 1); ix->R.n = (SA_TYPE)a->nsamp; ix->R.ratio = ratio;
}

void *ref_wrap(const ref_dir_arrays *bwd, const ref_dir_arrays *fwd, uint64_t n, uint32_t ratio) {
  ref_ctx *c = (ref_ctx *)calloc(1, sizeof(ref_ctx));
  c->n = n;
  c->wrapped = 1;
  bwt_init_replace_table(&c->cfg, "ACGT");
  wrap_dir(&c->backward, bwd, n, ratio);
  wrap_dir(&c->forward, fwd, n, ratio);
  reverse_strand_C(&c->backward_rev.C, &c->backward.C, &c->backward_rev.C1, &c->backward.C1, c->cfg);
  reverse_strand_O(&c->backward_rev.O, &c->backward.O, c->cfg);
  c->backward_rev.S = c->backward.S;
  c->backward_rev.R = c->backward.R;
  reverse_strand_C(&c->forward_rev.C, &c->forward.C, &c->forward_rev.C1, &c->forward.C1, c->cfg);
  reverse_strand_O(&c->forward_rev.O, &c->forward.O, c->cfg);
  c->forward_rev.S = c->forward.S;
  c->forward_rev.R = c->forward.R;
  return c;
}

void ref_free(void *h) {
  ref_ctx *c = (ref_ctx *)h;
  if (!c) return;
  if (c->wrapped) {
    bwt_index *all[4] = {&c->backward, &c->forward, &c->backward_rev, &c->forward_rev};
    for (int i = 0; i < 4; i++) {
      free(all[i]->C.vector);
      free(all[i]->C1.vector);
      free(all[i]->O.desp);
      free(all[i]->O.count);
    }
  } else {
    free_bwt_index(&c->backward_rev, &c->backward, true);
    free_bwt_index(&c->forward_rev, &c->forward, true);
    free(c->B.vector);
    free(c->Bi.vector);
  }
  bwt_free_replace_table(&c->cfg);
  free(c);
}

uint64_t ref_size_sa(void *h) { return size_SA(&((ref_ctx *)h)->backward); }

/* The BWT as written by direct_bwt (B.vec: n bytes without '$', dollar row). */
void ref_bwt(void *h, int fwd, uint8_t *out, uint64_t *dollar) {
  ref_ctx *c = (ref_ctx *)h;
  ref_vector *B = fwd ? &c->Bi : &c->B;
  memcpy(out, B->vector, B->n);
  *dollar = B->dollar;
}

uint64_t ref_words(void *h) { return ((ref_ctx *)h)->backward.O.m_desp; }
uint64_t ref_nsamp(void *h) { return ((ref_ctx *)h)->backward.S.n; }

void ref_arrays(void *h, int fwd, uint32_t *C, uint32_t *C1, uint32_t *desp,
                uint64_t *count, uint32_t *S, uint32_t *R) {
  ref_ctx *c = (ref_ctx *)h;
  bwt_index *ix = fwd ? &c->forward : &c->backward;
  uint64_t m = ix->O.m_desp;
  for (int b = 0; b < 4; b++) {
    C[b] = ix->C.vector[b];
    C1[b] = ix->C1.vector[b];
    memcpy(desp + (size_t)b * m, ix->O.desp[b], m * sizeof(uint32_t));
    memcpy(count + (size_t)b * m, ix->O.count[b], m * sizeof(uint64_t));
  }
  memcpy(S, ix->S.vector, ix->S.n * sizeof(uint32_t));
  memcpy(R, ix->R.vector, ix->R.n * sizeof(uint32_t));
}

static void to_hit(const result *r, ref_hit *h, uint8_t strand) {
  memset(h, 0, sizeof(*h));
  h->k = r->k; h->l = r->l;
  h->start = r->start; h->pos = r->pos; h->end = r->end;
  h->dir = r->dir ? 1 : 0;
  h->num_mismatches = r->num_mismatches;
  if (r->num_mismatches) {
    h->err_kind = r->err_kind[0];
    h->err_base = r->err_base[0];
    h->err_pos = r->err_pos[0];
  }
  h->strand = strand;
}

static void from_hit(const ref_hit *h, result *r) {
  init_result(r, h->dir);
  bound_result(r, h->start, h->end);
  change_result(r, h->k, h->l, h->pos);
}

/* BWExactSearchBackward (dir 0) / BWExactSearchForward (dir 1). */
void ref_exact(void *h, int view, int dir, const uint8_t *W, ref_hit *io, ref_hit *anchor) {
  ref_ctx *c = (ref_ctx *)h;
  result r, a;
  from_hit(io, &r);
  if (dir == 0) BWExactSearchBackward((uint8_t *)W, view_of(c, view), &r, anchor ? &a : NULL);
  else          BWExactSearchForward((uint8_t *)W, view_of(c, view), &r, anchor ? &a : NULL);
  to_hit(&r, io, 0);
  if (anchor) to_hit(&a, anchor, 0);
}

int16_t ref_exact_vector(void *h, int view, int dir, const uint8_t *W, int16_t start,
                         int16_t end, int64_t k, int64_t l, int64_t *vk, int64_t *vl) {
  ref_ctx *c = (ref_ctx *)h;
  intmax_t *a = (intmax_t *)vk, *b = (intmax_t *)vl;
  if (dir == 0) return BWExactSearchVectorBackward((uint8_t *)W, start, end, k, l, a, b, view_of(c, view));
  return BWExactSearchVectorForward((uint8_t *)W, start, end, k, l, a, b, view_of(c, view));
}

/* BWSearch1CPU with the strand pairing of bwt.c:1690-1701.  Returns the
 * number of records the reference kept (<= cap); *full is set when the
 * reference's add_result reported a full list. */
uint32_t ref_search1(void *h, int strand, const uint8_t *W, int16_t len, uint32_t cap,
                     ref_hit *out, ref_hit *anchors, int *full) {
  ref_ctx *c = (ref_ctx *)h;
  results_list rl, ra;
  new_results_list(&rl, cap ? cap : 1);
  rl.max_results = cap;
  new_results_list(&ra, 2);
  result res;
  bound_result(&res, 0, len - 1);
  change_result(&res, 0, size_SA(&c->backward) - 1, 0);
  bool f;
  if (strand == 0)
    f = BWSearch1CPU((uint8_t *)W, &c->backward, &c->forward, &res, &rl, anchors ? &ra : NULL, 4);
  else
    f = BWSearch1CPU((uint8_t *)W, &c->forward_rev, &c->backward_rev, &res, &rl, anchors ? &ra : NULL, 4);
  for (uintmax_t i = 0; i < rl.num_results; i++) to_hit(&rl.list[i], &out[i], (uint8_t)strand);
  if (anchors)
    for (uintmax_t i = 0; i < ra.num_results && i < 2; i++) to_hit(&ra.list[i], &anchors[i], (uint8_t)strand);
  if (full) *full = f ? 1 : 0;
  uint32_t nres = (uint32_t)rl.num_results;
  free(rl.list);
  free(ra.list);
  return nres;
}

/* The reference's text output of one read's 1-error search: BWSearch1CPU on
 * one strand (as ref_search1, list capacity cap), then write_results
 * (search/results.c:3-43) with contigs chr1.. of the given start/end, the
 * read text `mapping` and read_index.  Writes at most outcap - 1 bytes of
 * text to out (NUL-terminated) and returns the full length. */
size_t ref_write_results(void *h, int strand, const uint8_t *W, const char *mapping, int16_t len,
                         uint64_t read_index, uint32_t cap, uint64_t kcount,
                         const uint64_t *kstart, const uint64_t *kend, char *out, size_t outcap) {
  ref_ctx *c = (ref_ctx *)h;
  results_list rl;
  new_results_list(&rl, cap ? cap : 1);
  rl.max_results = cap;
  rl.read_index = read_index;
  result res;
  bound_result(&res, 0, len - 1);
  change_result(&res, 0, size_SA(&c->backward) - 1, 0);
  if (strand == 0)
    BWSearch1CPU((uint8_t *)W, &c->backward, &c->forward, &res, &rl, NULL, 4);
  else
    BWSearch1CPU((uint8_t *)W, &c->forward_rev, &c->backward_rev, &res, &rl, NULL, 4);
  exome *ex = exome_new();
  ex->offset[0] = 0;
  for (uint64_t i = 0; i < kcount; i++) {
    snprintf(ex->chromosome + i * IDMAX, IDMAX, "chr%llu", (unsigned long long)(i + 1));
    ex->start[i] = kstart[i];
    ex->end[i] = kend[i];
    ex->offset[i + 1] = ex->offset[i] + (kend[i] - kstart[i] + 1);
  }
  ex->size = kcount;
  intmax_t *kk = (intmax_t *)malloc((rl.num_results + 1) * sizeof(intmax_t));
  intmax_t *ll = (intmax_t *)malloc((rl.num_results + 1) * sizeof(intmax_t));
  char *buf = NULL;
  size_t blen = 0;
  FILE *fp = open_memstream(&buf, &blen);
  write_results(&rl, kk, ll, ex, &c->backward, &c->forward, (char *)mapping, (uintmax_t)len,
                strand == 0 ? 1 : 0, fp, &c->cfg);
  fclose(fp);
  size_t n = blen < outcap ? blen : (outcap ? outcap - 1 : 0);
  if (outcap) { memcpy(out, buf, n); out[n] = 0; }
  free(buf); free(kk); free(ll); free(rl.list);
  exome_free(ex);
  return blen;
}

/* The mapping layer's 1-error search of W[start..end] on one strand: the
 * interval vectors, then BWSearch1VectorHelper, as bwt.c:1940-1950 +
 * 1994-2003 (start = 0) and bwt.c:1447-1457 + 1465-1479 (a seed). */
uint32_t ref_search1_vector(void *h, int strand, const uint8_t *W, int16_t start, int16_t end,
                            uint32_t cap, ref_hit *out, int *full) {
  ref_ctx *c = (ref_ctx *)h;
  const int n1 = end - start + 1;
  intmax_t *v = (intmax_t *)malloc(4 * (size_t)n1 * sizeof(intmax_t));
  intmax_t *a_k = v, *a_l = v + n1, *b_k = v + 2 * n1, *b_l = v + 3 * n1;
  results_list rl;
  new_results_list(&rl, cap ? cap : 1);
  rl.max_results = cap;
  rl.num_results = 0;
  bool f;
  const intmax_t lsa = size_SA(&c->backward) - 1;
  if (strand == 0) {
    BWExactSearchVectorBackward((uint8_t *)W, start, end, 0, lsa, a_k, a_l, &c->backward);
    BWExactSearchVectorForward((uint8_t *)W, start, end, 0, lsa, b_k, b_l, &c->forward);
    f = BWSearch1VectorHelper((uint8_t *)W, start, end, a_k, a_l, b_k, b_l, &c->backward,
                              &c->forward, &rl, 4);
  } else {
    BWExactSearchVectorBackward((uint8_t *)W, start, end, 0, lsa, a_k, a_l, &c->forward_rev);
    BWExactSearchVectorForward((uint8_t *)W, start, end, 0, lsa, b_k, b_l, &c->backward_rev);
    f = BWSearch1VectorHelper((uint8_t *)W, start, end, a_k, a_l, b_k, b_l, &c->forward_rev,
                              &c->backward_rev, &rl, 4);
  }
  for (uintmax_t i = 0; i < rl.num_results; i++) to_hit(&rl.list[i], &out[i], (uint8_t)strand);
  if (full) *full = f ? 1 : 0;
  uint32_t nres = (uint32_t)rl.num_results;
  free(rl.list);
  free(v);
  return nres;
}

/* BWSearchCPU (search.c:1171) on one strand, as bwt_map_inexact_read_2
 * calls it (bwt.c:2430-2470): lists of list_cap records, anchors of
 * anchor_cap.  Records are copied as ref_hitk (up to 8 error records). */
typedef struct {
  int64_t k, l;
  int16_t start, pos, end;
  uint8_t dir, num_mismatches;
  uint8_t err_kind[8], err_base[8];
  int16_t err_pos[8];
  uint8_t pad[8];
} ref_hitk;

static void to_hitk(const result *r, ref_hitk *h) {
  memset(h, 0, sizeof(*h));
  h->k = r->k; h->l = r->l;
  h->start = r->start; h->pos = r->pos; h->end = r->end;
  h->dir = r->dir ? 1 : 0;
  h->num_mismatches = r->num_mismatches;
  for (int i = 0; i < r->num_mismatches && i < 8; i++) {
    h->err_kind[i] = r->err_kind[i];
    h->err_base[i] = r->err_base[i];
    h->err_pos[i] = r->err_pos[i];
  }
}

int ref_search_k(void *h, int strand, const uint8_t *W, int16_t len, int16_t fragsize,
                 uint32_t list_cap, uint32_t anchor_cap, ref_hitk *final_out, uint32_t *nfinal,
                 ref_hitk *anchors_out, uint32_t *nanchors) {
  ref_ctx *c = (ref_ctx *)h;
  results_list L[5], A;
  for (int i = 0; i < 5; i++) {
    new_results_list(&L[i], list_cap ? list_cap : 1);
    L[i].max_results = list_cap;
    L[i].num_results = 0;
  }
  new_results_list(&A, anchor_cap ? anchor_cap : 1);
  A.max_results = anchor_cap;
  A.num_results = 0;
  bool f;
  if (strand == 0)
    f = BWSearchCPU((uint8_t *)W, (uint64_t)len, &c->backward, &c->forward, &L[0], &L[1], &L[2],
                    &L[3], &L[4], anchors_out ? &A : NULL, fragsize, 1, 4);
  else
    f = BWSearchCPU((uint8_t *)W, (uint64_t)len, &c->forward_rev, &c->backward_rev, &L[0], &L[1],
                    &L[2], &L[3], &L[4], anchors_out ? &A : NULL, fragsize, 0, 4);
  for (uintmax_t i = 0; i < L[4].num_results; i++) to_hitk(&L[4].list[i], &final_out[i]);
  *nfinal = (uint32_t)L[4].num_results;
  if (anchors_out) {
    for (uintmax_t i = 0; i < A.num_results; i++) to_hitk(&A.list[i], &anchors_out[i]);
  }
  if (nanchors) *nanchors = anchors_out ? (uint32_t)A.num_results : 0;
  for (int i = 0; i < 5; i++) free(L[i].list);
  free(A.list);
  return f ? 1 : 0;
}

uint64_t ref_get_sa(void *h, int fwd, uint64_t m) {
  ref_ctx *c = (ref_ctx *)h;
  return get_SA((SA_TYPE)m, fwd ? &c->forward : &c->backward);
}

uint64_t ref_get_isa(void *h, int fwd, uint64_t m) {
  ref_ctx *c = (ref_ctx *)h;
  return get_ISA((SA_TYPE)m, fwd ? &c->forward : &c->backward);
}

uint64_t ref_binsearch(uint64_t *array, uint64_t size, uint64_t key) {
  return binsearch((uintmax_t *)array, size, key);
}

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* CPU baseline: BWSearch1CPU on both strands of every read, parallel over
 * reads as bwt_map_exact_batch does (bwt.c:1112,1134). */
double ref_search1_batch(void *h, const uint8_t *reads, const uint64_t *offsets,
                         uint64_t nreads, uint32_t cap, uint32_t *counts, int threads) {
  ref_ctx *c = (ref_ctx *)h;
  if (threads <= 0) threads = omp_get_max_threads();
  omp_set_num_threads(threads);
  long chunk = (long)(nreads / ((uint64_t)threads * 10));
  if (chunk < 1) chunk = 1;
  double t0 = now_s();
#pragma omp parallel
  {
    results_list rl;
    new_results_list(&rl, cap ? cap : 1);
    rl.max_results = cap;
#pragma omp for schedule(dynamic, chunk)
    for (long i = 0; i < (long)nreads; i++) {
      uint8_t *W = (uint8_t *)reads + offsets[i];
      int16_t len = (int16_t)(offsets[i + 1] - offsets[i]);
      for (int s = 0; s < 2; s++) {
        result res;
        rl.num_results = 0;
        bound_result(&res, 0, len - 1);
        change_result(&res, 0, size_SA(&c->backward) - 1, 0);
        if (s == 0) BWSearch1CPU(W, &c->backward, &c->forward, &res, &rl, NULL, 4);
        else        BWSearch1CPU(W, &c->forward_rev, &c->backward_rev, &res, &rl, NULL, 4);
        counts[2 * i + s] = (uint32_t)rl.num_results;
      }
    }
    free(rl.list);
  }
  return now_s() - t0;
}

/* CPU baseline for the k-error search: BWSearchCPU on both strands of every
 * read with its own fragment size, parallel over reads as bwt.c:1134; each
 * thread owns its lists (bwt_map_inexact_read_2 allocates them per read). */
double ref_search_k_batch(void *h, const uint8_t *reads, const uint64_t *offsets,
                          const int16_t *fragsize, uint64_t nreads, uint32_t list_cap,
                          uint32_t anchor_cap, uint32_t *nfinal, int threads) {
  ref_ctx *c = (ref_ctx *)h;
  if (threads <= 0) threads = omp_get_max_threads();
  omp_set_num_threads(threads);
  long chunk = (long)(nreads / ((uint64_t)threads * 10));
  if (chunk < 1) chunk = 1;
  double t0 = now_s();
#pragma omp parallel
  {
    results_list L[5], A;
    for (int i = 0; i < 5; i++) { new_results_list(&L[i], list_cap); }
    new_results_list(&A, anchor_cap);
#pragma omp for schedule(dynamic, chunk)
    for (long i = 0; i < (long)nreads; i++) {
      uint8_t *W = (uint8_t *)reads + offsets[i];
      uint64_t len = offsets[i + 1] - offsets[i];
      for (int s = 0; s < 2; s++) {
        for (int j = 0; j < 5; j++) L[j].num_results = 0;
        A.num_results = 0;
        if (s == 0)
          BWSearchCPU(W, len, &c->backward, &c->forward, &L[0], &L[1], &L[2], &L[3], &L[4], &A,
                      fragsize[i], 1, 4);
        else
          BWSearchCPU(W, len, &c->forward_rev, &c->backward_rev, &L[0], &L[1], &L[2], &L[3],
                      &L[4], &A, fragsize[i], 0, 4);
        nfinal[2 * i + s] = (uint32_t)L[4].num_results;
      }
    }
    for (int i = 0; i < 5; i++) free(L[i].list);
    free(A.list);
  }
  return now_s() - t0;
}

/* CPU baseline for exact search + locate: both strands as ref_exact_batch,
 * then get_SA(j, backward) over the first `cap` rows of each interval (the
 * loop after the search in bwt_map_exact_seed, bwt.c:1271-1277).  Per read
 * and strand: rows located and the sum of their positions. */
double ref_exact_locate_batch(void *h, const uint8_t *reads, const uint64_t *offsets,
                              uint64_t nreads, uint32_t cap, uint32_t *nrows, uint64_t *possum,
                              int threads) {
  ref_ctx *c = (ref_ctx *)h;
  if (threads <= 0) threads = omp_get_max_threads();
  omp_set_num_threads(threads);
  long chunk = (long)(nreads / ((uint64_t)threads * 10));
  if (chunk < 1) chunk = 1;
  double t0 = now_s();
#pragma omp parallel for schedule(dynamic, chunk)
  for (long i = 0; i < (long)nreads; i++) {
    uint8_t *W = (uint8_t *)reads + offsets[i];
    int16_t len = (int16_t)(offsets[i + 1] - offsets[i]);
    for (int s = 0; s < 2; s++) {
      result r;
      bound_result(&r, 0, len - 1);
      if (s == 0) {
        change_result(&r, 0, size_SA(&c->backward) - 1, len - 1);
        BWExactSearchBackward(W, &c->backward, &r, NULL);
      } else {
        change_result(&r, 0, size_SA(&c->backward) - 1, 0);
        BWExactSearchForward(W, &c->backward_rev, &r, NULL);
      }
      uint32_t m = 0;
      uint64_t sum = 0;
      for (intmax_t j = r.k; j <= r.l && m < cap; j++, m++)
        sum += (uint64_t)get_SA((SA_TYPE)j, &c->backward);
      nrows[2 * i + s] = m;
      possum[2 * i + s] = sum;
    }
  }
  return now_s() - t0;
}

/* CPU baseline for exact search: + strand backward on `backward`, - strand
 * forward on `backward_rev` (bwt.c:1222-1238). */
double ref_exact_batch(void *h, const uint8_t *reads, const uint64_t *offsets,
                       uint64_t nreads, int64_t *out_kl, int threads) {
  ref_ctx *c = (ref_ctx *)h;
  if (threads <= 0) threads = omp_get_max_threads();
  omp_set_num_threads(threads);
  long chunk = (long)(nreads / ((uint64_t)threads * 10));
  if (chunk < 1) chunk = 1;
  double t0 = now_s();
#pragma omp parallel for schedule(dynamic, chunk)
  for (long i = 0; i < (long)nreads; i++) {
    uint8_t *W = (uint8_t *)reads + offsets[i];
    int16_t len = (int16_t)(offsets[i + 1] - offsets[i]);
    result r;
    bound_result(&r, 0, len - 1);
    change_result(&r, 0, size_SA(&c->backward) - 1, len - 1);
    BWExactSearchBackward(W, &c->backward, &r, NULL);
    out_kl[4 * i + 0] = r.k; out_kl[4 * i + 1] = r.l;
    bound_result(&r, 0, len - 1);
    change_result(&r, 0, size_SA(&c->backward) - 1, 0);
    BWExactSearchForward(W, &c->backward_rev, &r, NULL);
    out_kl[4 * i + 2] = r.k; out_kl[4 * i + 3] = r.l;
  }
  return now_s() - t0;
}

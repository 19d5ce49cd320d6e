/*
 * hpg_search.h -- the reference's C surface of the FM-index search path,
 * over the MI355X engine of fmi.h.
 *
 * A C caller written against the reference's headers
 *   search/results.h  (result, results_list and their macros),
 *   search/runtime.h  (bwt_index, BWiteration, size_SA, get_SA, get_ISA,
 *                      load_bwt_index, free_bwt_index),
 *   search/search.h   (BWExactSearch*, BWSearch1CPU, BWSearch1VectorHelper,
 *                      BWSearchCPU),
 *   search/io.h       (binsearch)
 * includes this header instead and links libhpgfm.so: the names, argument
 * lists, field orders and per-read semantics are the reference's, and every
 * search runs on the GPU (a per-read call is a batch of one; the *_array
 * variants at the end run one batch for many reads).
 *
 * Conventions that differ from the reference (it calls exit() or reads out of
 * bounds in these cases):
 *  - a call whose arguments the engine cannot honour (a view pairing other
 *    than the two the mapping layer uses, an interval outside the index, a
 *    read longer than FMI_MAX_READ_LEN, lists of different capacities for
 *    BWSearchCPU) leaves its outputs untouched, returns false / 0 where the
 *    function returns a value, and records an fmi status that
 *    hpg_search_status() returns (fmi_last_error() has the text);
 *  - a base code above 3 (N) never occurs in the index: an LF step on it
 *    gives the empty interval (n + 1, n).
 * Every call is reentrant given distinct result lists: each calling thread
 * gets its own stream and device staging.
 */
#ifndef HPG_SEARCH_H
#define HPG_SEARCH_H

#include <inttypes.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdlib.h>

#include "fmi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Error kinds (bwt_commons.h:15-18). */
#ifndef MATCH
#define MATCH 0
#define DELETION 1
#define MISMATCH 2
#define INSERTION 3
#endif

/* ------------------------------------------------ search/results.h:11-103 */
#define MAX_MISMATCHES 20

/* Same fields, order and types as search/results.h:13-21 (112 bytes on
 * x86-64). */
typedef struct {
  bool dir;                      /* 0 - Backward, 1 - Forward */
  intmax_t k, l;
  int16_t start, pos, end;
  uint8_t num_mismatches;
  uint8_t err_kind[MAX_MISMATCHES];
  uint8_t err_base[MAX_MISMATCHES];
  int16_t err_pos[MAX_MISMATCHES];
} result;

/* search/results.h:23-28 */
typedef struct {
  result *list;
  uintmax_t num_results;
  uintmax_t max_results;
  uintmax_t read_index;
} results_list;

/* The list and record macros of search/results.h:30-103, with the same
 * effect, except that new_results_list leaves list NULL when the
 * allocation fails (the reference exits). */
#define new_results_list(_r_list, _max_results)                                  \
  do {                                                                           \
    (_r_list)->list = (result *)malloc((size_t)(_max_results) * sizeof(result));  \
    (_r_list)->num_results = 0;                                                  \
    (_r_list)->max_results = (_max_results);                                     \
  } while (0)

#define init_result(_r, _dir)   \
  do {                          \
    (_r)->dir = (_dir);         \
    (_r)->num_mismatches = 0;   \
  } while (0)

#define bound_result(_r, _start, _end) \
  do {                                 \
    (_r)->start = (_start);            \
    (_r)->end = (_end);                \
  } while (0)

#define change_result(_r, _k, _l, _pos) \
  do {                                  \
    (_r)->k = (_k);                     \
    (_r)->l = (_l);                     \
    (_r)->pos = (_pos);                 \
  } while (0)

#define add_mismatch(_r, _err_kind, _base, _position)           \
  do {                                                          \
    (_r)->err_kind[(_r)->num_mismatches] = (_err_kind);         \
    (_r)->err_base[(_r)->num_mismatches] = (_base);             \
    (_r)->err_pos[(_r)->num_mismatches] = (_position);          \
    (_r)->num_mismatches++;                                     \
  } while (0)

#define modify_last_mismatch3(_r, _err_kind, _base, _position)  \
  do {                                                          \
    (_r)->err_kind[(_r)->num_mismatches - 1] = (_err_kind);     \
    (_r)->err_base[(_r)->num_mismatches - 1] = (_base);         \
    (_r)->err_pos[(_r)->num_mismatches - 1] = (_position);      \
  } while (0)

#define modify_last_mismatch2(_r, _err_kind, _base)             \
  do {                                                          \
    (_r)->err_kind[(_r)->num_mismatches - 1] = (_err_kind);     \
    (_r)->err_base[(_r)->num_mismatches - 1] = (_base);         \
  } while (0)

#define modify_last_mismatch1(_r, _err_kind)                    \
  do {                                                          \
    (_r)->err_kind[(_r)->num_mismatches - 1] = (_err_kind);     \
  } while (0)

#define copy_result(_dest, _orig)                                                  \
  do {                                                                             \
    init_result((_dest), (_orig)->dir);                                            \
    bound_result((_dest), (_orig)->start, (_orig)->end);                           \
    change_result((_dest), (_orig)->k, (_orig)->l, (_orig)->pos);                  \
    for (uint8_t rr_ = 0; rr_ < (_orig)->num_mismatches; rr_++)                    \
      add_mismatch((_dest), (_orig)->err_kind[rr_], (_orig)->err_base[rr_],        \
                   (_orig)->err_pos[rr_]);                                         \
  } while (0)

/* Appends a copy of orig, or executes `return true` from the enclosing
 * function when the list is full (search/results.h:93-103). */
#define add_result(orig, _r_list)                                      \
  do {                                                                 \
    if ((_r_list)->num_results < (_r_list)->max_results) {             \
      result *dest_ = &(_r_list)->list[(_r_list)->num_results];        \
      (_r_list)->num_results++;                                        \
      copy_result(dest_, (orig));                                      \
    } else {                                                           \
      return true;                                                     \
    }                                                                  \
  } while (0)

/* ------------------------------------------------ search/runtime.h:14-124 */
/* The index configuration of bwt_commons.h:20-31 (load_bwt_index takes it
 * by value; the engine's alphabet is ACGT, so its fields are not read). */
typedef struct {
  uint8_t *table;
  char *nucleotides;
  char *reverse;
  uint8_t nA;
  uint8_t AA, CC, GG, TT;
  bool inverse_sa;
  bool duplicate_strand;
} bwt_config_t;

/* One of the four logical indexes of bwt_index_new (bwt.c:682-711): a view
 * (FMI_VIEW_*) of a device-resident engine index.  The reference's struct
 * holds the arrays themselves (C, C1, O, S, R); callers only pass pointers
 * to it and use the macros below, which is what this one supports. */
typedef struct {
  fmi_index *fmi;     /* the engine index, on the device        */
  int view;           /* FMI_VIEW_BACKWARD .. FMI_VIEW_FORWARD_REV */
  uint64_t siz;       /* size_SA: n + 1                           */
  void *owner;        /* set by load_bwt_index (its directory)    */
} bwt_index;

/* Fills *out with `view` of an index already on the device. */
int hpg_bwt_index_view(fmi_index *ix, int view, bwt_index *out);

/* load_bwt_index (search/runtime.h:90-124): reads the reference's index
 * directory (fmi_index_import_reference) once per directory, copies it to
 * the current device, and fills `index` with BACKWARD (direction 1) or
 * FORWARD (direction 0) and `index_rev` (may be NULL) with the matching
 * _REV view.  free_bwt_index releases a directory after its last view.
 * On failure the views are zeroed (fmi == NULL) and hpg_search_status()
 * says why. */
void load_bwt_index(bwt_index *index_rev, bwt_index *index, const char *directory,
                    int direction, bool inverse_sa, bwt_config_t config);
void free_bwt_index(bwt_index *index_rev, bwt_index *index, bool inverse_sa);

/* BWiteration (search/runtime.h:38-42): k_out = C1[b] + Occ(b, k_in),
 * l_out = C[b] + Occ(b, l_in + 1) on the view; one LF step on the GPU. */
void hpg_bw_iteration(intmax_t k_in, intmax_t l_in, intmax_t *k_out, intmax_t *l_out,
                      uint8_t b, const bwt_index *index);
#define BWiteration(k_in, l_in, k_out, l_out, b, index)                      \
  do {                                                                       \
    intmax_t hk_, hl_;                                                       \
    hpg_bw_iteration((k_in), (l_in), &hk_, &hl_, (uint8_t)(b), (index));    \
    (k_out) = hk_;                                                           \
    (l_out) = hl_;                                                           \
  } while (0)

/* size_SA, get_SA (getScompValue), get_ISA (getRcompValue):
 * search/runtime.h:86-88.  A _REV view shares its direction's samples
 * (runtime.h:109-121), so it gives the same values. */
uint64_t hpg_size_sa(const bwt_index *index);
uintmax_t hpg_get_sa(uintmax_t m, const bwt_index *index);
uintmax_t hpg_get_isa(uintmax_t m, const bwt_index *index);
#define size_SA(index) hpg_size_sa(index)
#define get_SA(m, index) hpg_get_sa((m), (index))
#define get_ISA(m, index) hpg_get_isa((m), (index))

/* --------------------------------------------------- search/io.h:43-64 */
/* The number of entries of array[0..size) that are <= key (an upper bound);
 * 0 when array is NULL. */
uintmax_t binsearch(uintmax_t *array, uintmax_t size, uintmax_t key);

/* ------------------------------------------------ search/search.h:9-207 */
/* BWExactSearchBackward / Forward (search/search.h:9-65): from r->k, r->l,
 * W[r->pos] down to W[r->start] (up to W[r->end]); r gets the final
 * interval and the position where the search stopped, r_anchor (may be
 * NULL) the last non-empty interval and its bounds. */
void BWExactSearchBackward(uint8_t *W, bwt_index *index, result *r, result *r_anchor);
void BWExactSearchForward(uint8_t *W, bwt_index *index, result *r, result *r_anchor);

/* BWExactSearchVectorBackward / Forward (search/search.c:1407-1489). */
int16_t BWExactSearchVectorBackward(uint8_t *W, int16_t start, int16_t end, intmax_t k,
                                    intmax_t l, intmax_t *vec_k, intmax_t *vec_l,
                                    bwt_index *index);
int16_t BWExactSearchVectorForward(uint8_t *W, int16_t start, int16_t end, intmax_t k,
                                   intmax_t l, intmax_t *vec_k, intmax_t *vec_l,
                                   bwt_index *index);

/* BWSearch1CPU (search/search.c:1876-1933): appends W's <=1-error records to
 * r_list in the reference's order and, with r_list_anchor, resets it and
 * adds the two anchors; returns true where the reference's add_result does
 * (its list filled while pass 1 was running, or the anchor list is full).
 * (backward, forward) is (BACKWARD, FORWARD) -- the + strand -- or
 * (FORWARD_REV, BACKWARD_REV) -- the - strand (bwt.c:1690-1701); res is
 * {k = 0, l = size_SA - 1, start = 0, end = len - 1} as the mapping layer
 * passes it. */
bool BWSearch1CPU(uint8_t *W, bwt_index *backward, bwt_index *forward, result *res,
                  results_list *r_list, results_list *r_list_anchor, uint8_t nA);

/* BWSearch1VectorHelper (search/search.c:1491-1709) over W[start..end] driven
 * by the caller's interval vectors (vec_k/vec_l from
 * BWExactSearchVectorBackward on `backward`, vec_ki/vec_li from
 * BWExactSearchVectorForward on `forward`, bwt.c:1940-2003); appends to
 * r_list and returns true when it filled up.  Windows of fewer than 3
 * bases give no records. */
bool BWSearch1VectorHelper(uint8_t *W, int16_t start, int16_t end, intmax_t *vec_k,
                           intmax_t *vec_l, intmax_t *vec_ki, intmax_t *vec_li,
                           bwt_index *backward, bwt_index *forward, results_list *r_list,
                           uint8_t nA);

/* BWSearchCPU (search/search.c:1171-1405): k-error search of W[0..nW) cut
 * into nW / fragsize fragments; rl_final receives the final list (it must
 * be empty on entry, as bwt.c:2435 leaves it) and rl_anchor (may be NULL)
 * the anchors; rl_prev / rl_next / rl_prev_i / rl_next_i are the working
 * lists and give the search their capacity, which must equal rl_final's
 * (bwt.c:2405-2409); they are left empty.  type 1 pairs with (BACKWARD,
 * FORWARD), type 0 with (FORWARD_REV, BACKWARD_REV) (bwt.c:2430-2470).
 * Returns true when the search stopped on a full anchor list. */
bool BWSearchCPU(uint8_t *W, uint64_t nW, bwt_index *backward, bwt_index *forward,
                 results_list *rl_prev, results_list *rl_next, results_list *rl_prev_i,
                 results_list *rl_next_i, results_list *rl_final, results_list *rl_anchor,
                 int16_t fragsize, bool type, uint8_t nA);

/* The fmi status of the calling thread's last call above (FMI_OK when it
 * was honoured). */
int hpg_search_status(void);

/* ------------------------------------------- many reads, one GPU batch */
/* BWSearch1CPU for every read i of a host batch (reads: codes, offsets:
 * nreads + 1 entries) with res = the whole read: r_lists[i] and (when
 * r_list_anchors is not NULL) r_list_anchors[i] as the per-read call leaves
 * them, ret[i] its return value.  Returns an fmi status. */
int hpg_search1_array(const uint8_t *reads, const uint64_t *offsets, uint32_t nreads,
                      bwt_index *backward, bwt_index *forward, results_list *r_lists,
                      results_list *r_list_anchors, bool *ret);
/* BWExactSearchBackward (direction 0) / Forward (direction 1) for every
 * read i: r[i] in and out as the per-read call, anchors[i] (may be NULL)
 * its anchor.  W of read i is reads + offsets[i]. */
int hpg_exact_array(const uint8_t *reads, const uint64_t *offsets, uint32_t nreads,
                    bwt_index *index, int direction, result *r, result *anchors);

#ifdef __cplusplus
}
#endif
#endif

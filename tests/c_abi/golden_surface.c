/* golden_surface.c -- a C caller of the reference-named search surface
 * (include/hpg_search.h), compiled with cc against the header and linked
 * with libhpgfm.so, that runs the reference's own call patterns on every
 * read of a golden set and writes what they leave in its results_list /
 * result structures:
 *
 *   BWSearch1CPU with r_list_anchor, both strand pairings (bwt.c:1686-1701)
 *   BWExactSearchBackward / Forward on the four views, with anchors
 *   get_SA / get_ISA on the backward and forward index
 *   BWExactSearchVector* + BWSearch1VectorHelper on whole reads and on one
 *     window per read (bwt.c:1940-2003, 1447-1479)
 *   BWSearchCPU with rl_anchor, both pairings (bwt.c:2430-2470)
 *
 * usage: golden_surface IN OUT  (IN: written by tests/test_hpg_search.py)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hpg_search.h"

static void *slurp(FILE *f, size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (!p || (bytes && fread(p, 1, bytes, f) != bytes)) {
    fprintf(stderr, "short input\n");
    exit(2);
  }
  return p;
}

static void put(FILE *f, const void *p, size_t bytes) {
  if (bytes && fwrite(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "write failed\n");
    exit(2);
  }
}

static void put_list(FILE *f, const results_list *rl) {
  uint32_t c = (uint32_t)rl->num_results;
  put(f, &c, 4);
  put(f, rl->list, (size_t)c * sizeof(result));
}

#define CHECK(cond, what)                                                      \
  do {                                                                         \
    if (!(cond)) {                                                             \
      fprintf(stderr, "%s failed: status %d: %s\n", what, hpg_search_status(), \
              fmi_last_error());                                               \
      return 1;                                                                \
    }                                                                          \
  } while (0)

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
    return 2;
  }
  FILE *in = fopen(argv[1], "rb");
  FILE *out = fopen(argv[2], "wb");
  if (!in || !out) return 2;
  uint64_t hdr[4];
  if (fread(hdr, 8, 4, in) != 4) return 2;
  const uint64_t n = hdr[0], nreads = hdr[2];
  const uint32_t ratio = (uint32_t)(hdr[1] & 0xFFFFFFFFu), cap = (uint32_t)(hdr[1] >> 32);
  const uint32_t kcap = (uint32_t)(hdr[3] & 0xFFFFFFFFu);
  uint8_t *genome = (uint8_t *)slurp(in, n);
  uint64_t *offsets = (uint64_t *)slurp(in, (nreads + 1) * 8);
  uint8_t *reads = (uint8_t *)slurp(in, offsets[nreads]);
  int16_t *bounds = (int16_t *)slurp(in, nreads * 4);
  int16_t *kfs = (int16_t *)slurp(in, nreads * 2);
  uint64_t nrows;
  if (fread(&nrows, 8, 1, in) != 1) return 2;
  uint64_t *rows = (uint64_t *)slurp(in, nrows * 8);
  fclose(in);

  fmi_index *ix = NULL;
  CHECK(fmi_index_build(genome, n, ratio, &ix) == FMI_OK, "fmi_index_build");
  CHECK(fmi_index_to_device(ix) == FMI_OK, "fmi_index_to_device");
  bwt_index backward, forward, backward_rev, forward_rev;
  CHECK(hpg_bwt_index_view(ix, FMI_VIEW_BACKWARD, &backward) == FMI_OK, "view");
  CHECK(hpg_bwt_index_view(ix, FMI_VIEW_FORWARD, &forward) == FMI_OK, "view");
  CHECK(hpg_bwt_index_view(ix, FMI_VIEW_BACKWARD_REV, &backward_rev) == FMI_OK, "view");
  CHECK(hpg_bwt_index_view(ix, FMI_VIEW_FORWARD_REV, &forward_rev) == FMI_OK, "view");
  bwt_index *views[4] = {&backward, &forward, &backward_rev, &forward_rev};
  const intmax_t lsa = (intmax_t)size_SA(&backward) - 1;

  results_list r_list, r_anchor;
  new_results_list(&r_list, cap);
  new_results_list(&r_anchor, 2);

  /* BWSearch1CPU, + strand (backward, forward) then - strand (forward_rev,
   * backward_rev), each with res = the whole read (bwt.c:1686-1701) */
  for (uint64_t i = 0; i < nreads; i++) {
    uint8_t *W = reads + offsets[i];
    const int16_t len = (int16_t)(offsets[i + 1] - offsets[i]);
    for (int s = 0; s < 2; s++) {
      result res;
      memset(&res, 0, sizeof res);
      init_result(&res, 0);
      change_result(&res, 0, lsa, len - 1);
      bound_result(&res, 0, len - 1);
      r_list.num_results = 0;
      bool full = s == 0 ? BWSearch1CPU(W, &backward, &forward, &res, &r_list, &r_anchor, 4)
                         : BWSearch1CPU(W, &forward_rev, &backward_rev, &res, &r_list, &r_anchor, 4);
      CHECK(hpg_search_status() == FMI_OK, "BWSearch1CPU");
      uint32_t f = full ? 1 : 0;
      put(out, &f, 4);
      put_list(out, &r_list);
      put_list(out, &r_anchor);
    }
  }

  /* exact searches on the four views from the whole index */
  for (int v = 0; v < 4; v++) {
    for (int d = 0; d < 2; d++) {
      for (uint64_t i = 0; i < nreads; i++) {
        uint8_t *W = reads + offsets[i];
        const int16_t len = (int16_t)(offsets[i + 1] - offsets[i]);
        result r, a;
        memset(&r, 0, sizeof r);
        memset(&a, 0, sizeof a);
        init_result(&r, d);
        bound_result(&r, 0, len - 1);
        change_result(&r, 0, lsa, d == 0 ? len - 1 : 0);
        if (d == 0)
          BWExactSearchBackward(W, views[v], &r, &a);
        else
          BWExactSearchForward(W, views[v], &r, &a);
        CHECK(hpg_search_status() == FMI_OK, "BWExactSearch");
        put(out, &r, sizeof r);
        put(out, &a, sizeof a);
      }
    }
  }

  /* get_SA / get_ISA */
  for (uint64_t j = 0; j < nrows; j++) {
    uint64_t v[4] = {get_SA(rows[j], &backward), get_SA(rows[j], &forward),
                     get_ISA(rows[j], &backward), get_ISA(rows[j], &forward)};
    CHECK(hpg_search_status() == FMI_OK, "get_SA / get_ISA");
    put(out, v, sizeof v);
  }

  /* BWSearch1VectorHelper on whole reads, then on each read's window */
  intmax_t *vec = (intmax_t *)malloc(4 * 1024 * sizeof(intmax_t));
  for (int pass = 0; pass < 2; pass++) {
    for (uint64_t i = 0; i < nreads; i++) {
      uint8_t *W = reads + offsets[i];
      const int16_t len = (int16_t)(offsets[i + 1] - offsets[i]);
      const int16_t start = pass ? bounds[2 * i] : 0, end = pass ? bounds[2 * i + 1] : len - 1;
      for (int s = 0; s < 2; s++) {
        r_list.num_results = 0;
        uint32_t f = 0;
        if (end - start >= 2) {
          intmax_t *ak = vec, *al = vec + 1024, *bk = vec + 2048, *bl = vec + 3072;
          bwt_index *B = s ? &forward_rev : &backward, *F = s ? &backward_rev : &forward;
          BWExactSearchVectorBackward(W, start, end, 0, lsa, ak, al, B);
          BWExactSearchVectorForward(W, start, end, 0, lsa, bk, bl, F);
          CHECK(hpg_search_status() == FMI_OK, "BWExactSearchVector");
          f = BWSearch1VectorHelper(W, start, end, ak, al, bk, bl, B, F, &r_list, 4) ? 1 : 0;
          CHECK(hpg_search_status() == FMI_OK, "BWSearch1VectorHelper");
        }
        put(out, &f, 4);
        put_list(out, &r_list);
      }
    }
  }

  /* BWSearchCPU, type 1 on (backward, forward), type 0 on (forward_rev,
   * backward_rev), lists of kcap records and anchor lists of 16 */
  results_list L[5], A;
  for (int j = 0; j < 5; j++) new_results_list(&L[j], kcap);
  new_results_list(&A, 16);
  for (uint64_t i = 0; i < nreads; i++) {
    uint8_t *W = reads + offsets[i];
    const uint64_t len = offsets[i + 1] - offsets[i];
    for (int s = 0; s < 2; s++) {
      L[4].num_results = 0;
      bool full = s == 0 ? BWSearchCPU(W, len, &backward, &forward, &L[0], &L[1], &L[2], &L[3],
                                       &L[4], &A, kfs[i], 1, 4)
                         : BWSearchCPU(W, len, &forward_rev, &backward_rev, &L[0], &L[1], &L[2],
                                       &L[3], &L[4], &A, kfs[i], 0, 4);
      CHECK(hpg_search_status() == FMI_OK, "BWSearchCPU");
      uint32_t f = full ? 1 : 0;
      put(out, &f, 4);
      put_list(out, &L[4]);
      put_list(out, &A);
    }
  }

  /* binsearch (search/io.h:43-64) on a small array */
  uintmax_t arr[5] = {0, 10, 10, 20, 35};
  uint64_t bs[6] = {binsearch(arr, 5, 0), binsearch(arr, 5, 9), binsearch(arr, 5, 10),
                    binsearch(arr, 5, 34), binsearch(arr, 5, 99), binsearch(NULL, 5, 3)};
  put(out, bs, sizeof bs);

  fclose(out);
  fmi_index_free(ix);
  printf("golden_surface: %llu reads done\n", (unsigned long long)nreads);
  return 0;
}

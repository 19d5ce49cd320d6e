/*
 * fmoracle.h -- CPU restatement of the reference FM-index search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in the product links or calls it.
 *
 * It keeps the reference's own data layout (per-symbol uint32 checkpoints
 * `desp` plus per-symbol uint64 occurrence bitmaps `count`, one word per 64
 * BWT rows, c/src/aligners/bwt/search/csafm.h:51-81 with FM_COMP_64) so that
 * its timing is representative of the reference CPU path.
 *
 * Every function cites the reference file:line it follows.  Paths are
 * relative to c/src/aligners/bwt/ of opencb/hpg-libs.
 *
 * Pinning (what checks each function against something other than itself):
 *   fmo_sa_naive / fmo_index_from_bwt : brute-force suffix sort; SHA of the
 *       reference builder's C/O/S/R files (golden); reference dbwt via _ref.
 *   fmo_exact / fmo_exact_vector : brute-force text scan (interval size =
 *       occurrence count, SA range), golden vectors, _ref.
 *   fmo_search1 : golden vectors from the compiled reference, _ref, and the
 *       Hamming-1 projection against a brute-force text scan.
 *   fmo_get_sa / fmo_get_isa : brute-force SA / ISA.
 *   fmo_binsearch / fmo_resolve : closed form on the karyotype table.
 */
#ifndef FMORACLE_H
#define FMORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error kinds, same codes as bwt_commons.h:15-18. */
#define FMO_MAX_LEN 1000   /* MAXLINE, bwt_commons.h:11 */
#define FMO_MATCH     0
#define FMO_DELETION  1
#define FMO_MISMATCH  2
#define FMO_INSERTION 3

/* One direction of the index (BWT of X, or BWT of reverse X). */
typedef struct {
  uint64_t n;          /* text length without $                          */
  uint64_t dollar;     /* BWT row whose symbol is $                      */
  uint32_t C[4];       /* # symbols < b           (preprocess.c:44-72)   */
  uint32_t C1[4];      /* C[b] + 1                                        */
  uint64_t words;      /* ceil((n+2)/64)          (preprocess.c:78-80)   */
  uint32_t *desp[4];   /* checkpoint counts       (preprocess.c:74-130)  */
  uint64_t *count[4];  /* bitmaps, row r at bit (r+1)%64 of word (r+1)/64 */
  uint32_t ratio;      /* SA / ISA sample rate                            */
  uint64_t nsamp;      /* ceil((n+1)/ratio)                               */
  uint32_t *S;         /* S[i/ratio] = SA[i] for rows i % ratio == 0     */
  uint32_t *R;         /* R[j/ratio] = ISA[j] for text pos j % ratio == 0 */
} fmo_dir;

typedef struct {
  fmo_dir bwd;         /* "backward": BWT of X                            */
  fmo_dir fwd;         /* "forward":  BWT of reverse(X)                   */
} fmo_index;

/* The four logical indexes of bwt.c:682-711: the *_rev views swap A<->T and
 * C<->G in C/C1/O (csafm.c:3-100), i.e. look up symbol 3-b. */
enum { FMO_VIEW_BACKWARD = 0, FMO_VIEW_FORWARD = 1,
       FMO_VIEW_BACKWARD_REV = 2, FMO_VIEW_FORWARD_REV = 3 };

/* Search record with the fields of `result` (search/results.h:13-21) that a
 * 1-error search can set (one error record at most). 32 bytes. */
typedef struct {
  int64_t k, l;
  int16_t start, pos, end, err_pos;
  uint8_t dir, num_mismatches, err_kind, err_base;
  uint8_t strand, pad[3];
} fmo_hit;

/* ---- builder ---------------------------------------------------------- */
/* Suffix array of text[0..n) + '$' by comparison sort (small texts only).
 * sa has n+1 entries; sa[0] = n (the '$' suffix). Returns 0 on success. */
int fmo_sa_naive(const uint8_t *text, uint64_t n, uint64_t *sa);
/* BWT without the $ position, plus dollar row, from a suffix array. */
void fmo_bwt_from_sa(const uint8_t *text, uint64_t n, const uint64_t *sa,
                     uint8_t *bwt_no_dollar, uint64_t *dollar);
/* calculate_C/O/S_and_R (preprocess.c:7-130) from a BWT given in the B.vec
 * convention (n bytes, the $ row removed, dollar row separately). */
int fmo_dir_from_bwt(fmo_dir *d, const uint8_t *bwt_no_dollar, uint64_t n,
                     uint64_t dollar, uint32_t ratio);
void fmo_dir_free(fmo_dir *d);
/* Builds both directions from a text (naive SA; small texts). */
int fmo_index_build_naive(fmo_index *ix, const uint8_t *text, uint64_t n, uint32_t ratio);
/* Builds both directions from externally computed BWTs (B of X, B of rev X). */
int fmo_index_from_bwts(fmo_index *ix, const uint8_t *bwt, uint64_t dollar,
                        const uint8_t *bwt_rev, uint64_t dollar_rev, uint64_t n,
                        uint32_t ratio);
void fmo_index_free(fmo_index *ix);
/* Installs S/R samples computed elsewhere (S[i] = SA[i * ratio],
 * R[i] = ISA[i * ratio], the layout calculate_S_and_R writes) into one
 * direction of an index built with ratio 0. */
int fmo_dir_set_samples(fmo_index *ix, int fwd, uint32_t ratio, const uint32_t *S,
                        const uint32_t *R, uint64_t nsamp);

/* Raw access for tests: copies the reference-layout arrays out. */
uint64_t fmo_dir_words(const fmo_index *ix, int fwd);
void fmo_dir_export(const fmo_index *ix, int fwd, uint32_t *C, uint32_t *C1,
                    uint32_t *desp /* 4*words */, uint64_t *count /* 4*words */,
                    uint32_t *S /* nsamp */, uint32_t *R /* nsamp */,
                    uint64_t *dollar, uint64_t *nsamp);

/* ---- runtime ---------------------------------------------------------- */
uint64_t fmo_size_sa(const fmo_index *ix);                       /* runtime.h:86 */
uint64_t fmo_occ(const fmo_index *ix, int view, unsigned b, uint64_t m);   /* csafm.h:166 */
void fmo_lf(const fmo_index *ix, int view, int64_t k, int64_t l, unsigned b,
            int64_t *ko, int64_t *lo);                            /* runtime.h:38 */
uint64_t fmo_get_sa(const fmo_index *ix, int fwd, uint64_t m);   /* csafm.h:208 */
uint64_t fmo_get_isa(const fmo_index *ix, int fwd, uint64_t m);  /* csafm.h:237 */

/* ---- search ----------------------------------------------------------- */
/* BWExactSearchBackward (dir=0) / BWExactSearchForward (dir=1),
 * search/search.h:9-65.  r holds k,l,start,pos,end on entry; on exit k,l,pos.
 * anchor may be NULL. */
void fmo_exact(const fmo_index *ix, int view, int dir, const uint8_t *W,
               fmo_hit *r, fmo_hit *anchor);
/* BWExactSearchVectorBackward/Forward, search/search.c:1407-1489. */
int16_t fmo_exact_vector(const fmo_index *ix, int view, int dir, const uint8_t *W,
                         int16_t start, int16_t end, int64_t k, int64_t l,
                         int64_t *vec_k, int64_t *vec_l);
/* BWSearch1CPU, search/search.c:1876-1933, with the strand pairing of
 * bwt.c:1999-2001: strand 0 = (backward, forward), strand 1 =
 * (forward_rev, backward_rev).  mismatch_only drops the deletion and
 * insertion branches (the {exact, MISMATCH} projection).  Records beyond
 * cap are counted but not stored.  anchors (may be NULL) receives the two
 * anchors of search.c:1899-1925.  Returns the total record count. */
uint32_t fmo_search1(const fmo_index *ix, int strand, int mismatch_only,
                     const uint8_t *W, int16_t len, fmo_hit *out, uint32_t cap,
                     fmo_hit *anchors);

/* BWSearch1VectorHelper (search.c:1491-1709) over W[start..end] on one
 * strand, with the interval vectors bwt.c:1940-1950 computes (+ strand:
 * backward/forward, - strand: forward_rev/backward_rev).  Records as
 * fmo_search1; mismatch_only keeps the exact record and MISMATCH records.
 * end - start < 2 gives no records (the reference reads outside its vectors
 * there; the mapping layer calls it for reads of 5 bases or more). */
uint32_t fmo_search1_vector(const fmo_index *ix, int strand, int mismatch_only,
                            const uint8_t *W, int16_t start, int16_t end, fmo_hit *out,
                            uint32_t cap);

/* ---- k-error search ---------------------------------------------------- */
/* A `result` (search/results.h:13-21) with room for FMO_KMAX error records.
 * 64 bytes; entries past num_mismatches are not part of the record. */
#define FMO_KMAX 8
typedef struct {
  int64_t k, l;
  int16_t start, pos, end;
  uint8_t dir, num_mismatches;
  uint8_t err_kind[FMO_KMAX], err_base[FMO_KMAX];
  int16_t err_pos[FMO_KMAX];
  uint8_t pad[8];
} fmo_hitk;

typedef struct { fmo_hitk *list; uint64_t n, cap; } klist;

/* BWSearchCPU (search.c:1171-1405) on one strand of W, as
 * bwt_map_inexact_read_2 calls it (bwt.c:2430-2470): fragments = len /
 * fragsize, up to fragments - 1 errors; the five working lists and the final
 * list hold list_cap records, the anchor list anchor_cap.  Copies the final
 * list and the anchors out; returns 1 when the search stopped because the
 * anchor list was full (the reference's return value), else 0. */
int fmo_search_k(const fmo_index *ix, int strand, const uint8_t *W, int16_t len, int16_t fragsize,
                 uint32_t list_cap, uint32_t anchor_cap, fmo_hitk *final_out, uint32_t *nfinal,
                 fmo_hitk *anchors_out, uint32_t *nanchors);

/* ---- seed mapping ------------------------------------------------------ */
/* One region of bwt_map_exact_seed (same layout as fmi_seed_region). */
typedef struct {
  uint64_t start, end;
  uint32_t chrom_idx, read;
  uint16_t seq_start, seq_end, seed_id;
  uint8_t strand, valid;
} fmo_seed_region;

/* bwt_map_exact_seed (bwt.c:1200-1294) for seeds [s, s + seed_len - 1],
 * s = 0, stride, ... of one read.  koff has kcount + 1 entries (contig text
 * offsets, offset[c+1] = offset[c] + length), kstart kcount entries.  Every
 * resolved row yields one region (valid = 0 when the contig-end check
 * fails, where the reference makes none).  info (may be NULL) gets 6 int64
 * per seed: k+, l+, k-, l-, rows+, rows-.  Returns the number of rows; at
 * most cap regions are stored. */
uint64_t fmo_seed_map(const fmo_index *ix, const uint8_t *W, uint32_t len, uint32_t seed_len,
                      uint32_t stride, int32_t filter, const uint64_t *koff,
                      const uint64_t *kstart, uint64_t kcount, uint32_t read,
                      fmo_seed_region *out, uint64_t cap, int64_t *info);

/* One region of bwt_map_inexact_seed (same layout as fmi_inexact_region). */
typedef struct {
  uint64_t start, end;
  uint32_t chrom_idx, read;
  uint16_t seq_start, seq_end, seq_len, seed_id;
  int32_t pos_err, type_err;
  uint8_t strand, pad[7];
} fmo_inexact_region;

/* bwt_map_inexact_seeds_seq (bwt.c:3262-3302) for one read: windows of
 * seed_size bases at 0, seed_size, ... and the remainder when it holds at
 * least min_seed_size bases; per window bwt_map_inexact_seed (bwt.c:
 * 1401-1644): the 1-error vector helper on both strands (+ then -), a
 * region per kept SA row (at most 50 rows, else rows k and k + 1), the
 * within-10-bp dedup of each strand's list and its reversed hand-over.
 * Regions go to out[(2 * w + s) * region_cap ...], their count to
 * nregions[2 * w + s] and status[2 * w + s] gets bit 0 when the strand's
 * records exceeded rec_cap, bit 1 when its regions exceeded region_cap
 * (the stored list is then not the reference's).  Returns the window
 * count; at most wcap windows are written. */
uint32_t fmo_map_inexact_seeds(const fmo_index *ix, const uint8_t *W, uint32_t len,
                               uint32_t seed_size, uint32_t min_seed_size, const uint64_t *koff,
                               const uint64_t *kstart, uint64_t kcount, uint32_t read,
                               uint32_t rec_cap, uint32_t region_cap, fmo_inexact_region *out,
                               uint32_t *nregions, uint8_t *status, uint32_t wcap);

/* fmo_map_inexact_seeds over a batch (reads of at most FMO_MAX_LEN bases),
 * OpenMP over reads; *nregions_total gets the regions kept.  Returns seconds. */
double fmo_map_inexact_seeds_batch(const fmo_index *ix, const uint8_t *reads,
                                   const uint64_t *offsets, uint64_t nreads, uint32_t seed_size,
                                   uint32_t min_seed_size, const uint64_t *koff,
                                   const uint64_t *kstart, uint64_t kcount, uint32_t rec_cap,
                                   uint32_t region_cap, uint64_t *nregions_total, int threads);

/* ---- read-level 1-error mapping ----------------------------------------- */
/* One alignment of __bwt_map_inexact_read (bwt.c:2158-2178): 24 bytes,
 * the layout of fmi_read_mapping. */
typedef struct {
  uint64_t position;    /* 0-based start[idx-1] + key - offset[idx-1] */
  uint32_t chromosome;  /* idx - 1 */
  uint32_t read;
  uint8_t strand;       /* !type: 0 +, 1 - */
  uint8_t secondary;    /* secondary_alignment after bwt.c:2217-2233 */
  uint16_t cigar_len;   /* the CIGAR "<cigar_len>M" (one operation) */
  uint32_t pad;
} fmo_read_mapping;

/* __bwt_map_inexact_read (bwt.c:1890-2236) for one read up to its list of
 * alignments (the anchor lists of bwt.c:2239-2269 for unmapped reads are not
 * built): see fmoracle.c.  status bit 0: filter_read_mappings exceeded,
 * bit 1: read under 5 bases.  At most cap stored; returns the total. */
uint32_t fmo_map_read(const fmo_index *ix, const uint8_t *W, uint32_t len, uint32_t filter,
                      const uint64_t *koff, const uint64_t *kstart, uint64_t kcount,
                      uint32_t read, fmo_read_mapping *out, uint32_t cap, uint8_t *status);

/* ---- position resolution ---------------------------------------------- */
/* binsearch, search/io.h:43-64: upper_bound of key in array[0..size). */
uint64_t fmo_binsearch(const uint64_t *array, uint64_t size, uint64_t key);

/* ---- batch drivers for the CPU baseline (OpenMP over reads as in
 * bwt.c:1112,1134: schedule(dynamic, max(1, n/(T*10)))) --------------- */
/* Runs fmo_search1 on both strands of every read; counts[2*i+s] gets the
 * record count; hits (may be NULL) gets cap records per (read, strand). */
double fmo_search1_batch(const fmo_index *ix, const uint8_t *reads,
                         const uint64_t *offsets, uint64_t nreads, int mismatch_only,
                         uint32_t cap, fmo_hit *hits, uint32_t *counts, int threads);
/* Exact backward search of every read on the backward index (+ strand) and
 * forward search on backward_rev (- strand), bwt.c:1200-1250 pairing.
 * out_kl[4*i + 0..3] = k+, l+, k-, l-. */
double fmo_exact_batch(const fmo_index *ix, const uint8_t *reads,
                       const uint64_t *offsets, uint64_t nreads, int64_t *out_kl,
                       int threads);
/* fmo_search_k on both strands of every read with fragsize[i]: task
 * t = 2 * read + strand gets final_out[t * list_cap ..], nfinal[t],
 * anchors_out[t * anchor_cap ..] (may be NULL), nanchors[t], full[t]. */
double fmo_search_k_batch(const fmo_index *ix, const uint8_t *reads, const uint64_t *offsets,
                          uint64_t nreads, const int16_t *fragsize, uint32_t list_cap,
                          uint32_t anchor_cap, fmo_hitk *final_out, uint32_t *nfinal,
                          fmo_hitk *anchors_out, uint32_t *nanchors, uint8_t *full, int threads);
/* The same searches, then get_SA(j, backward) over the first cap rows of
 * each interval: nrows / possum[2*i + strand]. */
double fmo_exact_locate_batch(const fmo_index *ix, const uint8_t *reads,
                              const uint64_t *offsets, uint64_t nreads, uint32_t cap,
                              uint32_t *nrows, uint64_t *possum, int threads);

#ifdef __cplusplus
}
#endif
#endif

/*
 * fmi.h -- C ABI of the MI355X FM-index read-search engine (libhpgfm.so).
 *
 * Drop-in boundary for the search path of opencb/hpg-libs
 * (c/src/aligners/bwt/search/ and the GPU glue gpu.h / gpu/gpu.cuh).  Each
 * batch entry point below computes, for every read of a batch, exactly what
 * the cited reference function computes for one read; the reference per-read
 * functions are CPU code called inside an OpenMP loop (bwt.c:1112,1134), the
 * engine runs the whole batch on one GPU.
 *
 * Conventions
 *  - Plain C types only.  Buffers named d_* are device (HBM) pointers
 *    allocated with fmi_device_alloc; everything else is host memory.
 *  - Bases are the codes 0..3 of the alphabet "ACGT" (the reference encodes a
 *    base as its position in config.txt's alphabet, bwt_commons.c:60-91).  A
 *    code > 3 (N) is a symbol that never occurs in the index: an LF step on it
 *    yields the empty interval (size_SA, size_SA - 1).  The reference leaves
 *    this case undefined.
 *  - Reads are concatenated; read i is reads[offsets[i] .. offsets[i+1]).
 *    Read length is at most FMI_MAX_READ_LEN (MAXLINE, bwt_commons.h:11).
 *  - Intervals are int64 [k, l], empty when k > l; positions are int16 like
 *    the reference's `result` (search/results.h:13-21).
 *  - Every call returns an fmi_status; nothing calls exit() (the reference's
 *    check_* macros do, bwt_commons.h:48-70).  Calls that take an
 *    fmi_stream are ordered on it and return before the GPU finishes;
 *    fmi_stream_sync waits.
 *  - The index is read-only after fmi_index_to_device and may be shared by
 *    streams; calls on one stream are serialised, calls on different streams
 *    may overlap.
 */
#ifndef FMI_H
#define FMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  FMI_OK = 0,
  FMI_EINVAL = -1,    /* bad argument                                   */
  FMI_ENOMEM = -2,    /* host or device allocation failed               */
  FMI_EIO = -3,       /* file could not be read or written              */
  FMI_EFORMAT = -4,   /* index file corrupt or of another version       */
  FMI_EDEVICE = -5,   /* HIP runtime error or no usable GPU             */
  FMI_ERANGE = -6,    /* a size is beyond what the engine supports      */
  FMI_ESTATE = -7,    /* call out of order (e.g. index not on device)   */
  FMI_EAGAIN = -8,    /* not now: fmi_server_wait timed out, or the
                         server's queue / staging is full (submit again) */
  FMI_ECANCELED = -9  /* the request was cancelled (fmi_server_cancel)  */
} fmi_status;

#define FMI_MAX_READ_LEN 1000

/* Error kinds: the codes of bwt_commons.h:15-18. */
#define FMI_MATCH 0
#define FMI_DELETION 1
#define FMI_MISMATCH 2
#define FMI_INSERTION 3

/* The four logical indexes of bwt_index_new (bwt.c:682-711): BACKWARD is
 * the BWT of the reference X, FORWARD the BWT of reverse(X); the *_REV views
 * swap A<->T and C<->G (reverse_strand_C/O, search/csafm.c:3-100). */
#define FMI_VIEW_BACKWARD 0
#define FMI_VIEW_FORWARD 1
#define FMI_VIEW_BACKWARD_REV 2
#define FMI_VIEW_FORWARD_REV 3

/* One search record: the fields of the reference's `result`
 * (search/results.h:13-21) that a search with at most one error sets.
 * num_mismatches is 0 for an exact record (err_* are then 0) and 1 for a
 * record with one error (err_kind, err_base, err_pos of err_*[0]).  strand
 * is 0 for the (backward, forward) pairing and 1 for (forward_rev,
 * backward_rev) (bwt.c:1693-1701). 32 bytes. */
typedef struct fmi_hit {
  int64_t k, l;
  int16_t start, pos, end, err_pos;
  uint8_t dir, num_mismatches, err_kind, err_base;
  uint8_t strand, pad[3];
} fmi_hit;

/* A `result` of the k-error search (search/results.h:13-21) with room for
 * FMI_KMAX error records; entries past num_mismatches are not part of the
 * record.  64 bytes. */
#define FMI_KMAX 8
typedef struct fmi_hitk {
  int64_t k, l;
  int16_t start, pos, end;
  uint8_t dir, num_mismatches;
  uint8_t err_kind[FMI_KMAX], err_base[FMI_KMAX];
  int16_t err_pos[FMI_KMAX];
  uint8_t pad[8];
} fmi_hitk;

typedef struct fmi_index fmi_index;   /* opaque: host arrays + device copy */
typedef void *fmi_stream;             /* a hipStream_t; NULL = default stream */

/* ---------------------------------------------------------------- errors */
const char *fmi_status_str(int status);
/* Text of the last error raised on the calling thread. */
const char *fmi_last_error(void);

/* ------------------------------------------------------------- the index */
/* Builds both directions of the index (BWT of text and of reverse(text),
 * C/C1, occurrence blocks, SA samples every sa_ratio rows, ISA samples every
 * sa_ratio text positions) from codes 0..3.  Replaces
 * bwt_generate_index_files (bwt.c:744-819; calculate_C/O/S_and_R,
 * search/preprocess.c:7-130) for an in-memory text.  n < 2^32 - 2. */
int fmi_index_build(const uint8_t *text, uint64_t n, uint32_t sa_ratio, fmi_index **out);
/* The same index, built on the current GPU (suffix sorting by prefix
 * doubling, blocks and samples written by kernels) and left on that device
 * (no fmi_index_to_device needed); host copies are kept for saving.  Needs
 * about 30 bytes of free device memory per base (3.1 Gbp: ~95 GB). */
int fmi_index_build_gpu(const uint8_t *text, uint64_t n, uint32_t sa_ratio, fmi_index **out);
/* Writes / reads the single-file index ("HPGFMI01" magic and format
 * version, then per direction the '$' row, C, the occurrence blocks and the
 * SA / ISA samples, then the karyotype and contig names, with an FNV-1a
 * checksum; files of format 2, without names, still load).  Replaces the reference's index
 * directory reader load_bwt_index (search/runtime.h:90-124). */
int fmi_index_save(const fmi_index *ix, const char *path);
int fmi_index_load(const char *path, fmi_index **out);
/* Reads an index directory written by the reference's builder
 * (bwt_generate_index_files, bwt.c:744-819) the way bwt_index_new
 * (bwt.c:682-712) does: config.txt (read_config, search/io.c:336-358),
 * C / C1, O / Oi (desp + count), S / R and Si / Ri (load_bwt_index,
 * search/runtime.h:90-124; read_vector, read_comp_vector, read_comp_matrix,
 * search/csafm.c:121-250) and the karyotype file `index` (load_exome_file,
 * search/io.c:26-81).  The occurrence bitmaps are re-laid out as this
 * engine's blocks; desp, C, C1 and the sample headers are checked against
 * them.  FMI_EIO when a file is missing, FMI_EFORMAT when the files are
 * inconsistent, use another alphabet than ACGT or a duplicated strand. */
int fmi_index_import_reference(const char *dir, fmi_index **out);
void fmi_index_free(fmi_index *ix);
/* Keeps the full suffix array and its inverse of both directions and the text
 * (2 bits per base) on the device, 16 bytes per base + n / 4 (3.1 Gbp:
 * ~50 GB), derived from the samples by LF walks on the GPU.  The 1-error
 * searches then continue every single-row interval by comparing the read
 * with the text at that row's suffix (one SA read to enter, one ISA read per
 * record) instead of one occurrence-block load per base; the records are
 * the same.  No reference counterpart (it trades HBM for dependent loads). */
int fmi_index_expand(fmi_index *ix);
/* A table of the interval of every string of k bases in each direction
 * (4^k entries of 8 bytes per direction: k = 12 takes 128 MiB, k = 13 512
 * MiB, k = 15 8 GiB), built on the device from the occurrence blocks.
 * fmi_search1_batch then takes the first k LF steps of each exact phase in
 * one read of the table; the records are the same.  k = 0 frees the table;
 * otherwise k is 4..15.  No reference counterpart (it trades HBM for the widest, most
 * expensive LF steps of every search). */
int fmi_index_kmer_table(fmi_index *ix, uint32_t k);
/* Copies one direction's table (4^k entries, k | l << 32, an empty
 * interval as (1, 0); entry index: base j of the string at bits 2j..2j+1,
 * in the order LF takes them) to the host; FMI_ESTATE without a table. */
int fmi_index_get_kmer_table(const fmi_index *ix, int forward, uint64_t *out);
/* Copies the expanded arrays of one direction (NULL skips one): sa, isa of
 * n + 1 entries; text (X, 16 bases per word, base j in bits 2(j%16)) of
 * (n + 15) / 16 words. */
int fmi_index_get_full(const fmi_index *ix, int forward, uint32_t *sa, uint32_t *isa,
                       uint32_t *text);
/* The sampled suffix array and inverse suffix array of one direction, as
 * the reference's S.vec / R.vec (Si / Ri for the forward direction) hold
 * them (calculate_S_and_R, search/preprocess.c:7-42): S[i] = SA[i * ratio],
 * R[j] = ISA[j * ratio], i, j < nsamp = ceil((n + 1) / ratio) entries each
 * (either pointer may be NULL). */
uint64_t fmi_index_sample_count(const fmi_index *ix);
int fmi_index_get_samples(const fmi_index *ix, int forward, uint32_t *S, uint32_t *R);
/* size_SA (search/runtime.h:86): n + 1. */
uint64_t fmi_index_size_sa(const fmi_index *ix);
uint32_t fmi_index_sa_ratio(const fmi_index *ix);
/* Bytes of the device copy by part: out[0] occurrence blocks and S / R
 * samples of both directions, out[1] the full SA / ISA and 2-bit text
 * (fmi_index_expand), out[2] the k-mer tables (fmi_index_kmer_table). */
int fmi_index_device_layout(const fmi_index *ix, uint64_t *out);
/* Bytes the device copy occupies. */
uint64_t fmi_index_device_bytes(const fmi_index *ix);
/* Copies the index into HBM of the current device (once). */
int fmi_index_to_device(fmi_index *ix);
/* The BWT of one direction in the reference's B.vec convention
 * (dbwt.c:732-763: n symbols with the '$' row left out, plus its row). */
int fmi_index_get_bwt(const fmi_index *ix, int forward, uint8_t *out_n_bytes, uint64_t *dollar);
/* Karyotype: contig c occupies text[offset[c] .. offset[c+1]) and starts at
 * coordinate start[c] of its chromosome (load_exome_file, search/io.c:26-80:
 * offset[c+1] = offset[c] + end[c] - start[c] + 1). */
int fmi_index_set_karyotype(fmi_index *ix, uint64_t count, const uint64_t *start,
                            const uint64_t *end);
/* Contig names (the `>name` of each karyotype line, search/io.c:47-56);
 * set_karyotype clears them. */
int fmi_index_set_contig_names(fmi_index *ix, uint64_t count, const char *const *names);
uint64_t fmi_index_contig_count(const fmi_index *ix);
/* Name of contig c ("" when unnamed), NULL past the end; valid while ix is. */
const char *fmi_index_contig_name(const fmi_index *ix, uint64_t c);

/* ------------------------------------------------------ device plumbing */
int fmi_device_count(int *count);
int fmi_device_alloc(void **d_ptr, size_t bytes);
int fmi_device_free(void *d_ptr);
int fmi_memcpy_h2d(void *d_dst, const void *src, size_t bytes, fmi_stream s);
int fmi_memcpy_d2h(void *dst, const void *d_src, size_t bytes, fmi_stream s);
/* On a stream that is being captured this records a fill kernel, not a
 * memset node. */
int fmi_memset_d(void *d_dst, int value, size_t bytes, fmi_stream s);
/* Name (as HIP reports it, NUL-terminated, at most size - 1 bytes) and
 * compute-unit count of the current device. */
int fmi_device_name(char *name, size_t size);
int fmi_device_cu_count(int *cus);
/* Free and total bytes of the current device (hipMemGetInfo). */
int fmi_device_mem_info(uint64_t *free_bytes, uint64_t *total_bytes);
int fmi_stream_create(fmi_stream *s);
/* Drains the stream, frees the work counters and scratch the library keeps
 * for it, then destroys it; FMI_ESTATE (nothing freed) while a graph
 * captured on it has not been destroyed. */
int fmi_stream_destroy(fmi_stream s);
int fmi_stream_sync(fmi_stream s);
/* Launch chains as HIP graphs: between fmi_graph_begin and fmi_graph_end
 * the batch calls made on stream s (one created with fmi_stream_create) are
 * recorded, not run; fmi_graph_launch replays the whole chain on a stream
 * with one call.  Each call must have run once on that stream before the
 * capture (it sizes the stream's work counters and scratch then), and a
 * replay reads and writes the buffers the calls named — their contents may
 * change between replays, not their addresses or sizes.  The library's
 * counter resets are kernels, never memset nodes.  While a graph captured on
 * a stream lives, a call on that stream that needs more scratch than the
 * stream holds returns FMI_ESTATE instead of moving it (the graph keeps the
 * address), and so does one that needs scratch during the capture.  No reference
 * counterpart (the reference launches synchronously, gpu.cu:118-203). */
typedef void *fmi_graph;
int fmi_graph_begin(fmi_stream s);
/* Orders stream waiter after the work issued so far on stream signaler
 * (both from fmi_stream_create): an event recorded on signaler, waited for
 * on waiter.  While signaler is being captured, a waiter that was not
 * capturing joins the capture (a fork: the calls made on it afterwards are
 * recorded into the same graph, and the graph holds its counters and
 * scratch as it holds the capturing stream's); it must be joined back with
 * fmi_stream_wait(capturing stream, waiter) before fmi_graph_end.  No
 * reference counterpart: it lets independent batch calls (say the three
 * error budgets of a mixed batch) run side by side in one graph. */
int fmi_stream_wait(fmi_stream waiter, fmi_stream signaler);
int fmi_graph_end(fmi_stream s, fmi_graph *g);
int fmi_graph_launch(fmi_graph g, fmi_stream s);
int fmi_graph_destroy(fmi_graph g);
/* Milliseconds the last kernel launched by the calling thread took on the
 * GPU (HIP events around that launch; valid after fmi_stream_sync). */
int fmi_last_kernel_ms(float *ms);
/* Timing region: between fmi_timing_begin and fmi_timing_end every search
 * kernel the calling thread launches is bracketed by its own pair of HIP
 * events on its stream; fmi_timing_end waits for them and returns the
 * number of launches and the sum / minimum / maximum of their durations. */
int fmi_timing_begin(void);
int fmi_timing_end(uint32_t *launches, double *sum_ms, float *min_ms, float *max_ms);

/* ------------------------------------------------------- exact search */
/* BWExactSearchBackward (direction 0, search/search.h:9-36) or
 * BWExactSearchForward (direction 1, search/search.h:38-65) of every whole
 * read on `view`, starting from [0, size_SA - 1]: the backward search
 * starts at pos = len-1, the forward one at pos = 0.  Per read, d_out gets the
 * final record (k, l, pos; start = 0, end = len - 1) and, when d_anchor is
 * not NULL, the anchor record (last non-empty interval).  This is the kernel
 * pair of gpu/gpu.cu:333-383 with the reference's per-read semantics.
 * Pairing for reverse-complement search: FORWARD direction on
 * BACKWARD_REV (bwt.c:1233-1237). */
int fmi_exact_batch(const fmi_index *ix, int view, int direction,
                    const uint8_t *d_reads, const uint64_t *d_offsets, uint32_t nreads,
                    fmi_hit *d_out, fmi_hit *d_anchor, fmi_stream s);
/* fmi_exact_batch (same arguments and results, synchronous on s) with the
 * launch's work counted, for the line model of the benchmark: work[0]
 * occurrence blocks read, [1] LF queries, [2] SA / ISA entries read, [3]
 * text compares (16 bases each, one text line), [4] k-mer table entries
 * read.  No reference counterpart. */
int fmi_exact_count_work(const fmi_index *ix, int view, int direction, const uint8_t *d_reads,
                         const uint64_t *d_offsets, uint32_t nreads, fmi_hit *d_out,
                         fmi_hit *d_anchor, fmi_stream s, uint64_t *work);

/* BWExactSearchVectorBackward / Forward (search/search.c:1407-1489) of every
 * whole read from [0, size_SA - 1]: d_vec_k/l[offsets[i] + j] is the interval
 * after the j-th step (same indexing as the reference's vec_k[j]);
 * d_last_good[i] is the return value. */
int fmi_exact_vector_batch(const fmi_index *ix, int view, int direction,
                           const uint8_t *d_reads, const uint64_t *d_offsets,
                           uint32_t nreads, int64_t *d_vec_k, int64_t *d_vec_l,
                           int16_t *d_last_good, fmi_stream s);

/* -------------------------------------------------- 1-error search */
#define FMI_MODE_EDIT 0      /* mismatches, insertions and deletions    */
#define FMI_MODE_MISMATCH 1  /* the {exact, MISMATCH} projection         */

#define FMI_STRAND_PLUS 1
#define FMI_STRAND_MINUS 2

/* BWSearch1CPU (search/search.c:1876-1933, with BWSimpleSearch1Backward/
 * Forward :1711-1874) on every read, for the strands in strand_mask: strand
 * 0 runs (backward, forward), strand 1 runs (forward_rev, backward_rev)
 * (bwt.c:1690-1701), each with res = {k=0, l=size_SA-1, start=0,
 * end=len-1}.  d_hits[(2*i + s) * cap ...] receives the records of read i,
 * strand s in the reference's order; d_counts[2*i + s] the total number of
 * records, which may exceed cap: then the first cap are stored, which is
 * the list the reference's results_list of capacity cap holds when its
 * add_result reports full (search/results.h:93-103).  MODE_MISMATCH keeps
 * only exact and MISMATCH records, in the same relative order.
 * Runs as one persistent launch with a device work queue. */
int fmi_search1_batch(const fmi_index *ix, int mode, int strand_mask,
                      const uint8_t *d_reads, const uint64_t *d_offsets, uint32_t nreads,
                      uint32_t cap, fmi_hit *d_hits, uint32_t *d_counts, fmi_stream s);

/* The anchor list BWSearch1CPU fills when it is given r_list_anchor
 * (search/search.c:1897-1925): per read i and strand s in strand_mask,
 * d_anchors[4*i + 2*s] is the anchor of pass 1 (BWExactSearchBackward of
 * W[len/2 .. len-1] on the pair's first index: dir 0, start = where the
 * extension stopped + 1, end = len - 1) and d_anchors[4*i + 2*s + 1] the
 * anchor of pass 2 (BWExactSearchForward of W[0 .. len/2 - 1] on the second:
 * dir 1, start 0, end = where it stopped - 1); each holds the last non-empty
 * interval (0, 0 when the first step already empties) and pos = the
 * position where the extension stopped.  Slots of strands not in the mask
 * are not written. */
int fmi_search1_anchors_batch(const fmi_index *ix, int strand_mask, const uint8_t *d_reads,
                              const uint64_t *d_offsets, uint32_t nreads, fmi_hit *d_anchors,
                              fmi_stream s);

/* fmi_search1_batch for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, records and
 * counts; the launch's work-queue counters live in a fixed 256-byte device
 * array of the library (one per GPU, held until the process ends) instead
 * of the per-stream workspace fmi_search1_batch allocates on a stream's
 * first call, and the call records no timing events.  The call calls no
 * hipMalloc / hipFree; every call on a GPU shares that GPU's array, so calls
 * that overlap on one GPU (on any streams) race: serialise them.  Its
 * scratch need does not depend on the batch (1 counter + 5 work counters);
 * the batch limits are fmi_search1_batch's. */
int fmi_search1_batch_noalloc(const fmi_index *ix, int mode, int strand_mask,
                              const uint8_t *d_reads, const uint64_t *d_offsets, uint32_t nreads,
                              uint32_t cap, fmi_hit *d_hits, uint32_t *d_counts, fmi_stream s);

/* The same results from the plain design of the reference's GPU glue
 * (gpu.cu:118-203): one thread per (read, strand), no work queue, one launch
 * per batch.  Kept as the comparison point for the persistent launch. */
int fmi_search1_batch_plain(const fmi_index *ix, int mode, int strand_mask,
                            const uint8_t *d_reads, const uint64_t *d_offsets,
                            uint32_t nreads, uint32_t cap, fmi_hit *d_hits,
                            uint32_t *d_counts, fmi_stream s);

/* The mapping layer's 1-error search: BWSearch1VectorHelper
 * (search/search.c:1491-1709) driven by the interval vectors of
 * BWExactSearchVectorBackward/Forward, per strand as __bwt_map_inexact_read
 * calls it (bwt.c:1940-1950, 1994-2003; + strand backward/forward, - strand
 * forward_rev/backward_rev).  d_bounds is NULL (the whole read, start 0) or
 * int16 [start, end] per read, relative to the read, as bwt_map_inexact_seed
 * passes a seed window (bwt.c:1447-1479).  Records and capacity as
 * fmi_search1_batch; windows shorter than 3 bases (or outside the read) and
 * reads longer than FMI_MAX_READ_LEN give no records (the reference reads
 * outside its vectors there; the mapping layer calls it for reads of 5 bases
 * or more).  Its scratch is a fixed slice per lane (at most 64 K lanes of
 * 8 KB), so the call neither reads the offsets on the host nor synchronises
 * and can be captured into a graph. */
int fmi_search1_vector_batch(const fmi_index *ix, int mode, int strand_mask,
                             const uint8_t *d_reads, const uint64_t *d_offsets,
                             const int16_t *d_bounds, uint32_t nreads, uint32_t cap,
                             fmi_hit *d_hits, uint32_t *d_counts, fmi_stream s);

/* Tuning of the persistent launch (the autotuner writes it): workgroups of
 * 256 lanes per CU, tasks a wavefront takes from the queue per atomic, and
 * the number of idle lanes (1..64) at which a wavefront refills them.  0
 * selects the default.  Results never depend on the tuning. */
int fmi_set_tuning(uint32_t wg_per_cu, uint32_t chunk, uint32_t refill_min);
int fmi_get_tuning(uint32_t *wg_per_cu, uint32_t *chunk, uint32_t *refill_min);
/* How the text-mode 1-error search issues an iteration's loads (the
 * autotuner writes it; results never depend on it): FMI_LOADS_ALL — every
 * lane issues the same batch, pointing the loads its step does not need at
 * data it reads anyway (fewer instructions); FMI_LOADS_OWN — each lane
 * issues only the loads its step needs (fewer memory requests, faster once
 * the index spreads over tens of GB); FMI_LOADS_AUTO (the default) — OWN for
 * an index of more than 16 GiB on the device, else ALL.  No reference
 * counterpart. */
#define FMI_LOADS_AUTO 0u
#define FMI_LOADS_ALL 1u
#define FMI_LOADS_OWN 2u
int fmi_set_load_mode(uint32_t mode);
int fmi_get_load_mode(uint32_t *mode);
/* Whether fmi_search1_batch stages reads (the default, 1): a streaming
 * kernel packs each read into a 96-byte record (2-bit codes, N flags) in
 * the stream's scratch (100 bytes per read, for batches of at most 2^27
 * reads) and every lane of the persistent search holds its read in LDS, so
 * no step reads it from memory again; reads longer than 256 bases are
 * searched afterwards without staging.  0 reads the read from memory at
 * every step.  Results never depend on it.  No reference counterpart. */
int fmi_set_staging(int enable);
int fmi_get_staging(int *enable);

/* ------------------------------------------------------------ locate */
/* get_SA (search/runtime.h:87, getScompValue search/csafm.h:208-235) for
 * every row of the intervals [d_k[i], d_l[i]] (empty intervals give no
 * rows), at most cap rows per interval starting at k: d_pos[i*cap + j] =
 * SA(k_i + j) of the BACKWARD (forward = 0) or FORWARD (forward = 1) index;
 * d_nrows[i] = min(l - k + 1, cap) (0 when empty). */
int fmi_locate_batch(const fmi_index *ix, int forward, const int64_t *d_k,
                     const int64_t *d_l, uint32_t nintervals, uint32_t cap,
                     uint64_t *d_pos, uint32_t *d_nrows, fmi_stream s);

/* The same rows for the interval of each record d_hits[i] (k, l of an
 * fmi_exact_batch or fmi_search1_batch record): the get_SA loop over [k, l]
 * that follows an exact search in bwt_map_exact_seq (bwt.c:927-952) and
 * bwt_map_exact_seed (bwt.c:1271-1277), without splitting the records into
 * k and l arrays first. */
int fmi_locate_hits(const fmi_index *ix, int forward, const fmi_hit *d_hits, uint32_t count,
                    uint32_t cap, uint64_t *d_pos, uint32_t *d_nrows, fmi_stream s);

/* get_ISA (search/runtime.h:88, getRcompValue search/csafm.h:237-272): the
 * BWT row of text position d_text_pos[i] (0..n) of the BACKWARD (forward =
 * 0) or FORWARD (forward = 1) index; positions beyond n give row UINT64_MAX. */
int fmi_inverse_locate_batch(const fmi_index *ix, int forward, const uint64_t *d_text_pos,
                             uint64_t count, uint64_t *d_rows, fmi_stream s);

/* encode_bases (bwt_commons.c:99-102) with the alphabet "ACGT": A/C/G/T in
 * either case become 0..3; every other byte becomes 4 (N; the reference
 * leaves those codes unset, bwt_commons.c:60-91).  Host function. */
int fmi_encode_bases(const char *src, uint64_t n, uint8_t *dst);

/* A read batch in the reference's fastq_batch_t layout
 * (bioformats/fastq/fastq_batch.h:24-34) on the device: d_seq holds the
 * sequences one after another, each followed by a NUL, and
 * d_data_indices[num_reads + 1] the start of each (read r has
 * data_indices[r+1] - data_indices[r] - 1 bases; indices increase by at
 * least one per read; the caller checks them).  Writes the engine's batch:
 * d_codes[offsets[r] ..] the codes of encode_bases (bwt_commons.c:99-102,
 * alphabet "ACGT" in either case; any other byte is N, 4) and
 * d_offsets[num_reads + 1] (offsets[r] = data_indices[r] - data_indices[0]
 * - r).  nbytes = data_indices[num_reads] - data_indices[0] (known to the
 * host that built the batch). */
int fmi_fastq_encode_batch(const char *d_seq, const int32_t *d_data_indices, uint32_t num_reads,
                           uint64_t nbytes, uint8_t *d_codes, uint64_t *d_offsets, fmi_stream s);
/* The codes of a batch packed 2 bits per base: base i of the concatenated
 * codes at bits 2(i % 16) of d_packed[i / 16], and (d_nmask may be NULL)
 * whether it is an N at bit i % 16 of d_nmask[i / 16] (an N packs as 0).
 * A quarter of the bytes for transfer or storage; no reference counterpart
 * (the reference keeps one byte per base, bwt_commons.c:99-102). */
int fmi_pack_codes(const uint8_t *d_codes, uint64_t nbases, uint32_t *d_packed,
                   uint16_t *d_nmask, fmi_stream s);

/* Position resolution of bwt.c:2158-2166 for 1-error records: for each hit
 * (as stored by fmi_search1_batch: hits[(2*i+s)*cap ...], counts) and each
 * SA row j of [k, l] (at most row_cap rows per hit), key =
 *   direction ? size_SA - SA_fwd(j) - len_calc - 1 : SA_bwd(j),
 * direction = strand ? !dir : dir, len_calc = len -1 for a DELETION and
 * len + 1 for an INSERTION; then idx = binsearch(offset, key) and the
 * mapping is kept when key + len_calc <= offset[idx], at coordinate
 * start[idx-1] + key - offset[idx-1].  Per stored hit h (in the flat order
 * of the stored records) d_chrom[h*row_cap + j] = idx - 1 or -1 when
 * rejected, d_coord[...] the coordinate. */
int fmi_resolve_hits(const fmi_index *ix, const fmi_hit *d_hits, const uint32_t *d_counts,
                     const uint64_t *d_offsets, uint32_t nreads, uint32_t cap,
                     uint32_t row_cap, int32_t *d_chrom, uint64_t *d_coord, fmi_stream s);
/* The same resolution with the length manage_single_result uses
 * (search/results.h:146-165): enW = end - start + 1 of the record, minus 1
 * for a DELETION, plus 1 for an INSERTION, in place of the read's length.
 * A record of the second pass of BWSearch1CPU carries start = half, so its
 * key moves by half against fmi_resolve_hits; these are the rows the
 * reference's text output prints (fmi_format_results). */
int fmi_resolve_hits_text(const fmi_index *ix, const fmi_hit *d_hits, const uint32_t *d_counts,
                          const uint64_t *d_offsets, uint32_t nreads, uint32_t cap,
                          uint32_t row_cap, int32_t *d_chrom, uint64_t *d_coord, fmi_stream s);

/* The reference's text lines for one (read, strand) record list
 * (write_results, search/results.c:3-43; manage_single_result,
 * search/results.h:125-187): records whose (k, l) repeats an earlier
 * record's are skipped; every other record's SA rows that lie inside a
 * contig give "read_<read_index> <+|-> <contig name> <coordinate>
 * <errors><mask> <read text>\n", the mask "_<pos>d", "_<pos>m<base>" or
 * "_<pos>i<base>".  hits are the list's records on the host (as
 * fmi_search1_batch stores them), chrom / coord their rows as
 * fmi_resolve_hits_text writes them (row_cap per record, chrom -1 for a row
 * that crosses a contig end), mapping the read text (its first nW characters are
 * printed).  Writes at most buf_size - 1 bytes and a NUL to buf (buf may be
 * NULL with buf_size 0), the full length to *len and the rows beyond
 * row_cap (not printed) to *rows_missing.  Host function. */
int fmi_format_results(const fmi_index *ix, uint64_t read_index, int strand, const char *mapping,
                       uint32_t nW, const fmi_hit *hits, uint32_t nhits, const int32_t *chrom,
                       const uint64_t *coord, uint32_t row_cap, char *buf, size_t buf_size,
                       size_t *len, uint64_t *rows_missing);

/* ------------------------------------------------------ seed mapping */
/* One region of bwt_map_exact_seed (bwt.c:1200-1294): an occurrence of a
 * seed on one strand that lies inside one contig. 32 bytes. */
typedef struct fmi_seed_region {
  uint64_t start;      /* start_mapping: start[idx-1] + key - offset[idx-1] + 1 (bwt.c:1271-1277) */
  uint64_t end;        /* start + len, len = seq_end - seq_start (bwt.c:1214, one short) */
  uint32_t chrom_idx;  /* idx as region_bwt_new receives it: binsearch result (contig + 1) */
  uint32_t read;       /* read index in the batch */
  uint16_t seq_start;  /* aux_seq_start: on the read (+) or on its reverse complement (-) */
  uint16_t seq_end;    /* aux_seq_end */
  uint16_t seed_id;    /* index of the seed in its read's tiling */
  uint8_t strand;      /* !type: 0 for +, 1 for - */
  uint8_t valid;       /* 0: the row's occurrence crosses a contig end (no region) */
} fmi_seed_region;

/* Per-seed summary: the exact intervals of both strands (+: backward search
 * on BACKWARD; -: forward search on BACKWARD_REV, bwt.c:1222-1238) and the
 * rows resolved for each after the cumulative filter (a strand is skipped
 * once the running count l - k + 1, summed over + then -, exceeds
 * filter_seed_mappings, bwt.c:1250-1269). 40 bytes. */
typedef struct fmi_seed_info {
  int64_t k_plus, l_plus, k_minus, l_minus;
  uint32_t rows_plus, rows_minus;
} fmi_seed_info;

/* bwt_map_exact_seed for every seed of every read: read i is tiled with
 * seeds [s, s + seed_len - 1] for s = 0, stride, 2*stride, ... while s +
 * seed_len <= len.  Seeds are numbered read after read, seed after seed;
 * d_seed_info (nseeds entries) is optional.  Every resolved SA row becomes
 * one entry of d_regions, in seed order, + strand rows then - strand rows,
 * each strand in row order (the reference's insertion order); entries past
 * region_cap are counted but not written.  *nseeds and *nrows are written
 * on the host: the call synchronises the stream (twice, for the seed and
 * row totals) and allocates its scratch (about 24 bytes per seed). */
int fmi_seed_map_batch(const fmi_index *ix, const uint8_t *d_reads, const uint64_t *d_offsets,
                       uint32_t nreads, uint32_t seed_len, uint32_t stride,
                       uint32_t filter_seed_mappings, fmi_seed_info *d_seed_info,
                       uint64_t seed_info_cap, fmi_seed_region *d_regions, uint64_t region_cap,
                       uint64_t *nseeds, uint64_t *nrows, fmi_stream s);

/* One region of bwt_map_inexact_seed (bwt.c:1401-1644; region_t,
 * bwt.h:186-198): an SA row of a 1-error record of a seed window that lies
 * inside one contig.  48 bytes. */
typedef struct fmi_inexact_region {
  uint64_t start;      /* start_mapping = start[idx-1] + key - offset[idx-1] + 1 (bwt.c:1573) */
  uint64_t end;        /* start + len - 1, len = seq_end - seq_start + 1 of the window */
  uint32_t chrom_idx;  /* idx as region_bwt_new receives it: binsearch result (contig + 1) */
  uint32_t read;       /* read index in the batch */
  uint16_t seq_start;  /* aux_seq_start: on the read (+) or its reverse complement (-) */
  uint16_t seq_end;    /* aux_seq_end */
  uint16_t seq_len;    /* read length */
  uint16_t seed_id;    /* window index in the read's tiling */
  int32_t pos_err;     /* final_pos (bwt.c:1508-1525; carried over where not assigned) */
  int32_t type_err;    /* final_err + seq_start: -1 mismatch, 1 SEED_INSERTION, 2 SEED_DELETION */
  uint8_t strand;      /* !type: 0 for +, 1 for - */
  uint8_t pad[7];
} fmi_inexact_region;

/* bwt_map_inexact_seeds_seq (bwt.c:3262-3302) over a batch: each read is
 * cut into windows of seed_size bases at 0, seed_size, ... plus the
 * remainder when it holds at least min_seed_size bases; each window runs
 * bwt_map_inexact_seed (bwt.c:1401-1644): BWSearch1VectorHelper on the +
 * strand (backward, forward) then the - strand (forward_rev, backward_rev),
 * one region per SA row of each record (all rows when l - k + 1 < 50, else
 * rows k and k + 1) whose occurrence lies inside one contig, the
 * within-10-bp dedup of each strand's list (bwt.c:1593-1615) and its
 * hand-over from last to first (bwt.c:1616-1626).  Window w of the batch
 * (reads in order, windows in read order) and strand s (0: +) own slot
 * t = 2 w + s: at most region_cap regions at d_regions[t * region_cap],
 * their count in d_nregions[t], and d_status[t] bit 0 when the strand had
 * more than rec_cap records (the reference keeps 20000), bit 1 when it had
 * more than region_cap regions before the dedup; a slot with a status bit
 * set is not the reference's list.  *nwindows gets the window count;
 * FMI_ERANGE (nothing written) when it exceeds windows_cap.  A window of
 * fewer than 3 bases has no records.  Synchronises the stream and
 * allocates its scratch (the window numbering and at most 1 GiB of interval
 * vectors). */
int fmi_map_inexact_seeds_batch(const fmi_index *ix, const uint8_t *d_reads,
                                const uint64_t *d_offsets, uint32_t nreads, uint32_t seed_size,
                                uint32_t min_seed_size, uint32_t rec_cap, uint32_t region_cap,
                                uint64_t windows_cap, fmi_inexact_region *d_regions,
                                uint32_t *d_nregions, uint8_t *d_status, uint64_t *nwindows,
                                fmi_stream s);

/* One alignment of __bwt_map_inexact_read (bwt.c:2158-2178): 24 bytes. */
typedef struct fmi_read_mapping {
  uint64_t position;    /* 0-based start[idx-1] + key - offset[idx-1] (bwt.c:2166) */
  uint32_t chromosome;  /* idx - 1 */
  uint32_t read;        /* read index in the batch */
  uint8_t strand;       /* !type: 0 for +, 1 for - */
  uint8_t secondary;    /* secondary_alignment as bwt.c:2217-2233 leaves it */
  uint16_t cigar_len;   /* the CIGAR is "<cigar_len>M", one operation */
  uint32_t pad;
} fmi_read_mapping;

/* __bwt_map_inexact_read (bwt.c:1890-2236) for every read, up to its list
 * of alignments: the error-free records of the read's vector-driven 1-error
 * search of each strand (+ then -; bwt.c:2038 skips every other record),
 * the running count of l - k over them checked against
 * filter_read_mappings before each record's rows (bwt.c:2145-2156: over it,
 * that strand's alignments are dropped and no later strand is searched),
 * one alignment per SA row inside a contig (bwt.c:2158-2178), then per
 * strand the dedup of alignments within 10 bp of an earlier one on the same
 * chromosome (the reference's int positions, bwt.c:2190-2215) and the
 * survivors appended from last to first, secondary unless first
 * (bwt.c:2217-2233).  Per read i: d_out[i * cap ..] the first cap
 * alignments, d_counts[i] their total, d_status[i] bit 0 when the filter
 * was exceeded, bit 1 when the read has fewer than 5 bases (the reference
 * makes one unmapped placeholder, bwt.c:1899-1915; none here), bit 2 when it
 * is longer than FMI_MAX_READ_LEN (not searched).  The anchor lists the
 * reference builds for reads left unmapped (bwt.c:2239-2269) are not made.
 * filter_read_mappings is 0..4096; the call allocates its scratch on the
 * stream (per lane, filter + 1 alignments of 16 bytes). */
int fmi_map_reads_batch(const fmi_index *ix, const uint8_t *d_reads, const uint64_t *d_offsets,
                        uint32_t nreads, uint32_t filter_read_mappings, uint32_t cap,
                        fmi_read_mapping *d_out, uint32_t *d_counts, uint8_t *d_status,
                        fmi_stream s);

/* ----------------------------------------------------------- serving */
/* A request loop over a device-resident index (no counterpart in the
 * reference, whose closest analog is the bounded work queue of
 * commons/workflow_scheduler.c:252-271 sized from free memory,
 * commons/system_utils.c:10-145).  A scheduler thread admits queued
 * requests first-come first-served into launches while the device arena
 * (device_budget_bytes) has room, coalesces up to max_batch_reads reads per
 * launch, keeps up to max_inflight launches in flight on separate streams
 * (requests that arrive meanwhile join the next launch as memory frees up),
 * compacts the stored records on the device and copies back only those;
 * fmi_server_wait / fmi_server_fetch hand each request its slice.
 * Each request gets exactly the records fmi_search1_batch gives for its reads
 * alone. */
typedef struct fmi_server fmi_server;

typedef struct fmi_server_config {
  uint64_t device_budget_bytes;  /* arena for in-flight launches (and as much pinned host) */
  uint32_t max_inflight;         /* launches in flight at once, 1..8          */
  uint32_t max_batch_reads;      /* reads coalesced into one launch           */
  uint32_t cap;                  /* records kept per (read, strand)           */
  int mode;                      /* FMI_MODE_EDIT or FMI_MODE_MISMATCH        */
  int strand_mask;               /* FMI_STRAND_PLUS | FMI_STRAND_MINUS        */
  uint32_t max_pending;          /* requests queued before submit refuses (0: 4096) */
  uint64_t staging_bytes;        /* pinned input staging (0: 2 x the budget)  */
} fmi_server_config;

typedef struct fmi_server_stats {
  uint64_t requests_done;
  uint64_t reads_done;
  uint64_t launches;
  uint64_t max_batch_reads_seen;  /* largest launch                         */
  uint32_t max_inflight_seen;     /* most launches in flight at once         */
  uint32_t pending;               /* requests waiting for admission now      */
  uint64_t refused;               /* submits refused with FMI_EAGAIN         */
  uint64_t cancelled;             /* requests cancelled before they finished */
} fmi_server_stats;

int fmi_server_create(const fmi_index *ix, const fmi_server_config *cfg, fmi_server **out);
/* Queues a request: reads/offsets as for fmi_search1_batch.  The reads and
 * offsets are copied, once, into the server's pinned staging before the
 * call returns (on the calling thread, so callers on several threads copy
 * in parallel).  FMI_ERANGE when the request alone exceeds the budget, the
 * staging or max_batch_reads (it is rejected, never queued); FMI_EAGAIN,
 * nothing queued, when max_pending requests are already queued or the
 * staging has no room now (backpressure: submit again after a wait). */
int fmi_server_submit(fmi_server *srv, const uint8_t *reads, const uint64_t *offsets,
                      uint32_t nreads, uint64_t *request_id);
/* The same request from a batch in the reference's fastq_batch_t layout
 * (see fmi_fastq_encode_batch; host pointers): seq and data_indices are
 * copied once into the staging, and the launch encodes them on the device. */
int fmi_server_submit_fastq(fmi_server *srv, const char *seq, const int32_t *data_indices,
                            uint32_t num_reads, uint64_t *request_id);
/* Cancels a request: one still queued is removed (it is never searched and
 * its id is released at once); one already on the GPU completes there and
 * its results are dropped; a finished one's results are released.  A
 * fmi_server_wait in progress on it returns FMI_ECANCELED; later calls with
 * the id return FMI_EINVAL. */
int fmi_server_cancel(fmi_server *srv, uint64_t request_id);
/* Waits up to timeout_us (< 0: forever) for a request; FMI_EAGAIN on time-out.
 * latency_ms: from submit to the moment its records reached host memory;
 * nrecords: the number of stored records (sum over its 2*nreads slots of
 * min(count, cap)). */
int fmi_server_wait(fmi_server *srv, uint64_t request_id, int64_t timeout_us,
                    double *latency_ms, uint64_t *nrecords);
/* Copies a finished request's results and releases them: counts[2*i + s]
 * as fmi_search1_batch gives, and the stored records of every (read,
 * strand) slot, slot after slot, densely (nrecords of them).  A finished
 * request's results stay in the server's memory (its launch's region of
 * the arena) until it is fetched or cancelled: fetch or cancel every
 * request, or the arena fills and admission waits. */
int fmi_server_fetch(fmi_server *srv, uint64_t request_id, uint32_t *counts, fmi_hit *records);
int fmi_server_get_stats(fmi_server *srv, fmi_server_stats *stats);
/* Finishes queued and in-flight work, then frees everything. */
int fmi_server_destroy(fmi_server *srv);

/* ------------------------------------------------- k-error search */
/* BWSearchCPU (search/search.c:1171-1405, helpers search.c:3-1169 and
 * search.h:66-198) on each strand of every read, as bwt_map_inexact_read_2
 * calls it (bwt.c:2430-2470; + strand backward/forward, - strand
 * forward_rev/backward_rev): read i is cut into len / d_fragsize[i]
 * fragments and searched with up to fragments - 1 errors (at most FMI_KMAX).
 * The four working lists and the final list hold list_cap records, the
 * anchor list anchor_cap (d_anchors may be NULL: no anchor list, as the
 * reference's NULL rl_anchor).  Per task t = 2 * read + strand:
 * d_final[t * list_cap ..] and d_nfinal[t] (the final list),
 * d_anchors[t * anchor_cap ..] and d_nanchors[t], d_full[t] = the
 * reference's return value (1: stopped on a full anchor list), or 2 when the
 * fragment size was out of range and the task was not searched.  Truncation
 * is the reference's: a full list ends the step that filled it. */
int fmi_search_k_batch(const fmi_index *ix, int strand_mask, const uint8_t *d_reads,
                       const uint64_t *d_offsets, const int16_t *d_fragsize, uint32_t nreads,
                       uint32_t list_cap, uint32_t anchor_cap, fmi_hitk *d_final,
                       uint32_t *d_nfinal, fmi_hitk *d_anchors, uint32_t *d_nanchors,
                       uint8_t *d_full, fmi_stream s);
/* fmi_search_k_batch (same arguments and results, synchronous on s) with
 * the launch's work counted, for the line model of the benchmark: work[0]
 * occurrence blocks read, [1] LF queries, [2] SA / ISA entries read, [3]
 * text compares (16 bases each, one text line), [4] k-mer table entries
 * read, [5] list records written or read (64 bytes each).  No reference
 * counterpart. */
int fmi_search_k_count_work(const fmi_index *ix, int strand_mask, const uint8_t *d_reads,
                            const uint64_t *d_offsets, const int16_t *d_fragsize, uint32_t nreads,
                            uint32_t list_cap, uint32_t anchor_cap, fmi_hitk *d_final,
                            uint32_t *d_nfinal, fmi_hitk *d_anchors, uint32_t *d_nanchors,
                            uint8_t *d_full, fmi_stream s, uint64_t *work);

/* ------------------------------------------- the plain comparison point */
/* The reference's GPU design (gpu/gpu.cu:118-203 glue, gpu.cu:333-451
 * kernels), re-created on this device and extended to the 1-error search:
 * the reference's occurrence layout (per-symbol desp[4][m] checkpoints and
 * count[4][m] bitmaps behind device pointer tables, search/csafm.h:51-81,
 * get_O csafm.h:166-181), reads padded to a 1000-byte stride, one thread per
 * read, one launch per strand, synchronous copies in and out.  Results equal
 * fmi_search1_batch's record for record.  Kept to measure the engine
 * against; it is not a product path. */
typedef struct fmi_plain fmi_plain;
/* Builds the reference layout on the device from an index on the device. */
int fmi_plain_create(const fmi_index *ix, fmi_plain **out);
/* One batch from HOST buffers: copy in, both launches, copy out.  counts and
 * hits as fmi_search1_batch (hits: nreads * 2 * cap records).  *ms_kernel is
 * the two launches, *ms_total the whole call. */
int fmi_plain_search1(fmi_plain *p, int mode, const uint8_t *reads, const uint64_t *offsets,
                      uint32_t nreads, uint32_t cap, fmi_hit *hits, uint32_t *counts,
                      double *ms_kernel, double *ms_total);
/* The layout of one direction: desp[b * words + w], count[b * words + w]
 * with words = ceil((n + 2) / 64) (calculate_O, search/preprocess.c:74-130). */
uint64_t fmi_plain_words(const fmi_plain *p);
int fmi_plain_get_layout(const fmi_plain *p, int forward, uint32_t *desp, uint64_t *count);
uint64_t fmi_plain_device_bytes(const fmi_plain *p);
int fmi_plain_free(fmi_plain *p);

/* ------------------------------------------------------ measurement */
/* Random gather of line_bytes-sized (16, 32, 64, 128 or 256) aligned lines
 * over d_buf (bytes >= 64 MiB), at least nlines loads: the random-access roofline of the path.
 * *ms is the kernel time and *lines_done the loads performed. */
int fmi_gather_bench(const void *d_buf, uint64_t bytes, uint32_t line_bytes,
                     uint64_t nlines, uint64_t *d_sink, float *ms, uint64_t *lines_done,
                     fmi_stream s);
/* The random-access roofline as a sweep point: 64-byte aligned lines at
 * random over d_buf (bytes >= 64 MiB), lines_in_flight (1, 2, 4, 8, 16)
 * loads outstanding per lane -- or, with cooperative, per quad of lanes that
 * load a 16-byte quarter of the same line each -- wg_per_cu workgroups of
 * 256 lanes per CU, at least nlines lines (*lines_done exactly).  *ms is the
 * kernel time.  No reference counterpart (measurement only). */
int fmi_gather_probe(const void *d_buf, uint64_t bytes, uint32_t lines_in_flight,
                     uint32_t wg_per_cu, int cooperative, uint64_t nlines, uint64_t *d_sink,
                     float *ms, uint64_t *lines_done, fmi_stream s);
/* Streaming copy of bytes (a multiple of 16) from d_src to d_dst: a kernel
 * whose HBM traffic is known (bytes read, bytes written), to calibrate
 * hardware fetch/write counters and to measure the streaming bandwidth.
 * *ms is the kernel time.  No reference counterpart (measurement only). */
int fmi_copy_bench(const void *d_src, void *d_dst, uint64_t bytes, float *ms, fmi_stream s);
/* Counts the distinct 64-byte occurrence blocks one fmi_search1_batch with
 * the same arguments fetches (the algorithmic bytes of the launch) and the
 * LF steps it takes. */
int fmi_search1_count_blocks(const fmi_index *ix, int mode, int strand_mask,
                             const uint8_t *d_reads, const uint64_t *d_offsets,
                             uint32_t nreads, uint64_t *blocks, uint64_t *lf_steps);
/* The same launch's whole work at capacity cap (work holds 5 counters):
 * work[0] occurrence blocks, work[1] LF queries, work[2] SA / ISA entries
 * read (text mode, including the ISA read of every stored record), work[3]
 * text runs (each compares up to 16 read bases with two 32-bit text words),
 * work[4] k-mer table lookups (8 bytes each). */
int fmi_search1_count_work(const fmi_index *ix, int mode, int strand_mask,
                           const uint8_t *d_reads, const uint64_t *d_offsets, uint32_t nreads,
                           uint32_t cap, uint64_t *work);

#ifdef __cplusplus
}
#endif
#endif

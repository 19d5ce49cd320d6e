// fmi_kernels.h -- host-side launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/fmi.h"
#include "fmi_internal.h"

namespace fmi {

// With work (5 counters), the counting instantiation: blocks, LF queries,
// SA / ISA entries, text compares, table entries.
hipError_t launch_exact(const DevIndex &ix, uint32_t dsel, uint32_t comp, uint32_t direction,
                        const uint8_t *reads, const uint64_t *offsets, uint32_t nreads,
                        fmi_hit *out, fmi_hit *anchor, hipStream_t s,
                        unsigned long long *work = nullptr);
hipError_t launch_vector(const DevIndex &ix, uint32_t dsel, uint32_t comp, uint32_t direction,
                         const uint8_t *reads, const uint64_t *offsets, uint32_t nreads,
                         int64_t *vk, int64_t *vl, int16_t *last_good, hipStream_t s);
// The persistent 1-error search.  With staged (device memory of
// search1_staged_scratch_bytes(nreads) bytes; persistent, not count) the
// reads are first packed into per-read records and every lane holds its
// read in LDS while it searches it; reads longer than kStagedMaxLen bases
// are searched afterwards by the unstaged kernel.  Records are the same
// either way.
constexpr uint32_t kStagedMaxLen = 256;
// A read of at most kStagedMaxLen bases packed for staging: 16 words of
// 2-bit codes (base j at bits 2(j % 16) of word j / 16) and 8 words of N
// flags (bit j % 32 of word 16 + j / 32).  launch_read_records writes one
// per read (kReadRecordWords words at rec + read * kReadRecordWords) and
// appends the longer reads to list (list_n, zeroed by the caller).
constexpr uint32_t kReadRecordWords = 24;
hipError_t launch_read_records(const uint8_t *reads, const uint64_t *offsets, uint32_t nreads,
                               uint32_t *rec, uint32_t *list, uint32_t *list_n, hipStream_t s);
constexpr uint32_t kStagedMaxTasks = 1u << 27;
uint64_t search1_staged_scratch_bytes(uint32_t nreads);
hipError_t launch_search1(const DevIndex &ix, bool mm_only, bool persistent, bool count,
                          uint32_t strand_mask, const uint8_t *reads, const uint64_t *offsets,
                          uint32_t nreads, uint32_t cap, fmi_hit *hits, uint32_t *counts,
                          uint32_t *queue, unsigned long long *stats, uint32_t grid_persistent,
                          uint32_t chunk, uint32_t refill_min, bool own_loads, hipStream_t s,
                          void *staged = nullptr);
hipError_t launch_zero(uint32_t *p, uint32_t n, hipStream_t s);
// Byte fill of bytes bytes with v (a kernel, so a captured graph holds no
// memset node).
hipError_t launch_fill(void *p, uint8_t v, uint64_t bytes, hipStream_t s);
// Frees the work counters and scratch the library keeps for stream s after
// draining it (fmi_api.cpp); FMI_ESTATE while a graph captured on s lives.
int release_stream_state(hipStream_t s);
// Dense copy of the stored records of nslots (read, strand) slots, slot after
// slot; *total (device) receives the record count.  scratch holds
// compact_scratch_bytes(nslots) bytes.
uint64_t compact_scratch_bytes(uint32_t nslots);
hipError_t launch_compact(const uint32_t *counts, const fmi_hit *hits, uint32_t nslots,
                          uint32_t cap, fmi_hit *out, unsigned long long *total, void *scratch,
                          hipStream_t s);
// BWSearch1VectorHelper on both strands of every read, `lanes` lanes (a
// multiple of 256, vhelper_lanes) striding over the tasks; scratch holds
// vhelper_scratch_bytes(lanes) bytes.
constexpr uint32_t kVhelperMaxLanes = 64u * 1024u;
uint32_t vhelper_lanes(uint32_t nreads);
uint64_t vhelper_scratch_bytes(uint32_t lanes);
hipError_t launch_vhelper(const DevIndex &ix, bool mm_only, uint32_t strand_mask,
                          const uint8_t *reads, const uint64_t *offsets, const int16_t *bounds,
                          uint32_t nreads, uint32_t cap, fmi_hit *hits, uint32_t *counts,
                          void *scratch, uint32_t lanes, hipStream_t s);
hipError_t launch_locate(const DevIndex &ix, uint32_t dsel, const int64_t *k, const int64_t *l,
                         uint32_t stride, uint32_t nintervals, uint32_t cap, uint64_t *pos,
                         uint32_t *nrows, hipStream_t s);
// Full SA / ISA of one direction from its samples (sa, isa: n + 1 entries).
hipError_t launch_expand(const DevDir &D, uint32_t n, uint32_t ratio, uint32_t *sa, uint32_t *isa,
                         hipStream_t s);
// X packed 2 bits per base ((n + 15) / 16 words) from the backward direction.
hipError_t launch_kmer_table(const DevDir &D, uint32_t n, uint32_t k, uint64_t *out, hipStream_t s);
hipError_t launch_pack_text(const DevDir &D, uint32_t n, const uint32_t *isa, uint32_t *text,
                            hipStream_t s);
hipError_t launch_isa(const DevIndex &ix, uint32_t dsel, const uint64_t *tpos, uint64_t count,
                      uint64_t *rows, hipStream_t s);
hipError_t launch_resolve(const DevIndex &ix, const fmi_hit *hits, const uint32_t *counts,
                          const uint64_t *offsets, uint32_t nreads, uint32_t cap,
                          uint32_t row_cap, const uint64_t *koff, const uint64_t *kstart,
                          uint32_t kcount, uint32_t record_len, int32_t *chrom, uint64_t *coord, hipStream_t s);
// BWSearchCPU per (read, strand) task (fmi_search_k.hip); scratch holds
// search_k_scratch_bytes(lanes, list_cap) (each lane's working lists and
// D), then, with front_lanes > 0 (at most kStagedMaxTasks reads),
// search_k_front_bytes(nreads) for the front kernel, which computes the
// first passes' exact searches and D of the reads of at most kStagedMaxLen
// bases beforehand on front_lanes lanes.  lanes and front_lanes are
// multiples of 256, best search_k_lanes_per_cu() / search_k_front_lanes_per_cu()
// x the CUs (all resident).
uint64_t search_k_lane_bytes(uint32_t list_cap);
uint64_t search_k_scratch_bytes(uint32_t lanes, uint32_t list_cap);
uint64_t search_k_front_bytes(uint32_t nreads);
uint32_t search_k_lanes_per_cu();
uint32_t search_k_front_lanes_per_cu();
hipError_t launch_search_k(const DevIndex &ix, const uint8_t *reads, const uint64_t *offsets,
                           const int16_t *fragsize, uint32_t nreads, uint32_t strand_mask,
                           uint32_t list_cap, uint32_t anchor_cap, fmi_hitk *fin, uint32_t *nfin,
                           fmi_hitk *anchors, uint32_t *nanchors, uint8_t *full, void *scratch,
                           uint32_t lanes, uint32_t front_lanes, uint32_t *counter, hipStream_t s,
                           unsigned long long *work = nullptr);
// The read-level mapping of __bwt_map_inexact_read (fmi_mapread.hip):
// `lanes` (map_reads_lanes) stride over the reads, each with a scratch slice
// of filter + 1 alignments (map_reads_scratch_bytes in all).
uint32_t map_reads_lanes(uint32_t nreads, uint32_t filter);
uint64_t map_reads_scratch_bytes(uint32_t lanes, uint32_t filter);
hipError_t launch_map_reads(const DevIndex &ix, const uint8_t *reads, const uint64_t *offsets,
                            uint32_t nreads, const uint64_t *koff, const uint64_t *kstart,
                            uint32_t kcount, uint32_t filter, uint32_t cap, fmi_read_mapping *out,
                            uint32_t *counts, uint8_t *status, void *scratch, uint32_t lanes,
                            hipStream_t s);
// fastq_batch_t ingestion (fmi_ingest.hip): nbytes = data_indices[nreads] -
// data_indices[0] bytes of seq become codes at base_off.. and offsets at
// read_off.. (values from 0); packing of nbases codes into 16-base words and
// 16-bit N masks.
hipError_t launch_fastq_encode(const char *seq, const int32_t *data_indices, uint32_t nreads,
                               uint64_t nbytes, uint64_t base_off, uint64_t read_off,
                               uint8_t *codes, uint64_t *offsets, hipStream_t s);
hipError_t launch_pack_codes(const uint8_t *codes, uint64_t nbases, uint32_t *packed,
                             uint16_t *nmask, hipStream_t s);
hipError_t launch_gather(const void *buf, uint64_t bytes, uint32_t line_bytes, uint64_t nloads,
                         unsigned long long *sink, uint32_t grid, hipStream_t s);
uint64_t gather_loads_done(uint64_t nloads, uint32_t grid);
// 64-byte random lines, ilp (1, 2, 4, 8, 16) in flight per lane (coop:
// per quad of lanes, each lane loading a quarter of the line); at least
// nlines lines, gather_probe_lines of them exactly.
hipError_t launch_gather_probe(const void *buf, uint64_t bytes, uint32_t ilp, bool coop,
                               uint64_t nlines, unsigned long long *sink, uint32_t grid,
                               hipStream_t s);
uint64_t gather_probe_lines(uint64_t nlines, uint32_t grid, uint32_t ilp, bool coop);
// bytes / 16 words copied from src to dst (both 16-byte aligned).
hipError_t launch_copy(const void *src, void *dst, uint64_t bytes, uint32_t grid, hipStream_t s);

}  // namespace fmi

// fmi_index.h -- the opaque fmi_index object behind the C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fmi.h"
#include "fmi_internal.h"

struct fmi_index {
  uint64_t n = 0;                 // text length (rows are 0..n)
  uint32_t ratio = 0;             // SA / ISA sample rate
  fmi::HostDir h[2];              // 0: BWT of X, 1: BWT of reverse(X)
  std::vector<uint64_t> kstart;   // karyotype: start coordinate of contig c
  std::vector<uint64_t> kend;     //            end coordinate of contig c
  std::vector<uint64_t> koffset;  //            text offset of contig c (count+1)
  std::vector<std::string> knames;  //          name of contig c (may be empty)

  // device copy
  bool on_device = false;
  int device = -1;
  void *d_mem = nullptr;          // one allocation for blocks + samples
  uint64_t *d_koffset = nullptr;  // karyotype offsets (count + 1)
  uint64_t *d_kstart = nullptr;
  uint64_t device_bytes = 0;
  void *d_full = nullptr;         // full SA / ISA of both directions + 2-bit text
  uint64_t full_bytes = 0;
  void *d_kmer = nullptr;         // k-mer interval tables of both directions
  uint64_t kmer_bytes = 0;
  fmi::DevIndex dev{};

  ~fmi_index();
};

namespace fmi {

// Builds both directions from codes 0..3; returns an fmi_status.
int build_index(const uint8_t *text, uint64_t n, uint32_t ratio, fmi_index *ix,
                std::string *err);
// The same on the current GPU; leaves the index on the device.
int build_index_gpu(const uint8_t *text, uint64_t n, uint32_t ratio, fmi_index *ix,
                    std::string *err);
// fmi_seed_map_batch (fmi_seed.hip); s is a hipStream_t.
int seed_map(const fmi_index *ix, const uint8_t *reads, const uint64_t *offsets, uint32_t nreads,
             uint32_t seed_len, uint32_t stride, uint32_t filter, fmi_seed_info *info,
             uint64_t info_cap, fmi_seed_region *regions, uint64_t region_cap, uint64_t *nseeds,
             uint64_t *nrows, void *s, std::string *err);
int map_inexact_seeds(const fmi_index *ix, const uint8_t *reads, const uint64_t *offsets,
                      uint32_t nreads, uint32_t seed_size, uint32_t min_seed_size,
                      uint32_t rec_cap, uint32_t region_cap, uint64_t windows_cap,
                      fmi_inexact_region *regions, uint32_t *nregions, uint8_t *status,
                      uint64_t *nwindows, void *s, std::string *err);
int format_results(const fmi_index *ix, uint64_t read_index, int strand, const char *mapping,
                   uint32_t nW, const fmi_hit *hits, uint32_t nhits, const int32_t *chrom,
                   const uint64_t *coord, uint32_t row_cap, std::string *out,
                   uint64_t *rows_missing);
int save_index(const fmi_index *ix, const char *path, std::string *err);
// Reads the reference's index directory (fmi_import.cpp).
int import_reference_dir(const char *dir, fmi_index *ix, std::string *err);
int load_index(const char *path, fmi_index *ix, std::string *err);
// The reference's GPU design on this device (fmi_refplain.hip).
struct RefPlain;
int refplain_create(const fmi_index *ix, RefPlain **out, std::string *err);
void refplain_free(RefPlain *rp);
int refplain_search1(RefPlain *rp, int mm_only, const uint8_t *reads, const uint64_t *offsets,
                     uint32_t nreads, uint32_t cap, fmi_hit *hits, uint32_t *counts,
                     double *ms_kernel, double *ms_total, std::string *err);
uint64_t refplain_device_bytes(const RefPlain *rp);
uint64_t refplain_words(const RefPlain *rp);
int refplain_layout(const RefPlain *rp, int forward, uint32_t *desp, uint64_t *count,
                    std::string *err);
// Symbol of BWT row r (0xFF for the '$' row).
uint32_t host_bwt_sym(const HostDir &d, uint64_t r);
// The 2-bit code stored at row r (the '$' row holds 0).
uint32_t host_bwt_sym_raw(const HostDir &d, uint64_t r);

}  // namespace fmi

// fmi_internal.h -- index layout shared by the builder, the HIP kernels and
// the C ABI of libhpgfm.so.
//
// Occurrence table: one 64-byte block per 192 BWT rows, so that an LF step
// touches one HBM line (two when k and l+1 fall in different blocks):
//
//   bytes  0..15  uint32 cnt[4]  occurrences of A,C,G,T in rows [0, 192*j)
//   bytes 16..39  uint64 lo[3]   low bit of the code of rows 192*j .. +191
//   bytes 40..63  uint64 hi[3]   high bit of the code
//                                row 192*j+t is bit t%64 of word t/64
//
// The two bit-planes make a rank three masked popcounts per plane word (one
// symbol: (lo^x) & (hi^y) & mask; all four: lo, hi, lo&hi).
//
// The '$' row is stored as code 0 (A) and corrected out of band: Occ(A, m)
// subtracts one when the '$' row lies in [192*j, m).  This replaces the
// reference's two per-symbol arrays (uint32 desp[4][m] checkpoints and
// uint64 count[4][m] bitmaps, search/csafm.h:51-81), whose get_O
// (csafm.h:166-181) needs two dependent loads from separate arrays per
// symbol per bound.
#pragma once

#include <cstdint>

namespace fmi {

constexpr uint32_t kBlockRows = 192;
constexpr uint32_t kBlockBytes = 64;
constexpr uint32_t kMaxReadLen = 1000;

struct alignas(64) OccBlock {
  uint32_t cnt[4];
  uint64_t lo[3];
  uint64_t hi[3];
};
static_assert(sizeof(OccBlock) == 64, "occurrence block must be one 64-byte line");

// One direction of the index on the device (passed by value to kernels).
struct DevDir {
  const OccBlock *blk;   // nblocks
  const uint32_t *S;     // S[i / ratio] = SA[i] for rows i % ratio == 0
  const uint32_t *R;     // R[j / ratio] = ISA[j] for text positions j % ratio == 0
  const uint32_t *sa;    // full SA (n + 1), or NULL: see fmi_index_expand
  const uint32_t *isa;   // full ISA (n + 1), or NULL
  uint32_t C[4];         // # symbols < b
  uint32_t dollar;       // row holding '$'
  uint32_t dollar_blk;   // dollar / 192
  uint32_t dollar_off;   // dollar % 192
  uint32_t pad;
};

struct DevIndex {
  DevDir d[2];           // 0: BWT of X ("backward"), 1: BWT of reverse X ("forward")
  const uint32_t *text;  // X, 2 bits per base, 16 bases per word (base j: word j/16,
                         // bits 2(j%16)..+1), or NULL; set with the full SA / ISA
  uint32_t n;            // text length; rows are 0..n
  uint32_t ratio;        // SA / ISA sample rate
  // Interval (k | l << 32) of every string of kmer_k symbols per direction
  // (index: symbol j fed to LF at bits 2j..2j+1), or NULL; kmer_k = 0
  // without a table.
  const uint64_t *kmer[2];
  uint32_t kmer_k;
  // The same for strings of kmer2_k = kmer_k - 4 symbols (kmer_k >= 8), or
  // NULL / 0: where a kmer_k-string's entry is empty or holds an N, the
  // searches that must find where their interval empties take the first
  // kmer2_k steps from it instead of one LF query each.
  const uint64_t *kmer2[2];
  uint32_t kmer2_k;
};

// Host-side view of one direction.
struct HostDir {
  uint64_t dollar = 0;
  uint32_t C[4] = {0, 0, 0, 0};
  OccBlock *blk = nullptr;
  uint64_t nblocks = 0;
  uint32_t *S = nullptr;
  uint32_t *R = nullptr;
  uint64_t nsamp = 0;
};

}  // namespace fmi

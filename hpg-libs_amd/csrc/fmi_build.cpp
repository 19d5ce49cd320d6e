// fmi_build.cpp -- index construction, file format, host-side helpers.
//
// The reference builds its index offline (bwt_generate_index_files,
// bwt.c:744-819): BWT by dbwt, C/C1 (preprocess.c:44-72), the per-symbol
// occurrence arrays (preprocess.c:74-130) and the S/R samples by an n-step
// LF walk (preprocess.c:7-42), for X and for reverse(X).  This builder
// computes the suffix array of each direction with SA-IS and writes the
// interleaved 64-byte occurrence blocks and the samples straight from it;
// the sampled values are the same (S[i/r] = SA[i], R[j/r] = ISA[j]).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "../../include/fmi.h"
#include "fmi_index.h"
#include "fmi_sais.h"

namespace fmi {

namespace {

template <typename T>
T *alloc_array(uint64_t count) {
  void *p = nullptr;
  if (count == 0) count = 1;
  if (posix_memalign(&p, 64, count * sizeof(T)) != 0) return nullptr;
  std::memset(p, 0, count * sizeof(T));
  return static_cast<T *>(p);
}

// Fills one direction from the suffix array SA[0..n) of T (the non-empty
// suffixes; row 0 of the BWT is the '$' suffix n).
template <typename I>
int fill_direction(const uint8_t *T, uint64_t n, const I *SA, uint32_t ratio, HostDir *d) {
  const uint64_t rows = n + 1;
  d->nblocks = (n + 2) / kBlockRows + 1;
  d->nsamp = rows / ratio + (rows % ratio ? 1 : 0);
  d->blk = alloc_array<OccBlock>(d->nblocks);
  d->S = alloc_array<uint32_t>(d->nsamp);
  d->R = alloc_array<uint32_t>(d->nsamp);
  if (!d->blk || !d->S || !d->R) return FMI_ENOMEM;

  uint64_t tot[4] = {0, 0, 0, 0};
  for (uint64_t i = 0; i < n; ++i) tot[T[i]]++;
  uint64_t acc = 0;
  for (int b = 0; b < 4; ++b) { d->C[b] = (uint32_t)acc; acc += tot[b]; }

  // Rows are split across threads at block boundaries: pass 1 counts each
  // slice, pass 2 writes blocks with the slice's starting counts.
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  const uint64_t nblk_rows = rows / kBlockRows + 1;
  const uint64_t per = (nblk_rows + nt - 1) / nt;
  std::vector<std::array<uint64_t, 4>> slice_cnt(nt);
  uint64_t dollar_row = 0;
  auto row_pos = [&](uint64_t r) -> uint64_t { return r == 0 ? n : (uint64_t)SA[r - 1]; };

  auto count_slice = [&](unsigned s) {
    std::array<uint64_t, 4> c{0, 0, 0, 0};
    uint64_t r0 = s * per * kBlockRows, r1 = std::min<uint64_t>(rows, (s + 1) * per * kBlockRows);
    for (uint64_t r = r0; r < r1; ++r) {
      uint64_t p = row_pos(r);
      if (p > 0) c[T[p - 1]]++;
    }
    slice_cnt[s] = c;
  };
  auto fill_slice = [&](unsigned s, std::array<uint64_t, 4> run) {
    uint64_t b0 = s * per, b1 = std::min<uint64_t>(nblk_rows, (s + 1) * per);
    for (uint64_t j = b0; j < b1; ++j) {
      OccBlock &B = d->blk[j];
      for (int b = 0; b < 4; ++b) B.cnt[b] = (uint32_t)run[b];
      for (int w = 0; w < 3; ++w) { B.lo[w] = 0; B.hi[w] = 0; }
      uint64_t r0 = j * kBlockRows, r1 = std::min<uint64_t>(rows, r0 + kBlockRows);
      for (uint64_t r = r0; r < r1; ++r) {
        uint64_t p = row_pos(r);
        uint32_t t = (uint32_t)(r - r0);
        if (r % ratio == 0) d->S[r / ratio] = (uint32_t)p;
        if (p % ratio == 0) d->R[p / ratio] = (uint32_t)r;
        if (p == 0) { dollar_row = r; continue; }   // '$' kept as code 0, never counted
        uint32_t sym = T[p - 1];
        B.lo[t >> 6] |= (uint64_t)(sym & 1u) << (t & 63);
        B.hi[t >> 6] |= (uint64_t)(sym >> 1) << (t & 63);
        run[sym]++;
      }
    }
  };

  std::vector<std::thread> th;
  for (unsigned s = 0; s < nt; ++s) th.emplace_back(count_slice, s);
  for (auto &x : th) x.join();
  th.clear();
  std::array<uint64_t, 4> run{0, 0, 0, 0};
  std::vector<std::array<uint64_t, 4>> starts(nt);
  for (unsigned s = 0; s < nt; ++s) {
    starts[s] = run;
    for (int b = 0; b < 4; ++b) run[b] += slice_cnt[s][b];
  }
  for (unsigned s = 0; s < nt; ++s) th.emplace_back(fill_slice, s, starts[s]);
  for (auto &x : th) x.join();
  // Blocks past the last row carry the totals.
  for (uint64_t j = nblk_rows; j < d->nblocks; ++j)
    for (int b = 0; b < 4; ++b) d->blk[j].cnt[b] = (uint32_t)tot[b];
  d->dollar = dollar_row;
  return FMI_OK;
}

template <typename I>
int build_direction(const uint8_t *T, uint64_t n, uint32_t ratio, HostDir *d) {
  std::vector<I> SA;
  try {
    SA.resize(n ? n : 1);
  } catch (const std::bad_alloc &) {
    return FMI_ENOMEM;
  }
  if (n) suffix_array<I>(T, SA.data(), (I)n, (I)4);
  return fill_direction<I>(T, n, SA.data(), ratio, d);
}

int build_direction_any(const uint8_t *T, uint64_t n, uint32_t ratio, HostDir *d) {
  if (n < (uint64_t)INT32_MAX - 8) return build_direction<int32_t>(T, n, ratio, d);
  return build_direction<int64_t>(T, n, ratio, d);
}

void free_dir(HostDir &d) {
  free(d.blk);
  free(d.S);
  free(d.R);
  d = HostDir{};
}

}  // namespace

uint32_t host_bwt_sym_raw(const HostDir &d, uint64_t r) {
  const OccBlock &B = d.blk[r / kBlockRows];
  uint32_t t = (uint32_t)(r % kBlockRows);
  return (uint32_t)((B.lo[t >> 6] >> (t & 63)) & 1u) | ((uint32_t)((B.hi[t >> 6] >> (t & 63)) & 1u) << 1);
}

uint32_t host_bwt_sym(const HostDir &d, uint64_t r) {
  if (r == d.dollar) return 0xFFu;
  const OccBlock &B = d.blk[r / kBlockRows];
  uint32_t t = (uint32_t)(r % kBlockRows);
  return (uint32_t)((B.lo[t >> 6] >> (t & 63)) & 1u) | ((uint32_t)((B.hi[t >> 6] >> (t & 63)) & 1u) << 1);
}

int build_index(const uint8_t *text, uint64_t n, uint32_t ratio, fmi_index *ix, std::string *err) {
  if (!text && n) { *err = "text is NULL"; return FMI_EINVAL; }
  if (ratio == 0) { *err = "sa_ratio must be >= 1"; return FMI_EINVAL; }
  if (n + 3 >= (uint64_t)UINT32_MAX) { *err = "text longer than 2^32 - 4 symbols"; return FMI_ERANGE; }
  for (uint64_t i = 0; i < n; ++i)
    if (text[i] > 3) { *err = "text holds a code > 3 (replace N before indexing)"; return FMI_EINVAL; }
  std::vector<uint8_t> rev;
  try {
    rev.resize(n ? n : 1);
  } catch (const std::bad_alloc &) {
    *err = "out of host memory";
    return FMI_ENOMEM;
  }
  for (uint64_t i = 0; i < n; ++i) rev[i] = text[n - 1 - i];   // Xi (bwt.c:791-793)
  ix->n = n;
  ix->ratio = ratio;
  int rc1 = FMI_OK;
  std::thread t1([&] { rc1 = build_direction_any(rev.data(), n, ratio, &ix->h[1]); });
  int rc0 = build_direction_any(text, n, ratio, &ix->h[0]);
  t1.join();
  if (rc0 != FMI_OK || rc1 != FMI_OK) {
    *err = "out of host memory while building the index";
    return FMI_ENOMEM;
  }
  // One contig spanning the text until a karyotype is set.
  ix->kstart.assign(1, 0);
  ix->kend.assign(1, n ? n - 1 : 0);
  ix->koffset = {0, n};
  ix->knames.assign(1, std::string());
  return FMI_OK;
}

// ---------------------------------------------------------------- file format
// Little-endian single file:
//   char magic[8] = "HPGFMI01"; uint32 version = 2; uint32 block_rows = 192;
//   uint64 n; uint32 ratio; uint32 reserved;
//   per direction: uint64 dollar; uint32 C[4]; uint64 nblocks; uint64 nsamp;
//   uint64 kcount; uint64 kstart[kcount]; uint64 kend[kcount];
//   (version 3) per contig: uint32 length; char name[length];
//   per direction: OccBlock blk[nblocks]; uint32 S[nsamp]; uint32 R[nsamp];
//   uint64 checksum (FNV-1a 64 over everything before it).
namespace {

const char kMagic[8] = {'H', 'P', 'G', 'F', 'M', 'I', '0', '1'};
// 2: bit-plane occurrence blocks (1 held 2-bit packed codes); 3: + contig names
const uint32_t kFormatVersion = 3;

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void *p, size_t n) {
    const uint8_t *b = static_cast<const uint8_t *>(p);
    for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
  }
};

struct Writer {
  FILE *f;
  Fnv sum;
  bool ok = true;
  void put(const void *p, size_t n) {
    if (!ok) return;
    if (n && fwrite(p, 1, n, f) != n) ok = false;
    sum.add(p, n);
  }
};

struct Reader {
  FILE *f;
  Fnv sum;
  bool ok = true;
  void get(void *p, size_t n) {
    if (!ok) return;
    if (n && fread(p, 1, n, f) != n) { ok = false; return; }
    sum.add(p, n);
  }
};

// The contents of one direction read from a file: every block's counts
// follow from the previous block's and its rows, the totals agree with C,
// the '$' row is stored as code 0, and every S / R sample names a row or a
// text position of the index (0..n).
bool check_direction(const HostDir &h, uint64_t n, uint32_t ratio, std::string *err) {
  const uint64_t rows = n + 1;
  uint64_t run[4] = {0, 0, 0, 0};
  if (host_bwt_sym_raw(h, h.dollar) != 0) {
    *err = "index '$' row is not stored as code 0";
    return false;
  }
  for (uint64_t j = 0; j < h.nblocks; ++j) {
    const OccBlock &B = h.blk[j];
    for (int b = 0; b < 4; ++b) {
      if (B.cnt[b] != run[b]) {
        *err = "index occurrence counts are inconsistent at block " + std::to_string(j);
        return false;
      }
    }
    const uint64_t r0 = j * kBlockRows;
    if (r0 >= rows) continue;
    const uint32_t m = (uint32_t)std::min<uint64_t>(kBlockRows, rows - r0);
    for (uint32_t w = 0; w < 3; ++w) {
      const uint32_t lo_t = w * 64;
      if (lo_t >= m) break;
      const uint32_t nb = std::min<uint32_t>(64, m - lo_t);
      const uint64_t mask = nb == 64 ? ~0ull : ((1ull << nb) - 1);
      const uint64_t lo = B.lo[w] & mask, hi = B.hi[w] & mask;
      run[1] += (uint64_t)__builtin_popcountll(lo & ~hi);
      run[2] += (uint64_t)__builtin_popcountll(~lo & hi & mask);
      run[3] += (uint64_t)__builtin_popcountll(lo & hi);
      run[0] += (uint64_t)__builtin_popcountll(~lo & ~hi & mask);
    }
    if (h.dollar >= r0 && h.dollar < r0 + m) run[0]--;   // the '$' row counts for nothing
  }
  uint64_t acc = 0;
  for (int b = 0; b < 4; ++b) {
    if (h.C[b] != acc) {
      *err = "index C does not match the occurrence counts";
      return false;
    }
    acc += run[b];
  }
  if (acc != n) {
    *err = "index occurrence counts do not add up to the text length";
    return false;
  }
  for (uint64_t i = 0; i < h.nsamp; ++i) {
    if (h.S[i] > n || h.R[i] > n) {
      *err = "index S / R sample " + std::to_string(i) + " lies outside 0..n";
      return false;
    }
  }
  (void)ratio;
  return true;
}

}  // namespace

int save_index(const fmi_index *ix, const char *path, std::string *err) {
  FILE *f = fopen(path, "wb");
  if (!f) { *err = std::string("cannot open ") + path; return FMI_EIO; }
  Writer w{f};
  uint32_t version = kFormatVersion, brows = kBlockRows, reserved = 0;
  w.put(kMagic, 8);
  w.put(&version, 4);
  w.put(&brows, 4);
  w.put(&ix->n, 8);
  w.put(&ix->ratio, 4);
  w.put(&reserved, 4);
  for (int d = 0; d < 2; ++d) {
    w.put(&ix->h[d].dollar, 8);
    w.put(ix->h[d].C, 16);
    w.put(&ix->h[d].nblocks, 8);
    w.put(&ix->h[d].nsamp, 8);
  }
  uint64_t kc = ix->kstart.size();
  w.put(&kc, 8);
  w.put(ix->kstart.data(), kc * 8);
  w.put(ix->kend.data(), kc * 8);
  for (uint64_t c = 0; c < kc; ++c) {
    const std::string nm = c < ix->knames.size() ? ix->knames[c] : std::string();
    uint32_t len = (uint32_t)nm.size();
    w.put(&len, 4);
    w.put(nm.data(), len);
  }
  for (int d = 0; d < 2; ++d) {
    w.put(ix->h[d].blk, ix->h[d].nblocks * sizeof(OccBlock));
    w.put(ix->h[d].S, ix->h[d].nsamp * 4);
    w.put(ix->h[d].R, ix->h[d].nsamp * 4);
  }
  uint64_t cs = w.sum.h;
  bool ok = w.ok && fwrite(&cs, 8, 1, f) == 1;
  ok = (fclose(f) == 0) && ok;
  if (!ok) { *err = std::string("write failed: ") + path; return FMI_EIO; }
  return FMI_OK;
}

int load_index(const char *path, fmi_index *ix, std::string *err) {
  FILE *f = fopen(path, "rb");
  if (!f) { *err = std::string("cannot open ") + path; return FMI_EIO; }
  Reader r{f};
  char magic[8];
  uint32_t version = 0, brows = 0, reserved = 0;
  r.get(magic, 8);
  r.get(&version, 4);
  r.get(&brows, 4);
  if (!r.ok || std::memcmp(magic, kMagic, 8) != 0 || (version != kFormatVersion && version != 2) ||
      brows != kBlockRows) {
    fclose(f);
    *err = "not an index file of this version";
    return FMI_EFORMAT;
  }
  r.get(&ix->n, 8);
  r.get(&ix->ratio, 4);
  r.get(&reserved, 4);
  for (int d = 0; d < 2; ++d) {
    r.get(&ix->h[d].dollar, 8);
    r.get(ix->h[d].C, 16);
    r.get(&ix->h[d].nblocks, 8);
    r.get(&ix->h[d].nsamp, 8);
  }
  uint64_t kc = 0;
  r.get(&kc, 8);
  const uint64_t rows = ix->n + 1;
  bool sane = r.ok && ix->ratio > 0 && ix->n + 3 < (uint64_t)UINT32_MAX && kc < (1ull << 24);
  for (int d = 0; sane && d < 2; ++d) {
    sane = ix->h[d].nblocks == (ix->n + 2) / kBlockRows + 1 &&
           ix->h[d].nsamp == rows / ix->ratio + (rows % ix->ratio ? 1 : 0) &&
           ix->h[d].dollar <= ix->n;
  }
  if (!sane) { fclose(f); *err = "index header is inconsistent"; return FMI_EFORMAT; }
  ix->kstart.resize(kc);
  ix->kend.resize(kc);
  r.get(ix->kstart.data(), kc * 8);
  r.get(ix->kend.data(), kc * 8);
  ix->knames.assign(kc, std::string());
  for (uint64_t c = 0; version >= 3 && c < kc && r.ok; ++c) {
    uint32_t len = 0;
    r.get(&len, 4);
    if (!r.ok || len > 4096) { fclose(f); *err = "index contig names are inconsistent"; return FMI_EFORMAT; }
    ix->knames[c].resize(len);
    r.get(&ix->knames[c][0], len);
  }
  for (int d = 0; d < 2; ++d) {
    ix->h[d].blk = alloc_array<OccBlock>(ix->h[d].nblocks);
    ix->h[d].S = alloc_array<uint32_t>(ix->h[d].nsamp);
    ix->h[d].R = alloc_array<uint32_t>(ix->h[d].nsamp);
    if (!ix->h[d].blk || !ix->h[d].S || !ix->h[d].R) { fclose(f); *err = "out of host memory"; return FMI_ENOMEM; }
    r.get(ix->h[d].blk, ix->h[d].nblocks * sizeof(OccBlock));
    r.get(ix->h[d].S, ix->h[d].nsamp * 4);
    r.get(ix->h[d].R, ix->h[d].nsamp * 4);
  }
  uint64_t want = r.sum.h, got = 0;
  bool ok = r.ok && fread(&got, 8, 1, f) == 1;
  fclose(f);
  if (!ok) { *err = "index file is truncated"; return FMI_EFORMAT; }
  if (got != want) { *err = "index checksum mismatch"; return FMI_EFORMAT; }
  // A file with a valid checksum can still be wrong inside; the kernels
  // index with these values, so check every one they trust.
  for (int d = 0; d < 2; ++d) {
    if (!check_direction(ix->h[d], ix->n, ix->ratio, err)) return FMI_EFORMAT;
  }
  ix->koffset.assign(kc + 1, 0);
  for (uint64_t c = 0; c < kc; ++c) {
    if (ix->kend[c] + 1 < ix->kstart[c] || ix->kend[c] - ix->kstart[c] + 1 > ix->n) {
      *err = "index karyotype has a contig that ends before it starts";
      return FMI_EFORMAT;
    }
    ix->koffset[c + 1] = ix->koffset[c] + (ix->kend[c] - ix->kstart[c] + 1);
  }
  if (kc && ix->koffset[kc] != ix->n) {
    *err = "index karyotype does not add up to the text length";
    return FMI_EFORMAT;
  }
  return FMI_OK;
}

}  // namespace fmi

fmi_index::~fmi_index() {
  fmi::free_dir(h[0]);
  fmi::free_dir(h[1]);
}

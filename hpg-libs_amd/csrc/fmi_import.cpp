// fmi_import.cpp -- reads an index directory written by the reference's
// builder (bwt_generate_index_files, bwt.c:744-819) into this engine's
// index: the same BWT rows, C and S / R samples, re-laid out as interleaved
// occurrence blocks.
//
// Files read (all little-endian, SA_TYPE = uint32, FM_COMP_64):
//   config.txt            alphabet line, duplicate-strand line (io.c:317-334)
//   C.vec, C1.vec         uint32 n; uint32 v[n]                (csafm.c:121-146)
//   O.desp / Oi.desp      uint32 siz, n_desp, m_desp; uint32 desp[n_desp][m_desp]
//   O.count / Oi.count    uint32 n_count, m_count; uint64 count[n_count][m_count]
//                                                              (csafm.c:182-246)
//   S/R.vec, Si/Ri.vec    uint32 siz, n, ratio; uint32 v[n]    (csafm.c:148-180)
//   index                 ">name start end" per contig          (io.c:26-81)
// Row r of a BWT is bit (r + 1) % 64 of word (r + 1) / 64 of its symbol's
// count array; the '$' row has no bit (preprocess.c:74-130).  desp[b][w]
// holds the number of b in rows [0, 64 w - 1) and is checked against the
// bitmaps, as are C / C1 and the sample headers, so a corrupt or foreign
// directory is rejected instead of producing wrong intervals.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fmi.h"
#include "fmi_index.h"

namespace fmi {

namespace {

struct File {
  FILE *f = nullptr;
  explicit File(const std::string &p) : f(fopen(p.c_str(), "rb")) {}
  ~File() {
    if (f) fclose(f);
  }
  bool read(void *p, size_t n) { return n == 0 || fread(p, 1, n, f) == n; }
  bool at_end() {
    int c = fgetc(f);
    return c == EOF;
  }
};

int open_err(const std::string &path, std::string *err) {
  *err = "cannot open " + path;
  return FMI_EIO;
}

int fmt_err(const std::string &what, std::string *err) {
  *err = what;
  return FMI_EFORMAT;
}

// read_vector: uint32 n, then n values.
int read_vec(const std::string &dir, const char *name, std::vector<uint32_t> &v, std::string *err) {
  std::string p = dir + "/" + name + ".vec";
  File f(p);
  if (!f.f) return open_err(p, err);
  uint32_t n = 0;
  if (!f.read(&n, 4) || n > (1u << 20)) return fmt_err(p + ": bad header", err);
  v.resize(n);
  if (!f.read(v.data(), (size_t)n * 4) || !f.at_end()) return fmt_err(p + ": wrong length", err);
  return FMI_OK;
}

// read_comp_vector: uint32 siz, n, ratio, then n values.
int read_comp_vec(const std::string &dir, const char *name, uint64_t rows, uint32_t *ratio,
                  uint32_t **out, uint64_t *nout, std::string *err) {
  std::string p = dir + "/" + name + ".vec";
  File f(p);
  if (!f.f) return open_err(p, err);
  uint32_t hdr[3];
  if (!f.read(hdr, 12)) return fmt_err(p + ": bad header", err);
  const uint32_t siz = hdr[0], n = hdr[1], r = hdr[2];
  if (siz != rows || r == 0 || n != rows / r + (rows % r ? 1 : 0))
    return fmt_err(p + ": header does not fit the index", err);
  if (*ratio && r != *ratio) return fmt_err(p + ": sample ratio differs between files", err);
  *ratio = r;
  void *mem = nullptr;
  if (posix_memalign(&mem, 64, (size_t)(n ? n : 1) * 4) != 0) {
    *err = "out of host memory";
    return FMI_ENOMEM;
  }
  *out = static_cast<uint32_t *>(mem);
  *nout = n;
  if (!f.read(*out, (size_t)n * 4) || !f.at_end()) return fmt_err(p + ": wrong length", err);
  return FMI_OK;
}

// One direction: the occurrence matrix O / Oi (desp + count) to blocks.
int read_direction(const std::string &dir, const char *oname, const uint32_t *C, uint64_t n,
                   HostDir *d, std::string *err) {
  const uint64_t rows = n + 1, siz = n + 2;
  const uint64_t words = siz / 64 + (siz % 64 ? 1 : 0);
  if (words > (1ull << 32)) return fmt_err("index too large", err);
  std::vector<std::vector<uint32_t>> desp(4);
  std::vector<std::vector<uint64_t>> cnt(4);
  {
    std::string p = dir + "/" + oname + ".desp";
    File f(p);
    if (!f.f) return open_err(p, err);
    uint32_t hdr[3];
    if (!f.read(hdr, 12)) return fmt_err(p + ": bad header", err);
    if (hdr[0] != siz || hdr[1] != 4 || hdr[2] != words)
      return fmt_err(p + ": header does not fit the index (4 symbols, n + 2 rows)", err);
    for (int b = 0; b < 4; ++b) {
      desp[b].resize(words);
      if (!f.read(desp[b].data(), words * 4)) return fmt_err(p + ": truncated", err);
    }
    if (!f.at_end()) return fmt_err(p + ": trailing bytes", err);
  }
  {
    std::string p = dir + "/" + oname + ".count";
    File f(p);
    if (!f.f) return open_err(p, err);
    uint32_t hdr[2];
    if (!f.read(hdr, 8)) return fmt_err(p + ": bad header", err);
    if (hdr[0] != 4 || hdr[1] != words) return fmt_err(p + ": header does not fit the index", err);
    for (int b = 0; b < 4; ++b) {
      cnt[b].resize(words);
      if (!f.read(cnt[b].data(), words * 8)) return fmt_err(p + ": truncated", err);
    }
    if (!f.at_end()) return fmt_err(p + ": trailing bytes", err);
  }

  d->nblocks = (n + 2) / kBlockRows + 1;
  void *mem = nullptr;
  if (posix_memalign(&mem, 64, d->nblocks * sizeof(OccBlock)) != 0) {
    *err = "out of host memory";
    return FMI_ENOMEM;
  }
  d->blk = static_cast<OccBlock *>(mem);
  std::memset(d->blk, 0, d->nblocks * sizeof(OccBlock));

  // Symbol of row r: the one count array holding bit r + 1; 4 = none ('$').
  auto sym_of = [&](uint64_t r, int *nset) -> uint32_t {
    const uint64_t g = r + 1, w = g >> 6;
    const uint64_t m = 1ull << (g & 63);
    uint32_t s = 4;
    int c = 0;
    for (uint32_t b = 0; b < 4; ++b)
      if (cnt[b][w] & m) { s = b; ++c; }
    *nset = c;
    return s;
  };

  // Blocks are filled in slices of whole blocks: pass 1 counts each slice,
  // pass 2 writes the blocks and checks desp from the slice's start counts.
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  const uint64_t nblk_rows = rows / kBlockRows + 1;
  const uint64_t per = (nblk_rows + nt - 1) / nt;
  std::vector<std::array<uint64_t, 5>> slice_cnt(nt);
  std::vector<int> bad(nt, 0);
  auto count_slice = [&](unsigned s) {
    std::array<uint64_t, 5> c{0, 0, 0, 0, 0};
    const uint64_t r0 = std::min<uint64_t>(rows, s * per * kBlockRows);
    const uint64_t r1 = std::min<uint64_t>(rows, (s + 1) * per * kBlockRows);
    for (uint64_t r = r0; r < r1; ++r) {
      int k;
      const uint32_t sym = sym_of(r, &k);
      if (k > 1) { bad[s] = 1; return; }
      c[sym]++;
    }
    slice_cnt[s] = c;
  };
  std::vector<uint64_t> dollar_of(nt, UINT64_MAX);
  auto fill_slice = [&](unsigned s, std::array<uint64_t, 5> run) {
    const uint64_t b0 = s * per, b1 = std::min<uint64_t>(nblk_rows, (s + 1) * per);
    for (uint64_t j = b0; j < b1; ++j) {
      OccBlock &B = d->blk[j];
      for (int b = 0; b < 4; ++b) B.cnt[b] = (uint32_t)run[b];
      const uint64_t r0 = j * kBlockRows, r1 = std::min<uint64_t>(rows, r0 + kBlockRows);
      for (uint64_t r = r0; r < r1; ++r) {
        if (((r + 1) & 63) == 0) {
          const uint64_t w = (r + 1) >> 6;
          for (int b = 0; b < 4; ++b)
            if (desp[b][w] != (uint32_t)run[b]) bad[s] = 2;
        }
        int k;
        const uint32_t sym = sym_of(r, &k);
        const uint32_t t = (uint32_t)(r - r0);
        if (sym == 4) {
          if (dollar_of[s] != UINT64_MAX) bad[s] = 3;
          dollar_of[s] = r;
          run[4]++;
          continue;
        }
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
  for (unsigned s = 0; s < nt; ++s)
    if (bad[s]) return fmt_err(std::string(oname) + ".count: a row holds two symbols", err);
  std::array<uint64_t, 5> run{0, 0, 0, 0, 0};
  std::vector<std::array<uint64_t, 5>> starts(nt);
  for (unsigned s = 0; s < nt; ++s) {
    starts[s] = run;
    for (int b = 0; b < 5; ++b) run[b] += slice_cnt[s][b];
  }
  if (run[4] != 1) return fmt_err(std::string(oname) + ".count: not exactly one '$' row", err);
  uint64_t acc = 0;
  for (int b = 0; b < 4; ++b) {
    if (C[b] != acc) return fmt_err("C.vec does not match the occurrence counts", err);
    acc += run[b];
  }
  for (unsigned s = 0; s < nt; ++s) th.emplace_back(fill_slice, s, starts[s]);
  for (auto &x : th) x.join();
  for (unsigned s = 0; s < nt; ++s)
    if (bad[s]) return fmt_err(std::string(oname) + ".desp does not match the bitmaps", err);
  for (uint64_t j = nblk_rows; j < d->nblocks; ++j)
    for (int b = 0; b < 4; ++b) d->blk[j].cnt[b] = (uint32_t)run[b];
  for (unsigned s = 0; s < nt; ++s)
    if (dollar_of[s] != UINT64_MAX) d->dollar = dollar_of[s];
  for (int b = 0; b < 4; ++b) d->C[b] = C[b];
  return FMI_OK;
}

// load_exome_file: ">name start end ..." lines; other lines are ignored.
int read_karyotype(const std::string &dir, fmi_index *ix, std::string *err) {
  std::string p = dir + "/index";
  FILE *f = fopen(p.c_str(), "r");
  if (!f) return open_err(p, err);
  std::vector<uint64_t> st, en;
  std::vector<std::string> names;
  char line[4096];
  bool ok = true;
  while (fgets(line, sizeof line, f)) {
    if (line[0] != '>') continue;
    const char *sp = strchr(line + 1, ' ');
    unsigned long long a = 0, b = 0;
    if (!sp || sscanf(sp + 1, "%llu %llu", &a, &b) != 2 || b < a) { ok = false; break; }
    names.emplace_back(line + 1, (size_t)(sp - (line + 1)));
    st.push_back(a);
    en.push_back(b);
  }
  fclose(f);
  if (!ok) return fmt_err(p + ": a contig line is not '>name start end'", err);
  uint64_t total = 0;
  for (size_t c = 0; c < st.size(); ++c) total += en[c] - st[c] + 1;
  if (!st.empty() && total != ix->n)
    return fmt_err(p + ": contig lengths do not add up to the text length", err);
  if (st.empty()) {
    st.assign(1, 0);
    en.assign(1, ix->n ? ix->n - 1 : 0);
    names.assign(1, "text");
  }
  ix->kstart = st;
  ix->kend = en;
  ix->knames = names;
  ix->koffset.assign(st.size() + 1, 0);
  for (size_t c = 0; c < st.size(); ++c) ix->koffset[c + 1] = ix->koffset[c] + (en[c] - st[c] + 1);
  return FMI_OK;
}

}  // namespace

int import_reference_dir(const char *dirc, fmi_index *ix, std::string *err) {
  if (!dirc) { *err = "directory is NULL"; return FMI_EINVAL; }
  const std::string dir(dirc);
  {
    std::string p = dir + "/config.txt";
    FILE *f = fopen(p.c_str(), "r");
    if (!f) return open_err(p, err);
    char a[256] = {0}, b[64] = {0};
    const bool got = fgets(a, sizeof a, f) != nullptr;
    if (got && !fgets(b, sizeof b, f)) b[0] = 0;
    fclose(f);
    std::string alpha(a);
    while (!alpha.empty() && (alpha.back() == '\n' || alpha.back() == '\r')) alpha.pop_back();
    if (!got || alpha != "ACGT")
      return fmt_err(p + ": alphabet is not ACGT (the engine's codes 0..3)", err);
    if (b[0] == '1')
      return fmt_err(p + ": duplicate-strand indexes are not supported (the engine searches "
                          "the reverse complement through the complement views)", err);
  }
  std::vector<uint32_t> C, C1;
  int rc = read_vec(dir, "C", C, err);
  if (rc) return rc;
  rc = read_vec(dir, "C1", C1, err);
  if (rc) return rc;
  if (C.size() != 4 || C1.size() != 4) return fmt_err("C.vec / C1.vec do not hold 4 symbols", err);
  for (int b = 0; b < 4; ++b)
    if (C1[b] != C[b] + 1) return fmt_err("C1.vec is not C.vec + 1", err);

  // The text length comes from S.vec's header (siz = n + 1).
  uint32_t hdr[3];
  {
    std::string p = dir + "/S.vec";
    File f(p);
    if (!f.f) return open_err(p, err);
    if (!f.read(hdr, 12) || hdr[0] == 0) return fmt_err(p + ": bad header", err);
  }
  const uint64_t n = (uint64_t)hdr[0] - 1;
  if (n + 3 >= (uint64_t)UINT32_MAX) return fmt_err("index too large", err);
  ix->n = n;
  uint32_t ratio = 0;
  const char *names[2][3] = {{"O", "S", "R"}, {"Oi", "Si", "Ri"}};
  for (int d = 0; d < 2; ++d) {
    rc = read_direction(dir, names[d][0], C.data(), n, &ix->h[d], err);
    if (rc) return rc;
    uint64_t ns = 0, nr = 0;
    rc = read_comp_vec(dir, names[d][1], n + 1, &ratio, &ix->h[d].S, &ns, err);
    if (rc) return rc;
    rc = read_comp_vec(dir, names[d][2], n + 1, &ratio, &ix->h[d].R, &nr, err);
    if (rc) return rc;
    ix->h[d].nsamp = ns;
    // S[dollar / ratio] = 0 and R[0] = dollar, as calculate_S_and_R writes them
    if (ix->h[d].R[0] != ix->h[d].dollar ||
        (ix->h[d].dollar % ratio == 0 && ix->h[d].S[ix->h[d].dollar / ratio] != 0))
      return fmt_err(std::string(names[d][1]) + ".vec / " + names[d][2] +
                         ".vec do not match the '$' row", err);
  }
  ix->ratio = ratio;
  return read_karyotype(dir, ix, err);
}

}  // namespace fmi

// host_check.cpp -- drives the host-side code of libhpgfm.so (the CPU index
// builder, the index file reader and writer, the reference-directory
// importer, the text formatter) under AddressSanitizer and
// UndefinedBehaviorSanitizer, built from the same sources with g++ (see
// Makefile target `check`).  No GPU code is involved.
//
//   host_check selftest SEED   build, save and load indexes of several
//                              texts; load every truncation of a file and
//                              files with flipped bytes (checksum kept and
//                              recomputed); format records of random fields
//   host_check load PATH       load one index file, print its status
//   host_check import DIR      import one reference directory, print status
//
// Exit status 0 unless a check fails (a sanitizer report aborts first).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../../include/fmi.h"
#include "../fmi_index.h"

namespace {

uint64_t fnv1a(const std::vector<uint8_t> &b, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}

std::vector<uint8_t> slurp(const std::string &p) {
  std::vector<uint8_t> v;
  FILE *f = fopen(p.c_str(), "rb");
  if (!f) return v;
  int c;
  while ((c = fgetc(f)) != EOF) v.push_back((uint8_t)c);
  fclose(f);
  return v;
}

bool spit(const std::string &p, const std::vector<uint8_t> &v, size_t n) {
  FILE *f = fopen(p.c_str(), "wb");
  if (!f) return false;
  const bool ok = fwrite(v.data(), 1, n, f) == n;
  return (fclose(f) == 0) && ok;
}

int load_status(const std::string &p) {
  fmi_index ix;
  std::string err;
  return fmi::load_index(p.c_str(), &ix, &err);
}

int fail(const char *what) {
  fprintf(stderr, "host_check: %s\n", what);
  return 1;
}

int selftest(uint32_t seed, const std::string &tmp) {
  std::mt19937_64 rng(seed);
  const uint64_t sizes[] = {1, 2, 97, 191, 192, 193, 5000, 40000};
  for (uint64_t n : sizes) {
    std::vector<uint8_t> text(n);
    for (auto &c : text) c = (uint8_t)(rng() & 3u);
    if (n > 1000) std::fill(text.begin() + 100, text.begin() + 700, (uint8_t)2);   // a run
    fmi_index ix;
    std::string err;
    if (fmi::build_index(text.data(), n, 1 + (uint32_t)(rng() % 32), &ix, &err) != FMI_OK)
      return fail("build_index");
    const std::string p = tmp + "/x.fmi";
    if (fmi::save_index(&ix, p.c_str(), &err) != FMI_OK) return fail("save_index");
    if (load_status(p) != FMI_OK) return fail("load of a good file");
    const std::vector<uint8_t> good = slurp(p);
    const std::string q = tmp + "/bad.fmi";
    // every truncation (a stride through large files)
    const size_t step = good.size() > 4096 ? good.size() / 1500 : 1;
    for (size_t cut = 0; cut < good.size(); cut += step) {
      if (!spit(q, good, cut)) return fail("write");
      if (load_status(q) == FMI_OK) return fail("a truncated file loaded");
    }
    // flipped bytes, with the checksum left and with it recomputed
    for (int t = 0; t < 300; ++t) {
      std::vector<uint8_t> b = good;
      const size_t at = (size_t)(rng() % (b.size() - 8));
      b[at] ^= (uint8_t)(1u << (rng() % 8));
      if (!spit(q, b, b.size())) return fail("write");
      if (load_status(q) == FMI_OK) return fail("a flipped byte loaded");
      const uint64_t h = fnv1a(b, b.size() - 8);
      std::memcpy(&b[b.size() - 8], &h, 8);
      if (!spit(q, b, b.size())) return fail("write");
      (void)load_status(q);   // may load (a valid file): must not misbehave
    }
    // text lines of random records
    std::vector<fmi_hit> hits(8);
    std::vector<int32_t> chrom(8 * 4);
    std::vector<uint64_t> coord(8 * 4);
    for (auto &h : hits) {
      std::memset(&h, 0, sizeof h);
      h.k = (int64_t)(rng() % (n + 1));
      h.l = h.k + (int64_t)(rng() % 3) - 1;
      h.start = (int16_t)(rng() % 50);
      h.end = (int16_t)(rng() % 120);
      h.err_pos = (int16_t)(rng() % 2000) - 1000;
      h.num_mismatches = (uint8_t)(rng() % 2);
      h.err_kind = (uint8_t)(rng() % 4);
      h.err_base = (uint8_t)(rng() % 5);
      h.dir = (uint8_t)(rng() % 2);
    }
    for (auto &c : chrom) c = (int32_t)(rng() % 3) - 1;
    for (auto &c : coord) c = rng();
    std::string out;
    uint64_t missing = 0;
    const char *read = "ACGTNACGTTGCA";
    for (uint32_t nw : {0u, 5u, 13u, 20u})
      if (fmi::format_results(&ix, rng(), (int)(rng() % 2), read, nw, hits.data(), 8,
                              chrom.data(), coord.data(), 4, &out, &missing) != FMI_OK)
        return fail("format_results");
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) return fail("usage: host_check selftest SEED [TMPDIR] | load PATH | import DIR");
  const std::string cmd = argv[1];
  if (cmd == "selftest") return selftest((uint32_t)strtoul(argv[2], nullptr, 10), argc > 3 ? argv[3] : "/tmp");
  if (cmd == "load") {
    printf("%d\n", load_status(argv[2]));
    return 0;
  }
  if (cmd == "import") {
    fmi_index ix;
    std::string err;
    printf("%d\n", fmi::import_reference_dir(argv[2], &ix, &err));
    return 0;
  }
  return fail("unknown command");
}

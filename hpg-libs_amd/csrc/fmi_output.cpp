// fmi_output.cpp -- the reference's text result lines (write_results,
// search/results.c:3-43, and manage_single_result, search/results.h:
// 105-187) from records whose SA rows were already resolved on the device
// by fmi_resolve_hits.  Host code: formatting only, no index access beyond
// the contig names.
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/fmi.h"
#include "fmi_index.h"

namespace fmi {

int format_results(const fmi_index *ix, uint64_t read_index, int strand, const char *mapping,
                   uint32_t nW, const fmi_hit *hits, uint32_t nhits, const int32_t *chrom,
                   const uint64_t *coord, uint32_t row_cap, std::string *out,
                   uint64_t *rows_missing) {
  static const char kBases[] = "ACGTN";
  const std::string search(mapping, strnlen(mapping, nW));   // strncat(search, mapping, nW)
  const char pm = strand ? '-' : '+';                        // plusminus[type], type = !strand
  std::vector<std::pair<int64_t, int64_t>> seen;             // write_results' k / l arrays
  char line[64];
  *rows_missing = 0;
  for (uint32_t h = 0; h < nhits; ++h) {
    const fmi_hit &r = hits[h];
    bool repeated = false;
    for (const auto &kl : seen)
      if (kl.first == r.k && kl.second == r.l) { repeated = true; break; }
    if (repeated) continue;
    seen.emplace_back(r.k, r.l);
    // concat_error_string (results.h:105-123)
    std::string mask;
    if (r.num_mismatches) {
      const char b = kBases[r.err_base < 4 ? r.err_base : 4];
      if (r.err_kind == FMI_DELETION) snprintf(line, sizeof line, "_%dd", r.err_pos);
      else if (r.err_kind == FMI_MISMATCH) snprintf(line, sizeof line, "_%dm%c", r.err_pos, b);
      else snprintf(line, sizeof line, "_%di%c", r.err_pos, b);
      mask = line;
    }
    if (r.k > r.l) continue;
    const uint64_t rows = (uint64_t)(r.l - r.k + 1);
    const uint64_t have = rows < row_cap ? rows : row_cap;
    *rows_missing += rows - have;
    for (uint64_t j = 0; j < have; ++j) {
      const int32_t c = chrom[(uint64_t)h * row_cap + j];
      if (c < 0) continue;                                   // crosses a contig end
      const char *name = (uint64_t)c < ix->knames.size() ? ix->knames[c].c_str() : "";
      snprintf(line, sizeof line, " %u", (unsigned)r.num_mismatches);
      *out += "read_" + std::to_string(read_index) + ' ' + pm + ' ' + name + ' ' +
              std::to_string(coord[(uint64_t)h * row_cap + j]) + line + mask + ' ' + search +
              '\n';
    }
  }
  return FMI_OK;
}

}  // namespace fmi

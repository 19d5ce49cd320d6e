/* check_oracle.c -- runs the oracle's search functions on random texts and
 * reads (repeats, N bases, empty and long reads, list capacities that
 * truncate) under AddressSanitizer and UndefinedBehaviorSanitizer (Makefile
 * target `check`).  Test infrastructure: nothing in the product uses it.
 *
 *   check_oracle SEED */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fmoracle.h"

static uint64_t st;
static uint32_t rnd(void) {
  st ^= st << 13; st ^= st >> 7; st ^= st << 17;
  return (uint32_t)(st >> 11);
}

int main(int argc, char **argv) {
  st = 0x9E3779B97F4A7C15ull ^ (argc > 1 ? strtoull(argv[1], NULL, 10) : 1);
  const uint64_t sizes[] = {1, 60, 3000, 30000};
  for (int si = 0; si < 4; si++) {
    const uint64_t n = sizes[si];
    uint8_t *text = (uint8_t *)malloc(n);
    for (uint64_t i = 0; i < n; i++) text[i] = (uint8_t)(rnd() & 3u);
    if (n > 1000) for (uint64_t i = 200; i < 800; i++) text[i] = text[200 + i % 7];
    fmo_index ix;
    if (fmo_index_build_naive(&ix, text, n, 8)) { fprintf(stderr, "build\n"); return 1; }
    uint64_t koff[3] = {0, n / 2, n}, kstart[2] = {0, 100};
    for (int r = 0; r < 200; r++) {
      const uint32_t len = rnd() % 160;
      uint8_t W[160];
      const uint64_t p = n > len ? rnd() % (n - len + 1) : 0;
      for (uint32_t i = 0; i < len; i++) {
        W[i] = (p + i < n) ? text[p + i] : (uint8_t)(rnd() & 3u);
        if (rnd() % 50 == 0) W[i] = (uint8_t)(rnd() % 5);
      }
      fmo_hit out[64], anc[2];
      for (int s = 0; s < 2; s++) {
        fmo_search1(&ix, s, (int)(rnd() & 1u), W, (int16_t)len, out, 1 + rnd() % 64, anc);
        if (len >= 3) {
          const int16_t a = (int16_t)(rnd() % (len - 2));
          const int16_t b = (int16_t)(a + 2 + rnd() % (len - a - 2));
          fmo_search1_vector(&ix, s, 0, W, a, b, out, 1 + rnd() % 64);
        }
        fmo_hitk fin[32], ak[8];
        uint32_t nf = 0, na = 0;
        const int16_t fs = (int16_t)(len / (2 + rnd() % 3));
        if (fs > 0) fmo_search_k(&ix, s, W, (int16_t)len, fs, 1 + rnd() % 32, 8, fin, &nf, ak, &na);
      }
      fmo_read_mapping mp[16];
      uint8_t status;
      fmo_map_read(&ix, W, len, rnd() % 40, koff, kstart, 2, (uint32_t)r, mp, 16, &status);
      fmo_seed_region sr[64];
      int64_t info[6 * 40];
      if (len >= 18)
        fmo_seed_map(&ix, W, len, 18, 5, 1000, koff, kstart, 2, (uint32_t)r, sr, 64, info);
    }
    fmo_index_free(&ix);
    free(text);
  }
  printf("check_oracle: ok\n");
  return 0;
}

// fmi_sais.h -- suffix array construction by induced sorting (SA-IS, Nong,
// Zhang & Chan 2009) for the index builder.
//
// The reference builds its BWT with dbwt (c/src/aligners/bwt/dbwt/dbwt.c:357,
// Okanohara-Sadakane direct BWT with SA-IS at the first level) and is limited
// to texts below 2^31 (int sa, dbwt.c:380).  This builder computes the full
// suffix array with a signed index type I (int32_t below 2^31, int64_t above)
// and derives the BWT, the SA samples and the ISA samples from it.
//
// The text is T[0..n) over [0, K) with a virtual sentinel after T[n-1] that
// is smaller than every symbol; the output SA[0..n) orders the n non-empty
// suffixes (the '$' suffix is the implicit row 0 of the BWT).
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

namespace fmi {

class TypeBits {
 public:
  explicit TypeBits(uint64_t n) : w_((n + 64) / 64, 0) {}
  bool s(uint64_t i) const { return (w_[i >> 6] >> (i & 63)) & 1u; }
  void set_s(uint64_t i) { w_[i >> 6] |= (uint64_t)1 << (i & 63); }

 private:
  std::vector<uint64_t> w_;
};

template <typename CharT, typename I>
static void sais_buckets(const CharT *T, I n, I K, I *bkt, bool ends) {
  std::fill(bkt, bkt + K, I(0));
  for (I i = 0; i < n; ++i) ++bkt[(I)T[i]];
  I sum = 0;
  for (I c = 0; c < K; ++c) {
    sum += bkt[c];
    bkt[c] = ends ? sum : sum - bkt[c];
  }
}

// L-type suffixes from the left, then S-type suffixes from the right.
template <typename CharT, typename I>
static void sais_induce(const CharT *T, I *SA, I n, I K, const TypeBits &t, I *bkt) {
  sais_buckets(T, n, K, bkt, false);
  // The suffix n-1 is L-type (it is followed by the sentinel) and is the
  // first one the sentinel induces.
  SA[bkt[(I)T[n - 1]]++] = n - 1;
  for (I i = 0; i < n; ++i) {
    I j = SA[i] - 1;
    if (SA[i] > 0 && !t.s((uint64_t)j)) SA[bkt[(I)T[j]]++] = j;
  }
  sais_buckets(T, n, K, bkt, true);
  for (I i = n - 1; i >= 0; --i) {
    I j = SA[i] - 1;
    if (SA[i] > 0 && t.s((uint64_t)j)) SA[--bkt[(I)T[j]]] = j;
  }
}

template <typename CharT, typename I>
static void sais_rec(const CharT *T, I *SA, I n, I K) {
  if (n <= 0) return;
  if (n == 1) { SA[0] = 0; return; }
  TypeBits t((uint64_t)n);
  // t[n-1] is L; t[i] is S when T[i] < T[i+1], or equal and t[i+1] is S.
  for (I i = n - 2; i >= 0; --i) {
    if (T[i] < T[i + 1] || (T[i] == T[i + 1] && t.s((uint64_t)(i + 1)))) t.set_s((uint64_t)i);
  }
  auto lms = [&](I i) -> bool { return i > 0 && t.s((uint64_t)i) && !t.s((uint64_t)(i - 1)); };

  std::vector<I> bkt((size_t)K);
  // 1. LMS suffixes at the ends of their buckets, then induce: this sorts
  //    the LMS substrings.
  std::fill(SA, SA + n, I(-1));
  sais_buckets(T, n, K, bkt.data(), true);
  for (I i = 1; i < n; ++i)
    if (lms(i)) SA[--bkt[(I)T[i]]] = i;
  sais_induce(T, SA, n, K, t, bkt.data());

  // 2. Compact the sorted LMS positions into SA[0..m).
  I m = 0;
  for (I i = 0; i < n; ++i)
    if (lms(SA[i])) SA[m++] = SA[i];
  // 3. Name the LMS substrings (equal substrings share a name); the name of
  //    the substring at p goes to SA[m + p/2] (LMS positions are >= 2 apart).
  std::fill(SA + m, SA + n, I(-1));
  I name = 0, prev = -1;
  for (I i = 0; i < m; ++i) {
    I p = SA[i];
    bool diff = (prev < 0);
    if (!diff) {
      for (I d = 0;; ++d) {
        if (p + d >= n || prev + d >= n) { diff = true; break; }   // hits the sentinel
        if (T[p + d] != T[prev + d] || t.s((uint64_t)(p + d)) != t.s((uint64_t)(prev + d))) {
          diff = true;
          break;
        }
        if (d > 0 && (lms(p + d) || lms(prev + d))) {
          diff = !(lms(p + d) && lms(prev + d));
          break;
        }
      }
    }
    if (diff) { ++name; prev = p; }
    SA[m + p / 2] = name - 1;
  }
  // Names in text order to the tail SA[n-m..n).
  for (I i = n - 1, j = n - 1; i >= m; --i)
    if (SA[i] >= 0) SA[j--] = SA[i];

  // 4. Sort the reduced string (recursively while names repeat).
  I *s1 = SA + n - m;
  I *sa1 = SA;
  if (name < m) {
    sais_rec<I, I>(s1, sa1, m, name);
  } else {
    for (I i = 0; i < m; ++i) sa1[s1[i]] = i;
  }
  // 5. Map reduced ranks back to LMS positions, seed the buckets in sorted
  //    order and induce the full suffix array.
  for (I i = 1, j = 0; i < n; ++i)
    if (lms(i)) s1[j++] = i;
  for (I i = 0; i < m; ++i) sa1[i] = s1[sa1[i]];
  std::fill(SA + m, SA + n, I(-1));
  sais_buckets(T, n, K, bkt.data(), true);
  for (I i = m - 1; i >= 0; --i) {
    I j = SA[i];
    SA[i] = -1;
    SA[--bkt[(I)T[j]]] = j;
  }
  sais_induce(T, SA, n, K, t, bkt.data());
}

// Suffix array of a text over symbols [0, K).  SA has n entries.
template <typename I>
inline void suffix_array(const uint8_t *T, I *SA, I n, I K) {
  sais_rec<uint8_t, I>(T, SA, n, K);
}

}  // namespace fmi

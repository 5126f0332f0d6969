// CPU BASELINE (test / measurement infrastructure, never the product): a C
// restatement of the reference's fast mode candidate generation
// (beguidedoff/fast/finder.py:163-277) over this directory's own exact-match
// index of the forward genome.
//
// Reference semantics restated here:
//  * the index text is the forward genome with non-ACGT bases as A
//    (fast/backend.py:28-49); anchors never span two contigs (separator
//    symbols, fast/fmindex.py:87-101);
//  * anchors: '+' guide[c0 : c0 + k], '-' rc(guide)[0 : k] (fast/finder.py:183-189),
//    with the cold region c0 .. c0 + k - 1 of regions.py:80-123 (9-20 at cutoff 5);
//  * at most max_hits occurrences per anchor (fast/finder.py:146-158); more set
//    the overflow flag;
//  * per occurrence: anchor and window inside one contig (:208-212), PAM test on
//    the decoded bases (:214-224, 41-49), 2-bit mismatch positions (:51-54),
//    N positions where the guide base equals the N symbol (A on '+', T on
//    '-') unioned in (:239-251), score = ascending sum of 1.2**p (scoring.py)
//    and kept when score <= cutoff (:253-255).
//
// Index: a direct-address table of 10-mers (4^10 buckets of ascending
// positions); a 12-mer anchor is its 10-mer bucket filtered on the last two
// bases.  The reference's FM-index returns the first max_hits occurrences in
// suffix-array order; this index returns the first max_hits in position order.
// Both flag the overflow, and the two differ only for anchors with more than
// max_hits occurrences.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define KB 10
#define NBUCKET (1u << (2 * KB))

typedef struct {
  const uint32_t* words;  // 2-bit forward genome (16 bases per word)
  int64_t total_bp;
  int n_contigs;
  int64_t* offsets;
  int64_t* lengths;
  uint64_t* start;  // [NBUCKET + 1]
  uint32_t* pos;    // ascending occurrences per bucket
} fast_index;

static inline uint32_t base_at(const uint32_t* w, int64_t i) {
  return (w[i >> 4] >> (2 * (i & 15))) & 3u;
}

// Build the index over contig-internal 10-mers.  Returns 0 or -1 (alloc).
int bgo_fast_build(const uint32_t* words, int64_t total_bp, const int64_t* offsets,
                   const int64_t* lengths, int n_contigs, int threads, void** out) {
  fast_index* ix = calloc(1, sizeof(fast_index));
  if (!ix) return -1;
  ix->words = words;
  ix->total_bp = total_bp;
  ix->n_contigs = n_contigs;
  ix->offsets = malloc(sizeof(int64_t) * n_contigs);
  ix->lengths = malloc(sizeof(int64_t) * n_contigs);
  memcpy(ix->offsets, offsets, sizeof(int64_t) * n_contigs);
  memcpy(ix->lengths, lengths, sizeof(int64_t) * n_contigs);
  if (threads <= 0) threads = omp_get_max_threads();
  uint32_t* cnt = calloc((size_t)threads * NBUCKET, sizeof(uint32_t));
  ix->start = malloc(sizeof(uint64_t) * (NBUCKET + 1));
  if (!cnt || !ix->start) return -1;
  const uint32_t kmask = NBUCKET - 1;
  // Pass 1: per-thread counts over equal slices of [0, total_bp).
#pragma omp parallel num_threads(threads)
  {
    const int t = omp_get_thread_num();
    const int64_t b = total_bp * t / threads, e = total_bp * (t + 1) / threads;
    uint32_t* c = cnt + (size_t)t * NBUCKET;
    int k = 0;
    while (k + 1 < n_contigs && offsets[k + 1] <= b) ++k;
    for (int64_t p = b; p < e;) {
      const int64_t cend = offsets[k] + lengths[k];
      const int64_t stop = e < cend - KB + 1 ? e : cend - KB + 1;
      if (p < stop) {
        uint32_t km = 0;
        for (int j = 0; j < KB - 1; ++j) km = (km << 2) | base_at(words, p + j);
        for (; p < stop; ++p) {
          km = ((km << 2) | base_at(words, p + KB - 1)) & kmask;
          ++c[km];
        }
      }
      p = cend > p ? cend : p;
      ++k;
      if (k >= n_contigs) break;
    }
  }
  // Bucket starts; per-thread write cursors (threads in slice order keep
  // every bucket ascending).
  uint64_t acc = 0;
  for (uint32_t km = 0; km < NBUCKET; ++km) {
    ix->start[km] = acc;
    for (int t = 0; t < threads; ++t) {
      const uint32_t v = cnt[(size_t)t * NBUCKET + km];
      cnt[(size_t)t * NBUCKET + km] = (uint32_t)(acc - ix->start[km]);
      acc += v;
    }
  }
  ix->start[NBUCKET] = acc;
  ix->pos = malloc(sizeof(uint32_t) * (acc ? acc : 1));
  if (!ix->pos) return -1;
  // Pass 2: fill.
#pragma omp parallel num_threads(threads)
  {
    const int t = omp_get_thread_num();
    const int64_t b = total_bp * t / threads, e = total_bp * (t + 1) / threads;
    uint32_t* c = cnt + (size_t)t * NBUCKET;
    int k = 0;
    while (k + 1 < n_contigs && offsets[k + 1] <= b) ++k;
    for (int64_t p = b; p < e;) {
      const int64_t cend = offsets[k] + lengths[k];
      const int64_t stop = e < cend - KB + 1 ? e : cend - KB + 1;
      if (p < stop) {
        uint32_t km = 0;
        for (int j = 0; j < KB - 1; ++j) km = (km << 2) | base_at(words, p + j);
        for (; p < stop; ++p) {
          km = ((km << 2) | base_at(words, p + KB - 1)) & kmask;
          ix->pos[ix->start[km] + c[km]++] = (uint32_t)p;
        }
      }
      p = cend > p ? cend : p;
      ++k;
      if (k >= n_contigs) break;
    }
  }
  free(cnt);
  *out = ix;
  return 0;
}

void bgo_fast_free(void* p) {
  fast_index* ix = p;
  if (!ix) return;
  free(ix->offsets);
  free(ix->lengths);
  free(ix->start);
  free(ix->pos);
  free(ix);
}

int64_t bgo_fast_bytes(void* p) {
  fast_index* ix = p;
  return (int64_t)(sizeof(uint64_t) * (NBUCKET + 1) + sizeof(uint32_t) * ix->start[NBUCKET]);
}

static int contig_of(const fast_index* ix, int64_t x) {
  int lo = 0, hi = ix->n_contigs;  // last offset <= x
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (ix->offsets[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// N mask: little-endian bit i of byte i / 8 marks base i (forward strand).
static inline int is_n(const uint8_t* nm, int64_t i) { return nm ? (nm[i >> 3] >> (i & 7)) & 1 : 0; }

typedef struct {
  int64_t leftmost;
  uint64_t site;  // 2-bit codes of the site, guide-oriented, base i at bits 2i
  uint32_t mm;    // bit p-1 = mismatch at 1-based position p
  int8_t strand;
  double score;
} fast_cand;

// Candidates of one guide (20 codes) and PAM; returns the count written (at
// most cap; the full count when out is NULL).
static int64_t search_one(const fast_index* ix, const uint8_t* nmask, const uint8_t* g,
                          const uint8_t* pam_mask, int pam_len, int c0, int k, double cutoff,
                          const double* tab, int64_t max_hits, fast_cand* out, int64_t cap,
                          int* overflow) {
  const int win = 20 + pam_len;
  uint8_t rc[20];
  for (int i = 0; i < 20; ++i) rc[i] = 3 - g[19 - i];
  int64_t n = 0;
  for (int sidx = 0; sidx < 2; ++sidx) {
    const uint8_t* anchor = sidx == 0 ? g + c0 : rc;
    uint32_t head = 0;
    for (int j = 0; j < KB; ++j) head = (head << 2) | anchor[j];
    const uint64_t b = ix->start[head], e = ix->start[head + 1];
    int64_t found = 0;
    for (uint64_t q = b; q < e; ++q) {
      const int64_t gpos = ix->pos[q];
      int ok = 1;
      for (int j = KB; j < k && ok; ++j)
        ok = gpos + j < ix->total_bp && base_at(ix->words, gpos + j) == anchor[j];
      if (!ok) continue;
      // the k-mer must lie inside one contig (a separator ends every contig)
      {
        const int kc = contig_of(ix, gpos);
        if (gpos + k > ix->offsets[kc] + ix->lengths[kc]) continue;
      }
      if (max_hits > 0 && found >= max_hits) { *overflow = 1; break; }
      ++found;
      const int64_t leftmost = sidx == 0 ? gpos - c0 : gpos - pam_len;
      if (leftmost < 0) continue;
      const int kc = contig_of(ix, leftmost);
      const int64_t coff = ix->offsets[kc], clen = ix->lengths[kc];
      if (gpos < coff || gpos + k > coff + clen) continue;
      if (leftmost < coff || leftmost + win > coff + clen) continue;
      uint8_t w[26];
      for (int j = 0; j < win; ++j) w[j] = (uint8_t)base_at(ix->words, leftmost + j);
      uint8_t site[20], pam[6];
      if (sidx == 0) {
        memcpy(site, w, 20);
        memcpy(pam, w + 20, pam_len);
      } else {
        for (int j = 0; j < pam_len; ++j) pam[j] = 3 - w[pam_len - 1 - j];
        for (int j = 0; j < 20; ++j) site[j] = 3 - w[pam_len + 19 - j];
      }
      int pam_ok = 1;
      for (int j = 0; j < pam_len && pam_ok; ++j) pam_ok = (pam_mask[j] >> pam[j]) & 1;
      if (!pam_ok) continue;
      uint32_t mm = 0;
      for (int i = 0; i < 20; ++i) mm |= (uint32_t)(site[i] != g[i]) << i;
      const int64_t gs = sidx == 0 ? leftmost : leftmost + pam_len;
      const uint8_t nsym = sidx == 0 ? 0 : 3;  // A on '+', T on '-'
      for (int i = 0; i < 20; ++i) {
        const int nn = sidx == 0 ? is_n(nmask, gs + i) : is_n(nmask, gs + 19 - i);
        if (nn && g[i] == nsym) mm |= 1u << i;
      }
      double sc = 0.0;
      for (int i = 0; i < 20; ++i)
        if ((mm >> i) & 1u) sc += tab[i + 1];
      if (sc > cutoff) continue;
      if (out && n < cap) {
        uint64_t sv = 0;
        for (int i = 0; i < 20; ++i) sv |= (uint64_t)site[i] << (2 * i);
        out[n].site = sv;
        out[n].leftmost = leftmost;
        out[n].mm = mm;
        out[n].strand = (int8_t)sidx;
        out[n].score = sc;
      }
      ++n;
    }
  }
  return n;
}

// Candidates of n guides (codes [n][20], PAM masks [n][6], lengths [n]).
// counts[i] receives guide i's candidate count; when out_* are non-NULL the
// candidates are written guide after guide at the offsets implied by counts
// (call once with NULL outputs to size them).  overflow[i] is set when an
// anchor had more than max_hits occurrences.
int bgo_fast_search(void* p, const uint8_t* nmask, const uint8_t* guides, const uint8_t* pam_masks,
                    const uint8_t* pam_lens, int n, int c0, int k, double cutoff,
                    const double* tab, int64_t max_hits, int threads, int64_t* counts,
                    int64_t* out_leftmost, uint64_t* out_site, uint32_t* out_mm,
                    int8_t* out_strand, double* out_score, uint8_t* overflow) {
  const fast_index* ix = p;
  if (k < KB || k > 20 || c0 < 0 || c0 + k > 20) return -1;
  if (threads <= 0) threads = omp_get_max_threads();
  int64_t* base = NULL;
  if (out_leftmost) {
    base = malloc(sizeof(int64_t) * (n + 1));
    base[0] = 0;
    for (int i = 0; i < n; ++i) base[i + 1] = base[i] + counts[i];
  }
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
  for (int i = 0; i < n; ++i) {
    int ovf = 0;
    if (!out_leftmost) {
      counts[i] = search_one(ix, nmask, guides + 20 * i, pam_masks + 6 * i, pam_lens[i], c0, k,
                             cutoff, tab, max_hits, NULL, 0, &ovf);
    } else {
      const int64_t cap = counts[i];
      fast_cand* tmp = malloc(sizeof(fast_cand) * (cap ? cap : 1));
      const int64_t got = search_one(ix, nmask, guides + 20 * i, pam_masks + 6 * i, pam_lens[i],
                                     c0, k, cutoff, tab, max_hits, tmp, cap, &ovf);
      for (int64_t j = 0; j < got && j < cap; ++j) {
        out_leftmost[base[i] + j] = tmp[j].leftmost;
        out_site[base[i] + j] = tmp[j].site;
        out_mm[base[i] + j] = tmp[j].mm;
        out_strand[base[i] + j] = tmp[j].strand;
        out_score[base[i] + j] = tmp[j].score;
      }
      free(tmp);
    }
    overflow[i] = (uint8_t)ovf;
  }
  free(base);
  return 0;
}

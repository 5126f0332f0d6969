/*
 * ORACLE — test infrastructure, not product code.
 *
 * CPU restatement of the reference extensive scan, used only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
 * Nothing in beguidedoff_amd/ links or calls this.
 *
 * It follows beguidedoff/cuda_search.py:163-314 (per-position predicate) and
 * :384-488 (driver range [0, total_bp - 26)) in the plainest form:
 *   - per base 2-bit decode of the packed array (encoding.py:36-46),
 *   - per base IUPAC allowed-set PAM test (pam.py:7-23; the reference's packed
 *     pattern/compare fields, cuda_search.py:101-160, are a re-expression of
 *     this test — pinned separately by exhaustive enumeration),
 *   - per base mismatch count over the first 20 guide bases,
 *   - the N correction: +1 for every N position of the 20-nt window whose
 *     guide base equals the N code of the strand (cuda_search.py:290-308).
 * Parity status: pinned by brute force over strings, the reference's
 * known-answer fixtures and invariants (see tests/test_oracle_*.py).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint32_t base_at(const uint32_t* w, int64_t i) {
  return (w[i >> 4] >> ((i & 15) * 2)) & 3u;
}

static inline int is_n(const uint8_t* m, int64_t nbytes, int64_t i) {
  if (!m || i < 0 || (i >> 3) >= nbytes) return 0;
  return (m[i >> 3] >> (i & 7)) & 1;
}

typedef struct {
  int32_t* g;
  int64_t* p;
  int8_t* mm;
  int64_t n, cap;
} hitbuf;

static int push(hitbuf* b, int32_t g, int64_t p, int8_t mm) {
  if (b->n == b->cap) {
    int64_t nc = b->cap ? 2 * b->cap : 1024;
    int32_t* ng = realloc(b->g, nc * sizeof(int32_t));
    if (!ng) return -1;
    b->g = ng;
    int64_t* np_ = realloc(b->p, nc * sizeof(int64_t));
    if (!np_) return -1;
    b->p = np_;
    int8_t* nm = realloc(b->mm, nc);
    if (!nm) return -1;
    b->mm = nm;
    b->cap = nc;
  }
  b->g[b->n] = g;
  b->p[b->n] = p;
  b->mm[b->n] = mm;
  b->n++;
  return 0;
}

/* One position, one guide: returns the N-corrected mismatch count if the
 * site is reported, else -1. */
static int site(const uint32_t* words, int64_t total_bp, const uint8_t* nmask,
                int64_t nmask_bytes, uint64_t guide, const uint8_t* pam_mask, int plen,
                int three_prime, int max_mm, uint32_t n_encoded, int64_t pos) {
  if (plen <= 0 || plen > 6) return -1; /* packed encoding disables the guide */
  int64_t gs, ps;
  if (three_prime) {
    gs = pos;
    ps = pos + 20;
  } else {
    gs = pos + plen;
    ps = pos;
    if (gs + 20 > total_bp || ps + plen > total_bp) return -1;
  }
  for (int i = 0; i < plen; ++i) {
    const uint32_t b = base_at(words, ps + i);
    if (!((1u << b) & (pam_mask[i] & 0xFu))) return -1;
  }
  int mm = 0;
  for (int i = 0; i < 20; ++i)
    if (base_at(words, gs + i) != ((guide >> (2 * i)) & 3u)) ++mm;
  if (mm > max_mm) return -1;
  if (nmask && nmask_bytes > 0) {
    for (int i = 0; i < 20; ++i)
      if (is_n(nmask, nmask_bytes, gs + i) && ((guide >> (2 * i)) & 3u) == n_encoded) ++mm;
    if (mm > max_mm) return -1;
  }
  return mm;
}

/*
 * Scan positions [pos_begin, pos_end) ∩ [0, total_bp - 26) of one strand
 * array.  Hits are returned ordered by (position, guide index).  Returns the
 * total number of hits (only the first `cap` are stored) or -1 on error.
 */
int64_t bgo_oracle_scan(const uint32_t* words, int64_t n_words, int64_t total_bp,
                        const uint8_t* nmask, int64_t nmask_bytes, const uint64_t* guides,
                        const uint8_t* pam_masks, const uint8_t* pam_lens,
                        const uint8_t* pam_3prime, int32_t n_guides, int32_t max_mm,
                        int32_t n_encoded, int64_t pos_begin, int64_t pos_end,
                        int32_t n_threads, int64_t cap, int32_t* out_guide, int64_t* out_pos,
                        int8_t* out_mm) {
  if (!words || total_bp < 0 || total_bp > n_words * 16 || n_guides < 0) return -1;
  const int64_t lo = pos_begin > 0 ? pos_begin : 0;
  const int64_t hi = pos_end < total_bp - 26 ? pos_end : total_bp - 26;
  if (hi <= lo || n_guides == 0) return 0;
  int T = n_threads > 0 ? n_threads : 1;
  hitbuf* bufs = calloc((size_t)T, sizeof(hitbuf));
  if (!bufs) return -1;
  int failed = 0;
  const int64_t span = hi - lo;
#pragma omp parallel num_threads(T) if (T > 1)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    const int64_t a = lo + span * t / T, b = lo + span * (t + 1) / T;
    for (int64_t pos = a; pos < b && !failed; ++pos) {
      for (int32_t g = 0; g < n_guides; ++g) {
        const int mm = site(words, total_bp, nmask, nmask_bytes, guides[g],
                            pam_masks + 6 * (int64_t)g, pam_lens[g], pam_3prime[g], max_mm,
                            (uint32_t)n_encoded, pos);
        if (mm >= 0 && push(&bufs[t], g, pos, (int8_t)mm)) failed = 1;
      }
    }
  }
  int64_t total = 0;
  for (int t = 0; t < T; ++t) {
    for (int64_t i = 0; i < bufs[t].n; ++i, ++total) {
      if (total < cap) {
        out_guide[total] = bufs[t].g[i];
        out_pos[total] = bufs[t].p[i];
        out_mm[total] = bufs[t].mm[i];
      }
    }
    free(bufs[t].g);
    free(bufs[t].p);
    free(bufs[t].mm);
  }
  free(bufs);
  return failed ? -1 : total;
}

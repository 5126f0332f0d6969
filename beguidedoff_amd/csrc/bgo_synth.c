/*
 * Synthetic genome generator (host, OpenMP).
 *
 * Writes the reference's genome artefact layout (beguidedoff/
 * genome_converter.py:15-108): 2-bit packed forward words, the per-contig
 * reverse complement packed the same way, and the N mask.  Bases are i.i.d.
 * with P(A)=P(T)=0.295, P(C)=P(G)=0.205 (41% GC) drawn from a counter-based
 * hash (splitmix64 of seed and base index), so any slice of the genome can be
 * regenerated independently and identically on any machine.
 *
 * Order of construction: hash bases -> N runs (code 0, mask bit set) ->
 * planted sites (already oriented forward) -> reverse complement per contig.
 */
#include <stdint.h>
#include <string.h>

static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* 16-bit thresholds of the cumulative base distribution A, C, G (T rest). */
#define THR_A 19333u /* 0.295 * 65536 */
#define THR_C 32768u /* 0.500 * 65536 */
#define THR_G 46203u /* 0.705 * 65536 */

static inline uint32_t code_of(uint32_t u16) {
  return (u16 >= THR_A) + (u16 >= THR_C) + (u16 >= THR_G);
}

/* 16 bases (one packed word) starting at base 16*w. */
static inline uint32_t hashed_word(uint64_t seed, uint64_t w) {
  uint32_t out = 0;
  for (int q = 0; q < 4; ++q) {
    const uint64_t h = splitmix64(seed ^ splitmix64(4 * w + q));
    for (int k = 0; k < 4; ++k) {
      const uint32_t u = (uint32_t)(h >> (16 * k)) & 0xFFFFu;
      out |= code_of(u) << (2 * (4 * q + k));
    }
  }
  return out;
}

static inline uint32_t get_base(const uint32_t* words, int64_t i) {
  return (words[i >> 4] >> ((i & 15) * 2)) & 3u;
}

static inline void set_base(uint32_t* words, int64_t i, uint32_t code) {
  const int sh = (int)(i & 15) * 2;
  words[i >> 4] = (words[i >> 4] & ~(3u << sh)) | (code << sh);
}

/*
 * Fill the genome.  Arrays are caller-allocated:
 *   fwd_words, rc_words   uint32[ceil(total_bp / 16)]   (rc_words may be NULL)
 *   nmask_fwd, nmask_rc   uint8[ceil(total_bp / 8)]     little-endian bits (NULL ok)
 *   nmask_big             uint8[ceil(total_bp / 8)]     np.packbits order (NULL ok)
 * N runs: half-open [n_start[k], n_end[k]) global forward coordinates.
 * Plants: plant_codes holds the concatenated forward-oriented 2-bit codes
 * of every plant, plant k of length plant_len[k] starting at plant_pos[k].
 * Returns 0, or -1 on bad arguments.
 */
int bgo_synth_fill(uint64_t seed, int64_t total_bp, const int64_t* contig_off,
                   const int64_t* contig_len, int32_t n_contigs, const int64_t* n_start,
                   const int64_t* n_end, int64_t n_runs, const int64_t* plant_pos,
                   const uint8_t* plant_codes, const int32_t* plant_len, int64_t n_plants,
                   uint32_t* fwd_words, uint32_t* rc_words, uint8_t* nmask_fwd,
                   uint8_t* nmask_rc, uint8_t* nmask_big) {
  if (total_bp < 0 || !fwd_words || n_contigs < 0) return -1;
  const int64_t n_words = (total_bp + 15) / 16;
  const int64_t n_bytes = (total_bp + 7) / 8;

#pragma omp parallel for schedule(static)
  for (int64_t w = 0; w < n_words; ++w) fwd_words[w] = hashed_word(seed, (uint64_t)w);
  /* bases past total_bp read as A (0), like the reference's zero padding */
  if (total_bp % 16) fwd_words[n_words - 1] &= (1u << (2 * (total_bp % 16))) - 1u;

  for (int64_t k = 0; k < n_runs; ++k) {
    const int64_t a = n_start[k] < 0 ? 0 : n_start[k];
    const int64_t b = n_end[k] > total_bp ? total_bp : n_end[k];
    if (b <= a) continue;
#pragma omp parallel for schedule(static)
    for (int64_t w = a >> 4; w <= (b - 1) >> 4; ++w) {
      const int64_t lo = (w << 4) > a ? (w << 4) : a;
      const int64_t hi = ((w + 1) << 4) < b ? ((w + 1) << 4) : b;
      for (int64_t i = lo; i < hi; ++i) set_base(fwd_words, i, 0u);
    }
  }

  int64_t off = 0;
  for (int64_t k = 0; k < n_plants; ++k) {
    for (int32_t j = 0; j < plant_len[k]; ++j) {
      const int64_t i = plant_pos[k] + j;
      if (i >= 0 && i < total_bp) set_base(fwd_words, i, plant_codes[off + j] & 3u);
    }
    off += plant_len[k];
  }

  if (nmask_fwd) memset(nmask_fwd, 0, (size_t)n_bytes);
  if (nmask_big) memset(nmask_big, 0, (size_t)n_bytes);
  if (nmask_rc) memset(nmask_rc, 0, (size_t)n_bytes);
  for (int64_t k = 0; k < n_runs; ++k) {
    const int64_t a = n_start[k] < 0 ? 0 : n_start[k];
    const int64_t b = n_end[k] > total_bp ? total_bp : n_end[k];
    for (int64_t i = a; i < b; ++i) {
      if (nmask_fwd) nmask_fwd[i >> 3] |= (uint8_t)(1u << (i & 7));
      if (nmask_big) nmask_big[i >> 3] |= (uint8_t)(0x80u >> (i & 7));
    }
  }

  for (int32_t c = 0; c < n_contigs; ++c) {
    const int64_t o = contig_off[c], len = contig_len[c];
    if (rc_words) {
      /* rc[o + j] = 3 - fwd[o + len - 1 - j]; parallel over whole rc words
         inside the contig, edges done serially. */
      const int64_t first_w = (o + 15) >> 4, last_w = (o + len) >> 4; /* full words */
      for (int64_t i = o; i < (first_w << 4) && i < o + len; ++i)
        set_base(rc_words, i, 3u - get_base(fwd_words, o + len - 1 - (i - o)));
#pragma omp parallel for schedule(static)
      for (int64_t w = first_w; w < last_w; ++w) {
        uint32_t v = 0;
        for (int q = 0; q < 16; ++q) {
          const int64_t i = (w << 4) + q;
          v |= (3u - get_base(fwd_words, o + len - 1 - (i - o))) << (2 * q);
        }
        rc_words[w] = v;
      }
      for (int64_t i = (last_w << 4) > (first_w << 4) ? (last_w << 4) : (first_w << 4);
           i < o + len; ++i)
        set_base(rc_words, i, 3u - get_base(fwd_words, o + len - 1 - (i - o)));
    }
    if (nmask_rc && nmask_fwd) {
      for (int64_t k = 0; k < n_runs; ++k) {
        const int64_t a = n_start[k] > o ? n_start[k] : o;
        const int64_t b = n_end[k] < o + len ? n_end[k] : o + len;
        for (int64_t i = a; i < b; ++i) {
          const int64_t r = o + len - 1 - (i - o);
          nmask_rc[r >> 3] |= (uint8_t)(1u << (r & 7));
        }
      }
    }
  }
  if (rc_words && total_bp % 16) rc_words[n_words - 1] &= (1u << (2 * (total_bp % 16))) - 1u;
  return 0;
}

/* Decode bases [start, start + n) of a packed array into 2-bit codes. */
void bgo_synth_extract(const uint32_t* words, int64_t start, int64_t n, uint8_t* out) {
#pragma omp parallel for schedule(static) if (n > 1000000)
  for (int64_t i = 0; i < n; ++i) out[i] = (uint8_t)get_base(words, start + i);
}

/*
 * Artefact N mask (np.packbits big bit order, bit i = forward position i)
 * -> little-endian masks of the forward array and of the reverse-complement
 * array (reversed within each contig), as beguidedoff/finder.py:310-345
 * builds them.  Output arrays hold ceil(total_bp / 8) bytes (callers pad).
 * Returns the number of N positions.
 */
int64_t bgo_nmask_convert(const uint8_t* big, int64_t total_bp, const int64_t* contig_off,
                          const int64_t* contig_len, int32_t n_contigs, uint8_t* le_fwd,
                          uint8_t* le_rc) {
  const int64_t n_bytes = (total_bp + 7) / 8;
  int64_t n_set = 0;
#pragma omp parallel for schedule(static) reduction(+ : n_set)
  for (int64_t b = 0; b < n_bytes; ++b) {
    uint8_t v = big[b], r = 0;
    for (int k = 0; k < 8; ++k)
      if (v & (0x80u >> k)) r |= (uint8_t)(1u << k);
    if (b == n_bytes - 1 && (total_bp & 7)) r &= (uint8_t)((1u << (total_bp & 7)) - 1u);
    le_fwd[b] = r;
    n_set += __builtin_popcount(r);
  }
  if (le_rc) {
    memset(le_rc, 0, (size_t)n_bytes);
    for (int32_t c = 0; c < n_contigs; ++c) {
      const int64_t o = contig_off[c], len = contig_len[c];
      /* bytes fully inside the contig in rc coordinates are written in
         parallel; the two edge bytes are shared with neighbours */
      const int64_t first_b = (o + 7) >> 3, last_b = (o + len) >> 3;
      for (int64_t i = o; i < (first_b << 3) && i < o + len; ++i) {
        const int64_t f = o + len - 1 - (i - o);
        if ((le_fwd[f >> 3] >> (f & 7)) & 1u) le_rc[i >> 3] |= (uint8_t)(1u << (i & 7));
      }
#pragma omp parallel for schedule(static)
      for (int64_t b = first_b; b < last_b; ++b) {
        uint8_t r = 0;
        for (int k = 0; k < 8; ++k) {
          const int64_t i = (b << 3) + k;
          const int64_t f = o + len - 1 - (i - o);
          r |= (uint8_t)(((le_fwd[f >> 3] >> (f & 7)) & 1u) << k);
        }
        le_rc[b] = r;
      }
      for (int64_t i = (last_b << 3) > (first_b << 3) ? (last_b << 3) : (first_b << 3);
           i < o + len; ++i) {
        const int64_t f = o + len - 1 - (i - o);
        if ((le_fwd[f >> 3] >> (f & 7)) & 1u) le_rc[i >> 3] |= (uint8_t)(1u << (i & 7));
      }
    }
  }
  return n_set;
}

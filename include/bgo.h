/*
 * bgo.h — C ABI of libbgo_hip.so, the MI355X (gfx950) extensive-mode
 * off-target scan.
 *
 * Plain C types only (no torch, no HIP types).  Every entry point returns an
 * int status: BGO_OK (0) on success, a negative BGO_ERR_* code on failure, in
 * which case bgo_last_error() (thread-local) holds a message.  Python maps any
 * non-zero status to RuntimeError, which is what the reference raises for an
 * unusable GPU (beguidedoff/finder.py:154-155, 283-290) and for too many
 * guides (beguidedoff/cuda_search.py:411-417).
 *
 * Coordinates: a "strand array" is the reference's 2-bit packed genome
 * (16 nt per uint32, base i at bits 2(i%16)..+1 of word i/16, A0 C1 G2 T3,
 * N encoded as A; beguidedoff/encoding.py:36-46) of either the forward genome
 * or the per-contig reverse complement (genome_converter.py:66-71,92-95).
 * N masks are little-endian bit masks, bit i = position i of that strand
 * array (beguidedoff/finder.py:310-345).  Positions are 0-based offsets into
 * the scanned strand array.
 */
#ifndef BGO_H
#define BGO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BGO_OK 0
#define BGO_ERR_ARG (-1)      /* invalid argument (sizes, pointers, limits) */
#define BGO_ERR_HIP (-2)      /* a HIP runtime call failed */
#define BGO_ERR_NODEV (-3)    /* no usable gfx950 device */
#define BGO_ERR_INJECTED (-4) /* failure forced by bgo_debug_fail_next() */

/* Limits of the plain (drop-in) kernel; cuda_search.py:40-41. */
#define BGO_PLAIN_MAX_GUIDES 2048
#define BGO_MAX_PAM_LEN 6

/* ---------------------------------------------------------------- library */

/* ABI version of this header (bumped on any incompatible change). */
int bgo_abi_version(void);

/* Number of visible HIP devices (0 when none; never fails on a host without
 * a GPU).  Replaces the torch.cuda.is_available() probe of
 * beguidedoff/finder.py:279-290. */
int bgo_device_count(int32_t* n_out);

/* Select the device used by every later call of this thread and create the
 * library's per-device state.  Idempotent. */
int bgo_init(int32_t device);

/* Properties of a device: compute units, peak engine clock (kHz), HBM bytes
 * and the gfx architecture name (NUL-terminated, truncated to arch_len). */
int bgo_device_info(int32_t device, int32_t* cu_count, int32_t* clock_khz,
                    int64_t* hbm_bytes, char* arch, int32_t arch_len);

/* Message of the last failed call on this thread ("" when none). */
const char* bgo_last_error(void);

/* Test hook (no reference counterpart): the next `n_calls`-th entry point
 * that would touch the GPU fails with BGO_ERR_INJECTED before doing any work.
 * Lets CPU-only tests check the Python RuntimeError mapping. 0 disarms. */
int bgo_debug_fail_next(int32_t n_calls);

/* ------------------------------------------------------ one strand array */

typedef struct bgo_strand bgo_strand;

/* Upload one strand array and its optional N mask.  Replaces the identity-
 * keyed upload caches _genome_to_device / _n_mask_to_device /
 * _prepare_n_mask_device (beguidedoff/cuda_search.py:321-381): the caller
 * owns the handle instead of a module-level cache.  nmask_le may be NULL (or
 * nmask_bytes 0): then the scan has no N gate, like an empty mask in the
 * reference.  The mask is zero-padded on the device to (total_bp >> 3) + 16
 * bytes. */
int bgo_strand_create(const uint32_t* words, int64_t n_words,
                      const uint8_t* nmask_le, int64_t nmask_bytes,
                      int64_t total_bp, bgo_strand** out);
int bgo_strand_destroy(bgo_strand* s);

/* Plain drop-in scan of one strand array == search_genome_cuda
 * (beguidedoff/cuda_search.py:384-488) with the kernel of
 * beguidedoff/cuda_search.py:163-314: one lane per position, the guide table
 * in LDS, per guide a packed PAM test, a 40-bit XOR-fold-popcount and the
 * N correction, a per-guide atomic append.  Launched per `chunk_size`
 * positions over [0, total_bp - 26).
 *   guides     uint64[n]    2-bit packed guide (first 20 nt are compared)
 *   pam_masks  uint8[n][6]  IUPAC allowed-base nibbles (exotic-PAM fallback)
 *   pats, cmps uint32[n]    encode_pam_packed() fields (cuda_search.py:124-160)
 *   n_guides   <= BGO_PLAIN_MAX_GUIDES, else BGO_ERR_ARG
 *   rpg        results recorded per guide (the count keeps going past it)
 *   n_encoded  2-bit code an N reads as on this strand (0 on '+', 3 on '-')
 * Outputs (host memory, caller-allocated): counts int32[n] exact raw match
 * counts; positions int64[n][rpg] and mismatches int8[n][rpg], of which
 * entries [0, min(count, rpg)) of each row are written. */
int bgo_scan_plain(bgo_strand* s, const uint64_t* guides,
                   const uint8_t* pam_masks, const uint32_t* pats,
                   const uint32_t* cmps, int32_t n_guides,
                   int32_t max_mismatches, int64_t chunk_size, int64_t rpg,
                   int32_t n_encoded, int32_t* counts, int64_t* positions,
                   int8_t* mismatches);

/* Milliseconds the kernels of the last bgo_scan_plain on this thread took
 * (HIP events on the scan stream). */
int bgo_scan_plain_last_ms(float* ms_out);

/* --------------------------------------------------- two-strand genome */

typedef struct bgo_genome bgo_genome;

/* Upload a genome for the batched engine: the forward array and (optional)
 * the reverse-complement array with their N masks, all as above.  The
 * library also derives the bit-plane layout the batched kernel reads (two
 * uint32 planes per 32 bases: bit0 and bit1 of each base code).  Replaces
 * the device copies of beguidedoff/finder.py:299-308 / cuda_search.py:325-362.
 * total_bp must be below 2^32 - 64. */
int bgo_genome_create(const uint32_t* words_fwd, const uint32_t* words_rc,
                      int64_t n_words, const uint8_t* nmask_fwd,
                      const uint8_t* nmask_rc, int64_t nmask_bytes,
                      int64_t total_bp, bgo_genome** out);
int bgo_genome_destroy(bgo_genome* g);
/* Device bytes held by the genome (for admission control). */
int bgo_genome_device_bytes(const bgo_genome* g, int64_t* bytes_out);

/* One guide of a batched scan.  Guides of any PAM (IUPAC, 1-6 nt, 3' of the
 * protospacer), editor and max-mismatch budget may be mixed in one launch;
 * the library groups them by PAM family. */
typedef struct bgo_guide_desc {
  uint64_t seq2bit;      /* encode_guide_64bit(encode_dna(seq)); 20 nt used */
  uint8_t pam_mask[6];   /* encode_pam_mask(pam) (pam.py:38-45) */
  uint8_t pam_len;       /* 1..6 */
  uint8_t max_mm;        /* mismatch budget incl. the N correction, <= 20 */
  uint32_t edit_mask;    /* bit i set: window index i (0-based) is checked */
  uint8_t editor;        /* 0: editable base 'A' (ABE); 1: 'C' (CBE/CGE) */
  uint8_t pad[3];
} bgo_guide_desc;        /* 24 bytes */

/* One reported site.  Fields follow the extensive finder's per-hit values:
 * pos and strand as the kernel reports them (cuda_search.py:310-314), n_mm
 * with the N correction (cuda_search.py:290-308), mm_mask and score from the
 * 2-bit window (finder.py:435-439, scoring.py:5-24), editable as
 * editable_base_in_window() (base_editing.py:9-32) for the guide's editor. */
typedef struct bgo_hit {
  uint32_t guide;    /* index into the launch's guide array */
  uint32_t pos;      /* 0-based position in the scanned strand array */
  uint32_t x0;       /* bit0 planes of bases pos..pos+25 (window + PAM) */
  uint32_t x1;       /* bit1 planes of the same bases */
  uint32_t mm_mask;  /* bit i: window base i differs from guide base i */
  uint8_t strand;    /* 0 '+', 1 '-' */
  uint8_t n_mm;      /* N-corrected mismatch count */
  uint8_t editable;  /* 1 when an editable base lies in the edit window */
  uint8_t pad;
  double score;      /* CCTOP: sum of table[i+1] over set mm_mask bits */
} bgo_hit;           /* 32 bytes */

/* Kernel selection and tiling of the batched engine (0 = library default /
 * autotuned value). */
typedef struct bgo_tuning {
  int32_t chunk_positions;   /* positions per work item (multiple of 64) */
  int32_t waves_per_block;   /* 1, 2 or 4 (engine 4 runs one 16-wave block per CU) */
  int32_t blocks_per_cu;     /* resident blocks requested per CU */
  int32_t engine;            /* 0 auto (VALU, fp4 MFMA or seed, whichever a
                                per-family cost model of the guide counts and
                                PAM pass rates favours), 1 VALU xor/popcount,
                                2 int8 MFMA, 3 fp4 MFMA, 4 seed-and-verify
                                (VALU over LDS tables of the guides' 4-base
                                blocks; every budget must be <= 4) */
  int32_t guide_unroll;      /* output: guides per compare step (8 VALU, 32 MFMA) */
  int32_t reserved[3];       /* [0]: 1 = work items alternate strands;
                                [1]: VALU engine, PAM families of at most this
                                many guides compare their windows in place
                                instead of through the window buffer (1..8;
                                -1 = never; 0 = 8 for PAMs that pass >= 5% of
                                the positions of a 41% GC genome, 4 for
                                sparser ones) */
} bgo_tuning;

typedef struct bgo_plan bgo_plan;

/* Prepare a batched scan: sort/group the guides by PAM family on the host,
 * upload the guide table, allocate a hit buffer of hit_capacity records and a
 * private HIP stream.  score_table[0..20]: the per-position CCTOP penalty,
 * score_table[p] = 1.2 ** p for p = 1..20 as computed by the host.
 * tuning may be NULL. */
int bgo_plan_create(bgo_genome* g, const bgo_guide_desc* guides,
                    int32_t n_guides, const double* score_table,
                    const bgo_tuning* tuning, int64_t hit_capacity,
                    bgo_plan** out);
int bgo_plan_destroy(bgo_plan* p);

/* Replace the guide set of a plan (any size; device tables grow as needed).
 * Waits for the plan's previous launch, then uploads on the plan's stream.
 * Used by the serving loop to admit and retire requests between launches. */
int bgo_plan_set_guides(bgo_plan* p, const bgo_guide_desc* guides,
                        int32_t n_guides);

/* Replace the hit buffer with one of hit_capacity records (waits for the
 * previous launch; the old records are discarded). */
int bgo_plan_set_hit_capacity(bgo_plan* p, int64_t hit_capacity);

/* Positions >= total_bp - end_guard are not scanned.  The default, 26,
 * reproduces the reference's scan range (cuda_search.py:211-213: the
 * extensive mode skips the genome's last 26 positions).  The fast mode sets
 * 20 + its shortest PAM, so every window that fits in the genome is scanned;
 * windows that run past the genome end then read zero padding and are
 * dropped by its contig-internal filter.  Range [21, 64]. */
int bgo_plan_set_end_guard(bgo_plan* p, int32_t end_guard);

/* Enqueue one scan of positions [pos_begin, pos_end) (clamped to
 * [0, total_bp - 26), cuda_search.py:211-213) on the strands of strand_mask
 * (bit0 '+', bit1 '-') on the plan's stream.  Returns without waiting. */
int bgo_plan_launch(bgo_plan* p, int64_t pos_begin, int64_t pos_end,
                    int32_t strand_mask);

/* Wait for the last launch; n_hits_out = exact number of sites found (may
 * exceed the capacity: then only the first `capacity` were stored and the
 * caller re-plans with a larger buffer); kernel_ms_out = kernel time
 * measured with HIP events on the plan's stream (may be NULL). */
int bgo_plan_wait(bgo_plan* p, int64_t* n_hits_out, float* kernel_ms_out);

/* Copy min(n_hits, capacity, cap) records of the last launch to host memory. */
int bgo_plan_fetch(bgo_plan* p, bgo_hit* out, int64_t cap, int64_t* n_copied);

/* The first min(cap, sites, capacity) records (cap < 0: no limit), copied into
 * a page-locked buffer the plan owns (grown as needed; direct DMA, no staging
 * copy).  *out points to it and stays valid until the plan's next
 * fetch_pinned call or its destruction.  No counterpart in the reference: the
 * zero-copy fetch serves loops that consume the records right away (serving,
 * benchmark steps). */
int bgo_plan_fetch_pinned(bgo_plan* p, int64_t cap, const bgo_hit** out, int64_t* n_copied);

/* Capture [counter reset + scan of both strands over the full genome] in a
 * hipGraph once, then replay it with bgo_plan_graph_launch (same wait/fetch).
 * No memset nodes are recorded: counters are cleared by a kernel. */
int bgo_plan_graph_launch(bgo_plan* p, int32_t strand_mask);

/* Resolved tuning of a plan (after defaults/autotune are applied). */
int bgo_plan_get_tuning(const bgo_plan* p, bgo_tuning* out);

/* Number of PAM families of a plan (families are numbered in order of first
 * appearance in the guide array) and, for the first `cap` of them, the number
 * of PAM-passing windows the last launch compared and the number of guides
 * (for roofline accounting). */
int bgo_plan_family_stats(const bgo_plan* p, int32_t cap, int32_t* n_fam_out,
                          int64_t* windows_out /*[cap]*/,
                          int32_t* guides_out /*[cap]*/);

/* Seed engine (tuning engine 4): for the first `cap` families (the families
 * of bgo_plan_family_stats; a PAM family of more than 4,096 guides counts as
 * several), the number of guide-table entries the last launch verified
 * against windows, i.e. (window, guide) pairs that share an exact 4-base
 * block (roofline accounting).  0 for the other engines.  No counterpart in
 * the reference. */
int bgo_plan_seed_stats(const bgo_plan* p, int32_t cap, int64_t* verified_out /*[cap]*/);

/* ------------------------------------------------ table assembly on the device
 * The extensive finder's per-hit epilogue and table order
 * (beguidedoff/finder.py:193-244, 405-472, 545-566) over the last launch's
 * records where they lie in HBM, so only table rows cross PCIe.  Per record
 * (site) of scan key k: forward leftmost and reported position (for '-'
 * records the reverse-complement array is reversed per contig), the
 * on-target exclusion, and, for each input guide mapped to k, the
 * editable-base test; rows are ordered per guide id by (score, chromosome
 * name, reported position, strand, key, input) and cut to the top n_top
 * unless return_all. */
typedef struct bgo_asm_key {
  uint32_t g0, g1;        /* guide planes of the key's first 20 bases (checks mm_mask) */
  int64_t ot_local;       /* on-target local leftmost (0-based), if ot_contig >= 0 */
  int32_t ot_contig;      /* on-target contig index; -1: none (finder.py:208-218) */
  int32_t ot_strand;      /* 0 '+', 1 '-' */
  uint32_t first_input;   /* the key's inputs: inputs[first_input, + n_inputs) */
  uint32_t n_inputs;
  uint32_t pam_len;       /* 1..6 */
  uint32_t pad;
} bgo_asm_key;            /* 40 bytes */

typedef struct bgo_asm_input {
  uint32_t id;            /* guide_id index, numbered by first appearance */
  uint32_t window;        /* edit window mask: bit i = site base i (0-based) */
  uint32_t target;        /* editable base: 0 'A' (ABE), 1 'C' (CBE/CGE) */
  uint32_t order;         /* rank of this (key, input) pair, key-major (< 2^31) */
} bgo_asm_input;          /* 16 bytes */

typedef struct bgo_asm_row {
  bgo_hit hit;            /* the site */
  uint32_t input;         /* index into the inputs array */
  uint32_t rank;          /* 0-based rank within its guide id */
  uint32_t contig;        /* contig of the reported position */
  uint32_t local;         /* 0-based reported position in that contig */
} bgo_asm_row;            /* 48 bytes */

/* keys[n_keys] (n_keys = the plan's guides, in its guide order); inputs
 * grouped by key; contig_name_rank[c] = rank of contig c's name among the
 * sorted distinct names.  threshold is used when use_threshold != 0:
 * below_out[id] = editable sites with score <= threshold (finder.py:466-470).
 * When a key has more than strand_ceiling sites on one strand, nothing is
 * assembled: *max_strand_sites_out > strand_ceiling tells the caller to take
 * the truncating host path.  Rows come ordered by (id, rank); when there
 * are more than rows_cap, *n_rows_out is their number and only rows_cap are
 * stored.  The last launch must have stored all its sites (BGO_ERR_ARG
 * otherwise). */
int bgo_plan_assemble(bgo_plan* p, const bgo_asm_key* keys, int32_t n_keys,
                      const bgo_asm_input* inputs, int32_t n_inputs, int32_t n_ids,
                      const int64_t* contig_offsets, const int64_t* contig_lengths,
                      const uint32_t* contig_name_rank, int32_t n_contigs, int32_t n_top,
                      int32_t return_all, double threshold, int32_t use_threshold,
                      int64_t strand_ceiling, bgo_asm_row* rows, int64_t rows_cap,
                      int64_t* n_rows_out, int64_t* below_out, int64_t* max_strand_sites_out);

/* One-shot convenience: create a plan, scan both strands of the whole
 * genome, fetch, destroy.  Returns the exact count in n_hits_out; stores
 * min(n_hits, out_capacity) records. */
int bgo_scan_batched(bgo_genome* g, const bgo_guide_desc* guides,
                     int32_t n_guides, const double* score_table,
                     const bgo_tuning* tuning, bgo_hit* out,
                     int64_t out_capacity, int64_t* n_hits_out);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same scan, arguments and results as bgo_scan_batched,
 * but every device buffer its kernels use (guide tables, counters, hit
 * records) lives in a fixed array of the library's, 9 MiB per device, held
 * until the process ends; the call allocates no device or page-locked
 * memory.  Requires bgo_init(device) of the genome's device first (it
 * resolves that device's array).  Limits: n_guides <= BGO_NOALLOC_MAX_GUIDES
 * in <= BGO_NOALLOC_MAX_FAMILIES PAM families, out_capacity <=
 * BGO_NOALLOC_MAX_HITS; larger requests fail with BGO_ERR_ARG.  Calls on one
 * device share its array, which keeps the last call's contents: the library
 * serialises them with a lock, so threads that call at once wait for each
 * other instead of racing on it. */
#define BGO_NOALLOC_MAX_GUIDES 4096
#define BGO_NOALLOC_MAX_FAMILIES 64
#define BGO_NOALLOC_MAX_HITS 262144
int bgo_scan_batched_noalloc(bgo_genome* g, const bgo_guide_desc* guides,
                             int32_t n_guides, const double* score_table,
                             const bgo_tuning* tuning, bgo_hit* out,
                             int64_t out_capacity, int64_t* n_hits_out);

/* -------------------------------------------------------- fast-mode index
 * Anchor index of a resident genome for the fast mode; replaces the
 * reference's FM-index (fast/fmindex.py, fast/finder.py:149-159).  Every
 * forward position is keyed by its 12 bases (2-bit codes, first base in bits
 * 22-23, N read as A); positions are grouped by key on the GPU.  The index
 * keeps a pointer to `g`, which must outlive it. */
typedef struct bgo_fast_index bgo_fast_index;

/* One anchor of one guide (fast/finder.py:186-189): its occurrences are the
 * positions with keys in [key_lo, key_hi); each occurrence's window is
 * checked on the GPU against the whole cold block (site positions
 * cold_first..19 equal to a0/a1), the PAM, and score <= cutoff over the
 * 2-bit mismatches against g0/g1 plus the N positions in n_code. */
typedef struct bgo_fast_anchor {
  uint32_t key_lo, key_hi;  /* key range of the anchor's first bases */
  uint32_t a0, a1;          /* anchor planes on site positions cold_first..19 */
  uint32_t g0, g1;          /* guide planes, site positions 0..19 */
  uint32_t n_code;          /* site positions whose guide base is the strand's N code */
  uint32_t pam_nibs;        /* allowed-base nibble of PAM position i at bits 4i */
  uint32_t guide;           /* caller's guide index, copied to the hit record */
  uint8_t strand;           /* 0 '+', 1 '-' */
  uint8_t pam_len;          /* 1..6 */
  uint8_t cold_first;       /* first cold site position (0-based) */
  uint8_t pad;
} bgo_fast_anchor;          /* 40 bytes */

/* Build the index on g's device (contigs must tile the genome). */
int bgo_fast_index_create(bgo_genome* g, const int64_t* contig_offsets,
                          const int64_t* contig_lengths, int32_t n_contigs,
                          bgo_fast_index** out);
int bgo_fast_index_destroy(bgo_fast_index* x);
int bgo_fast_index_bytes(const bgo_fast_index* x, int64_t* bytes_out);

/* Candidates of n_anchors anchors as bgo_hit records: pos and x0/x1 on the
 * scanned strand (as the batched scan reports them), mm_mask = 2-bit
 * mismatches plus N positions, n_mm its popcount, score its CCTOP score
 * (score_table[p] = 1.2**p, summed in ascending p).  Returns the exact count
 * in n_hits_out and stores min(count, out_capacity) records (order
 * unspecified).  The index keeps its query buffers between calls (grown on
 * demand, freed by bgo_fast_index_destroy), so queries on one index must not
 * overlap. */
int bgo_fast_index_query(bgo_fast_index* x, const bgo_fast_anchor* anchors,
                         int32_t n_anchors, double cutoff, const double* score_table,
                         bgo_hit* out, int64_t out_capacity, int64_t* n_hits_out,
                         float* kernel_ms);

/* The index's bucket table, copied to host memory: out[key] = first entry
 * of key's positions, out[2^24] = number of positions (uint32[2^24 + 1]).
 * With bgo_fast_index_positions it lists an anchor's occurrences, which the
 * reference's FM-index gives as an SA interval (csrc/fmindex_binding.cpp:
 * 141-176). */
int bgo_fast_index_offsets(bgo_fast_index* x, uint32_t* out);
/* Entries [begin, end) of the index's position array (forward genome
 * positions grouped by key; order within a key unspecified). */
int bgo_fast_index_positions(bgo_fast_index* x, int64_t begin, int64_t end, uint32_t* out);

/* --------------------------------------------- anchor suffix-array order
 * The reference's fast mode keeps the first max_hits_per_anchor entries of
 * an anchor's suffix-array interval in its FM-index text
 * (csrc/fmindex_binding.cpp:164-176, fast/finder.py:149-159): contigs as
 * symbols A C G T = 1 2 3 4 (N and other bytes = 1), a separator 6 between
 * contigs, a terminal 0 (fast/fmindex.py:63-128, fast/backend.py:28-49).
 * bgo_sa_text holds a 2-bit forward genome (caller-owned words, which must
 * outlive it; contigs non-empty and tiling the genome) and its maximal runs
 * of >= 64 equal bases.  Host code: no GPU is used. */
typedef struct bgo_sa_text bgo_sa_text;
int bgo_sa_text_create(const uint32_t* words, int64_t n_words, int64_t total_bp,
                       const int64_t* contig_offsets, const int64_t* contig_lengths,
                       int32_t n_contigs, bgo_sa_text** out);
int bgo_sa_text_destroy(bgo_sa_text* t);
/* Number of maximal runs of >= 64 equal bases in the genome. */
int bgo_sa_text_long_runs(const bgo_sa_text* t, int64_t* n_out);
/* Of `positions`, the occurrences of the m-base anchor (codes 0..3, m <= 32)
 * that lie inside one contig (n_occ_out of them); if there are more than
 * cap (>= 1), the cap occurrences whose suffixes of the reference's text are
 * smallest, else all.  out (capacity cap) receives them in ascending
 * position order, n_out their number. */
int bgo_sa_text_select(const bgo_sa_text* t, const uint32_t* positions, int64_t n,
                       const uint8_t* anchor, int32_t m, int64_t cap, uint32_t* out,
                       int64_t* n_occ_out, int64_t* n_out);

/* ---------------------------------------------------------- genome encoder
 * FASTA sequence bytes -> genome artefact arrays on `device`; replaces the
 * per-base loops of genome_converter.convert (genome_converter.py:15-108).
 * seq: the contigs' bytes back to back (any case); contig k occupies
 * [contig_offsets[k], +contig_lengths[k]), and the contigs must tile
 * [0, n_bytes).  Outputs (host arrays, each may be NULL):
 *   words_fwd[ceil(n/16)]: upper-cased A C G T -> 0 1 2 3, other bytes -> 0,
 *                          base i of a word at bits 2i (encoding.py:36-46);
 *   words_rc[ceil(n/16)]:  per-contig reverse complement (3 - code);
 *   n_mask[ceil(n/8)]:     np.packbits of (upper-cased byte == 'N').
 * kernel_ms (may be NULL) receives the encode kernels' time. */
int bgo_encode_genome(int32_t device, const uint8_t* seq, int64_t n_bytes,
                      const int64_t* contig_offsets, const int64_t* contig_lengths,
                      int32_t n_contigs, uint32_t* words_fwd, uint32_t* words_rc,
                      uint8_t* n_mask, float* kernel_ms);

#ifdef __cplusplus
}
#endif

#endif /* BGO_H */

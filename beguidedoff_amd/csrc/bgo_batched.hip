// Batched persistent scan engine for gfx950.
//
// Computes the same sites as the reference extensive kernel
// (beguidedoff/cuda_search.py:163-314) for any mix of guides, PAMs, editors
// and mismatch budgets in one launch, and fuses the finder's per-hit
// epilogue (beguidedoff/finder.py:435-470: 2-bit mismatch positions, CCTOP
// score, editable-window test).
//
// Layout.  Each strand array is re-laid out once as bit planes: for every
// 32-base block b, planes[b] = {P0, P1} where bit j of P0 (P1) is bit 0
// (bit 1) of the 2-bit code of base 32b + j.  A 20-nt window then is two
// 20-bit words, and a guide/window mismatch mask is (W0^g0)|(W1^g1) -- one
// v_xor + one v_bitop3 -- followed by one v_bcnt.  The N mask is used as
// uint32 words of the reference's little-endian bytes.
//
// Work decomposition.  Persistent waves pull work items (a strand and a run
// of `chunk` positions) from an atomic counter.  Each lane of a wave owns a
// 32-position block: from the plane words it tests the IUPAC PAM of every PAM
// family of the launch for all 32 positions at once (one bit-select per PAM
// base), and the passing windows are appended to per-family rings in LDS
// (every lane writes its next window in each step; slots by ballot + mbcnt).
// Whenever a ring holds a batch, the wave compares the batch against all
// guides of that family with one of three engines: VALU xor/bitop3/popcount
// (8 wave-uniform guides per scalar load, min-reduced), or the matrix cores
// over a tetrahedral +-1 code of the bases (int8 or fp4 MFMA, sign-bit
// reduction).  Candidates take the exact path: N correction, per-guide
// budget, fused epilogue and a wave-aggregated append of hit records.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <mutex>
#include <vector>

#include "bgo_common.h"

namespace bgo {
namespace bk {

constexpr int kMaxFam = 8;     // PAM families per kernel launch
constexpr int kUnroll = 8;     // guides per compare step
constexpr int kDirectMax = 8;  // guides of a family compared in place (VALU engine)
constexpr int kDirectSparse = 4;  // default limit for PAMs passing < 5% of positions
// Records a wave of an MFMA engine stages in LDS before one atomic reserves
// their slots (same-address device atomics serialise at ~12 ns each, and a
// 1.5 M-site scan would otherwise take one per candidate event).
constexpr uint32_t kRecStage = 64;
constexpr int kMaxBlock = 256; // threads per block (4 waves)
constexpr uint32_t kWin = 0xFFFFFu;
// Work items are handed out by kWorkCtrs counters per launch group, each in a
// 256-byte slot of its own (kCtrStride 32-bit words): same-address device
// atomics serialise at about 12 ns each on gfx950, so one counter for the
// whole grid costs ~0.1 ms to start 8192 waves and bounds a scan to one item
// per 12 ns.
constexpr uint32_t kWorkCtrs = 16;
constexpr uint32_t kCtrStride = 64;

struct Fam {
  uint32_t nibs;            // allowed-base nibble of PAM position i at bits 4i..4i+3 (N past the PAM)
  uint32_t plen;            // PAM length 1..6
  uint32_t g_begin, g_end;  // padded guide range
  uint32_t kmax;            // largest budget among the family's guides
  uint32_t g_live;          // end of the family's real (unpadded) guides
  uint32_t direct;          // 1: windows compared in place, without the buffer
  uint32_t seed_start;      // seed engine: word offset of the key starts in the block's table
  uint32_t seed_ent;        // seed engine: word offset of the entries in the block's table
  uint32_t seed_n;          // seed engine: the family's table entries (bucket pads included)
};

struct Args {
  const uint2* planes0;
  const uint2* planes1;
  const uint32_t* nwords0;
  const uint32_t* nwords1;
  uint32_t slot_strand0, slot_strand1;  // strand scanned by enumeration slot 0/1
  uint32_t valid_begin, valid_end;
  uint32_t chunk, chunks_per_slot, n_items, n_fam;
  uint32_t n_slots, interleave;  // interleave: consecutive items alternate strands
  uint32_t btab_bytes;           // size of the MFMA guide operand table (seed engine: of the group's table)
  uint32_t fam_total;            // families of the plan: per-family verify counts follow the window counts
  unsigned long long hit_cap;
};

// Read-only tables and outputs.  They reach the kernel as separate
// __restrict__ parameters so wave-uniform reads of the guide and family
// tables compile to scalar (s_load) loads.
struct Tab {
  const Fam* __restrict__ fam;           // [n_fam]
  const uint32_t* __restrict__ g01;      // [2 * n_pad]: g0, g1 planes per guide
  const int32_t* __restrict__ gk;        // [n_pad] budget, -1 for padding
  const uint32_t* __restrict__ gaux;     // [n_pad] edit window mask | editor << 24
  const uint32_t* __restrict__ gid;      // [n_pad] caller's guide index
  const uint32_t* __restrict__ gnm;      // [2 * n_pad] guide positions equal to the N code
  const double* __restrict__ score_tab;  // [21]
  bgo_hit* __restrict__ hits;
  unsigned long long* __restrict__ hit_count;
  const int4* __restrict__ btab;         // MFMA engines: guide operand tiles
  bgo_hit* rec_stage;                    // MFMA and seed engines: the wave's staged records (LDS), else null
  uint32_t* rec_cnt;                     // records staged (LDS)
  const uint32_t* runs;                  // [2][2] (LDS): position base and strand of the
                                         // runs whose windows the buffers hold, by entry flag
  const uint32_t* stab;                  // seed engine: the block's seed table (LDS)
  const uint32_t* __restrict__ sgix;     // seed engine: guide of each table word
};

__device__ inline uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Per-family window buffers of a wave: windows are appended at the end and
// compared in batches of ring_flush(engine) windows from the front (64 = one
// per lane for the VALU compare and the int8 engine's two A tiles; 128 = four
// fp4 A tiles, so each guide operand load serves 128 windows).  A buffer
// holds one batch plus one append pass (64 windows); after a batch the at
// most 64 windows behind it move to the front, so indices never wrap.  A slot
// is {b0 window bits | offset bits 0..11 << 20, b1 window bits | offset bits
// 12..22 << 20 | run flag << 31}, the offset counted from the 32-aligned start
// of the run (work item) that appended it.  A wave carries the windows left in
// its buffers at the end of a run into its next run, so the entries of two
// consecutive runs can share a batch: the flag (the parity of the wave's run
// count) picks the run's position base and strand from Tab::runs.
__host__ __device__ constexpr uint32_t ring_flush(int engine) { return engine == 3 ? 128 : 64; }
__host__ __device__ constexpr uint32_t ring_size(int engine) { return ring_flush(engine) + 64; }

// Bytes of LDS one wave uses with nf families (see the kernel's layout).
__host__ __device__ constexpr size_t wave_lds_bytes(int nf, int engine) {
  return static_cast<size_t>(nf) * ring_size(engine) * 8 + kMaxFam * 8 +
         static_cast<size_t>(nf) * 64 * 4 + kMaxFam * 4 + 16 + 64 * 8 +
         (engine >= 2 ? kRecStage * sizeof(bgo_hit) + 16 : 0) +
         (engine == 4 ? static_cast<size_t>(nf) * 64 * 4 : 0);
}

__device__ inline uint32_t slot_pos(uint2 e) { return (e.x >> 20) | (((e.y >> 20) & 0x7FFu) << 12); }

// Position and strand of a buffered window.
__device__ inline uint32_t entry_pos(const uint32_t* runs, uint2 e) {
  return runs[2 * (e.y >> 31)] + slot_pos(e);
}
__device__ inline uint32_t entry_strand(const uint32_t* runs, uint2 e) {
  return runs[2 * (e.y >> 31) + 1];
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The fully scored record of a site (the finder's per-hit epilogue,
// finder.py:435-470): window and PAM bit planes, 2-bit mismatch mask, CCTOP
// score and the editable flag.
__device__ inline bgo_hit make_record(const Args& a, const Tab& t, uint32_t s, uint32_t gi,
                                      uint32_t pos, uint32_t W0, uint32_t W1, int nmm) {
  const uint32_t g0 = t.g01[2 * gi], g1 = t.g01[2 * gi + 1];
  const uint2* planes = s ? a.planes1 : a.planes0;
  const uint32_t w = pos >> 5, sh = pos & 31u;
  const uint2 qa = planes[w], qb = planes[w + 1];
  bgo_hit h;
  h.guide = t.gid[gi];
  h.pos = pos;
  h.x0 = funnel(qb.x, qa.x, sh) & 0x3FFFFFFu;
  h.x1 = funnel(qb.y, qa.y, sh) & 0x3FFFFFFu;
  uint32_t mm = ((W0 ^ g0) | (W1 ^ g1)) & kWin;
  h.mm_mask = mm;
  h.strand = static_cast<uint8_t>(s);
  h.n_mm = static_cast<uint8_t>(nmm);
  const uint32_t aux = t.gaux[gi];
  const uint32_t amask = ~(W0 | W1) & kWin;  // window bases == A
  const uint32_t cmask = W0 & ~W1 & kWin;    // window bases == C
  h.editable = (((aux >> 24) & 1u ? cmask : amask) & aux & kWin) != 0;
  h.pad = 0;
  double sc = 0.0;  // ascending positions, as sum(1.2 ** p for p in positions)
  while (mm) {
    const int i = __builtin_ctz(mm);
    mm &= mm - 1;
    sc += t.score_tab[i + 1];
  }
  h.score = sc;
  return h;
}

// Append the wave's staged records (all lanes call this): one atomic reserves
// their slots.  Records past the buffer's capacity are counted, not stored.
__device__ inline void flush_records(const Args& a, const Tab& t, uint32_t lane) {
  wave_sync();
  const uint32_t n = uniform(*t.rec_cnt);
  if (n == 0) return;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(t.hit_count, static_cast<unsigned long long>(n));
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(base), 0);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(base >> 32), 0);
  base = (static_cast<unsigned long long>(hi) << 32) | lo;
  if (lane < n && base + lane < a.hit_cap) t.hits[base + lane] = t.rec_stage[lane];
  wave_sync();
  if (lane == 0) *t.rec_cnt = 0;
  wave_sync();
}

// Exact path for a (window, guide) pair whose 2-bit mismatch count c is within
// the guide's budget: add the N correction (cuda_search.py:290-308), re-test,
// and append a fully scored record.  All lanes of the wave call this.  The
// MFMA engines stage records in LDS (kRecStage per wave); the VALU engine
// reserves slots with one atomic per call.
__device__ inline void record_candidates(const Args& a, const Tab& t, uint32_t s, uint32_t gi,
                                         uint32_t pos, uint32_t W0, uint32_t W1, uint32_t c,
                                         bool cand, uint32_t lane) {
  const int k = t.gk[gi];
  int nmm = static_cast<int>(c);
  bool rec = false;
  if (cand) {
    const uint32_t* nw = s ? a.nwords1 : a.nwords0;
    if (nw) {
      const uint32_t w = pos >> 5, sh = pos & 31u;
      const uint32_t nb = funnel(nw[w + 1], nw[w], sh) & kWin;
      if (nb) nmm += __popc(nb & t.gnm[2 * gi + s]);
    }
    rec = nmm <= k;
  }
  const uint64_t bal = ballot(rec);
  if (!bal) return;
  if (t.rec_stage) {
    const uint32_t n = static_cast<uint32_t>(__popcll(bal));
    uint32_t cnt = uniform(*t.rec_cnt);
    if (cnt + n > kRecStage) {
      flush_records(a, t, lane);
      cnt = 0;
    }
    if (rec) t.rec_stage[cnt + lanes_below(bal)] = make_record(a, t, s, gi, pos, W0, W1, nmm);
    wave_sync();
    if (lane == 0) *t.rec_cnt = cnt + n;
    wave_sync();
    return;
  }
  const uint32_t leader = static_cast<uint32_t>(__builtin_ctzll(bal));
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(t.hit_count, static_cast<unsigned long long>(__popcll(bal)));
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(base), leader);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(base >> 32), leader);
  base = (static_cast<unsigned long long>(hi) << 32) | lo;
  if (!rec) return;
  const unsigned long long idx = base + lanes_below(bal);
  if (idx >= a.hit_cap) return;
  t.hits[idx] = make_record(a, t, s, gi, pos, W0, W1, nmm);
}

// Compare the first n (<= 64) windows of a family ring with all guides of
// the family.  One window per lane; guides are wave-uniform.
__device__ inline void compare_list(const Args& a, const Tab& t, const Fam& F, const uint2* L,
                                    uint32_t n, uint32_t lane) {
  if (lane >= n) return;
  const uint2 e = L[lane];
  const uint32_t W0 = e.x & kWin, W1 = e.y & kWin;
  const uint32_t kmax = F.kmax;
  const uint32_t* __restrict__ g01 = t.g01;
  for (uint32_t g = F.g_begin; g < F.g_end; g += kUnroll) {
    // 8 guides = 64 aligned bytes, one wide scalar load.
    const uint4* __restrict__ gq = reinterpret_cast<const uint4*>(g01 + 2 * g);
    const uint4 q0 = gq[0], q1 = gq[1], q2 = gq[2], q3 = gq[3];
    uint32_t c[kUnroll];
    c[0] = __popc((W0 ^ q0.x) | (W1 ^ q0.y));
    c[1] = __popc((W0 ^ q0.z) | (W1 ^ q0.w));
    c[2] = __popc((W0 ^ q1.x) | (W1 ^ q1.y));
    c[3] = __popc((W0 ^ q1.z) | (W1 ^ q1.w));
    c[4] = __popc((W0 ^ q2.x) | (W1 ^ q2.y));
    c[5] = __popc((W0 ^ q2.z) | (W1 ^ q2.w));
    c[6] = __popc((W0 ^ q3.x) | (W1 ^ q3.y));
    c[7] = __popc((W0 ^ q3.z) | (W1 ^ q3.w));
    uint32_t m = min(min(min(c[0], c[1]), min(c[2], c[3])), min(min(c[4], c[5]), min(c[6], c[7])));
    if (ballot(m <= kmax)) {
      // Rare: recompute per guide so the exact path exists once in the code.
#pragma nounroll
      for (uint32_t i = 0; i < kUnroll; ++i) {
        const uint32_t gi = g + i;
        const uint32_t ci = __popc((W0 ^ g01[2 * gi]) | (W1 ^ g01[2 * gi + 1]));
        const bool cand = static_cast<int>(ci) <= t.gk[gi];
        // the window's position and strand only here (they cost a read of
        // the run slots)
        if (ballot(cand))
          record_candidates(a, t, entry_strand(t.runs, e), gi, entry_pos(t.runs, e), W0, W1, ci,
                            cand, lane);
      }
    }
  }
}

// Compare this lane's passing windows (bits of pw in block bb) in place with
// the NG (<= kDirectMax = kUnroll) guides of family F: the guides take wide
// scalar loads (the family's table is padded to kUnroll; guides 5..8 move to
// VGPRs, which the VALU engine has to spare, rather than add SGPRs), and each
// pass tests the next window of every lane that has one against them, so
// the family needs no window buffer, appends or padded compares.  Returns
// the number of windows compared.
template <int NG>
__device__ inline uint32_t direct_passes(const Args& a, const Tab& t, const Fam& F, uint32_t s,
                                         uint32_t pw, uint32_t bb, uint2 qa, uint2 qb,
                                         uint32_t lane) {
  const uint32_t kmax = F.kmax, g_begin = F.g_begin;
  const uint4* __restrict__ gq = reinterpret_cast<const uint4*>(t.g01 + 2 * g_begin);
  const uint4 q0 = gq[0], q1 = gq[1];
  uint4 q2 = make_uint4(0, 0, 0, 0), q3 = q2;
  if constexpr (NG > 4) {
    q2 = gq[2];
    q3 = gq[3];
    asm volatile("" : "+v"(q2.x), "+v"(q2.y), "+v"(q2.z), "+v"(q2.w));
    asm volatile("" : "+v"(q3.x), "+v"(q3.y), "+v"(q3.z), "+v"(q3.w));
  }
  const uint32_t gx[8] = {q0.x, q0.z, q1.x, q1.z, q2.x, q2.z, q3.x, q3.z};
  const uint32_t gy[8] = {q0.y, q0.w, q1.y, q1.w, q2.y, q2.w, q3.y, q3.w};
  uint32_t n = 0;
  for (uint64_t act = ballot(pw != 0); act; act = ballot(pw != 0)) {
    n += static_cast<uint32_t>(__popcll(act));
    // A lane without a window reads the one at bit 31 (pw | bit 31 is never
    // empty); the exact path below records windows of lanes with pw != 0 only.
    const uint32_t j = static_cast<uint32_t>(__builtin_ctz(pw | 0x80000000u));
    const uint32_t W0 = funnel(qb.x, qa.x, j) & kWin, W1 = funnel(qb.y, qa.y, j) & kWin;
    uint32_t m = __popc((W0 ^ gx[0]) | (W1 ^ gy[0]));
#pragma unroll
    for (int i = 1; i < NG; ++i)
      m = min(m, static_cast<uint32_t>(__popc((W0 ^ gx[i]) | (W1 ^ gy[i]))));
    if (ballot(m <= kmax)) {  // rare: per-guide budgets and the exact path
      const bool has = pw != 0;
      const uint32_t pos = (bb << 5) + j;
#pragma nounroll
      for (uint32_t i = 0; i < NG; ++i) {
        const uint32_t g = g_begin + i;
        const uint32_t ci = __popc((W0 ^ t.g01[2 * g]) | (W1 ^ t.g01[2 * g + 1]));
        const bool cand = has && static_cast<int>(ci) <= t.gk[g];
        if (ballot(cand)) record_candidates(a, t, s, g, pos, W0, W1, ci, cand, lane);
      }
    }
    pw &= pw - 1;
  }
  return n;
}

__device__ inline uint32_t direct_family(const Args& a, const Tab& t, const Fam& F, uint32_t s,
                                         uint32_t pw, uint32_t bb, uint2 qa, uint2 qb,
                                         uint32_t lane) {
  static_assert(kDirectMax == 8 && kUnroll == 8, "direct_family dispatches 1..8 guides");
  switch (F.g_live - F.g_begin) {
    case 1: return direct_passes<1>(a, t, F, s, pw, bb, qa, qb, lane);
    case 2: return direct_passes<2>(a, t, F, s, pw, bb, qa, qb, lane);
    case 3: return direct_passes<3>(a, t, F, s, pw, bb, qa, qb, lane);
    case 4: return direct_passes<4>(a, t, F, s, pw, bb, qa, qb, lane);
    case 5: return direct_passes<5>(a, t, F, s, pw, bb, qa, qb, lane);
    case 6: return direct_passes<6>(a, t, F, s, pw, bb, qa, qb, lane);
    case 7: return direct_passes<7>(a, t, F, s, pw, bb, qa, qb, lane);
    default: return direct_passes<8>(a, t, F, s, pw, bb, qa, qb, lane);
  }
}

// Bases of a PAM position whose 2-bit code (x1 x0) is in the allowed set
// `nib` (bit b = code b allowed), for 32 positions at once.  nib is
// wave-uniform; the select x1 ? (x0 ? n3 : n2) : (x0 ? n1 : n0) over its bit
// masks needs no branches.
__device__ inline uint32_t allowed(uint32_t nib, uint32_t x0, uint32_t x1) {
  const uint32_t n0 = 0u - (nib & 1u), n1 = 0u - ((nib >> 1) & 1u);
  const uint32_t n2 = 0u - ((nib >> 2) & 1u), n3 = 0u - ((nib >> 3) & 1u);
  const uint32_t lo = n0 ^ (x0 & (n0 ^ n1));
  const uint32_t hi = n2 ^ (x0 & (n2 ^ n3));
  return (x1 & hi) | (~x1 & lo);
}

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

// MFMA engines of compare_list.  Windows and guides are coded as 64-element
// vectors over the tetrahedral code of the 2-bit bases: one element for each
// of b1, b0 and b0^b1 of every base (elements 0..19, 20..39, 40..59).  Two
// bases agree in all three elements or differ in exactly two, so the codes of
// a window and a guide differ in 2 x (2-bit mismatches) elements.  Elements
// 60..63 carry a bias that turns "within budget" into the sign of the
// product.  One wave multiplies up to ring_flush(ENGINE) buffered windows
// (32-row A tiles, row = lane % 32) by 32-guide B tiles (column = lane % 32);
// a sign-bit AND-reduction over the lane's accumulators settles the common
// "no site" case, and non-negative entries take the exact path.
//  * ENGINE 2, int8 (v_mfma_i32_32x32x32_i8, two steps of K = 32): elements
//    are +1 (bit 0) / -1 (bit 1); guide elements 60..63 hold kmax - 15 and
//    window ones +1, so a product is 4 * (kmax - mismatches).
//  * ENGINE 3, fp4 e2m1 (v_mfma_scale_f32_32x32x64_f8f6f4, K = 64 in one
//    step, twice the int8 rate): every code element contributes
//    0.5 * bit * (+1 if the guide's bit is 1, else -1).  With window bits a
//    and guide bits b, sum(a (2b - 1)) / 2 = |b| / 2 - mm and |b| = 2 x
//    (non-A guide bases), so the code part is (non-A guide bases) - mm; four
//    bias elements add budget - (non-A guide bases), and the f32 product is
//    exactly budget - mm.  Element i of lane group g sits in dword i % 4 at
//    nibble i / 4; the window's element is its bit at nibble bit i % 4 (value
//    0.5, 1.0 or 2.0; i % 4 = 3 is shifted down to 0.5), the guide's weight
//    +-1 / value, stored doubled (+-2, +-1, +-0.5) under a B scale of 2^-1.
//    The bias elements are 2.0 in the window, so the guide's four stored
//    values sum to the bias.
// Padding guides get a negative bias, so they never reach the exact path.
// Lane layouts (checked on gfx950): A/B row or column = lane % 32; lane group
// lane / 32 holds half of K and the hardware pairs equal (group, slot) of A
// and B; D register r of lane l is row 8*(r/4) + 4*(l/32) + r%4, column
// l % 32.

// 16 bits -> 16 int8 of +1 (bit 0) / -1 (bit 1), 4 per dword.
__device__ inline v4i expand16(uint32_t bits) {
  v4i r;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t spread = (((bits >> (4 * q)) & 0xFu) * 0x00204081u) & 0x01010101u;
    r[q] = static_cast<int>(__builtin_amdgcn_perm(0u, 0x0000FF01u, spread));
  }
  return r;
}

// fp4 A operand dword d of a 32-bit code word c, OR-ed with the pad
// pattern: nibble k holds bit 4k + d of c, in place for d = 0, 1, 2 (e2m1
// 0.5, 1.0 and 2.0) and shifted to the nibble's low bit for d = 3 (0.5), so
// three of the four dwords take one AND each.
__device__ inline uint32_t fp4_dword(uint32_t c, uint32_t d, uint32_t pad) {
  if (d == 3) return ((c >> 3) & 0x11111111u) | pad;
  return (c & (0x11111111u << d)) | pad;
}

// Exact check of the (window, guide) pairs of one guide tile whose
// accumulator is non-negative: bit r (16 + r) of `bits` = register r of the
// product with A tile abase (abase + 1).
template <int ENGINE>
__device__ inline void mfma_exact(const Args& a, const Tab& t, const uint2* L, uint32_t n,
                                  uint32_t lane, uint32_t bits, uint32_t gt, uint32_t abase) {
  const uint32_t m = lane & 31u, g = lane >> 5;
  while (ballot(bits != 0)) {
    bool cand = false;
    uint32_t gi = 0, pos = 0, s = 0, W0 = 0, W1 = 0, c = 0;
    if (bits) {
      const uint32_t bit = __builtin_ctz(bits);
      bits &= bits - 1;
      const uint32_t r = bit & 15u, tile = bit >> 4;
      const uint32_t row = 8u * (r >> 2) + 4u * g + (r & 3u);
      const uint32_t widx = 32u * (abase + tile) + row;
      gi = gt + m;
      if (widx < n) {
        const uint2 e = L[widx];
        W0 = e.x & kWin;
        W1 = e.y & kWin;
        pos = entry_pos(t.runs, e);
        s = entry_strand(t.runs, e);
        c = __popc((W0 ^ t.g01[2 * gi]) | (W1 ^ t.g01[2 * gi + 1]));
        cand = static_cast<int>(c) <= t.gk[gi];
      }
    }
    if (ballot(cand)) record_candidates(a, t, s, gi, pos, W0, W1, c, cand, lane);
  }
  // Leave no memory operation of this rare path outstanding: the common path
  // keeps a guide operand load in flight across this join, and a pending
  // load here would make the compiler wait for everything at the join.
  __builtin_amdgcn_s_waitcnt(0);
}

// Sign-bit test of a guide tile's products with A tiles abase and abase + 1
// (as 32-bit patterns): the common case returns after one AND-reduction and
// one ballot.
template <int ENGINE>
__device__ inline void mfma_signs(const Args& a, const Tab& t, const uint2* L, uint32_t n,
                                  uint32_t lane, const v16i& r0, const v16i& r1, uint32_t gt,
                                  uint32_t abase) {
  uint32_t all = static_cast<uint32_t>(r0[0]);
#pragma unroll
  for (int r = 1; r < 16; ++r) all &= static_cast<uint32_t>(r0[r]);
#pragma unroll
  for (int r = 0; r < 16; ++r) all &= static_cast<uint32_t>(r1[r]);
  if (!ballot(static_cast<int>(all) >= 0)) return;
  // Sign bits shifted in from the top register down, so bit r (16 + r) is
  // register r of r0 (r1).  A shift-in chain needs only inline constants;
  // OR-ing 1 << r per register would keep 32 such constants in VGPRs for the
  // whole kernel (VOP3 takes no literals here).
  uint32_t neg = 0;
#pragma unroll
  for (int r = 15; r >= 0; --r) neg = (neg << 1) | (static_cast<uint32_t>(r1[r]) >> 31);
#pragma unroll
  for (int r = 15; r >= 0; --r) neg = (neg << 1) | (static_cast<uint32_t>(r0[r]) >> 31);
  mfma_exact<ENGINE>(a, t, L, n, lane, ~neg, gt, abase);
}

// int8: one 32-guide tile against the (up to) two 32-window A tiles.
__device__ inline void i8_tile(const Args& a, const Tab& t, const uint2* L, uint32_t n,
                               uint32_t lane, bool two,
                               const v4i (&A)[2][2], int4 b0, int4 b1, uint32_t gt) {
  const v16i zero = {};
  const v4i B0 = {b0.x, b0.y, b0.z, b0.w}, B1 = {b1.x, b1.y, b1.z, b1.w};
  v16i acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][0], B0, zero, 0, 0, 0);
  acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][1], B1, acc0, 0, 0, 0);
  v16i acc1 = zero - 1;
  if (two) {
    acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[1][0], B0, zero, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[1][1], B1, acc1, 0, 0, 0);
  }
  mfma_signs<2>(a, t, L, n, lane, acc0, acc1, gt, 0);
}

// E8M0 scales of the fp4 products, one per byte of a single register (the
// instruction's op_sel picks the byte): 127 = 2^0 (the A scale and field 0
// of the three-field test), 134 = 2^7 and 141 = 2^14 (fields 1 and 2), 126 =
// 2^-1 (the sign test: guide weights are stored doubled).
constexpr int kScaleBytes = 127 | (134 << 8) | (141 << 16) | (126 << 24);

// fp4 product of one A tile with one guide tile (B scale 2^-1), as 32-bit
// patterns for the sign test.
__device__ inline v16i fp4_mma(const v8i& A, v4i b) {
  const v16f zero = {};
  const v8i B = {b[0], b[1], b[2], b[3], 0, 0, 0, 0};
  return __builtin_bit_cast(v16i, __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                                      A, B, zero, 4, 4, 0, kScaleBytes, 3, kScaleBytes));
}

// fp4 product of one A tile with one guide tile added to c, with B scale
// 2^(7 Q): field Q of the three-field test of the full-batch compare.
template <int Q>
__device__ inline v16f fp4_acc(const v8i& A, v4i b, const v16f& c) {
  static_assert(Q >= 0 && Q <= 2, "three fields");
  const v8i B = {b[0], b[1], b[2], b[3], 0, 0, 0, 0};
  return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A, B, c, 4, 4, 0, kScaleBytes, Q,
                                                         kScaleBytes);
}

// Start value and test bits of the three-field accumulator (see the full-batch
// compare): 2^23 + 64 (1 + 2^7 + 2^14), and bit 6 of each 7-bit field.
constexpr float kField3Start = 9445440.0f;
constexpr uint32_t kField3Mask = (1u << 6) | (1u << 13) | (1u << 20);
constexpr uint32_t kField2Mask = (1u << 6) | (1u << 13);  // chains of two products
constexpr uint32_t kField1Mask = 1u << 6;                 // a single product

// OR of a three-field accumulator's 16 elements as 32-bit patterns (one
// v_or_b32 and seven v_or3_b32).  The whole vector is cast: hipcc 7.2 lowers
// __builtin_bit_cast of a single element of a float vector to element 0.
__device__ inline uint32_t or16(const v16f& pf) {
  const v16i p = __builtin_bit_cast(v16i, pf);
  uint32_t x = static_cast<uint32_t>(p[0]);
#pragma unroll
  for (int r = 1; r < 16; ++r) x |= static_cast<uint32_t>(p[r]);
  return x;
}

// AND of a product's 16 elements as 32-bit patterns: the sign bit is clear
// iff some element is >= 0.
__device__ inline uint32_t and16(const v16i& p) {
  uint32_t x = static_cast<uint32_t>(p[0]);
#pragma unroll
  for (int r = 1; r < 16; ++r) x &= static_cast<uint32_t>(p[r]);
  return x;
}

// fp4: one 32-guide tile against the nat (1..4) 32-window A tiles in use.
__device__ inline void fp4_tile(const Args& a, const Tab& t, const uint2* L, uint32_t n,
                                uint32_t lane,
                                uint32_t nat, const v8i (&A)[4], v4i b, uint32_t gt) {
  v16i p0 = fp4_mma(A[0], b);
  v16i p1 = v16i{} - 1;
  if (nat > 1) p1 = fp4_mma(A[1], b);
  mfma_signs<3>(a, t, L, n, lane, p0, p1, gt, 0);
  if (nat > 2) {
    p0 = fp4_mma(A[2], b);
    p1 = v16i{} - 1;
    if (nat > 3) p1 = fp4_mma(A[3], b);
    mfma_signs<3>(a, t, L, n, lane, p0, p1, gt, 2);
  }
}

// Three-field test of the products of one or two guide tiles (b0, then b1)
// with the KAT window tiles, in that order, chained three at a time: field q
// of a chain is the pair (window 32 at + m, guide n) of its q-th product.  A
// chain of one or two products tests only its own fields (the others stay
// at 64).  Returns whether a test bit is set anywhere in the wave.
template <uint32_t KAT, uint32_t NP>
__device__ inline bool leftover_chains(const v8i (&A)[4], v4i b0, v4i b1, const v16f& k3) {
  uint32_t any = 0;
  v16f c = k3;
#pragma unroll
  for (uint32_t p = 0; p < NP; ++p) {
    const v8i& at = A[p % KAT];
    const v4i b = p < KAT ? b0 : b1;
    if (p % 3 == 0) c = fp4_acc<0>(at, b, k3);
    else if (p % 3 == 1) c = fp4_acc<1>(at, b, c);
    else c = fp4_acc<2>(at, b, c);
    if (p % 3 == 2 || p == NP - 1) {
      const uint32_t mask = p % 3 == 2 ? kField3Mask : p % 3 == 1 ? kField2Mask : kField1Mask;
      any |= or16(c) & mask;
      asm volatile("" : "+v"(any));
    }
  }
  return ballot(any != 0) != 0;
}

// Compare the first n (<= ring_flush(ENGINE)) windows of a family ring with
// all guides of the family on the matrix cores.  The guide operand tables end
// with spare all-zero tiles, so the look-ahead loads of later tiles stay in
// bounds.
template <int ENGINE>
__device__ inline void compare_list_mfma(const Args& a, const Tab& t, const Fam& F,
                                         const uint2* L, uint32_t n, uint32_t lane) {
  constexpr uint32_t kAT = ring_flush(ENGINE) / 32;  // A tiles
  const uint32_t m = lane & 31u, g = lane >> 5;
  if constexpr (ENGINE == 3) {
    static_assert(kAT >= 1 && kAT <= 4, "fp4_tile takes up to four window tiles");
    // Lane group g's 32 code bits of window row 32 tile + m: the b0 (g = 0)
    // or b1 (g = 1) plane of the 20 bases, then bits 0..11 (g = 0) or
    // 12..19 (g = 1) of b0^b1 at bits 20 and up; bits 28..31 of group 1
    // are the bias elements.
    const uint32_t xmask = g ? 0xFF000u : 0xFFFu, xshift = g ? 8u : 20u;
    const uint32_t pad = g ? 0x40000000u : 0u;  // bias elements = 2.0
    v8i A[4] = {};  // tiles kAT.. unused
#pragma unroll
    for (int tile = 0; tile < static_cast<int>(kAT); ++tile) {
      const uint32_t widx = 32u * tile + m;
      uint2 e = make_uint2(0u, 0u);
      if (widx < n) e = L[widx];
      const uint32_t c = ((g ? e.y : e.x) & kWin) | (((e.x ^ e.y) & xmask) << xshift);
      A[tile] = v8i{static_cast<int>(fp4_dword(c, 0, pad)), static_cast<int>(fp4_dword(c, 1, pad)),
                    static_cast<int>(fp4_dword(c, 2, pad)), static_cast<int>(fp4_dword(c, 3, pad)),
                    0, 0, 0, 0};
    }
    const uint32_t nat = (n + 31) / 32;
    // Two operand registers, each reloaded two tiles ahead (no copies); the
    // table ends with four spare tiles for the look-ahead loads.
    // Guide tiles through a buffer resource: the tile offset in an SGPR, the
    // lane offset in a VGPR, so a load needs no VALU address arithmetic.
    const __amdgpu_buffer_rsrc_t brs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int4*>(t.btab), 0, static_cast<int>(a.btab_bytes), 0x00020000);
    const uint32_t fbase = (F.g_begin / 32) * 1024u;
    const uint32_t l16 = lane * 16u;
    auto btile = [&](uint32_t tile) {
      return __builtin_bit_cast(
          v4i, __builtin_amdgcn_raw_buffer_load_b128(brs, l16, fbase + tile * 1024u, 0));
    };
    const uint32_t n_tiles = (F.g_end - F.g_begin) / 32;
    v4i b0 = btile(0), b1 = btile(1);
    if (n == kAT * 32) {
      // Full flush, every A row a window.  Three guide tiles share one
      // accumulator and one test: the products of an A tile with guide
      // tiles ti, ti + 1, ti + 2 are chained with B scales 2^0, 2^7, 2^14
      // (instead of 2^-1), so tile q adds 2 (budget - mm) 128^q to each
      // element, onto a start value of 2^23 + 64 (1 + 128 + 128^2).  Every
      // partial sum is an integer in [2^23, 2^24), exact in f32 in any
      // order of accumulation, so the mantissa holds three 7-bit fields
      // 64 + 2 (budget - mm) in [16, 104] (padding guides: budget - mm =
      // -24), and bit 6 of a field is set iff mm <= budget.  One OR over the
      // 16 elements, masked with those bits, tests 3 x 1024 pairs where a
      // sign test covered 1024.  A triple with a set bit (rare) is redone per
      // tile with the sign test.
      uint32_t ti = 0;
      // The kernel's waves run at priority 1 except here: the compare keeps
      // many independent products and reductions in flight, the scan and
      // append phases are chains of dependent steps, so they are given the
      // issue slots first (measured 11.73 -> 11.45 ms; the other way round,
      // 12.23 ms).
      __builtin_amdgcn_s_setprio(0);
      v4i b2 = btile(2);
      const v16f k3 = v16f{} + kField3Start;
      for (; ti + 2 < n_tiles; ti += 3) {
        uint32_t any = 0;
        // two window tiles per step, their chains interleaved, so the matrix
        // core has independent products back to back
        static_assert(kAT % 2 == 0, "window tiles in pairs");
#pragma unroll
        for (int at = 0; at < static_cast<int>(kAT); at += 2) {
          v16f c = fp4_acc<0>(A[at], b0, k3);
          v16f d = fp4_acc<0>(A[at + 1], b0, k3);
          c = fp4_acc<1>(A[at], b1, c);
          d = fp4_acc<1>(A[at + 1], b1, d);
          if (at + 2 == static_cast<int>(kAT)) b0 = btile(ti + 3);
          c = fp4_acc<2>(A[at], b2, c);
          d = fp4_acc<2>(A[at + 1], b2, d);
          if (at + 2 == static_cast<int>(kAT)) {
            b1 = btile(ti + 4);
            b2 = btile(ti + 5);
          }
          // (no barrier here: the compiler then issues this pair's field
          // tests between the next pair's products)
          any |= or16(c) | or16(d);
        }
        if (ballot((any & kField3Mask) != 0)) {
          // Rare: the guide tiles whose field has a set bit somewhere in the
          // wave are redone from reloaded operands (the registers already
          // hold the next triple).
#pragma nounroll
          for (uint32_t q = 0; q < 3; ++q)
            if (ballot((any >> (6 + 7 * q)) & 1u))
              fp4_tile(a, t, L, n, lane, kAT, A, btile(ti + q), F.g_begin + 32 * (ti + q));
        }
      }
      // the last one or two guide tiles (b0, b1 hold tiles ti, ti + 1)
      const uint32_t left = n_tiles - ti;
      if (left == 1) {
        if (leftover_chains<kAT, kAT>(A, b0, b0, k3))
          fp4_tile(a, t, L, n, lane, kAT, A, b0, F.g_begin + 32 * ti);
      } else if (left == 2) {
        if (leftover_chains<kAT, 2 * kAT>(A, b0, b1, k3)) {
          fp4_tile(a, t, L, n, lane, kAT, A, b0, F.g_begin + 32 * ti);
          fp4_tile(a, t, L, n, lane, kAT, A, b1, F.g_begin + 32 * ti + 32);
        }
      }
      __builtin_amdgcn_s_setprio(1);
      return;
    }
    for (uint32_t ti = 0; ti < n_tiles; ti += 2) {
      fp4_tile(a, t, L, n, lane, nat, A, b0, F.g_begin + 32 * ti);
      b0 = btile(ti + 2);
      if (ti + 1 < n_tiles)
        fp4_tile(a, t, L, n, lane, nat, A, b1, F.g_begin + 32 * ti + 32);
      b1 = btile(ti + 3);
    }
  } else {
    uint32_t lo[kAT], hi[kAT];  // code elements 0..31 and 32..63 of window rows 32 i + m
#pragma unroll
    for (uint32_t tile = 0; tile < kAT; ++tile) {
      const uint32_t widx = 32u * tile + m;
      uint2 e = make_uint2(0u, 0u);
      if (widx < n) e = L[widx];
      const uint32_t W0 = e.x & kWin, W1 = e.y & kWin;
      lo[tile] = W1 | (W0 << 20);
      hi[tile] = (W0 >> 12) | ((W0 ^ W1) << 8);
    }
    const bool two = n > 32;
    v4i A[2][2];
#pragma unroll
    for (int tile = 0; tile < 2; ++tile) {
      A[tile][0] = expand16((lo[tile] >> (16 * g)) & 0xFFFFu);
      A[tile][1] = expand16((hi[tile] >> (16 * g)) & 0xFFFFu);
    }
    const int4* __restrict__ bp = t.btab + (F.g_begin / 32) * 128;
    int4 b0 = bp[lane], b1 = bp[64 + lane];
    for (uint32_t gt = F.g_begin; gt < F.g_end; gt += 32, bp += 128) {
      const int4 c0 = b0, c1 = b1;
      b0 = bp[128 + lane];
      b1 = bp[192 + lane];
      i8_tile(a, t, L, n, lane, two, A, c0, c1, gt);
    }
  }
}

// Seed engine (ENGINE 4): a window within 4 mismatches of a guide matches it
// exactly in at least one of the five 4-base blocks of the 20 compared bases
// (pigeonhole), so each window looks up, per block, the family's guides with
// the same 4 bases there and verifies only those.  Per family, the seed table
// (shared by the waves of a block, in LDS) holds for each block j 257 key
// starts (the key is the 8 plane bits of the block's 4 bases), then the
// entries of all blocks: per block and key, the guides with that key, padded
// to an even count, two guides per pair of words: the first word holds both
// guides' b0 plane bits of their other 16 bases (first guide low), the second
// their b1 plane bits.  A start is the entry index of its key's first guide,
// with bit 0 set when the key's guide count is odd (the last entry is then a
// pad).  Entry e's guide is in sgix at the same offset.  Only families whose
// budgets are all <= 4 use it.
constexpr uint32_t kSeedBlocks = 5;
constexpr uint32_t kSeedKeys = 256;
constexpr uint32_t kSeedStep = 8;  // most table entries (word pairs x 2) read per step of a walk
constexpr uint32_t kSeedPad = kSeedStep;  // spare words after a family's entries

// The 16 bits of a 20-bit plane word outside bases 4j..4j+3, packed in order.
__host__ __device__ inline uint32_t seed_other(uint32_t w, uint32_t j) {
  return (w & ((1u << (4 * j)) - 1u)) | ((w >> (4 * j + 4)) << (4 * j));
}

// Exact path of a seed candidate: the guide behind table word widx (an entry
// of the window's key in block j), its full 2-bit count, and the pair only if
// block j is the first block in which it matches exactly, so a pair matching
// in several blocks is recorded once.
__device__ inline void seed_exact(const Args& a, const Tab& t, uint2 e, uint32_t lane,
                                  uint32_t W0, uint32_t W1, uint32_t j, uint32_t widx, bool cand) {
  uint32_t gi = 0, c = 0;
  bool rec = false;
  if (cand) {
    gi = t.sgix[widx];
    const uint32_t m = ((W0 ^ t.g01[2 * gi]) | (W1 ^ t.g01[2 * gi + 1])) & kWin;
    uint32_t nz = m | (m >> 1);
    nz |= nz >> 2;  // bit 4b: block b has a mismatch
    const uint32_t upto = 0x11111u & ((1u << (4 * j + 1)) - 1u);  // blocks 0..j
    const uint32_t below = 0x11111u & ((1u << (4 * j)) - 1u);     // blocks 0..j-1
    c = static_cast<uint32_t>(__popc(m));
    rec = (nz & upto) == below && static_cast<int>(c) <= t.gk[gi];
  }
  if (ballot(rec))
    record_candidates(a, t, entry_strand(t.runs, e), gi, entry_pos(t.runs, e), W0, W1, c, rec,
                      lane);
}

// One window's walk through the guides of its key in block j (entries
// [beg, end) of the family's table), STEP entries (word pairs of two) per
// step, read together.  A lane whose key's entries are done goes on reading
// the following ones while the wave's longest walk lasts; only the rare
// candidates are checked to lie in the lane's own key (i + q < end), and the
// reads stay in the family's table (clamped to its end, which is followed by
// kSeedPad spare words; a clamped read is never in the lane's key).
template <uint32_t STEP>
__device__ inline void seed_walk(const Args& a, const Tab& t, const Fam& F, uint2 e, uint32_t lane,
                                 uint32_t W0, uint32_t W1, uint32_t j, uint32_t wl, uint32_t wh,
                                 uint32_t beg, uint32_t end, const uint32_t* en, uint32_t last,
                                 uint32_t kmax) {
  static_assert(STEP % 2 == 0 && STEP <= kSeedPad, "steps of word pairs within the spare words");
  const uint2* pend = reinterpret_cast<const uint2*>(en + end);
  const uint2* plast = reinterpret_cast<const uint2*>(en + last);
  for (const uint2* pi = reinterpret_cast<const uint2*>(en + beg); ballot(pi < pend);
       pi += STEP / 2) {
    const uint2* pb = pi < plast ? pi : plast;  // entries 2 (pb - en) .. (even)
    // per word pair, each half's count: the high half's, and the whole
    // word's less it
    uint32_t p[STEP];
    uint32_t lo = 32;
#pragma unroll
    for (uint32_t q = 0; q < STEP / 2; ++q) {
      const uint2 pr = pb[q];  // entries 2q, 2q + 1 of the step
      const uint32_t m = (pr.x ^ wl) | (pr.y ^ wh);
      const uint32_t h = static_cast<uint32_t>(__popc(m >> 16));
      p[2 * q] = static_cast<uint32_t>(__popc(m)) - h;
      p[2 * q + 1] = h;
      lo = min(lo, min(p[2 * q], h));
    }
    if (ballot(lo <= kmax)) {
      const uint32_t b = static_cast<uint32_t>(reinterpret_cast<const uint32_t*>(pb) - en);
      const uint32_t i = static_cast<uint32_t>(reinterpret_cast<const uint32_t*>(pi) - en);
#pragma nounroll
      for (uint32_t q = 0; q < STEP; ++q) {
        const bool cand = p[q] <= kmax && i + q < end;
        if (ballot(cand)) seed_exact(a, t, e, lane, W0, W1, j, F.seed_ent + b + q, cand);
      }
    }
  }
}

// Compare the first n (<= 64) windows of a family ring with the family's
// guides through the seed table: one window per lane; per block, each lane
// walks its own key's guides, the wave as long as its longest walk.
template <uint32_t STEP>
__device__ inline void seed_compare(const Args& a, const Tab& t, const Fam& F, const uint2* L,
                                    uint32_t n, uint32_t lane, uint32_t* vcnt) {
  uint32_t verified = 0;
  const bool has = lane < n;
  const uint2 e = has ? L[lane] : make_uint2(0u, 0u);
  const uint32_t W0 = e.x & kWin, W1 = e.y & kWin;
  const uint32_t kmax = F.kmax;
  const uint32_t* st = t.stab + F.seed_start;
  const uint32_t* en = t.stab + F.seed_ent;
  const uint32_t last = F.seed_n;  // reads of entries last.. hit the kSeedPad spare words
  // every block's key starts, read together up front
  uint32_t s0s[kSeedBlocks], s1s[kSeedBlocks];
#pragma unroll
  for (uint32_t j = 0; j < kSeedBlocks; ++j) {
    const uint32_t key = ((W0 >> (4 * j)) & 0xFu) | (((W1 >> (4 * j)) & 0xFu) << 4);
    const uint32_t* sj = st + j * (kSeedKeys + 1);
    s0s[j] = sj[key];
    s1s[j] = sj[key + 1];
  }
#pragma unroll
  for (uint32_t j = 0; j < kSeedBlocks; ++j) {
    // the window's other 16 bases, per plane, in both halves of a word
    const uint32_t wl = seed_other(W0, j) * 0x10001u, wh = seed_other(W1, j) * 0x10001u;
    const uint32_t s0 = s0s[j], s1 = s1s[j];
    const uint32_t beg = s0 & ~1u;
    const uint32_t end = has ? (s1 & ~1u) - (s0 & 1u) : beg;
    verified += end - beg;
    seed_walk<STEP>(a, t, F, e, lane, W0, W1, j, wl, wh, beg, end, en, last, kmax);
  }
  vcnt[lane] += verified;
}

// Per-wave LDS state of the scan loop.
struct WaveState {
  uint2* lists;                  // [NF][ring_size] window buffers
  uint32_t* pws;                 // [NF][64] pass words of the current blocks
  uint32_t* counts;              // [kMaxFam] windows in each buffer (carried-over ones << 16)
  uint32_t* runs;                // [2][2] position base and strand of the runs by entry flag
  unsigned long long* wstats;    // [kMaxFam] windows compared
  uint2* stage;                  // [64] plane words of the next step's blocks
  uint32_t* vcnt;                // seed engine: [NF][64] table entries each lane verified
  template <int ENGINE>
  __device__ uint2* ring(uint32_t f) const {
    return lists + f * ring_size(ENGINE);
  }
};

// Compare the batch at the front of family f's buffer (ring_flush windows)
// and move the at most 64 windows behind it to the front.
template <int ENGINE, int SS = 8>
__device__ inline void flush_front(const Args& a, const Tab& t, const WaveState& w, uint32_t f,
                                   uint2* L, uint32_t lane) {
  constexpr uint32_t kFlush = ring_flush(ENGINE);
  wave_sync();
  if constexpr (ENGINE == 2 || ENGINE == 3)
    compare_list_mfma<ENGINE>(a, t, t.fam[f], L, kFlush, lane);
  else if constexpr (ENGINE == 4)
  {
    __builtin_amdgcn_s_setprio(0);
    seed_compare<SS>(a, t, t.fam[f], L, kFlush, lane, w.vcnt + f * 64);
    __builtin_amdgcn_s_setprio(1);
  }
  else
  {
    // VALU batch compare at wave priority 0, like the fp4 full-batch compare
    __builtin_amdgcn_s_setprio(0);
    compare_list(a, t, t.fam[f], L, kFlush, lane);
    __builtin_amdgcn_s_setprio(1);
  }
  const uint2 e = L[kFlush + lane];  // LDS operations of a wave complete in order
  L[lane] = e;
  if (lane == 0) w.wstats[f] += kFlush;
  wave_sync();
}

// Append this lane's passing windows (bits of pw in block bb) to family f's
// ring, comparing every ring_flush(ENGINE) buffered windows.  All lanes call
// this.  Each step appends the next window of every lane that has one (slots
// by ballot + mbcnt), so all lanes write in parallel; the ring always keeps
// room for one such step.
template <int ENGINE, int SS = 8>
__device__ inline void append_family(const Args& a, const Tab& t, const WaveState& w, uint32_t f,
                                     uint32_t pw, uint32_t bb, uint2 qa, uint2 qb, uint32_t cb,
                                     uint32_t flag, uint32_t lane) {
  constexpr uint32_t kSize = ring_size(ENGINE), kFlush = ring_flush(ENGINE);
  uint2* L = w.ring<ENGINE>(f);
  // counts[f]: windows in the buffer (low half), of which the first (high
  // half) were carried over from the wave's previous run
  const uint32_t cw = uniform(w.counts[f]);
  uint32_t cnt = cw & 0xFFFFu, old = cw >> 16;
  // Window offsets are from the 32-aligned base cb, so the block's offset
  // is a multiple of 32 and the entry's offset fields are block constants
  // plus the bit index.
  const uint32_t off0 = (bb << 5) - cb;
  const uint32_t x0 = off0 << 20, y0 = ((off0 >> 12) << 20) | (flag << 31);
  // Each pass appends the next window of every lane that has one; passes run
  // back to back while the buffer has room for a full pass (64 slots), and a
  // batch is compared when it has not.
  uint64_t act = ballot(pw != 0);
  while (act) {
    if (cnt > kSize - 64) {
      flush_front<ENGINE, SS>(a, t, w, f, L, lane);
      cnt -= kFlush;
      old = old > kFlush ? old - kFlush : 0u;
      continue;
    }
    do {
      // Lanes with a window store it at their rank among them; pw & (pw - 1)
      // leaves an empty word empty.
      if (pw) {
        const uint32_t j = static_cast<uint32_t>(__builtin_ctz(pw));
        const uint32_t slot = lanes_below(act, cnt);
        L[slot] = make_uint2((funnel(qb.x, qa.x, j) & kWin) | lshl_or(j, 20, x0),
                             (funnel(qb.y, qa.y, j) & kWin) | y0);
      }
      pw &= pw - 1;
      cnt += static_cast<uint32_t>(__popcll(act));
      act = ballot(pw != 0);
    } while (act && cnt <= kSize - 64);
  }
  if (lane == 0) w.counts[f] = cnt | (old << 16);
  wave_sync();
}

// Stream positions [cs, ce) of strand s.  Each lane owns a 32-position block:
// the PAM test of every family runs on whole plane words (one allowed-set op
// per PAM position for 32 positions), then the passing windows are appended
// to the family's ring in LDS; every ring_flush(ENGINE) buffered windows are
// compared against the family's guides.
// Compare every window in family f's buffer and empty it.
template <int ENGINE, int SS = 8>
__device__ inline void drain_family(const Args& a, const Tab& t, const WaveState& w, uint32_t f,
                                    uint32_t lane) {
  uint32_t c = uniform(w.counts[f]) & 0xFFFFu;
  if (c > ring_flush(ENGINE)) {
    flush_front<ENGINE, SS>(a, t, w, f, w.ring<ENGINE>(f), lane);
    c -= ring_flush(ENGINE);
  }
  if (c) {
    wave_sync();
    if constexpr (ENGINE == 2 || ENGINE == 3)
      compare_list_mfma<ENGINE>(a, t, t.fam[f], w.ring<ENGINE>(f), c, lane);
    else if constexpr (ENGINE == 4)
    {
      __builtin_amdgcn_s_setprio(0);
      seed_compare<SS>(a, t, t.fam[f], w.ring<ENGINE>(f), c, lane, w.vcnt + f * 64);
      __builtin_amdgcn_s_setprio(1);
    }
    else
    {
      __builtin_amdgcn_s_setprio(0);
      compare_list(a, t, t.fam[f], w.ring<ENGINE>(f), c, lane);
      __builtin_amdgcn_s_setprio(1);
    }
    if (lane == 0) w.wstats[f] += c;
  }
  wave_sync();
  if (lane == 0) w.counts[f] = 0;
  wave_sync();
}

template <int NF, int ENGINE, int SS = 8>
__device__ inline void scan_run(const Args& a, const Tab& t, const WaveState& w, uint32_t s,
                                uint32_t cs, uint32_t ce, uint32_t flag, uint32_t lane) {
  const uint2* __restrict__ planes = s ? a.planes1 : a.planes0;
  uint32_t nibs[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    nibs[f] = t.fam[f].nibs;
  }
  const uint32_t b_first = cs >> 5, b_last = (ce - 1) >> 5;
  const uint32_t cb = b_first << 5;  // ring entry offsets are from here
  // This run's entries carry `flag`; the windows left in the buffers by the
  // wave's previous run (the other flag) are marked as carried over.
  if (lane == 0) {
    w.runs[2 * flag] = cb;
    w.runs[2 * flag + 1] = s;
  }
  if (lane < kMaxFam) w.counts[lane] = (w.counts[lane] & 0xFFFFu) * 0x10001u;
  wave_sync();
  // Each lane loads one plane word per step, of block min(b, b_last + 1) (the
  // plane array extends past every scanned block), two steps ahead: the word
  // loaded in step k is stored to the wave's LDS stage in step k + 1 and read
  // back in step k + 2, so one register set carries it and no load is waited
  // for in the step that issued it.  The following block's word comes from
  // lane + 1 (DPP wave_shl:1); lane 63 reads it from the stage (lane 0's
  // entry of the next step's blocks).  LDS operations of a wave complete in
  // order, so each read returns the entry before the store that follows it.
  const uint32_t b_end = b_last + 1;
  uint2 nxt = planes[min(b_first + lane, b_end)];
  w.stage[lane] = nxt;
  nxt = planes[min(b_first + 64 + lane, b_end)];
  for (uint32_t b0 = b_first; b0 <= b_last; b0 += 64) {
    const uint32_t bb = b0 + lane;
    const bool inb = bb <= b_last;
    const uint2 qa = w.stage[lane];
    w.stage[lane] = nxt;
    nxt = planes[min(b0 + 128 + lane, b_end)];
    const uint2 e = w.stage[0];
    const uint2 qb = make_uint2(
        static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(e.x),
                                                          static_cast<int>(qa.x), 0x130, 0xF,
                                                          0xF, false)),
        static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(e.y),
                                                          static_cast<int>(qa.y), 0x130, 0xF,
                                                          0xF, false)));
    uint32_t vmask = 0;
    if (inb) {
      const uint32_t p0 = bb << 5;
      const uint32_t lo = cs > p0 ? cs - p0 : 0u;
      const uint32_t hi = ce - p0 >= 32 ? 32u : ce - p0;
      vmask = (hi == 32 ? ~0u : ((1u << hi) - 1u)) & (~0u << lo);
    }
    // plane bits of PAM position i for the block's 32 positions
    uint32_t f0[6], f1[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      f0[i] = funnel(qb.x, qa.x, 20 + i);
      f1[i] = funnel(qb.y, qa.y, 20 + i);
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      uint32_t pw = vmask;
      // The family's nibbles stay in one SGPR; their masks are rederived
      // here (a few scalar ops) rather than hoisted out of the loop, which
      // would take 4 SGPRs per PAM base.
      uint32_t fn = nibs[f];
      asm volatile("" : "+s"(fn));
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const uint32_t nib = (fn >> (4 * i)) & 0xFu;
        if (nib != 0xFu) pw &= allowed(nib, f0[i], f1[i]);
      }
      w.pws[f * 64 + lane] = pw;
    }
    wave_sync();
#pragma nounroll
    for (uint32_t f = 0; f < NF; ++f) {
      const uint32_t pw = w.pws[f * 64 + lane];
      if (ballot(pw != 0)) {
        if (ENGINE == 1 && t.fam[f].direct) {
          const uint32_t n = direct_family(a, t, t.fam[f], s, pw, bb, qa, qb, lane);
          if (lane == 0) w.wstats[f] += n;
        } else {
          append_family<ENGINE, SS>(a, t, w, f, pw, bb, qa, qb, cb, flag, lane);
        }
      }
    }
  }
  // A buffer that still holds windows of the previous run is compared now
  // (the next run reuses that run's flag); the others carry their windows
  // into the wave's next run, where they fill whole batches.
#pragma nounroll
  for (uint32_t f = 0; f < NF; ++f)
    if (uniform(w.counts[f]) >> 16) drain_family<ENGINE, SS>(a, t, w, f, lane);
  wave_sync();
}

template <int NF, int ENGINE, int SS = 8>
__device__ __forceinline__ void scan_kernel_body(
    uint2* lds, Args a, const Fam* __restrict__ fam, const uint32_t* __restrict__ g01,
    const int32_t* __restrict__ gk, const uint32_t* __restrict__ gaux,
    const uint32_t* __restrict__ gid, const uint32_t* __restrict__ gnm,
    const double* __restrict__ score_tab, bgo_hit* __restrict__ hits,
    unsigned long long* __restrict__ hit_count, unsigned int* __restrict__ work_counter,
    unsigned long long* __restrict__ fam_windows, const int4* __restrict__ btab) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t n_waves = blockDim.x >> 6;
  // Seed engine: the block's seed table (a.btab_bytes, a multiple of 16, at
  // btab, followed by its guide words) is copied to the front of the LDS by
  // the whole block; the waves' own regions follow it.
  size_t lds_front = 0;
  const uint32_t* stab = nullptr;
  const uint32_t* sgix = nullptr;
  if constexpr (ENGINE == 4) {
    const uint32_t n16 = a.btab_bytes / 16;
    uint4* dst = reinterpret_cast<uint4*>(lds);
    const uint4* src = reinterpret_cast<const uint4*>(btab);
    for (uint32_t i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
    stab = reinterpret_cast<const uint32_t*>(lds);
    sgix = reinterpret_cast<const uint32_t*>(btab) + a.btab_bytes / 4;
    lds_front = (static_cast<size_t>(a.btab_bytes) + 255) / 256 * 256;
  }
  // LDS per wave: NF window buffers (ring_size x 8 B), kMaxFam stats (8 B),
  // NF*64 pass words, kMaxFam buffer counts (4 B each), 64 staged plane
  // words (8 B each).
  constexpr size_t kRings = NF * ring_size(ENGINE) * 8;
  char* my = reinterpret_cast<char*>(lds) + lds_front + wave * wave_lds_bytes(NF, ENGINE);
  WaveState w;
  w.lists = reinterpret_cast<uint2*>(my);
  w.wstats = reinterpret_cast<unsigned long long*>(my + kRings);
  w.pws = reinterpret_cast<uint32_t*>(my + kRings + kMaxFam * 8);
  w.counts = w.pws + NF * 64;
  w.runs = w.counts + kMaxFam;
  w.stage = reinterpret_cast<uint2*>(w.runs + 4);
  // MFMA engines: kRecStage staged records and their count after the plane words
  bgo_hit* const rec_stage = ENGINE >= 2 ? reinterpret_cast<bgo_hit*>(w.stage + 64) : nullptr;
  uint32_t* const rec_cnt =
      ENGINE >= 2 ? reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(w.stage + 64) +
                                                kRecStage * sizeof(bgo_hit))
                  : nullptr;
  // seed engine: per-lane verify counters after the record stage
  w.vcnt = ENGINE == 4 ? reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(w.stage + 64) +
                                                     kRecStage * sizeof(bgo_hit) + 16)
                       : nullptr;
  if constexpr (ENGINE == 4) {
#pragma unroll
    for (int f = 0; f < NF; ++f) w.vcnt[f * 64 + lane] = 0;
  }
  const Tab t{fam,       g01,       gk,        gaux,    gid,    gnm,  score_tab, hits,
              hit_count, btab,      rec_stage, rec_cnt, w.runs, stab, sgix};
  // fp4 and VALU: priority 1 outside the batch compares, which run at 0 (the
  // scan, append and in-place passes are dependent chains; a compare has
  // many independent steps in flight).  VALU: 22 / 80 / 1024 mixed guides
  // 3.38 / 5.83 / 37.8 -> 3.27 / 5.43 / 35.8 ms.
  if constexpr (ENGINE == 3 || ENGINE == 1 || ENGINE == 4) __builtin_amdgcn_s_setprio(1);
  if (lane < kMaxFam) {
    w.wstats[lane] = 0;
    w.counts[lane] = 0;
  }
  if constexpr (ENGINE >= 2) {
    if (lane == 0) *rec_cnt = 0;
  }
  wave_sync();
  // Counter k hands out items k, k + kWorkCtrs, k + 2 kWorkCtrs, ...  Global
  // wave g takes from counter (g + g / kWorkCtrs) % kWorkCtrs: each aligned
  // group of kWorkCtrs waves covers every counter (the host launches at least
  // one group), and the per-group rotation gives every counter waves on all
  // XCDs (blocks are dealt to the XCDs round-robin), so a counter's items are
  // balanced across the chip.  Each subset spreads evenly over the scanned
  // range, so the subsets end together, and a wave exits after one failed
  // take on its own counter: a launch costs two takes per wave, kWorkCtrs-fold
  // fewer per address than with one counter.
  const uint32_t gw = blockIdx.x * n_waves + wave;
  const uint32_t k = __builtin_amdgcn_readfirstlane((gw + gw / kWorkCtrs) % kWorkCtrs);
  unsigned int* const my_counter = work_counter + k * kCtrStride;
  uint32_t n_run = 0;  // runs of this wave so far; the parity flags their buffer entries
  for (;;) {
    uint32_t r = 0;
    if (lane == 0) r = atomicAdd(my_counter, 1u);
    r = __builtin_amdgcn_readlane(r, 0);
    const uint64_t item64 = static_cast<uint64_t>(r) * kWorkCtrs + k;
    if (item64 >= a.n_items) break;
    const uint32_t item = static_cast<uint32_t>(item64);
    uint32_t slot, run;
    if (a.interleave) {
      slot = item % a.n_slots;
      run = item / a.n_slots;
    } else {
      slot = item / a.chunks_per_slot;
      run = item - slot * a.chunks_per_slot;
    }
    const uint32_t s = slot ? a.slot_strand1 : a.slot_strand0;
    const uint32_t cs = a.valid_begin + run * a.chunk;
    const uint32_t ce = min(cs + a.chunk, a.valid_end);
    scan_run<NF, ENGINE, SS>(a, t, w, s, cs, ce, n_run & 1u, lane);
    ++n_run;
  }
  // the windows the last runs left in the buffers
#pragma nounroll
  for (uint32_t f = 0; f < NF; ++f)
    if (uniform(w.counts[f]) & 0xFFFFu) drain_family<ENGINE, SS>(a, t, w, f, lane);
  if constexpr (ENGINE >= 2) flush_records(a, t, lane);
  wave_sync();
  if (lane < NF && w.wstats[lane]) atomicAdd(&fam_windows[lane], w.wstats[lane]);
  if constexpr (ENGINE == 4) {
    // one atomic per wave and family: the lanes' counts summed across the wave
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      unsigned long long v = w.vcnt[f * 64 + lane];
#pragma unroll
      for (int sh = 32; sh >= 1; sh >>= 1) v += __shfl_xor(v, sh, 64);
      if (lane == 0 && v) atomicAdd(&fam_windows[a.fam_total + f], v);
    }
  }
}

template <int NF, int ENGINE>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(4))) void batched_scan_kernel(
    Args a, const Fam* __restrict__ fam, const uint32_t* __restrict__ g01,
    const int32_t* __restrict__ gk, const uint32_t* __restrict__ gaux,
    const uint32_t* __restrict__ gid, const uint32_t* __restrict__ gnm,
    const double* __restrict__ score_tab, bgo_hit* __restrict__ hits,
    unsigned long long* __restrict__ hit_count, unsigned int* __restrict__ work_counter,
    unsigned long long* __restrict__ fam_windows, const int4* __restrict__ btab) {
  extern __shared__ uint2 lds[];
  scan_kernel_body<NF, ENGINE>(lds, a, fam, g01, gk, gaux, gid, gnm, score_tab, hits, hit_count,
                               work_counter, fam_windows, btab);
}

// The seed engine's kernel: one block of kSeedBlock threads per CU shares the
// seed table in LDS.
constexpr int kSeedBlock = 1024;
template <int NF, int SS>
__global__ __launch_bounds__(kSeedBlock) void seed_scan_kernel(
    Args a, const Fam* __restrict__ fam, const uint32_t* __restrict__ g01,
    const int32_t* __restrict__ gk, const uint32_t* __restrict__ gaux,
    const uint32_t* __restrict__ gid, const uint32_t* __restrict__ gnm,
    const double* __restrict__ score_tab, bgo_hit* __restrict__ hits,
    unsigned long long* __restrict__ hit_count, unsigned int* __restrict__ work_counter,
    unsigned long long* __restrict__ fam_windows, const int4* __restrict__ btab) {
  extern __shared__ uint2 lds[];
  scan_kernel_body<NF, 4, SS>(lds, a, fam, g01, gk, gaux, gid, gnm, score_tab, hits, hit_count,
                              work_counter, fam_windows, btab);
}

using KernelFn = void (*)(Args, const Fam*, const uint32_t*, const int32_t*, const uint32_t*,
                         const uint32_t*, const uint32_t*, const double*, bgo_hit*,
                         unsigned long long*, unsigned int*, unsigned long long*, const int4*);

template <int ENGINE>
inline KernelFn kernel_for_nf(int nf) {
  switch (nf) {
    case 1: return batched_scan_kernel<1, ENGINE>;
    case 2: return batched_scan_kernel<2, ENGINE>;
    case 3: return batched_scan_kernel<3, ENGINE>;
    case 4: return batched_scan_kernel<4, ENGINE>;
    case 5: return batched_scan_kernel<5, ENGINE>;
    case 6: return batched_scan_kernel<6, ENGINE>;
    case 7: return batched_scan_kernel<7, ENGINE>;
    default: return batched_scan_kernel<8, ENGINE>;
  }
}

// Seed engine launches group at most kSeedMaxFam families (their rings and
// the shared table share one block's LDS).
constexpr int kSeedMaxFam = 4;

template <int SS>
inline KernelFn seed_kernel_for(int nf) {
  switch (nf) {
    case 1: return seed_scan_kernel<1, SS>;
    case 2: return seed_scan_kernel<2, SS>;
    case 3: return seed_scan_kernel<3, SS>;
    default: return seed_scan_kernel<4, SS>;
  }
}

// seed_step: table entries per step of the seed engine's walks (4 or 8).
inline KernelFn kernel_for(int nf, int engine, int seed_step = 8) {
  if (engine == 4) return seed_step == 4 ? seed_kernel_for<4>(nf) : seed_kernel_for<8>(nf);
  return engine == 3 ? kernel_for_nf<3>(nf) : engine == 2 ? kernel_for_nf<2>(nf) : kernel_for_nf<1>(nf);
}

__global__ void reset_counters_kernel(unsigned long long* c, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) c[i] = 0;
}

// 2-bit words -> bit planes, one 32-base block per lane.
__global__ void planes_kernel(const uint32_t* __restrict__ words, uint64_t n_words,
                              uint2* __restrict__ planes, uint64_t n_blocks) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= n_blocks) return;
  const uint32_t w0 = (2 * b < n_words) ? words[2 * b] : 0u;
  const uint32_t w1 = (2 * b + 1 < n_words) ? words[2 * b + 1] : 0u;
  planes[b] = make_uint2(even_bits(w0) | (even_bits(w1) << 16),
                         even_bits(w0 >> 1) | (even_bits(w1 >> 1) << 16));
}

}  // namespace bk
}  // namespace bgo

using namespace bgo;
using bgo::bk::Fam;

struct bgo_plan {
  bgo_genome* g = nullptr;
  int device = 0;
  int32_t end_guard = 26;             // positions >= total_bp - end_guard are not scanned
  int32_t n_guides = 0;
  int64_t hit_cap = 0;
  double score[21] = {};
  size_t mem_bytes = 0;               // size of d_mem
  // host mirrors of the device tables (alive until the next upload completes)
  std::vector<uint32_t> h_g01, h_gaux, h_gid, h_gnm;
  std::vector<int32_t> h_gk;
  bgo_tuning tune{};
  bool engine_auto = false;  // engine chosen per guide set
  int direct_max = 0;        // VALU engine: families of <= this many guides compare in
                             // place (0: 8 for PAMs passing >= 5% of positions, else 4)
  bool direct_off = false;   // no family compares in place
  // families, grouped into launches of <= kMaxFam
  std::vector<Fam> fams;              // all families (g ranges into the padded table)
  std::vector<int32_t> fam_guides;    // real guides per family
  std::vector<int> group_first;       // first family of each launch group
  // device memory
  void* d_mem = nullptr;
  Fam* d_fam = nullptr;
  uint32_t* d_g01 = nullptr;
  int32_t* d_gk = nullptr;
  uint32_t* d_gaux = nullptr;
  uint32_t* d_gid = nullptr;
  uint32_t* d_gnm = nullptr;
  double* d_score = nullptr;
  // [1 hit count][fam windows][to 256 B][groups x kWorkCtrs work counters, 256 B each]
  unsigned long long* d_counters = nullptr;
  int8_t* d_btab = nullptr;                  // MFMA guide operand tiles (engine 2)
  std::vector<int8_t> h_btab;
  uint32_t n_counter_words = 0;
  uint32_t work_word = 0;  // 64-bit word of the first work counter
  bgo_hit* d_hits = nullptr;
  bgo_hit* h_pinned = nullptr;  // page-locked copy of the records (fetch_pinned)
  int64_t pinned_cap = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_mask = -1;
  int n_cu = 0;
  int resolved_blocks_per_cu = 0;
  std::vector<int> group_blocks;     // persistent blocks per launch group
  std::vector<size_t> group_lds;     // dynamic LDS bytes per launch group
  std::vector<size_t> group_img_off;    // seed engine: byte offset of each group's table in d_btab
  std::vector<size_t> group_img_bytes;  // seed engine: bytes of each group's table (LDS copy)
  std::vector<int> group_step;          // seed engine: table entries per walk step (4 or 8)
  int threads = 0;
  // Tables and hit buffer in memory the plan does not own (the no-alloc
  // entry point's scratch array): never grown, never freed.
  bool ext_mem = false;
  std::mutex mu;
};

namespace {

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

int build_genome_strand(bgo_genome* G, int s, const uint32_t* words, const uint8_t* nmask,
                        int64_t nmask_bytes) {
  uint32_t* d_words = nullptr;
  BGO_HIP(hipMalloc(&d_words, G->n_words * sizeof(uint32_t)));
  if (hipMemcpy(d_words, words, G->n_words * sizeof(uint32_t), hipMemcpyHostToDevice) !=
      hipSuccess) {
    (void)hipFree(d_words);
    return fail(BGO_ERR_HIP, "upload genome words");
  }
  int rc = BGO_OK;
  if (hipMalloc(&G->d_planes[s], G->n_blocks * sizeof(uint2)) != hipSuccess) {
    rc = fail(BGO_ERR_HIP, "hipMalloc planes");
  } else {
    const uint64_t grid = (G->n_blocks + 255) / 256;
    hipLaunchKernelGGL(bk::planes_kernel, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, 0,
                       d_words, static_cast<uint64_t>(G->n_words), G->d_planes[s], G->n_blocks);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
      rc = fail(BGO_ERR_HIP, "planes kernel");
  }
  (void)hipFree(d_words);
  if (rc != BGO_OK) return rc;
  G->bytes += G->n_blocks * sizeof(uint2);
  if (nmask && nmask_bytes > 0) {
    const size_t nbytes = G->n_blocks * sizeof(uint32_t);
    BGO_HIP(hipMalloc(&G->d_nwords[s], nbytes));
    BGO_HIP(hipMemset(G->d_nwords[s], 0, nbytes));
    BGO_HIP(hipMemcpy(G->d_nwords[s], nmask, std::min<size_t>(nmask_bytes, nbytes),
                      hipMemcpyHostToDevice));
    G->bytes += nbytes;
  }
  return BGO_OK;
}

}  // namespace

extern "C" int bgo_genome_create(const uint32_t* words_fwd, const uint32_t* words_rc,
                                 int64_t n_words, const uint8_t* nmask_fwd,
                                 const uint8_t* nmask_rc, int64_t nmask_bytes, int64_t total_bp,
                                 bgo_genome** out) {
  if (!out || !words_fwd || n_words <= 0 || total_bp < 0 || total_bp > n_words * 16)
    return fail(BGO_ERR_ARG, "bgo_genome_create: bad arguments");
  if (total_bp >= (int64_t(1) << 32) - 64)
    return fail(BGO_ERR_ARG, "bgo_genome_create: total_bp must be < 2^32 - 64");
  if (nmask_bytes < 0 || (nmask_bytes > 0 && !nmask_fwd) ||
      (nmask_bytes > 0 && words_rc && !nmask_rc))
    return fail(BGO_ERR_ARG, "bgo_genome_create: bad N mask");
  BGO_TRY(injected());
  int device = 0;
  BGO_HIP(hipGetDevice(&device));
  auto* G = new bgo_genome();
  G->device = device;
  G->total_bp = total_bp;
  G->n_words = n_words;
  // +4 zero blocks: the kernel reads planes[w + 1] for w <= (total_bp + 63) / 32.
  G->n_blocks = static_cast<uint64_t>((n_words + 1) / 2) + 4;
  G->n_strands = words_rc ? 2 : 1;
  int rc = build_genome_strand(G, 0, words_fwd, nmask_fwd, nmask_bytes);
  if (rc == BGO_OK && words_rc) rc = build_genome_strand(G, 1, words_rc, nmask_rc, nmask_bytes);
  if (rc != BGO_OK) {
    bgo_genome_destroy(G);
    return rc;
  }
  *out = G;
  return BGO_OK;
}

extern "C" int bgo_genome_destroy(bgo_genome* G) {
  if (!G) return BGO_OK;
  for (int s = 0; s < 2; ++s) {
    if (G->d_planes[s]) (void)hipFree(G->d_planes[s]);
    if (G->d_nwords[s]) (void)hipFree(G->d_nwords[s]);
  }
  delete G;
  return BGO_OK;
}

extern "C" int bgo_genome_device_bytes(const bgo_genome* G, int64_t* bytes_out) {
  if (!G || !bytes_out) return fail(BGO_ERR_ARG, "null pointer");
  *bytes_out = G->bytes;
  return BGO_OK;
}

namespace {

// Family key: PAM length + allowed-base nibbles.
uint64_t fam_key(const bgo_guide_desc& d) {
  uint64_t k = d.pam_len;
  for (int i = 0; i < d.pam_len && i < 6; ++i) k |= static_cast<uint64_t>(d.pam_mask[i] & 0xF) << (4 + 4 * i);
  return k;
}

Fam make_fam(const bgo_guide_desc& d) {
  Fam F{};
  // positions past the PAM read as N (no constraint)
  for (int i = 0; i < 6; ++i) F.nibs |= (i < d.pam_len ? d.pam_mask[i] & 0xFu : 0xFu) << (4 * i);
  F.plen = d.pam_len;
  return F;
}

// Fraction of the positions of a 41% GC genome whose bases pass the PAM.
double pam_pass_rate(uint32_t nibs) {
  static constexpr double kBaseFreq[4] = {0.295, 0.205, 0.205, 0.295};  // A C G T
  double pass = 1.0;
  for (int i = 0; i < 6; ++i) {
    const uint32_t nib = (nibs >> (4 * i)) & 0xFu;
    double p = 0.0;
    for (int b = 0; b < 4; ++b) p += (nib >> b) & 1u ? kBaseFreq[b] : 0.0;
    pass *= p;
  }
  return pass;
}

// Guides up to which a family compares its windows in place (VALU engine):
// in-place compares cost less than appends and a padded compare.  By default
// 8 for PAMs passing >= 5% of the positions of a 41% GC genome (NG 20%,
// NNNRRT 7.4%: appends are their main cost) and 4 for sparser ones (NGG
// 4.2%, NNGRRT 1.5%), the crossovers measured on a 3.09 Gbp genome of that
// composition.
int direct_limit(const bgo_plan* P, double pass) {
  if (P->direct_off) return 0;
  if (P->direct_max) return P->direct_max;
  return pass >= 0.05 ? bk::kDirectMax : bk::kDirectSparse;
}

// Four fp4 (e2m1) codes summing to x, |x| <= 22 (magnitudes 6, 4, 3, 2, 1).
bool fp4_sum4(int x, uint8_t out[4]) {
  const uint8_t sign = x < 0 ? 0x8 : 0x0;
  int r = x < 0 ? -x : x;
  for (int i = 0; i < 4; ++i) {
    const int v = r >= 6 ? 6 : (r == 5 ? 4 : r);
    static const uint8_t code[7] = {0x0, 0x2, 0x4, 0x5, 0x6, 0x0, 0x7};  // 0 1 2 3 4 - 6
    out[i] = v ? static_cast<uint8_t>(code[v] | sign) : 0;
    r -= v;
  }
  return r == 0;
}

int resolve_tuning(bgo_plan* P, const bgo_tuning* t) {
  bgo_tuning r{};
  if (t) r = *t;
  if (r.chunk_positions <= 0) r.chunk_positions = 131072;
  r.chunk_positions = std::max(64, (r.chunk_positions + 63) / 64 * 64);
  // ring entries hold 23-bit offsets from the 32-aligned chunk start
  if (r.chunk_positions > (1 << 23) - 64)
    return fail(BGO_ERR_ARG, "chunk_positions must be <= 2^23 - 64");
  if (r.waves_per_block <= 0) r.waves_per_block = 4;
  if (r.waves_per_block != 1 && r.waves_per_block != 2 && r.waves_per_block != 4)
    return fail(BGO_ERR_ARG, "waves_per_block must be 1, 2 or 4");
  // reserved[1]: 0 = default, -1 = never, n > 0 = at most n guides
  // 0 = by PAM density (below), n > 0 = n for every family (at most kDirectMax)
  P->direct_max = r.reserved[1] == 0 ? 0 : std::clamp(r.reserved[1], 0, bk::kDirectMax);
  P->direct_off = r.reserved[1] < 0;
  P->engine_auto = r.engine <= 0;
  if (r.engine <= 0) r.engine = 1;
  if (r.engine < 1 || r.engine > 4) return fail(BGO_ERR_ARG, "engine %d not available", r.engine);
  r.guide_unroll = r.engine == 4 ? 1 : r.engine >= 2 ? 32 : bk::kUnroll;
  P->tune = r;
  return BGO_OK;
}

int plan_configure_launch(bgo_plan* P) {
  int n_cu = 0;
  BGO_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, P->device));
  const bool seed = P->tune.engine == 4;
  P->threads = seed ? bk::kSeedBlock : P->tune.waves_per_block * 64;
  P->group_blocks.clear();
  P->group_lds.clear();
  int min_occ = 1 << 30;
  for (size_t gi = 0; gi < P->group_first.size(); ++gi) {
    const int f_end = gi + 1 < P->group_first.size() ? P->group_first[gi + 1]
                                                     : static_cast<int>(P->fams.size());
    const int nf = f_end - P->group_first[gi];
    const size_t per_wave = bk::wave_lds_bytes(nf, P->tune.engine);
    const size_t lds = seed ? align256(P->group_img_bytes[gi]) + per_wave * (bk::kSeedBlock / 64)
                            : per_wave * P->tune.waves_per_block;
    const int step = seed ? P->group_step[gi] : 8;
    if (seed)
      BGO_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(bk::kernel_for(nf, 4, step)),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(lds)));
    int occ = 0;
    BGO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &occ, bk::kernel_for(nf, P->tune.engine, step), P->threads, lds));
    if (occ <= 0) occ = 1;
    if (P->tune.blocks_per_cu > 0) occ = std::min(occ, P->tune.blocks_per_cu);
    P->n_cu = n_cu;
    min_occ = std::min(min_occ, occ);
    // Wave w takes items from work counter w % kWorkCtrs only, so the grid
    // needs a wave for every counter (blocks past occupancy start later and
    // still drain their counter).
    const int wpb = P->threads / 64;
    const int min_blocks = (static_cast<int>(bk::kWorkCtrs) + wpb - 1) / wpb;
    P->group_blocks.push_back(std::max(occ * n_cu, min_blocks));
    P->group_lds.push_back(lds);
  }
  if (!P->group_first.empty()) P->resolved_blocks_per_cu = min_occ;
  return BGO_OK;
}

}  // namespace

extern "C" int bgo_plan_destroy(bgo_plan* P) {
  if (!P) return BGO_OK;
  if (P->graph) (void)hipGraphExecDestroy(P->graph);
  if (P->d_mem && !P->ext_mem) (void)hipFree(P->d_mem);
  if (P->d_hits && !P->ext_mem) (void)hipFree(P->d_hits);
  if (P->h_pinned) (void)hipHostFree(P->h_pinned);
  if (P->ev0) (void)hipEventDestroy(P->ev0);
  if (P->ev1) (void)hipEventDestroy(P->ev1);
  if (P->stream) (void)hipStreamDestroy(P->stream);
  delete P;
  return BGO_OK;
}

namespace {

int validate_guides(const bgo_guide_desc* guides, int32_t n_guides) {
  if (n_guides < 0 || (n_guides > 0 && !guides)) return fail(BGO_ERR_ARG, "bad guide array");
  for (int32_t i = 0; i < n_guides; ++i) {
    const auto& d = guides[i];
    if (d.pam_len < 1 || d.pam_len > BGO_MAX_PAM_LEN)
      return fail(BGO_ERR_ARG, "guide %d: pam_len %d not in 1..6", i, d.pam_len);
    if (d.max_mm > 20) return fail(BGO_ERR_ARG, "guide %d: max_mm %d > 20", i, d.max_mm);
    if (d.editor > 1) return fail(BGO_ERR_ARG, "guide %d: editor %d", i, d.editor);
  }
  return BGO_OK;
}

// Group guides by PAM family (order of first appearance), build the padded
// tables, grow the device allocation if needed and upload on the plan stream.
// The previous launch must be complete (callers synchronise first).
// Seed engine tables.  A family part holds at most kSeedMaxGuides guides,
// so that its table (5 x 257 key starts, 5 entries per guide, at most 5 x 256
// pad entries, 8 spare words and alignment: at most 92,236 bytes at 4,096 guides) and one
// family's wave buffers for a 1,024-thread block fit the 160 KiB of LDS of a
// CU.
constexpr size_t kSeedMaxGuides = 4096;
constexpr size_t kSeedLdsBytes = 160 * 1024;

// Largest table of a family part of n_guides guides (every key's guide count
// padded to even).
size_t seed_image_bytes(uint32_t n_guides) {
  return (bk::kSeedBlocks * (bk::kSeedKeys + 1) + 3 +
          bk::kSeedBlocks * (size_t{n_guides} + bk::kSeedKeys) + bk::kSeedPad + 3) * 4;
}

// Guide counts of the parts a family of g guides is split into for the seed
// engine: as few parts as hold kSeedMaxGuides each, of equal sizes (the
// first g % n parts one guide larger).
std::vector<uint32_t> seed_part_sizes(size_t g) {
  const size_t n = std::max<size_t>(1, (g + kSeedMaxGuides - 1) / kSeedMaxGuides);
  std::vector<uint32_t> sizes(n, static_cast<uint32_t>(g / n));
  for (size_t p = 0; p < g % n; ++p) ++sizes[p];
  return sizes;
}

// A family's guides (indices into guides, in order) split into seed-engine
// parts of seed_part_sizes.  A window's walk in a part lasts as long as the
// longest key list among the wave's windows, so each guide goes to the part
// (with room left) whose lists for the guide's five block keys are shortest
// so far (on a tie, the part with fewer guides), which evens out the parts'
// list lengths.
std::vector<std::vector<int32_t>> seed_split(const bgo_guide_desc* guides,
                                             const std::vector<int32_t>& members) {
  const std::vector<uint32_t> sizes = seed_part_sizes(members.size());
  const size_t n = sizes.size();
  std::vector<std::vector<int32_t>> parts(n);
  if (n == 1) {
    parts[0] = members;
    return parts;
  }
  std::vector<uint32_t> len(n * bk::kSeedBlocks * bk::kSeedKeys, 0u);
  for (int32_t i : members) {
    uint32_t key[bk::kSeedBlocks];
    for (uint32_t j = 0; j < bk::kSeedBlocks; ++j) {
      uint32_t k = 0;
      for (uint32_t b = 0; b < 4; ++b) {  // the key's plane bits, as the table builds it
        const uint32_t code = static_cast<uint32_t>(guides[i].seq2bit >> (2 * (4 * j + b))) & 3u;
        k |= ((code & 1u) << b) | ((code >> 1) << (4 + b));
      }
      key[j] = k;
    }
    size_t best = n;
    uint32_t best_load = 0;
    for (size_t p = 0; p < n; ++p) {
      if (parts[p].size() >= sizes[p]) continue;
      uint32_t load = 0;
      for (uint32_t j = 0; j < bk::kSeedBlocks; ++j)
        load += len[(p * bk::kSeedBlocks + j) * bk::kSeedKeys + key[j]];
      if (best == n || load < best_load ||
          (load == best_load && parts[p].size() < parts[best].size()))
        best = p, best_load = load;
    }
    parts[best].push_back(i);
    for (uint32_t j = 0; j < bk::kSeedBlocks; ++j)
      ++len[(best * bk::kSeedBlocks + j) * bk::kSeedKeys + key[j]];
  }
  return parts;
}

// LDS of a seed-engine block: the table (its start rounded to 256 bytes for
// the wave buffers after it) and the buffers of its 16 waves.
size_t seed_block_lds(size_t image_bytes, int nf) {
  return (image_bytes + 255) / 256 * 256 + bk::wave_lds_bytes(nf, 4) * (bk::kSeedBlock / 64);
}

// First family of each seed-engine launch group, for family parts of these
// guide counts in order: greedily, as many parts (at most kSeedMaxFam) as
// fit one block's LDS with their tables.
std::vector<int> seed_groups(const std::vector<uint32_t>& part_guides) {
  std::vector<int> first;
  const int n = static_cast<int>(part_guides.size());
  int f = 0;
  while (f < n) {
    first.push_back(f);
    size_t img = seed_image_bytes(part_guides[f]);
    int nf = 1;
    while (f + nf < n && nf < bk::kSeedMaxFam) {
      const size_t more = img + seed_image_bytes(part_guides[f + nf]);
      if (seed_block_lds(more, nf + 1) > kSeedLdsBytes) break;
      img = more;
      ++nf;
    }
    f += nf;
  }
  return first;
}

// Expected VALU of a wave's walks through one block's key lists (sizes padded
// to even) in steps of `step` entries: the wave walks as long as the longest
// list among its 64 windows, E[max] = sum over t >= 0 of 1 - P(steps <= t)^64,
// with the window keys drawn with the base odds of a 41% GC genome; a step of
// 4 entries issues 18 VALU, a step of 8 entries 32.
double seed_walk_cost(const std::vector<uint32_t>& sizes, uint32_t step) {
  static const double kBase[4] = {0.295, 0.205, 0.205, 0.295};  // A C G T
  std::vector<double> p_steps;
  for (uint32_t k = 0; k < bk::kSeedKeys; ++k) {
    double pk = 1.0;
    for (uint32_t b = 0; b < 4; ++b) pk *= kBase[((k >> b) & 1u) | (((k >> (4 + b)) & 1u) << 1)];
    const uint32_t n = (sizes[k] + step - 1) / step;
    if (p_steps.size() <= n) p_steps.resize(n + 1, 0.0);
    p_steps[n] += pk;
  }
  double e = 0.0, cdf = 0.0;
  for (size_t t = 0; t + 1 < p_steps.size(); ++t) {
    cdf += p_steps[t];
    e += 1.0 - std::pow(cdf, 64.0);
  }
  return e * (step == 4 ? 18.0 : 32.0);
}

// Per launch group: the image copied to LDS (per family the layout described
// at kSeedBlocks: 5 x 257 key starts, then the entries of blocks 0..4, keyed
// by the 8-bit key of bases 4j..4j+3, as word pairs of two guides' other 16
// bases per plane), padded to 16 bytes, followed by as many words naming the
// guide behind each entry.  Fam::seed_start / seed_ent are word offsets into
// the group's image; Fam::seed_n counts the family's entries.  Each group's
// walk step (4 or 8 entries) is the one with the lower seed_walk_cost summed
// over its families' blocks.
int build_seed_tables(bgo_plan* P) {
  P->h_btab.clear();
  for (size_t gi = 0; gi < P->group_first.size(); ++gi) {
    const int f0 = P->group_first[gi];
    const int f_end = gi + 1 < P->group_first.size() ? P->group_first[gi + 1]
                                                     : static_cast<int>(P->fams.size());
    std::vector<uint32_t> img, gix;
    double cost4 = 0.0, cost8 = 0.0;  // the group's walks' expected VALU by step
    for (int f = f0; f < f_end; ++f) {
      Fam& F = P->fams[f];
      F.seed_start = static_cast<uint32_t>(img.size());
      // entries start 16-byte aligned (word pairs are read as 8-byte words)
      F.seed_ent = (F.seed_start + bk::kSeedBlocks * (bk::kSeedKeys + 1) + 3) / 4 * 4;
      img.resize(F.seed_ent, 0u);
      uint32_t cur = 0;  // entries so far (always even)
      for (uint32_t j = 0; j < bk::kSeedBlocks; ++j) {
        std::vector<std::vector<uint32_t>> bucket(bk::kSeedKeys);
        for (uint32_t g = F.g_begin; g < F.g_live; ++g) {  // guide order within a key
          const uint32_t p0 = P->h_g01[2 * g], p1 = P->h_g01[2 * g + 1];
          bucket[((p0 >> (4 * j)) & 0xFu) | (((p1 >> (4 * j)) & 0xFu) << 4)].push_back(g);
        }
        for (uint32_t k = 0; k <= bk::kSeedKeys; ++k) {
          const uint32_t odd = k < bk::kSeedKeys ? static_cast<uint32_t>(bucket[k].size() & 1u) : 0u;
          img[F.seed_start + j * (bk::kSeedKeys + 1) + k] = cur | odd;
          if (k == bk::kSeedKeys) break;
          auto& b = bucket[k];
          // The pad entry repeats the key's last guide, so that it passes the
          // mismatch test exactly when that guide does, in the same step (it is
          // never in its key's range).
          if (odd) b.push_back(b.back());
          for (size_t e = 0; e < b.size(); e += 2) {
            uint32_t lo = 0, hi = 0;
            for (uint32_t h = 0; h < 2; ++h) {
              const uint32_t g = b[e + h];
              lo |= bk::seed_other(P->h_g01[2 * g], j) << (16 * h);
              hi |= bk::seed_other(P->h_g01[2 * g + 1], j) << (16 * h);
            }
            img.push_back(lo);
            img.push_back(hi);
            gix.resize(img.size(), 0u);
            gix[F.seed_ent + cur + e] = b[e];
            gix[F.seed_ent + cur + e + 1] = b[e + 1];
          }
          cur += static_cast<uint32_t>(b.size());
        }
        std::vector<uint32_t> sizes(bk::kSeedKeys);
        for (uint32_t k = 0; k < bk::kSeedKeys; ++k) sizes[k] = static_cast<uint32_t>(bucket[k].size());
        cost4 += seed_walk_cost(sizes, 4);
        cost8 += seed_walk_cost(sizes, 8);
      }
      F.seed_n = cur;
      img.resize((img.size() + bk::kSeedPad + 3) / 4 * 4, 0u);
      gix.resize(img.size(), 0u);
    }
    img.resize((img.size() + 3) / 4 * 4, 0u);
    gix.resize(img.size(), 0u);
    const size_t off = (P->h_btab.size() + 255) / 256 * 256;
    const size_t bytes = img.size() * 4;
    if (seed_block_lds(bytes, f_end - f0) > kSeedLdsBytes)
      return fail(BGO_ERR_ARG, "seed table of %zu bytes does not fit a block's LDS", bytes);
    P->h_btab.resize(off + 2 * bytes, 0);
    std::memcpy(P->h_btab.data() + off, img.data(), bytes);
    std::memcpy(P->h_btab.data() + off + bytes, gix.data(), bytes);
    P->group_img_off.push_back(off);
    P->group_step.push_back(cost4 < cost8 ? 4 : 8);
    P->group_img_bytes.push_back(bytes);
  }
  return BGO_OK;
}

int plan_set_guides(bgo_plan* P, const bgo_guide_desc* guides, int32_t n_guides) {
  BGO_TRY(validate_guides(guides, n_guides));
  std::map<uint64_t, int> fam_index;
  std::vector<std::vector<int32_t>> members;
  std::vector<Fam> proto;
  for (int32_t i = 0; i < n_guides; ++i) {
    const uint64_t key = fam_key(guides[i]);
    auto it = fam_index.find(key);
    if (it == fam_index.end()) {
      it = fam_index.emplace(key, static_cast<int>(members.size())).first;
      members.emplace_back();
      proto.push_back(make_fam(guides[i]));
    }
    members[it->second].push_back(i);
  }
  int n_fam = static_cast<int>(members.size());
  if (P->engine_auto) {
    // The engine with the lower modelled cost.  Per family, in ms per unit of
    // PAM pass rate on a 3.09 Gbp genome (fitted to single-family and mixed
    // launches on MI355X): VALU in place 0.86 + 0.94 G, VALU through the
    // window buffer 6.7 + 3.7 ceil(G / 8), fp4 8.1 + 4.0 ceil(G / 32), seed
    // (budgets <= 4 only) 36 + 27 G / 1024, plus 0.9 ms for each launch the
    // seed tables need beyond the first (a launch streams the genome once).
    // The streaming all engines share drops out.
    double valu = 0.0, fp4 = 0.0, seed = 0.0;
    bool seed_ok = true;
    std::vector<uint32_t> parts;
    for (int f = 0; f < n_fam; ++f) {
      const double pass = pam_pass_rate(proto[f].nibs);
      const int g = static_cast<int>(members[f].size());
      valu += pass * (g <= direct_limit(P, pass) ? 0.86 + 0.94 * g : 6.7 + 3.7 * ((g + 7) / 8));
      fp4 += pass * (8.1 + 4.0 * ((g + 31) / 32));
      seed += pass * (36.0 + 27.0 * g / 1024.0);
      for (int32_t i : members[f]) seed_ok = seed_ok && guides[i].max_mm <= 4;
      for (uint32_t n : seed_part_sizes(static_cast<size_t>(g))) parts.push_back(n);
    }
    seed += 0.9 * (static_cast<double>(seed_groups(parts).size()) - 1.0);
    P->tune.engine = fp4 < valu ? 3 : 1;
    if (seed_ok && seed < std::min(fp4, valu)) P->tune.engine = 4;
    P->tune.guide_unroll = P->tune.engine == 4 ? 1 : P->tune.engine >= 2 ? 32 : bk::kUnroll;
  }
  if (P->tune.engine == 4) {
    // The seed engine's pigeonhole (five exact 4-base blocks) covers budgets
    // up to 4; a family's table holds at most kSeedMaxGuides guides, so
    // larger families are split (the parts scan in launches of their own).
    for (int32_t i = 0; i < n_guides; ++i)
      if (guides[i].max_mm > 4)
        return fail(BGO_ERR_ARG, "engine 4 (seed) needs budgets <= 4; guide %d has %d", i,
                    guides[i].max_mm);
    std::vector<std::vector<int32_t>> parts;
    std::vector<Fam> part_proto;
    for (int f = 0; f < n_fam; ++f)
      for (auto& part : seed_split(guides, members[f])) {
        parts.push_back(std::move(part));
        part_proto.push_back(proto[f]);
      }
    members.swap(parts);
    proto.swap(part_proto);
  }
  const int n_fam_parts = static_cast<int>(members.size());
  P->fams.clear();
  P->fam_guides.clear();
  P->group_first.clear();
  P->h_g01.clear();
  P->h_gaux.clear();
  P->h_gid.clear();
  P->h_gnm.clear();
  P->h_gk.clear();
  auto& g01 = P->h_g01;
  auto& gaux = P->h_gaux;
  auto& gid = P->h_gid;
  auto& gnm = P->h_gnm;
  auto& gk = P->h_gk;
  for (int f = 0; f < n_fam_parts; ++f) {
    Fam F = proto[f];
    F.g_begin = static_cast<uint32_t>(gk.size());
    uint32_t kmax = 0;
    for (int32_t i : members[f]) {
      const auto& d = guides[i];
      uint32_t p0 = 0, p1 = 0, eqA = 0, eqT = 0;
      for (int b = 0; b < 20; ++b) {
        const uint32_t code = static_cast<uint32_t>(d.seq2bit >> (2 * b)) & 3u;
        p0 |= (code & 1u) << b;
        p1 |= (code >> 1) << b;
        if (code == 0) eqA |= 1u << b;
        if (code == 3) eqT |= 1u << b;
      }
      g01.push_back(p0);
      g01.push_back(p1);
      gk.push_back(d.max_mm);
      gaux.push_back((d.edit_mask & bk::kWin) | (static_cast<uint32_t>(d.editor & 1) << 24));
      gid.push_back(static_cast<uint32_t>(i));
      gnm.push_back(eqA);
      gnm.push_back(eqT);
      kmax = std::max<uint32_t>(kmax, d.max_mm);
    }
    const size_t pad_to = P->tune.engine == 4 ? 1 : P->tune.engine >= 2 ? 32 : bk::kUnroll;
    while ((gk.size() - F.g_begin) % pad_to) {  // padding guides never match
      g01.push_back(0xFFF00000u);
      g01.push_back(0xFFF00000u);
      gk.push_back(-1);
      gaux.push_back(0);
      gid.push_back(0xFFFFFFFFu);
      gnm.push_back(0);
      gnm.push_back(0);
    }
    F.g_end = static_cast<uint32_t>(gk.size());
    F.g_live = F.g_begin + static_cast<uint32_t>(members[f].size());
    F.kmax = kmax;
    F.direct = P->tune.engine == 1 &&
               static_cast<int>(members[f].size()) <= direct_limit(P, pam_pass_rate(F.nibs));
    P->fams.push_back(F);
    P->fam_guides.push_back(static_cast<int32_t>(members[f].size()));
  }
  if (P->tune.engine == 4) {
    // Launch groups of families whose seed tables and wave buffers fit one
    // block's LDS together.
    std::vector<uint32_t> sizes;
    for (const Fam& F : P->fams) sizes.push_back(F.g_live - F.g_begin);
    P->group_first = seed_groups(sizes);
  } else {
    for (int f = 0; f < n_fam_parts; f += bk::kMaxFam) P->group_first.push_back(f);
  }
  n_fam = n_fam_parts;
  const size_t n_pad = gk.size();
  P->n_guides = n_guides;

  // MFMA guide operands (the tables end with spare all-zero tiles for the
  // kernel's look-ahead loads: one for int8, four for fp4).
  //  * int8: per 32-guide tile, per k-half j and lane l (column l % 32, lane
  //    group l / 32) the 16 int8 of k = 32j + 16(l/32) .. +15 of the guide's
  //    tetrahedral code; k 60..63 hold the family's kmax - 15; padding
  //    guides hold -127 there (never a candidate).
  //  * fp4: per 32-guide tile and lane l the 32 nibbles of elements
  //    32(l/32) .. +31 (element 32(l/32) + i in dword i%4 at bits 4(i/4), the
  //    kernel's A layout); elements 60..63 sum to budget - (non-A bases);
  //    padding guides are 0 with -6 x 4 there.
  P->h_btab.clear();
  if (P->tune.engine == 2) {
    P->h_btab.assign((n_pad / 32 + 1) * 2 * 64 * 16, 0);
    for (const Fam& F : P->fams) {
      const int8_t bias = static_cast<int8_t>(static_cast<int>(F.kmax) - 15);
      for (size_t gi = F.g_begin; gi < F.g_end; ++gi) {
        const bool real = gk[gi] >= 0;
        const uint64_t g0 = g01[2 * gi], g1 = g01[2 * gi + 1];
        const uint64_t bits = real ? (g1 | (g0 << 20) | ((g0 ^ g1) << 40)) : 0;
        const size_t tile = gi / 32, col = gi % 32;
        for (int j = 0; j < 2; ++j)
          for (int grp = 0; grp < 2; ++grp) {
            int8_t* dst = P->h_btab.data() + ((tile * 2 + j) * 64 + grp * 32 + col) * 16;
            for (int b = 0; b < 16; ++b) {
              const int kk = 32 * j + 16 * grp + b;
              if (kk >= 60)
                dst[b] = real ? bias : static_cast<int8_t>(-127);
              else
                dst[b] = real ? (((bits >> kk) & 1u) ? -1 : 1) : 0;
            }
          }
      }
    }
  } else if (P->tune.engine == 3) {
    P->h_btab.assign((n_pad / 32 + 4) * 64 * 16, 0);
    for (const Fam& F : P->fams) {
      for (size_t gi = F.g_begin; gi < F.g_end; ++gi) {
        uint8_t nib[64] = {};
        if (gk[gi] >= 0) {
          const uint32_t g0 = g01[2 * gi] & bk::kWin, g1 = g01[2 * gi + 1] & bk::kWin;
          const uint32_t x = g0 ^ g1;
          // the window code words of lane groups 0 and 1 (bias bits clear)
          const uint32_t code[2] = {g0 | ((x & 0xFFFu) << 20), g1 | ((x & 0xFF000u) << 8)};
          // stored weight +-1 / (window element value) for i % 4 = 0, 1, 2, 3
          static const uint8_t pos_w[4] = {0x4, 0x2, 0x1, 0x4};  // 2, 1, 0.5, 2
          for (int k = 0; k < 60; ++k) {
            const int grp = k / 32, i = k % 32;
            const uint8_t w = pos_w[i % 4];
            nib[k] = ((code[grp] >> i) & 1u) ? w : static_cast<uint8_t>(w | 0x8);
          }
          const int bias = gk[gi] - __builtin_popcount(static_cast<uint32_t>(g0 | g1) & bk::kWin);
          if (!fp4_sum4(bias, nib + 60)) return fail(BGO_ERR_ARG, "guide bias %d out of range", bias);
        } else {
          for (int k = 60; k < 64; ++k) nib[k] = 0xF;  // -6
        }
        const size_t tile = gi / 32, col = gi % 32;
        for (int grp = 0; grp < 2; ++grp) {
          uint32_t w[4] = {0, 0, 0, 0};
          for (int i = 0; i < 32; ++i) w[i % 4] |= static_cast<uint32_t>(nib[32 * grp + i]) << (4 * (i / 4));
          std::memcpy(P->h_btab.data() + (tile * 64 + grp * 32 + col) * 16, w, 16);
        }
      }
    }
  }
  P->group_img_off.clear();
  P->group_step.clear();
  P->group_img_bytes.clear();
  if (P->tune.engine == 4) BGO_TRY(build_seed_tables(P));
  const size_t btab_bytes = P->h_btab.size();

  // [hit count][windows per family][seed verifies per family][pad to 256 B][work counters]
  P->work_word = (1 + 2 * n_fam + 31) / 32 * 32;
  P->n_counter_words = P->work_word + static_cast<uint32_t>(P->group_first.size()) *
                                         bk::kWorkCtrs * (bk::kCtrStride / 2);
  size_t off = 0;
  const size_t o_fam = off; off = align256(off + std::max<size_t>(n_fam, 1) * sizeof(Fam));
  const size_t o_g01 = off; off = align256(off + std::max<size_t>(n_pad, 1) * 8);
  const size_t o_gk = off; off = align256(off + std::max<size_t>(n_pad, 1) * 4);
  const size_t o_gaux = off; off = align256(off + std::max<size_t>(n_pad, 1) * 4);
  const size_t o_gid = off; off = align256(off + std::max<size_t>(n_pad, 1) * 4);
  const size_t o_gnm = off; off = align256(off + std::max<size_t>(n_pad, 1) * 8);
  const size_t o_score = off; off = align256(off + 21 * sizeof(double));
  const size_t o_cnt = off; off = align256(off + P->n_counter_words * 8);
  const size_t o_btab = off; off = align256(off + std::max<size_t>(btab_bytes, 16));
  if (off > P->mem_bytes && P->ext_mem)
    return fail(BGO_ERR_ARG,
                "guide tables need %zu bytes; the no-alloc scratch array holds %zu "
                "(at most %d guides in at most %d PAM families)",
                off, P->mem_bytes, BGO_NOALLOC_MAX_GUIDES, BGO_NOALLOC_MAX_FAMILIES);
  if (off > P->mem_bytes) {
    // At least 4 MB (tables of ~16K padded guides): freeing device memory waits
    // for the device's running kernels, which would stall a serving loop
    // that swaps guide sets while another plan's scan runs.
    const size_t want = std::max({off, P->mem_bytes * 3 / 2, size_t{4} << 20});
    if (P->d_mem) (void)hipFree(P->d_mem);
    P->d_mem = nullptr;
    P->mem_bytes = 0;
    BGO_HIP(hipMalloc(&P->d_mem, want));
    P->mem_bytes = want;
  }
  char* b = static_cast<char*>(P->d_mem);
  P->d_fam = reinterpret_cast<Fam*>(b + o_fam);
  P->d_g01 = reinterpret_cast<uint32_t*>(b + o_g01);
  P->d_gk = reinterpret_cast<int32_t*>(b + o_gk);
  P->d_gaux = reinterpret_cast<uint32_t*>(b + o_gaux);
  P->d_gid = reinterpret_cast<uint32_t*>(b + o_gid);
  P->d_gnm = reinterpret_cast<uint32_t*>(b + o_gnm);
  P->d_score = reinterpret_cast<double*>(b + o_score);
  P->d_counters = reinterpret_cast<unsigned long long*>(b + o_cnt);
  P->d_btab = reinterpret_cast<int8_t*>(b + o_btab);
  hipStream_t st = P->stream;
  if (btab_bytes)
    BGO_HIP(hipMemcpyAsync(P->d_btab, P->h_btab.data(), btab_bytes, hipMemcpyHostToDevice, st));
  if (n_fam) BGO_HIP(hipMemcpyAsync(P->d_fam, P->fams.data(), n_fam * sizeof(Fam), hipMemcpyHostToDevice, st));
  if (n_pad) {
    BGO_HIP(hipMemcpyAsync(P->d_g01, g01.data(), n_pad * 8, hipMemcpyHostToDevice, st));
    BGO_HIP(hipMemcpyAsync(P->d_gk, gk.data(), n_pad * 4, hipMemcpyHostToDevice, st));
    BGO_HIP(hipMemcpyAsync(P->d_gaux, gaux.data(), n_pad * 4, hipMemcpyHostToDevice, st));
    BGO_HIP(hipMemcpyAsync(P->d_gid, gid.data(), n_pad * 4, hipMemcpyHostToDevice, st));
    BGO_HIP(hipMemcpyAsync(P->d_gnm, gnm.data(), n_pad * 8, hipMemcpyHostToDevice, st));
  }
  BGO_HIP(hipMemcpyAsync(P->d_score, P->score, 21 * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(bk::reset_counters_kernel, dim3(1), dim3(256), 0, st, P->d_counters,
                     P->n_counter_words);
  BGO_HIP(hipGetLastError());
  if (P->graph) {  // kernel arguments changed
    (void)hipGraphExecDestroy(P->graph);
    P->graph = nullptr;
    P->graph_mask = -1;
  }
  return plan_configure_launch(P);
}

}  // namespace

extern "C" int bgo_plan_create(bgo_genome* G, const bgo_guide_desc* guides, int32_t n_guides,
                               const double* score_table, const bgo_tuning* tuning,
                               int64_t hit_capacity, bgo_plan** out) {
  if (!G || !out || !score_table || hit_capacity < 0)
    return fail(BGO_ERR_ARG, "bgo_plan_create: bad arguments");
  BGO_TRY(validate_guides(guides, n_guides));
  BGO_TRY(injected());
  auto* P = new bgo_plan();
  auto bail = [&](int rc) {
    bgo_plan_destroy(P);
    return rc;
  };
  P->g = G;
  P->device = G->device;
  P->hit_cap = hit_capacity;
  std::memcpy(P->score, score_table, sizeof(P->score));
  if (hipSetDevice(G->device) != hipSuccess) return bail(fail(BGO_ERR_HIP, "hipSetDevice"));
  if (int rc = resolve_tuning(P, tuning); rc != BGO_OK) return bail(rc);
  if (hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&P->ev0) != hipSuccess || hipEventCreate(&P->ev1) != hipSuccess)
    return bail(fail(BGO_ERR_HIP, "stream/event create"));
  if (hit_capacity > 0 && hipMalloc(&P->d_hits, hit_capacity * sizeof(bgo_hit)) != hipSuccess)
    return bail(fail(BGO_ERR_HIP, "hipMalloc hit buffer (%lld records)", (long long)hit_capacity));
  if (int rc = plan_set_guides(P, guides, n_guides); rc != BGO_OK) return bail(rc);
  if (hipStreamSynchronize(P->stream) != hipSuccess) return bail(fail(BGO_ERR_HIP, "upload plan tables"));
  *out = P;
  return BGO_OK;
}

extern "C" int bgo_plan_set_guides(bgo_plan* P, const bgo_guide_desc* guides, int32_t n_guides) {
  if (!P) return fail(BGO_ERR_ARG, "null plan");
  BGO_TRY(injected());
  std::lock_guard<std::mutex> lock(P->mu);
  BGO_HIP(hipSetDevice(P->device));
  BGO_HIP(hipStreamSynchronize(P->stream));
  return plan_set_guides(P, guides, n_guides);
}

extern "C" int bgo_plan_set_end_guard(bgo_plan* P, int32_t end_guard) {
  if (!P || end_guard < 21 || end_guard > 64)
    return fail(BGO_ERR_ARG, "end_guard must be in [21, 64]");
  BGO_TRY(injected());
  std::lock_guard<std::mutex> lock(P->mu);
  BGO_HIP(hipSetDevice(P->device));
  BGO_HIP(hipStreamSynchronize(P->stream));
  P->end_guard = end_guard;
  if (P->graph) {
    (void)hipGraphExecDestroy(P->graph);
    P->graph = nullptr;
    P->graph_mask = -1;
  }
  return BGO_OK;
}

extern "C" int bgo_plan_set_hit_capacity(bgo_plan* P, int64_t hit_capacity) {
  if (!P || hit_capacity < 0) return fail(BGO_ERR_ARG, "bad arguments");
  BGO_TRY(injected());
  std::lock_guard<std::mutex> lock(P->mu);
  BGO_HIP(hipSetDevice(P->device));
  BGO_HIP(hipStreamSynchronize(P->stream));
  if (P->d_hits) (void)hipFree(P->d_hits);
  P->d_hits = nullptr;
  P->hit_cap = 0;
  if (hit_capacity > 0) BGO_HIP(hipMalloc(&P->d_hits, hit_capacity * sizeof(bgo_hit)));
  P->hit_cap = hit_capacity;
  if (P->graph) {
    (void)hipGraphExecDestroy(P->graph);
    P->graph = nullptr;
    P->graph_mask = -1;
  }
  return BGO_OK;
}

namespace {

// Enqueue counter reset + one kernel per family group on P->stream.
int enqueue_scan(bgo_plan* P, int64_t pos_begin, int64_t pos_end, int32_t strand_mask) {
  bgo_genome* G = P->g;
  hipLaunchKernelGGL(bk::reset_counters_kernel, dim3(1), dim3(256), 0, P->stream, P->d_counters,
                     P->n_counter_words);
  BGO_HIP(hipGetLastError());
  const int64_t hard_end = G->total_bp - P->end_guard;  // 26: cuda_search.py:211-213
  const int64_t vb = std::max<int64_t>(pos_begin, 0);
  const int64_t ve = std::min<int64_t>(pos_end, hard_end);
  uint32_t slots[2];
  uint32_t n_slots = 0;
  if (strand_mask & 1) slots[n_slots++] = 0;
  if ((strand_mask & 2) && G->n_strands > 1) slots[n_slots++] = 1;
  if (ve <= vb || n_slots == 0 || P->fams.empty()) return BGO_OK;
  const uint32_t chunk = static_cast<uint32_t>(P->tune.chunk_positions);
  const uint64_t cps = (static_cast<uint64_t>(ve - vb) + chunk - 1) / chunk;
  if (cps * n_slots >= 0xFFFFFFF0ull) return fail(BGO_ERR_ARG, "too many work items");
  const bool seed = P->tune.engine == 4;
  for (size_t gi = 0; gi < P->group_first.size(); ++gi) {
    const int f0 = P->group_first[gi];
    const int f_end = gi + 1 < P->group_first.size() ? P->group_first[gi + 1]
                                                     : static_cast<int>(P->fams.size());
    const int nf = f_end - f0;
    bk::Args a{};
    a.planes0 = G->d_planes[0];
    a.planes1 = G->d_planes[1];
    a.nwords0 = G->d_nwords[0];
    a.nwords1 = G->d_nwords[1];
    a.slot_strand0 = slots[0];
    a.slot_strand1 = n_slots > 1 ? slots[1] : slots[0];
    a.valid_begin = static_cast<uint32_t>(vb);
    a.valid_end = static_cast<uint32_t>(ve);
    a.chunk = chunk;
    a.chunks_per_slot = static_cast<uint32_t>(cps);
    a.n_items = static_cast<uint32_t>(cps * n_slots);
    a.n_slots = n_slots;
    a.interleave = P->tune.reserved[0] ? 1u : 0u;
    a.n_fam = static_cast<uint32_t>(nf);
    a.fam_total = static_cast<uint32_t>(P->fams.size());
    a.btab_bytes = static_cast<uint32_t>(seed ? P->group_img_bytes[gi] : P->h_btab.size());
    a.hit_cap = static_cast<unsigned long long>(P->hit_cap);
    const int8_t* btab = seed ? P->d_btab + P->group_img_off[gi] : P->d_btab;
    hipLaunchKernelGGL(bk::kernel_for(nf, P->tune.engine, seed ? P->group_step[gi] : 8),
                       dim3(P->group_blocks[gi]), dim3(P->threads),
                       P->group_lds[gi], P->stream, a, P->d_fam + f0, P->d_g01, P->d_gk, P->d_gaux, P->d_gid,
                       P->d_gnm, P->d_score, P->d_hits, P->d_counters,
                       reinterpret_cast<unsigned int*>(P->d_counters + P->work_word) +
                           gi * bk::kWorkCtrs * bk::kCtrStride,
                       P->d_counters + 1 + f0,
                       reinterpret_cast<const int4*>(btab));
    BGO_HIP(hipGetLastError());
  }
  return BGO_OK;
}

}  // namespace

extern "C" int bgo_plan_launch(bgo_plan* P, int64_t pos_begin, int64_t pos_end,
                               int32_t strand_mask) {
  if (!P) return fail(BGO_ERR_ARG, "null plan");
  BGO_TRY(injected());
  std::lock_guard<std::mutex> lock(P->mu);
  BGO_HIP(hipSetDevice(P->device));
  BGO_HIP(hipEventRecord(P->ev0, P->stream));
  BGO_TRY(enqueue_scan(P, pos_begin, pos_end, strand_mask));
  BGO_HIP(hipEventRecord(P->ev1, P->stream));
  return BGO_OK;
}

extern "C" int bgo_plan_graph_launch(bgo_plan* P, int32_t strand_mask) {
  if (!P) return fail(BGO_ERR_ARG, "null plan");
  BGO_TRY(injected());
  std::lock_guard<std::mutex> lock(P->mu);
  BGO_HIP(hipSetDevice(P->device));
  if (!P->graph || P->graph_mask != strand_mask) {
    if (P->graph) {
      (void)hipGraphExecDestroy(P->graph);
      P->graph = nullptr;
    }
    hipGraph_t graph = nullptr;
    BGO_HIP(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_scan(P, 0, P->g->total_bp, strand_mask);
    const hipError_t ce = hipStreamEndCapture(P->stream, &graph);
    if (rc != BGO_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (ce != hipSuccess) return fail(BGO_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
    const hipError_t ie = hipGraphInstantiate(&P->graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(BGO_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    P->graph_mask = strand_mask;
  }
  BGO_HIP(hipEventRecord(P->ev0, P->stream));
  BGO_HIP(hipGraphLaunch(P->graph, P->stream));
  BGO_HIP(hipEventRecord(P->ev1, P->stream));
  return BGO_OK;
}

extern "C" int bgo_plan_wait(bgo_plan* P, int64_t* n_hits_out, float* kernel_ms_out) {
  if (!P || !n_hits_out) return fail(BGO_ERR_ARG, "null pointer");
  BGO_TRY(injected());
  std::lock_guard<std::mutex> lock(P->mu);
  BGO_HIP(hipSetDevice(P->device));
  BGO_HIP(hipStreamSynchronize(P->stream));
  unsigned long long n = 0;
  BGO_HIP(hipMemcpyAsync(&n, P->d_counters, sizeof(n), hipMemcpyDeviceToHost, P->stream));
  BGO_HIP(hipStreamSynchronize(P->stream));
  *n_hits_out = static_cast<int64_t>(n);
  if (kernel_ms_out) {
    float ms = 0.f;
    BGO_HIP(hipEventElapsedTime(&ms, P->ev0, P->ev1));
    *kernel_ms_out = ms;
  }
  return BGO_OK;
}

extern "C" int bgo_plan_fetch(bgo_plan* P, bgo_hit* out, int64_t cap, int64_t* n_copied) {
  if (!P || !n_copied || (cap > 0 && !out)) return fail(BGO_ERR_ARG, "bad arguments");
  BGO_TRY(injected());
  std::lock_guard<std::mutex> lock(P->mu);
  BGO_HIP(hipSetDevice(P->device));
  BGO_HIP(hipStreamSynchronize(P->stream));
  unsigned long long n = 0;
  BGO_HIP(hipMemcpyAsync(&n, P->d_counters, sizeof(n), hipMemcpyDeviceToHost, P->stream));
  BGO_HIP(hipStreamSynchronize(P->stream));
  const int64_t m = std::min<int64_t>({static_cast<int64_t>(n), P->hit_cap, cap});
  if (m > 0) {
    BGO_HIP(hipMemcpyAsync(out, P->d_hits, m * sizeof(bgo_hit), hipMemcpyDeviceToHost, P->stream));
    BGO_HIP(hipStreamSynchronize(P->stream));
  }
  *n_copied = std::max<int64_t>(m, 0);
  return BGO_OK;
}

extern "C" int bgo_plan_fetch_pinned(bgo_plan* P, int64_t cap, const bgo_hit** out,
                                     int64_t* n_copied) {
  if (!P || !out || !n_copied) return fail(BGO_ERR_ARG, "bad arguments");
  BGO_TRY(injected());
  std::lock_guard<std::mutex> lock(P->mu);
  BGO_HIP(hipSetDevice(P->device));
  BGO_HIP(hipStreamSynchronize(P->stream));
  unsigned long long n = 0;
  BGO_HIP(hipMemcpyAsync(&n, P->d_counters, sizeof(n), hipMemcpyDeviceToHost, P->stream));
  BGO_HIP(hipStreamSynchronize(P->stream));
  int64_t m = std::min<int64_t>(static_cast<int64_t>(n), P->hit_cap);
  if (cap >= 0) m = std::min(m, cap);
  m = std::max<int64_t>(m, 0);
  if (m > P->pinned_cap) {
    // grows at least 2x from 64K records: freeing page-locked memory can wait
    // for the device's running kernels
    const int64_t want = std::max<int64_t>({m, P->pinned_cap * 2, int64_t{1} << 16});
    if (P->h_pinned) (void)hipHostFree(P->h_pinned);
    P->h_pinned = nullptr;
    P->pinned_cap = 0;
    BGO_HIP(hipHostMalloc(reinterpret_cast<void**>(&P->h_pinned), want * sizeof(bgo_hit),
                          hipHostMallocDefault));
    P->pinned_cap = want;
  }
  if (m > 0) {
    BGO_HIP(hipMemcpyAsync(P->h_pinned, P->d_hits, m * sizeof(bgo_hit), hipMemcpyDeviceToHost,
                           P->stream));
    BGO_HIP(hipStreamSynchronize(P->stream));
  }
  *out = P->h_pinned;
  *n_copied = m;
  return BGO_OK;
}

extern "C" int bgo_plan_get_tuning(const bgo_plan* P, bgo_tuning* out) {
  if (!P || !out) return fail(BGO_ERR_ARG, "null pointer");
  *out = P->tune;
  if (out->blocks_per_cu <= 0) out->blocks_per_cu = P->resolved_blocks_per_cu;
  out->reserved[1] = P->direct_off ? -1 : P->direct_max;
  return BGO_OK;
}

extern "C" int bgo_plan_family_stats(const bgo_plan* P, int32_t cap, int32_t* n_fam_out,
                                     int64_t* windows_out, int32_t* guides_out) {
  if (!P || !n_fam_out) return fail(BGO_ERR_ARG, "null pointer");
  const int n_fam = static_cast<int>(P->fams.size());
  *n_fam_out = n_fam;
  if (cap <= 0) return BGO_OK;
  BGO_TRY(injected());
  std::vector<unsigned long long> w(std::max(n_fam, 1));
  if (n_fam) {
    BGO_HIP(hipSetDevice(P->device));
    BGO_HIP(hipStreamSynchronize(P->stream));
    BGO_HIP(hipMemcpyAsync(w.data(), P->d_counters + 1,
                           n_fam * sizeof(unsigned long long), hipMemcpyDeviceToHost, P->stream));
    BGO_HIP(hipStreamSynchronize(P->stream));
  }
  for (int f = 0; f < std::min(cap, n_fam); ++f) {
    if (windows_out) windows_out[f] = static_cast<int64_t>(w[f]);
    if (guides_out) guides_out[f] = P->fam_guides[f];
  }
  return BGO_OK;
}

extern "C" int bgo_plan_seed_stats(const bgo_plan* P, int32_t cap, int64_t* verified_out) {
  if (!P || (cap > 0 && !verified_out)) return fail(BGO_ERR_ARG, "null pointer");
  const int n_fam = static_cast<int>(P->fams.size());
  if (cap <= 0 || n_fam == 0) return BGO_OK;
  BGO_TRY(injected());
  std::vector<unsigned long long> v(n_fam);
  BGO_HIP(hipSetDevice(P->device));
  BGO_HIP(hipStreamSynchronize(P->stream));
  BGO_HIP(hipMemcpyAsync(v.data(), P->d_counters + 1 + n_fam, n_fam * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, P->stream));
  BGO_HIP(hipStreamSynchronize(P->stream));
  for (int f = 0; f < std::min(cap, n_fam); ++f) verified_out[f] = static_cast<int64_t>(v[f]);
  return BGO_OK;
}

extern "C" int bgo_scan_batched(bgo_genome* G, const bgo_guide_desc* guides, int32_t n_guides,
                                const double* score_table, const bgo_tuning* tuning, bgo_hit* out,
                                int64_t out_capacity, int64_t* n_hits_out) {
  if (!n_hits_out || out_capacity < 0 || (out_capacity > 0 && !out))
    return fail(BGO_ERR_ARG, "bgo_scan_batched: bad arguments");
  bgo_plan* P = nullptr;
  BGO_TRY(bgo_plan_create(G, guides, n_guides, score_table, tuning, out_capacity, &P));
  int rc = bgo_plan_launch(P, 0, G->total_bp, 3);
  int64_t n = 0, copied = 0;
  if (rc == BGO_OK) rc = bgo_plan_wait(P, &n, nullptr);
  if (rc == BGO_OK) rc = bgo_plan_fetch(P, out, out_capacity, &copied);
  bgo_plan_destroy(P);
  if (rc == BGO_OK) *n_hits_out = n;
  return rc;
}

namespace bgo {

// The last launch's records on the device (for the table assembly):
// stored = min(sites, capacity).  Waits for the launch.
int plan_device_hits(bgo_plan* P, const bgo_hit** hits, int64_t* stored, int64_t* total,
                     hipStream_t* stream, int32_t* n_guides) {
  std::lock_guard<std::mutex> lock(P->mu);
  BGO_HIP(hipSetDevice(P->device));
  BGO_HIP(hipStreamSynchronize(P->stream));
  unsigned long long n = 0;
  BGO_HIP(hipMemcpyAsync(&n, P->d_counters, sizeof(n), hipMemcpyDeviceToHost, P->stream));
  BGO_HIP(hipStreamSynchronize(P->stream));
  *hits = P->d_hits;
  *total = static_cast<int64_t>(n);
  *stored = std::min<int64_t>(static_cast<int64_t>(n), P->hit_cap);
  *stream = P->stream;
  *n_guides = P->n_guides;
  return BGO_OK;
}

}  // namespace bgo

// ------------------------------------------------------------ no-alloc scan
//
// bgo_scan_batched_noalloc runs bgo_scan_batched's kernels with every device
// buffer they use (guide tables, family and work counters, hit records) in
// `scratch`, a file-scope device array of a fixed size, so a call allocates
// no device or page-locked memory.  Each device that loads the library holds
// its own copy of the array until the process ends.  The host-side name
// `scratch` is only a stand-in for that memory: bgo_init resolves the device
// address once per device with hipGetSymbolAddress, and only that address
// is used afterwards (kernel arguments, copies).

namespace bgo {
namespace na {

constexpr size_t kTableBytes = size_t{1} << 20;  // guide tables + counters
constexpr size_t kHitBytes = static_cast<size_t>(BGO_NOALLOC_MAX_HITS) * sizeof(bgo_hit);
constexpr size_t kScratchBytes = kTableBytes + kHitBytes;
constexpr int kMaxDevices = 64;

struct Device {
  std::mutex mu;        // serialises the calls on this device
  bool ready = false;
  char* ptr = nullptr;  // device address of this device's `scratch`
  bgo_plan* plan = nullptr;  // stream, events and host tables, reused per call
};

Device& device_state(int device) {
  static Device devs[kMaxDevices];
  return devs[device];
}

}  // namespace na
}  // namespace bgo

static __device__ __align__(256) unsigned char scratch[bgo::na::kScratchBytes];

namespace bgo {

// Called by bgo_init with `device` current: resolve this device's copy of
// `scratch` and build the reusable plan (stream and timing events).
int noalloc_setup(int device) {
  if (device < 0 || device >= na::kMaxDevices)
    return fail(BGO_ERR_ARG, "device %d: the no-alloc scan supports devices 0..%d", device,
                na::kMaxDevices - 1);
  na::Device& d = na::device_state(device);
  std::lock_guard<std::mutex> lock(d.mu);
  if (d.ready) return BGO_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(BGO_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                hipGetErrorString(e));
  auto* P = new bgo_plan();
  P->device = device;
  P->ext_mem = true;
  P->d_mem = p;
  P->mem_bytes = na::kTableBytes;
  P->d_hits = reinterpret_cast<bgo_hit*>(static_cast<char*>(p) + na::kTableBytes);
  if (hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&P->ev0) != hipSuccess || hipEventCreate(&P->ev1) != hipSuccess) {
    bgo_plan_destroy(P);
    return fail(BGO_ERR_HIP, "no-alloc stream/event create on device %d", device);
  }
  d.ptr = static_cast<char*>(p);
  d.plan = P;
  d.ready = true;
  return BGO_OK;
}

}  // namespace bgo

extern "C" int bgo_scan_batched_noalloc(bgo_genome* G, const bgo_guide_desc* guides,
                                        int32_t n_guides, const double* score_table,
                                        const bgo_tuning* tuning, bgo_hit* out,
                                        int64_t out_capacity, int64_t* n_hits_out) {
  if (!G || !score_table || !n_hits_out || out_capacity < 0 || (out_capacity > 0 && !out))
    return fail(BGO_ERR_ARG, "bgo_scan_batched_noalloc: bad arguments");
  if (out_capacity > BGO_NOALLOC_MAX_HITS)
    return fail(BGO_ERR_ARG,
                "bgo_scan_batched_noalloc: out_capacity %lld exceeds the scratch array's %d "
                "records",
                (long long)out_capacity, BGO_NOALLOC_MAX_HITS);
  if (n_guides > BGO_NOALLOC_MAX_GUIDES)
    return fail(BGO_ERR_ARG, "bgo_scan_batched_noalloc: %d guides exceed the limit of %d",
                n_guides, BGO_NOALLOC_MAX_GUIDES);
  BGO_TRY(validate_guides(guides, n_guides));
  {
    std::set<uint64_t> fams;
    for (int32_t i = 0; i < n_guides; ++i) fams.insert(fam_key(guides[i]));
    if (fams.size() > static_cast<size_t>(BGO_NOALLOC_MAX_FAMILIES))
      return fail(BGO_ERR_ARG, "bgo_scan_batched_noalloc: %zu PAM families exceed the limit of %d",
                  fams.size(), BGO_NOALLOC_MAX_FAMILIES);
  }
  BGO_TRY(injected());
  if (G->device < 0 || G->device >= na::kMaxDevices || !na::device_state(G->device).ready)
    return fail(BGO_ERR_ARG, "bgo_scan_batched_noalloc: call bgo_init(%d) first", G->device);
  na::Device& d = na::device_state(G->device);
  std::lock_guard<std::mutex> lock(d.mu);
  bgo_plan* P = d.plan;
  std::lock_guard<std::mutex> plock(P->mu);
  BGO_HIP(hipSetDevice(G->device));
  P->g = G;
  P->end_guard = 26;
  P->hit_cap = out_capacity;
  std::memcpy(P->score, score_table, sizeof(P->score));
  BGO_TRY(resolve_tuning(P, tuning));
  // tables into the scratch array (error when they do not fit); the kernel
  // that follows clears every counter, and the scan clears them again
  BGO_TRY(plan_set_guides(P, guides, n_guides));
  BGO_HIP(hipEventRecord(P->ev0, P->stream));
  BGO_TRY(enqueue_scan(P, 0, G->total_bp, 3));
  BGO_HIP(hipEventRecord(P->ev1, P->stream));
  unsigned long long n = 0;
  BGO_HIP(hipMemcpyAsync(&n, P->d_counters, sizeof(n), hipMemcpyDeviceToHost, P->stream));
  BGO_HIP(hipStreamSynchronize(P->stream));
  const int64_t m = std::min<int64_t>(static_cast<int64_t>(n), out_capacity);
  if (m > 0) {
    BGO_HIP(hipMemcpyAsync(out, P->d_hits, m * sizeof(bgo_hit), hipMemcpyDeviceToHost, P->stream));
    BGO_HIP(hipStreamSynchronize(P->stream));
  }
  *n_hits_out = static_cast<int64_t>(n);
  return BGO_OK;
}

// Anchor index of a resident genome for the fast mode, and its query.
//
// The fast mode (beguidedoff/fast/finder.py:163-277) looks up the exact
// "cold" anchor of each guide, then checks every occurrence's window.  The
// index here replaces the reference's FM-index: every forward position p has
// a key, the 12 bases p..p+11 as 2-bit codes with base p in bits 22-23 (N read
// as A, bases past the genome end as A).  Positions are grouped by key by a
// three-pass MSD radix placement on the GPU, 8 bits per pass, with per-block
// histograms in LDS (no global atomics) and each block round sorted by digit
// in LDS before it is written, so stores leave as runs.  An anchor of m
// cold bases is then a contiguous key range (its first min(m, 12) bases as a
// prefix), and the query checks each occurrence's window on the GPU: window
// inside one contig, the whole cold block, the PAM, the 2-bit mismatches plus
// the N positions that count as mismatches, and score <= cutoff.
//
// Memory: 4 bytes per position (12.4 GB at 3.09 Gbp) plus 64 MiB of bucket
// offsets kept; the build holds 7 more bytes per position (positions and
// key bits between passes) and < 100 MiB of histograms and block tables.
// Runs of one key in a wave (N runs, repeats) take one LDS atomic.
#include <algorithm>
#include <vector>

#include "bgo_common.h"

struct bgo_fast_index {
  bgo_genome* g = nullptr;
  int device = 0;
  uint64_t n_pos = 0;
  uint32_t* d_offsets = nullptr;  // [kBuckets + 1] start of each key's positions
  uint32_t* d_pos = nullptr;      // [n_pos] positions grouped by key
  int64_t* d_contigs = nullptr;   // [2 * n_contigs] offsets, then lengths
  int32_t n_contigs = 0;
  int64_t bytes = 0;
  // query buffers, kept between queries and grown on demand
  bgo_fast_anchor* d_anchors = nullptr;
  int64_t anchors_cap = 0;
  bgo_hit* d_hits = nullptr;
  int64_t hits_cap = 0;
  double* d_tab = nullptr;                // [21]
  unsigned long long* d_count = nullptr;  // [1]
  uint32_t* d_chunks = nullptr;           // [anchors] occurrence chunks per anchor
  int64_t chunks_cap = 0;
  uint64_t* d_first = nullptr;            // [anchors + 1] first chunk of each anchor
  int64_t first_cap = 0;
  std::vector<uint32_t> h_chunks;
  std::vector<uint64_t> h_first;
  hipEvent_t ev[2] = {nullptr, nullptr};  // query kernel timing
};

namespace bgo {
namespace fi {

constexpr uint32_t kBuckets = 1u << 24;
constexpr uint32_t kScanChunk = 4096;  // keys per block in the scan (256 threads x 16)
constexpr uint32_t kWin = 0xFFFFFu;
constexpr uint32_t kQueryChunk = 1024;  // occurrences per workgroup of the query

// 12 bits -> even bit positions (bit i -> bit 2i).
__device__ inline uint32_t spread12(uint32_t x) {
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  return (x | (x << 1)) & 0x55555555u;
}

// Key of forward position p: base p + i in bits 2(11 - i) .. +1.
__device__ inline uint32_t key_at(const uint2* __restrict__ planes, uint64_t p) {
  const uint64_t b = p >> 5;
  const uint32_t sh = static_cast<uint32_t>(p & 31);
  const uint2 a = planes[b], c = planes[b + 1];
  const uint32_t x0 = funnel(c.x, a.x, sh) & 0xFFFu, x1 = funnel(c.y, a.y, sh) & 0xFFFu;
  const uint32_t r0 = __builtin_bitreverse32(x0) >> 20, r1 = __builtin_bitreverse32(x1) >> 20;
  return spread12(r0) | (spread12(r1) << 1);
}

__device__ inline uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Exclusive scan of n u32 values in place (n a multiple of kScanChunk, the
// array n + 1 long: data[n] receives the total) in three steps: per-chunk scan
// (chunk totals out), scan of the chunk totals (one block), add the bases.
__global__ void scan_chunks_kernel(uint32_t* __restrict__ data, uint32_t* __restrict__ totals) {
  __shared__ uint32_t part[256];
  const uint32_t t = threadIdx.x;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanChunk + t * 16;
  uint32_t v[16], s = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    v[i] = data[base + i];
    s += v[i];
  }
  part[t] = s;
  __syncthreads();
  for (uint32_t d = 1; d < 256; d <<= 1) {  // inclusive scan of the thread sums
    const uint32_t x = t >= d ? part[t - d] : 0u;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  uint32_t run = part[t] - s;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    data[base + i] = run;
    run += v[i];
  }
  if (t == 255) totals[blockIdx.x] = part[255];
}

// One block of 256 threads; totals[n] receives the grand total.
__global__ void scan_totals_kernel(uint32_t* __restrict__ totals, uint32_t n) {
  __shared__ uint32_t part[256];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (n + 255) / 256;
  const uint32_t b = min(n, t * per), e = min(n, b + per);
  uint32_t s = 0;
  for (uint32_t j = b; j < e; ++j) s += totals[j];
  part[t] = s;
  __syncthreads();
  for (uint32_t d = 1; d < 256; d <<= 1) {
    const uint32_t x = t >= d ? part[t - d] : 0u;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  uint32_t run = part[t] - s;
  for (uint32_t j = b; j < e; ++j) {
    const uint32_t v = totals[j];
    totals[j] = run;
    run += v;
  }
  if (t == 255) totals[n] = part[255];
}

__global__ void add_bases_kernel(uint32_t* __restrict__ data, const uint32_t* __restrict__ totals,
                                 uint64_t n, uint32_t n_chunks) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) data[i] += totals[i / kScanChunk];
  if (i == n) data[n] = totals[n_chunks];
}

inline hipError_t exclusive_scan(uint32_t* data, uint64_t n, uint32_t* totals) {
  const uint64_t chunks = n / kScanChunk;
  hipLaunchKernelGGL(scan_chunks_kernel, dim3(static_cast<uint32_t>(chunks)), dim3(256), 0, 0,
                     data, totals);
  hipLaunchKernelGGL(scan_totals_kernel, dim3(1), dim3(256), 0, 0, totals,
                     static_cast<uint32_t>(chunks));
  hipLaunchKernelGGL(add_bases_kernel, dim3(static_cast<uint32_t>((n + 256) / 256)), dim3(256), 0,
                     0, data, totals, n, static_cast<uint32_t>(chunks));
  return hipGetLastError();
}

// Radix placement: three MSD passes of 8 bits of the key (bits 16-23, 8-15,
// 0-7).  A pass works on segments (entries of equal higher digits; pass 1:
// all positions), each cut into blocks of at most kTile entries.  Pass 1
// writes each position and the key's low 16 bits, pass 2 the position and the
// low 8 bits, pass 3 the position alone.
constexpr uint32_t kDigitBins = 256;
constexpr uint32_t kTile = 1u << 20;
constexpr uint32_t kBuildThreads = 256;
constexpr uint32_t kRound = 4096;  // entries staged in LDS per block round
constexpr uint32_t kPerThread = kRound / kBuildThreads;

struct PassBlock {
  uint32_t begin, end, seg, pad;
};

// LDS slot of this lane's element with digit d among the block's elements of
// that digit so far (hist[d] counts them): lanes of a wave that share the
// first lane's digit take their slots with one atomic, the others one each.
__device__ inline uint32_t lds_slot(uint32_t* hist, uint32_t d, bool ok) {
  const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
  const bool same = ok && d == d0;
  const uint64_t m = ballot64(same);
  uint32_t base = 0;
  if (same && lanes_below(m) == 0) base = atomicAdd(&hist[d0], static_cast<uint32_t>(__popcll(m)));
  const int leader = m ? static_cast<int>(__builtin_ctzll(m)) : 0;
  base = static_cast<uint32_t>(__shfl(static_cast<int>(base), leader, 64));
  if (same) return base + lanes_below(m);
  if (ok) return atomicAdd(&hist[d], 1u);
  return 0u;
}

// Element i of pass P: its position, digit and the key bits below the digit.
template <int P>
__device__ inline void pass_elem(uint64_t i, const uint2* __restrict__ planes,
                                 const uint32_t* __restrict__ pos_in,
                                 const uint16_t* __restrict__ r16_in,
                                 const uint8_t* __restrict__ r8_in, uint32_t& pos, uint32_t& digit,
                                 uint32_t& rest) {
  if constexpr (P == 1) {
    const uint32_t key = key_at(planes, i);
    pos = static_cast<uint32_t>(i);
    digit = key >> 16;
    rest = key & 0xFFFFu;
  } else if constexpr (P == 2) {
    const uint32_t v = r16_in[i];
    pos = pos_in[i];
    digit = v >> 8;
    rest = v & 0xFFu;
  } else {
    pos = pos_in[i];
    digit = r8_in[i];
    rest = 0;
  }
}

template <int P>
__device__ inline uint32_t pass_digit(uint64_t i, const uint2* __restrict__ planes,
                                      const uint16_t* __restrict__ r16_in,
                                      const uint8_t* __restrict__ r8_in) {
  if constexpr (P == 1) return key_at(planes, i) >> 16;
  else if constexpr (P == 2) return static_cast<uint32_t>(r16_in[i]) >> 8;
  else return r8_in[i];
}

struct PassArgs {
  const uint2* planes;
  const uint32_t* pos_in;
  const uint16_t* r16_in;
  const uint8_t* r8_in;
  uint32_t* pos_out;
  uint16_t* r16_out;
  uint8_t* r8_out;
  const PassBlock* blocks;
  const uint32_t* first;      // [segments + 1] first block of each segment
  const uint32_t* seg_start;  // [segments + 1] first entry of each segment
  const uint32_t* S;          // exclusive scan of the counts H (digit-major, stride nbp)
  uint32_t nbp;               // block columns of H (a multiple of 16)
};

// Counts: H[d * nbp + k] = entries of block k with digit d.
template <int P>
__global__ __launch_bounds__(kBuildThreads) void pass_hist_kernel(PassArgs a, uint32_t* __restrict__ H) {
  __shared__ uint32_t hist[kDigitBins];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const PassBlock w = a.blocks[blockIdx.x];
  // every wave runs the same rounds (lds_slot is wave-wide)
  for (uint64_t q = w.begin; q < w.end; q += kBuildThreads) {
    const uint64_t i = q + threadIdx.x;
    const bool ok = i < w.end;
    (void)lds_slot(hist, ok ? pass_digit<P>(i, a.planes, a.r16_in, a.r8_in) : 0u, ok);
  }
  __syncthreads();
  H[static_cast<uint64_t>(threadIdx.x) * a.nbp + blockIdx.x] = hist[threadIdx.x];
}

// Exclusive scan of v over the block's 256 threads (x: LDS scratch of 256).
__device__ inline uint32_t block_scan256(uint32_t* x, uint32_t v) {
  const uint32_t t = threadIdx.x;
  x[t] = v;
  __syncthreads();
  for (uint32_t d = 1; d < 256; d <<= 1) {
    const uint32_t y = t >= d ? x[t - d] : 0u;
    __syncthreads();
    x[t] += y;
    __syncthreads();
  }
  const uint32_t r = x[t] - v;
  __syncthreads();
  return r;
}

// Placement: block k of segment s writes its entries of digit d from
// seg_start[s] + (entries of s with a smaller digit) + (entries of digit d in
// the earlier blocks of s).  Each round of kRound entries is sorted by digit
// in LDS first, so the stores go out as runs per digit.
template <int P>
__global__ __launch_bounds__(kBuildThreads) void pass_place_kernel(PassArgs a) {
  __shared__ uint32_t gbase[kDigitBins], cnt[kDigitBins], lstart[kDigitBins], tmp[kDigitBins];
  __shared__ uint32_t st_pos[kRound];
  __shared__ uint16_t st_rest[kRound];
  __shared__ uint8_t st_dig[kRound];
  const uint32_t t = threadIdx.x;
  const PassBlock w = a.blocks[blockIdx.x];
  {
    const uint64_t row = static_cast<uint64_t>(t) * a.nbp;
    const uint32_t f = a.first[w.seg], f1 = a.first[w.seg + 1];
    const uint32_t in_seg = a.S[row + f1] - a.S[row + f];
    const uint32_t before = a.S[row + blockIdx.x] - a.S[row + f];
    gbase[t] = a.seg_start[w.seg] + block_scan256(tmp, in_seg) + before;
  }
  for (uint64_t q = w.begin; q < w.end; q += kRound) {
    cnt[t] = 0;
    __syncthreads();
    uint32_t pos[kPerThread], dig[kPerThread], rest[kPerThread], rank[kPerThread];
#pragma unroll
    for (uint32_t j = 0; j < kPerThread; ++j) {
      const uint64_t i = q + j * kBuildThreads + t;
      const bool ok = i < w.end;
      pos[j] = dig[j] = rest[j] = 0;
      if (ok) pass_elem<P>(i, a.planes, a.pos_in, a.r16_in, a.r8_in, pos[j], dig[j], rest[j]);
      rank[j] = lds_slot(cnt, dig[j], ok);
    }
    __syncthreads();
    lstart[t] = block_scan256(tmp, cnt[t]);
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kPerThread; ++j) {
      const uint64_t i = q + j * kBuildThreads + t;
      if (i < w.end) {
        const uint32_t slot = lstart[dig[j]] + rank[j];
        st_pos[slot] = pos[j];
        st_rest[slot] = static_cast<uint16_t>(rest[j]);
        st_dig[slot] = static_cast<uint8_t>(dig[j]);
      }
    }
    __syncthreads();
    const uint32_t m = static_cast<uint32_t>(w.end - q < kRound ? w.end - q : kRound);
    for (uint32_t e = t; e < m; e += kBuildThreads) {
      const uint32_t d = st_dig[e];
      const uint32_t out = gbase[d] + (e - lstart[d]);
      a.pos_out[out] = st_pos[e];
      if constexpr (P == 1) a.r16_out[out] = st_rest[e];
      if constexpr (P == 2) a.r8_out[out] = static_cast<uint8_t>(st_rest[e]);
    }
    __syncthreads();
    gbase[t] += cnt[t];
    __syncthreads();
  }
}

// First entry of each (segment, digit) after a pass: out[s * 256 + d], and
// out[segments * 256] = n.  One block per segment.
__global__ __launch_bounds__(kBuildThreads) void seg_starts_kernel(PassArgs a, uint64_t n,
                                                                   uint32_t n_seg,
                                                                   uint32_t* __restrict__ out) {
  __shared__ uint32_t tmp[kDigitBins];
  const uint32_t s = blockIdx.x, t = threadIdx.x;
  const uint64_t row = static_cast<uint64_t>(t) * a.nbp;
  const uint32_t in_seg = a.S[row + a.first[s + 1]] - a.S[row + a.first[s]];
  out[static_cast<uint64_t>(s) * kDigitBins + t] = a.seg_start[s] + block_scan256(tmp, in_seg);
  if (s == 0 && t == 0) out[static_cast<uint64_t>(n_seg) * kDigitBins] = static_cast<uint32_t>(n);
}

// Occurrence chunks of each anchor's key range.
__global__ void chunk_count_kernel(const bgo_fast_anchor* __restrict__ anchors, int32_t n,
                                   const uint32_t* __restrict__ offsets,
                                   uint32_t* __restrict__ chunks) {
  const int32_t i = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) {
    const uint32_t occ = offsets[anchors[i].key_hi] - offsets[anchors[i].key_lo];
    chunks[i] = (occ + kQueryChunk - 1) / kQueryChunk;
  }
}

struct QueryArgs {
  const uint2* planes0;
  const uint2* planes1;
  const uint32_t* nwords0;  // N words per strand; null when the genome has no N
  const uint32_t* nwords1;
  const uint32_t* offsets;
  const uint32_t* pos;
  const int64_t* contigs;  // offsets, then lengths
  int32_t n_contigs;
  const bgo_fast_anchor* anchors;
  int32_t n_anchors;
  const uint64_t* first;  // [n_anchors + 1] first chunk of each anchor, then the total
  uint64_t n_chunks;
  double cutoff;
  const double* score_tab;  // [21]
  bgo_hit* hits;
  unsigned long long* n_hits;
  uint64_t cap;
};

__device__ inline int32_t contig_of(const int64_t* __restrict__ offs, int32_t n, int64_t p) {
  int32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (offs[mid] <= p) lo = mid; else hi = mid;
  }
  return lo;
}

// The anchor owning chunk c: the last a with first[a] <= c.
__device__ inline int32_t anchor_of(const uint64_t* __restrict__ first, int32_t n, uint64_t c) {
  int32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (first[mid] <= c) lo = mid; else hi = mid;
  }
  return lo;
}

// One workgroup per chunk of kQueryChunk occurrences of one anchor (grid-
// stride over the chunks); its threads take one occurrence each.
__global__ void query_kernel(QueryArgs q) {
  for (uint64_t c = blockIdx.x; c < q.n_chunks; c += gridDim.x) {
    const int32_t ai = anchor_of(q.first, q.n_anchors, c);
    const bgo_fast_anchor A = q.anchors[ai];
    const uint32_t key_begin = q.offsets[A.key_lo], key_end = q.offsets[A.key_hi];
    const uint64_t from = key_begin + (c - q.first[ai]) * kQueryChunk;
    const uint32_t begin = static_cast<uint32_t>(from);
    const uint32_t end = static_cast<uint32_t>(min<uint64_t>(from + kQueryChunk, key_end));
    const uint32_t plen = A.pam_len, window = 20 + plen;
    const uint32_t cold = (kWin >> A.cold_first) << A.cold_first;
    const uint2* __restrict__ planes = A.strand ? q.planes1 : q.planes0;
    const uint32_t* __restrict__ nwords = A.strand ? q.nwords1 : q.nwords0;
    const int64_t* offs = q.contigs;
    const int64_t* lens = q.contigs + q.n_contigs;
    for (uint32_t i = begin + threadIdx.x; i < end; i += blockDim.x) {
      const int64_t gpos = q.pos[i];
      const int64_t left = A.strand ? gpos - plen : gpos - A.cold_first;
      if (left < 0) continue;
      const int32_t k = contig_of(offs, q.n_contigs, left);
      const int64_t c_off = offs[k], c_end = offs[k] + lens[k];
      if (left < c_off || left + window > c_end) continue;
      // the window on the scanned strand (the reverse complement keeps contig order)
      const int64_t s = A.strand ? c_off + (c_end - left) - window : left;
      const uint64_t b = static_cast<uint64_t>(s) >> 5;
      const uint32_t sh = static_cast<uint32_t>(s & 31);
      const uint2 w0 = planes[b], w1 = planes[b + 1];
      const uint32_t x0 = funnel(w1.x, w0.x, sh) & 0x3FFFFFFu;
      const uint32_t x1 = funnel(w1.y, w0.y, sh) & 0x3FFFFFFu;
      if (((x0 ^ A.a0) | (x1 ^ A.a1)) & cold) continue;
      bool pam_ok = true;
      for (uint32_t j = 0; j < plen; ++j) {
        const uint32_t code = ((x0 >> (20 + j)) & 1u) | (((x1 >> (20 + j)) & 1u) << 1);
        pam_ok &= ((A.pam_nibs >> (4 * j + code)) & 1u) != 0;
      }
      if (!pam_ok) continue;
      // a genome without N carries no N words (null)
      const uint32_t nb = nwords ? funnel(nwords[b + 1], nwords[b], sh) & kWin : 0u;
      const uint32_t mm = ((((x0 ^ A.g0) | (x1 ^ A.g1)) & kWin) | (nb & A.n_code));
      double score = 0.0;
      for (uint32_t m = mm; m; m &= m - 1) score += q.score_tab[__builtin_ctz(m) + 1];
      if (!(score <= q.cutoff)) continue;
      const unsigned long long idx = atomicAdd(q.n_hits, 1ull);
      if (idx < q.cap) {
        bgo_hit h;
        h.guide = A.guide;
        h.pos = static_cast<uint32_t>(s);
        h.x0 = x0;
        h.x1 = x1;
        h.mm_mask = mm;
        h.strand = A.strand;
        h.n_mm = static_cast<uint8_t>(__popc(mm));
        h.editable = 0;
        h.pad = 0;
        h.score = score;
        q.hits[idx] = h;
      }
    }
  }
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Grow a device buffer to at least `need` elements (contents not kept).
template <typename T>
hipError_t reserve(T** buf, int64_t* cap, int64_t need) {
  if (need <= *cap) return hipSuccess;
  if (*buf) (void)hipFree(*buf);
  *buf = nullptr;
  *cap = 0;
  const hipError_t e = hipMalloc(reinterpret_cast<void**>(buf), static_cast<size_t>(need) * sizeof(T));
  if (e == hipSuccess) *cap = need;
  return e;
}

void free_query_buffers(bgo_fast_index* X) {
  if (X->d_chunks) (void)hipFree(X->d_chunks);
  if (X->d_first) (void)hipFree(X->d_first);
  X->d_chunks = nullptr;
  X->d_first = nullptr;
  X->chunks_cap = X->first_cap = 0;
  for (hipEvent_t& ev : X->ev) {
    if (ev) (void)hipEventDestroy(ev);
    ev = nullptr;
  }
  if (X->d_anchors) (void)hipFree(X->d_anchors);
  if (X->d_hits) (void)hipFree(X->d_hits);
  if (X->d_tab) (void)hipFree(X->d_tab);
  if (X->d_count) (void)hipFree(X->d_count);
  X->d_anchors = nullptr;
  X->d_hits = nullptr;
  X->d_tab = nullptr;
  X->d_count = nullptr;
  X->anchors_cap = X->hits_cap = 0;
}

}  // namespace fi
}  // namespace bgo

using namespace bgo;

namespace bgo {
namespace fi {

// Positions 0 .. n - 1 grouped by key into pos_out, bucket starts into
// offsets[0 .. kBuckets] (the three radix passes above); the DevBufs hold the
// temporaries.
hipError_t build_passes(const uint2* planes, uint64_t n, uint32_t* offsets, uint32_t* pos_out,
                        DevBuf& h, DevBuf& totals, DevBuf& pos_tmp, DevBuf& r16, DevBuf& r8,
                        DevBuf& blocks_d, DevBuf& first_d, DevBuf& starts_a, DevBuf& starts_b) {
  // blocks of a pass: at most n / kTile + one per segment (65536 in pass 3)
  const uint64_t nb_max = n / kTile + 1 + (1u << 16);
  const uint64_t nbp_max = (nb_max + 15) / 16 * 16;
  hipError_t e = hipMalloc(&h.p, (kDigitBins * nbp_max + 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&totals.p, (nbp_max / 16 + 2) * 4);
  if (e == hipSuccess) e = hipMalloc(&pos_tmp.p, n * 4);
  if (e == hipSuccess) e = hipMalloc(&r16.p, n * 2);
  if (e == hipSuccess) e = hipMalloc(&r8.p, n);
  if (e == hipSuccess) e = hipMalloc(&blocks_d.p, nb_max * sizeof(PassBlock));
  if (e == hipSuccess) e = hipMalloc(&first_d.p, ((1u << 16) + 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&starts_a.p, (kDigitBins + 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&starts_b.p, ((1u << 16) + 1) * 4);
  if (e != hipSuccess) return e;
  auto* H = static_cast<uint32_t*>(h.p);
  auto* tot = static_cast<uint32_t*>(totals.p);
  // pass outputs: positions to pos_out, then pos_tmp, then pos_out
  uint32_t* pos_buf[3] = {pos_out, static_cast<uint32_t*>(pos_tmp.p), pos_out};
  // segment starts of the current pass, host and device
  std::vector<uint32_t> st = {0u, static_cast<uint32_t>(n)};
  const uint32_t zero_n[2] = {0u, static_cast<uint32_t>(n)};
  uint32_t* seg_d = static_cast<uint32_t*>(starts_b.p);
  if ((e = hipMemcpy(seg_d, zero_n, sizeof(zero_n), hipMemcpyHostToDevice)) != hipSuccess) return e;
  std::vector<PassBlock> blocks;
  std::vector<uint32_t> first;
  uint32_t n_seg = 1;
  for (int pass = 1; pass <= 3; ++pass) {
    blocks.clear();
    first.assign(n_seg + 1, 0);
    for (uint32_t sg = 0; sg < n_seg; ++sg) {
      first[sg] = static_cast<uint32_t>(blocks.size());
      for (uint64_t b = st[sg]; b < st[sg + 1]; b += kTile)
        blocks.push_back({static_cast<uint32_t>(b),
                          static_cast<uint32_t>(std::min<uint64_t>(b + kTile, st[sg + 1])), sg, 0});
    }
    first[n_seg] = static_cast<uint32_t>(blocks.size());
    const uint32_t nb = static_cast<uint32_t>(blocks.size());
    if (nb == 0 || nb > nb_max) return hipErrorInvalidValue;
    const uint32_t nbp = (nb + 15) / 16 * 16;
    if ((e = hipMemcpy(blocks_d.p, blocks.data(), nb * sizeof(PassBlock), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(first_d.p, first.data(), first.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(H, 0, static_cast<size_t>(kDigitBins) * nbp * 4)) != hipSuccess)
      return e;
    PassArgs a{};
    a.planes = planes;
    a.pos_in = pass == 1 ? nullptr : pos_buf[pass - 2];
    a.r16_in = static_cast<const uint16_t*>(r16.p);
    a.r8_in = static_cast<const uint8_t*>(r8.p);
    a.pos_out = pos_buf[pass - 1];
    a.r16_out = static_cast<uint16_t*>(r16.p);
    a.r8_out = static_cast<uint8_t*>(r8.p);
    a.blocks = static_cast<const PassBlock*>(blocks_d.p);
    a.first = static_cast<const uint32_t*>(first_d.p);
    a.seg_start = seg_d;
    a.S = H;
    a.nbp = nbp;
    auto hist = pass == 1 ? pass_hist_kernel<1> : pass == 2 ? pass_hist_kernel<2> : pass_hist_kernel<3>;
    auto place = pass == 1 ? pass_place_kernel<1> : pass == 2 ? pass_place_kernel<2> : pass_place_kernel<3>;
    hipLaunchKernelGGL(hist, dim3(nb), dim3(kBuildThreads), 0, 0, a, H);
    if ((e = exclusive_scan(H, static_cast<uint64_t>(kDigitBins) * nbp, tot)) != hipSuccess) return e;
    hipLaunchKernelGGL(place, dim3(nb), dim3(kBuildThreads), 0, 0, a);
    // starts of the next pass's segments (after pass 3: the bucket offsets)
    uint32_t* next = pass == 3 ? offsets : pass == 1 ? static_cast<uint32_t*>(starts_a.p)
                                                      : static_cast<uint32_t*>(starts_b.p);
    hipLaunchKernelGGL(seg_starts_kernel, dim3(n_seg), dim3(kBuildThreads), 0, 0, a, n, n_seg, next);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    n_seg *= kDigitBins;
    if (pass < 3) {
      st.resize(n_seg + 1);
      if ((e = hipMemcpy(st.data(), next, st.size() * 4, hipMemcpyDeviceToHost)) != hipSuccess) return e;
      seg_d = next;
    }
  }
  return hipSuccess;
}

hipError_t build_index(const uint2* planes, uint64_t n, uint32_t* offsets, uint32_t* pos_out) {
  DevBuf h, totals, pos_tmp, r16, r8, blocks_d, first_d, starts_a, starts_b;
  const hipError_t e = build_passes(planes, n, offsets, pos_out, h, totals, pos_tmp, r16, r8,
                                    blocks_d, first_d, starts_a, starts_b);
  // no kernel may still use the temporaries when they are freed
  const hipError_t s = hipDeviceSynchronize();
  return e != hipSuccess ? e : s;
}

}  // namespace fi
}  // namespace bgo

extern "C" int bgo_fast_index_create(bgo_genome* g, const int64_t* contig_offsets,
                                     const int64_t* contig_lengths, int32_t n_contigs,
                                     bgo_fast_index** out) {
  BGO_TRY(injected());
  if (!g || !out || n_contigs <= 0 || !contig_offsets || !contig_lengths)
    return fail(BGO_ERR_ARG, "bad arguments");
  if (g->total_bp <= 0 || g->total_bp >= (int64_t{1} << 32))
    return fail(BGO_ERR_ARG, "genome length %lld outside (0, 2^32)",
                static_cast<long long>(g->total_bp));
  int64_t next = 0;
  for (int32_t k = 0; k < n_contigs; ++k) {
    if (contig_offsets[k] != next || contig_lengths[k] < 0)
      return fail(BGO_ERR_ARG, "contig %d does not tile the genome", k);
    next += contig_lengths[k];
  }
  if (next != g->total_bp) return fail(BGO_ERR_ARG, "contig lengths do not sum to the genome length");
  BGO_HIP(hipSetDevice(g->device));
  auto* X = new bgo_fast_index();
  X->g = g;
  X->device = g->device;
  X->n_pos = static_cast<uint64_t>(g->total_bp);
  X->n_contigs = n_contigs;
  auto cleanup = [&](int rc) {
    fi::free_query_buffers(X);
    if (X->d_offsets) (void)hipFree(X->d_offsets);
    if (X->d_pos) (void)hipFree(X->d_pos);
    if (X->d_contigs) (void)hipFree(X->d_contigs);
    delete X;
    return rc;
  };
  hipError_t e = hipMalloc(&X->d_offsets, (fi::kBuckets + 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&X->d_pos, X->n_pos * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&X->d_contigs, 2 * static_cast<size_t>(n_contigs) * 8);
  if (e == hipSuccess) e = hipMalloc(&X->d_tab, 21 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&X->d_count, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipEventCreate(&X->ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&X->ev[1]);
  if (e == hipSuccess)
    e = hipMemcpy(X->d_contigs, contig_offsets, n_contigs * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(X->d_contigs + n_contigs, contig_lengths, n_contigs * 8, hipMemcpyHostToDevice);
  if (e != hipSuccess)
    return cleanup(fail(BGO_ERR_HIP, "fast index allocation failed: %s", hipGetErrorString(e)));
  e = fi::build_index(g->d_planes[0], X->n_pos, X->d_offsets, X->d_pos);
  if (e != hipSuccess)
    return cleanup(fail(BGO_ERR_HIP, "fast index build failed: %s", hipGetErrorString(e)));
  X->bytes = static_cast<int64_t>((fi::kBuckets + 1) * 4 + X->n_pos * 4 + 16 * n_contigs);
  *out = X;
  return BGO_OK;
}

extern "C" int bgo_fast_index_destroy(bgo_fast_index* X) {
  if (!X) return BGO_OK;
  (void)hipSetDevice(X->device);
  fi::free_query_buffers(X);
  if (X->d_offsets) (void)hipFree(X->d_offsets);
  if (X->d_pos) (void)hipFree(X->d_pos);
  if (X->d_contigs) (void)hipFree(X->d_contigs);
  delete X;
  return BGO_OK;
}

extern "C" int bgo_fast_index_bytes(const bgo_fast_index* X, int64_t* bytes_out) {
  if (!X || !bytes_out) return fail(BGO_ERR_ARG, "null pointer");
  *bytes_out = X->bytes;
  return BGO_OK;
}

extern "C" int bgo_fast_index_offsets(bgo_fast_index* X, uint32_t* out) {
  BGO_TRY(injected());
  if (!X || !out) return fail(BGO_ERR_ARG, "null pointer");
  BGO_HIP(hipSetDevice(X->device));
  BGO_HIP(hipMemcpy(out, X->d_offsets, (fi::kBuckets + 1) * sizeof(uint32_t),
                    hipMemcpyDeviceToHost));
  return BGO_OK;
}

extern "C" int bgo_fast_index_positions(bgo_fast_index* X, int64_t begin, int64_t end,
                                        uint32_t* out) {
  BGO_TRY(injected());
  if (!X || begin < 0 || end < begin || end > static_cast<int64_t>(X->n_pos) ||
      (end > begin && !out))
    return fail(BGO_ERR_ARG, "bgo_fast_index_positions: bad range [%lld, %lld)", (long long)begin,
                (long long)end);
  if (end == begin) return BGO_OK;
  BGO_HIP(hipSetDevice(X->device));
  BGO_HIP(hipMemcpy(out, X->d_pos + begin, (end - begin) * sizeof(uint32_t),
                    hipMemcpyDeviceToHost));
  return BGO_OK;
}

extern "C" int bgo_fast_index_query(bgo_fast_index* X, const bgo_fast_anchor* anchors,
                                    int32_t n_anchors, double cutoff, const double* score_table,
                                    bgo_hit* out, int64_t out_capacity, int64_t* n_hits_out,
                                    float* kernel_ms) {
  BGO_TRY(injected());
  if (!X || n_anchors < 0 || (n_anchors > 0 && !anchors) || !score_table || !n_hits_out ||
      out_capacity < 0 || (out_capacity > 0 && !out))
    return fail(BGO_ERR_ARG, "bad arguments");
  for (int32_t i = 0; i < n_anchors; ++i) {
    const bgo_fast_anchor& A = anchors[i];
    if (A.key_lo >= A.key_hi || A.key_hi > fi::kBuckets || A.pam_len < 1 || A.pam_len > 6 ||
        A.cold_first > 19 || A.strand > 1)
      return fail(BGO_ERR_ARG, "anchor %d out of range", i);
  }
  *n_hits_out = 0;
  if (kernel_ms) *kernel_ms = 0.f;
  if (n_anchors == 0) return BGO_OK;
  bgo_genome* g = X->g;
  if (g->n_strands < 2) return fail(BGO_ERR_ARG, "the genome has no reverse-complement strand");
  BGO_HIP(hipSetDevice(X->device));
  BGO_HIP(fi::reserve(&X->d_anchors, &X->anchors_cap, n_anchors));
  BGO_HIP(fi::reserve(&X->d_hits, &X->hits_cap, std::max<int64_t>(out_capacity, 1)));
  BGO_HIP(fi::reserve(&X->d_chunks, &X->chunks_cap, n_anchors));
  BGO_HIP(fi::reserve(&X->d_first, &X->first_cap, n_anchors + 1));
  BGO_HIP(hipMemcpy(X->d_anchors, anchors, n_anchors * sizeof(bgo_fast_anchor),
                    hipMemcpyHostToDevice));
  BGO_HIP(hipMemcpy(X->d_tab, score_table, 21 * sizeof(double), hipMemcpyHostToDevice));
  BGO_HIP(hipMemset(X->d_count, 0, sizeof(unsigned long long)));
  // chunks of each anchor's occurrences, and their prefix on the host
  hipLaunchKernelGGL(fi::chunk_count_kernel, dim3((n_anchors + 255) / 256), dim3(256), 0, 0,
                     X->d_anchors, n_anchors, X->d_offsets, X->d_chunks);
  BGO_HIP(hipGetLastError());
  X->h_chunks.resize(n_anchors);
  X->h_first.resize(n_anchors + 1);
  BGO_HIP(hipMemcpy(X->h_chunks.data(), X->d_chunks, n_anchors * sizeof(uint32_t),
                    hipMemcpyDeviceToHost));
  uint64_t total = 0;
  for (int32_t i = 0; i < n_anchors; ++i) {
    X->h_first[i] = total;
    total += X->h_chunks[i];
  }
  X->h_first[n_anchors] = total;
  BGO_HIP(hipMemcpy(X->d_first, X->h_first.data(), (n_anchors + 1) * sizeof(uint64_t),
                    hipMemcpyHostToDevice));
  fi::QueryArgs q{g->d_planes[0], g->d_planes[1], g->d_nwords[0], g->d_nwords[1],
                  X->d_offsets, X->d_pos, X->d_contigs, X->n_contigs, X->d_anchors, n_anchors,
                  X->d_first, total, cutoff, X->d_tab, X->d_hits, X->d_count,
                  static_cast<uint64_t>(out_capacity)};
  // at most 2^20 workgroups; more chunks are taken grid-stride
  const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(total, 1u << 20)));
  hipEvent_t e0 = X->ev[0], e1 = X->ev[1];
  hipError_t e = hipEventRecord(e0, 0);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(fi::query_kernel, dim3(grid), dim3(256), 0, 0, q);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  BGO_HIP(e);
  unsigned long long n = 0;
  BGO_HIP(hipMemcpy(&n, X->d_count, sizeof(n), hipMemcpyDeviceToHost));
  const uint64_t copy = std::min<uint64_t>(n, static_cast<uint64_t>(out_capacity));
  if (copy) BGO_HIP(hipMemcpy(out, X->d_hits, copy * sizeof(bgo_hit), hipMemcpyDeviceToHost));
  *n_hits_out = static_cast<int64_t>(n);
  if (kernel_ms) *kernel_ms = ms;
  return BGO_OK;
}

// Anchor index of a resident genome for the fast mode, and its query.
//
// The fast mode (beguidedoff/fast/finder.py:163-277) looks up the exact
// "cold" anchor of each guide, then checks every occurrence's window.  The
// index here replaces the reference's FM-index: every forward position p has
// a key, the 12 bases p..p+11 as 2-bit codes with base p in bits 22-23 (N read
// as A, bases past the genome end as A).  Positions are grouped by key by a
// counting sort on the GPU: count, exclusive scan, scatter.  An anchor of m
// cold bases is then a contiguous key range (its first min(m, 12) bases as a
// prefix), and the query checks each occurrence's window on the GPU: window
// inside one contig, the whole cold block, the PAM, the 2-bit mismatches plus
// the N positions that count as mismatches, and score <= cutoff.
//
// Memory: 4 bytes per position (12.4 GB at 3.09 Gbp) plus 64 MiB of bucket
// offsets kept, and 2 x 64 MiB of counts and cursors during the build.  Runs
// of one key (N runs, repeats) are counted and placed with one atomic per wave.
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

// Count keys; a wave whose active positions share its first lane's key adds
// them with one atomic.
__global__ void count_kernel(const uint2* __restrict__ planes, uint64_t n_pos,
                             uint32_t* __restrict__ counts) {
  const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool ok = p < n_pos;
  const uint32_t key = ok ? key_at(planes, p) : 0u;
  const uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
  const bool same = ok && key == k0;
  const uint64_t m = ballot64(same);
  if (same) {
    if (lanes_below(m) == 0) atomicAdd(&counts[k0], static_cast<uint32_t>(__popcll(m)));
  } else if (ok) {
    atomicAdd(&counts[key], 1u);
  }
}

// Exclusive scan of kBuckets counts in three steps: per-chunk scan (chunk
// totals out), scan of the chunk totals (one block), add the chunk bases.
__global__ void scan_chunks_kernel(const uint32_t* __restrict__ counts, uint32_t* __restrict__ out,
                                   uint32_t* __restrict__ totals) {
  __shared__ uint32_t part[256];
  const uint32_t t = threadIdx.x;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanChunk + t * 16;
  uint32_t v[16], s = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    v[i] = counts[base + i];
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
    out[base + i] = run;
    run += v[i];
  }
  if (t == 255) totals[blockIdx.x] = part[255];
}

__global__ void scan_totals_kernel(uint32_t* __restrict__ totals, uint32_t n) {
  // one block of 256 threads; n = kBuckets / kScanChunk = 4096 totals, 16 per thread
  __shared__ uint32_t part[256];
  const uint32_t t = threadIdx.x;
  uint32_t v[16], s = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const uint32_t j = t * 16 + i;
    v[i] = j < n ? totals[j] : 0u;
    s += v[i];
  }
  part[t] = s;
  __syncthreads();
  for (uint32_t d = 1; d < 256; d <<= 1) {
    const uint32_t x = t >= d ? part[t - d] : 0u;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  uint32_t run = part[t] - s;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const uint32_t j = t * 16 + i;
    if (j < n) totals[j] = run;
    run += v[i];
  }
}

__global__ void add_bases_kernel(uint32_t* __restrict__ out, const uint32_t* __restrict__ totals,
                                 uint32_t* __restrict__ cursor, uint64_t n_pos) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < kBuckets) {
    const uint32_t v = out[i] + totals[i / kScanChunk];
    out[i] = v;
    cursor[i] = v;
  }
  if (i == 0) out[kBuckets] = static_cast<uint32_t>(n_pos);
}

// Place positions; runs of one key within a wave take one atomic.
__global__ void scatter_kernel(const uint2* __restrict__ planes, uint64_t n_pos,
                               uint32_t* __restrict__ cursor, uint32_t* __restrict__ pos_out) {
  const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool ok = p < n_pos;
  const uint32_t key = ok ? key_at(planes, p) : 0u;
  const uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
  const bool same = ok && key == k0;
  const uint64_t m = ballot64(same);
  uint32_t base = 0;
  if (same && lanes_below(m) == 0) base = atomicAdd(&cursor[k0], static_cast<uint32_t>(__popcll(m)));
  // the leader's base to every lane of the run: the lowest set lane of m
  const int leader = m ? static_cast<int>(__builtin_ctzll(m)) : 0;
  base = static_cast<uint32_t>(__shfl(static_cast<int>(base), leader, 64));
  if (same) {
    pos_out[base + lanes_below(m)] = static_cast<uint32_t>(p);
  } else if (ok) {
    pos_out[atomicAdd(&cursor[key], 1u)] = static_cast<uint32_t>(p);
  }
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
  fi::DevBuf counts, cursor, totals;
  hipError_t e = hipMalloc(&X->d_offsets, (fi::kBuckets + 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&X->d_pos, X->n_pos * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&X->d_contigs, 2 * static_cast<size_t>(n_contigs) * 8);
  if (e == hipSuccess) e = hipMalloc(&X->d_tab, 21 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&X->d_count, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipEventCreate(&X->ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&X->ev[1]);
  if (e == hipSuccess) e = hipMalloc(&counts.p, fi::kBuckets * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&cursor.p, fi::kBuckets * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&totals.p, (fi::kBuckets / fi::kScanChunk) * sizeof(uint32_t));
  if (e != hipSuccess)
    return cleanup(fail(BGO_ERR_HIP, "fast index allocation failed: %s", hipGetErrorString(e)));
  e = hipMemcpy(X->d_contigs, contig_offsets, n_contigs * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(X->d_contigs + n_contigs, contig_lengths, n_contigs * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(counts.p, 0, fi::kBuckets * sizeof(uint32_t));
  const uint2* planes = g->d_planes[0];
  const uint64_t pos_blocks = (X->n_pos + 255) / 256;
  auto* c = static_cast<uint32_t*>(counts.p);
  auto* cur = static_cast<uint32_t*>(cursor.p);
  auto* tot = static_cast<uint32_t*>(totals.p);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(fi::count_kernel, dim3(static_cast<uint32_t>(pos_blocks)), dim3(256), 0, 0,
                       planes, X->n_pos, c);
    hipLaunchKernelGGL(fi::scan_chunks_kernel, dim3(fi::kBuckets / fi::kScanChunk), dim3(256), 0, 0,
                       c, X->d_offsets, tot);
    hipLaunchKernelGGL(fi::scan_totals_kernel, dim3(1), dim3(256), 0, 0, tot,
                       fi::kBuckets / fi::kScanChunk);
    hipLaunchKernelGGL(fi::add_bases_kernel, dim3(fi::kBuckets / 256), dim3(256), 0, 0,
                       X->d_offsets, tot, cur, X->n_pos);
    hipLaunchKernelGGL(fi::scatter_kernel, dim3(static_cast<uint32_t>(pos_blocks)), dim3(256), 0,
                       0, planes, X->n_pos, cur, X->d_pos);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
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

// Plain drop-in scan: the reference's work decomposition on gfx950.
//
// Behaviour follows beguidedoff/cuda_search.py:163-314 (kernel) and
// :384-488 (driver): one lane per genome position of one strand array, the
// guide table in LDS, per guide the packed PAM test (or the per-base IUPAC
// fallback for exotic codes), a 40-bit XOR / fold / popcount over the 20-nt
// window, the N correction, and a per-guide atomic append into a
// [guides x results_per_guide] buffer.  Positions >= total_bp - 26 are never
// scanned (cuda_search.py:211-213) and windows may span contig junctions
// (the kernel knows nothing of contigs).
//
// This is the baseline the batched engine is measured against, so it keeps
// that decomposition; it is written for wave64 (256-lane blocks = 4 waves).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "bgo_common.h"

struct bgo_strand {
  int device = 0;
  uint32_t* d_words = nullptr;
  int64_t n_words = 0;
  uint8_t* d_nmask = nullptr;  // little-endian bits, zero padded
  int has_nmask = 0;
  int64_t total_bp = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // per-call scratch, grown on demand
  void* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  std::mutex mu;
};

namespace bgo {
namespace {

constexpr int kPlainBlock = 256;
constexpr uint64_t kMask40 = 0x5555555555ull;

// 64 genome bits starting at base `start` (bit i = genome bit 2*start + i);
// words past the end read as 0 (cuda_search.py:61-87).  Bits above the
// requested width are neighbouring bases; callers mask them.
__device__ inline uint64_t pack_bits(const uint32_t* __restrict__ g, uint64_t n_words,
                                     uint64_t start) {
  const uint64_t b0 = start >> 4;
  const uint32_t off = static_cast<uint32_t>(start & 15) * 2;
  const uint64_t w0 = g[b0];
  const uint64_t w1 = (b0 + 1 < n_words) ? g[b0 + 1] : 0;
  uint64_t v = (w0 | (w1 << 32)) >> off;
  if (off > 24 && b0 + 2 < n_words) v |= static_cast<uint64_t>(g[b0 + 2]) << (64 - off);
  return v;
}

__device__ inline uint32_t base_at(const uint32_t* __restrict__ g, uint64_t pos) {
  return (g[pos >> 4] >> ((pos & 15) * 2)) & 3u;
}

// Five N-mask bytes starting at byte (pos >> 3), shifted to bit `pos`.
__device__ inline uint64_t nmask_bits(const uint8_t* __restrict__ m, uint64_t pos) {
  const uint64_t b = pos >> 3;
  const uint64_t w = static_cast<uint64_t>(m[b]) | (static_cast<uint64_t>(m[b + 1]) << 8) |
                     (static_cast<uint64_t>(m[b + 2]) << 16) |
                     (static_cast<uint64_t>(m[b + 3]) << 24) |
                     (static_cast<uint64_t>(m[b + 4]) << 32);
  return w >> (pos & 7);
}

__global__ __launch_bounds__(kPlainBlock) void plain_scan_kernel(
    const uint32_t* __restrict__ genome, uint64_t n_words, const uint8_t* __restrict__ nmask,
    int has_nmask, int64_t total_bp, const uint64_t* __restrict__ guides, int n_guides,
    const uint8_t* __restrict__ pam_masks, const uint32_t* __restrict__ pats,
    const uint32_t* __restrict__ cmps, int max_mm, int64_t chunk_start, int64_t chunk_end,
    int64_t rpg, uint32_t n_encoded, int32_t* __restrict__ counts,
    int64_t* __restrict__ hit_pos, int8_t* __restrict__ hit_mm) {
  __shared__ uint64_t s_guides[BGO_PLAIN_MAX_GUIDES];
  __shared__ uint32_t s_pats[BGO_PLAIN_MAX_GUIDES];
  __shared__ uint32_t s_cmps[BGO_PLAIN_MAX_GUIDES];
  for (int i = threadIdx.x; i < n_guides; i += blockDim.x) {
    s_guides[i] = guides[i];
    s_pats[i] = pats[i];
    s_cmps[i] = cmps[i];
  }
  __syncthreads();

  const int64_t pos = chunk_start + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (pos >= chunk_end || pos >= total_bp - 26) return;
  const uint64_t upos = static_cast<uint64_t>(pos);

  // 26-bp gate: is any N in [pos, pos + 26)?  (cuda_search.py:228-237)
  const bool need_n = has_nmask && ((nmask_bits(nmask, upos) & 0x3FFFFFFull) != 0);
  const uint64_t window3 = pack_bits(genome, n_words, upos);
  const uint64_t pam6 = pack_bits(genome, n_words, upos + 20);

  for (int g = 0; g < n_guides; ++g) {
    const uint32_t info = s_pats[g];
    const int plen = static_cast<int>((info >> 12) & 7u);
    if (plen == 0) continue;
    uint64_t guide_start, pam_start, window, pam;
    if ((info >> 16) & 1u) {  // 3' PAM
      guide_start = upos;
      pam_start = upos + 20;
      window = window3;
      pam = pam6;
    } else {  // 5' PAM
      guide_start = upos + plen;
      pam_start = upos;
      if (static_cast<int64_t>(guide_start) + 20 > total_bp ||
          static_cast<int64_t>(pam_start) + plen > total_bp)
        continue;
      window = pack_bits(genome, n_words, guide_start);
      pam = pack_bits(genome, n_words, pam_start);
    }
    if ((info >> 15) & 1u) {  // exotic IUPAC code: per-base allowed-set test
      bool ok = true;
      for (int i = 0; i < plen; ++i) {
        const uint32_t nuc = base_at(genome, pam_start + i);
        if (((1u << nuc) & pam_masks[g * BGO_MAX_PAM_LEN + i]) == 0) {
          ok = false;
          break;
        }
      }
      if (!ok) continue;
    } else if (((pam ^ info) & s_cmps[g]) != 0) {
      continue;
    }

    const uint64_t gval = s_guides[g];
    const uint64_t x = window ^ gval;
    const int n_mm = __popcll((x | (x >> 1)) & kMask40);
    if (n_mm > max_mm) continue;

    int n_final = n_mm;
    if (need_n) {
      // N positions whose guide base equals the N code read as matches;
      // count them as mismatches (cuda_search.py:290-308).
      uint64_t nb = nmask_bits(nmask, guide_start) & 0xFFFFFull;
      while (nb) {
        const int i = __builtin_ctzll(nb);
        nb &= nb - 1;
        if (((gval >> (2 * i)) & 3u) == n_encoded) ++n_final;
      }
      if (n_final > max_mm) continue;
    }

    const int idx = atomicAdd(&counts[g], 1);
    if (idx < rpg) {
      hit_pos[g * rpg + idx] = pos;
      hit_mm[g * rpg + idx] = static_cast<int8_t>(n_final);
    }
  }
}

// Compact the recorded prefix of every guide row into consecutive slots.
__global__ void plain_gather_kernel(const int64_t* __restrict__ hit_pos,
                                    const int8_t* __restrict__ hit_mm, int64_t rpg,
                                    const int64_t* __restrict__ row_off,
                                    const int32_t* __restrict__ row_len, int n_rows,
                                    int64_t* __restrict__ out_pos, int8_t* __restrict__ out_mm) {
  const int row = blockIdx.y;
  if (row >= n_rows) return;
  const int64_t len = row_len[row];
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < len;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    out_pos[row_off[row] + i] = hit_pos[row * rpg + i];
    out_mm[row_off[row] + i] = hit_mm[row * rpg + i];
  }
}

struct PlainLast {
  float ms = 0.f;
};
PlainLast& plain_last() {
  static thread_local PlainLast v;
  return v;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace
}  // namespace bgo

using namespace bgo;

extern "C" int bgo_strand_create(const uint32_t* words, int64_t n_words, const uint8_t* nmask_le,
                                 int64_t nmask_bytes, int64_t total_bp, bgo_strand** out) {
  if (!out || !words || n_words <= 0 || total_bp < 0 || total_bp > n_words * 16)
    return fail(BGO_ERR_ARG, "bgo_strand_create: bad arguments (n_words=%lld total_bp=%lld)",
                (long long)n_words, (long long)total_bp);
  if (nmask_bytes < 0 || (nmask_bytes > 0 && !nmask_le))
    return fail(BGO_ERR_ARG, "bgo_strand_create: bad N mask");
  BGO_TRY(injected());
  int dev = 0;
  BGO_HIP(hipGetDevice(&dev));
  auto* s = new bgo_strand();
  s->device = dev;
  s->n_words = n_words;
  s->total_bp = total_bp;
  auto cleanup = [&](int rc) {
    bgo_strand_destroy(s);
    return rc;
  };
  if (hipMalloc(&s->d_words, n_words * sizeof(uint32_t)) != hipSuccess)
    return cleanup(fail(BGO_ERR_HIP, "hipMalloc genome words (%lld)", (long long)n_words));
  if (hipMemcpy(s->d_words, words, n_words * sizeof(uint32_t), hipMemcpyHostToDevice) !=
      hipSuccess)
    return cleanup(fail(BGO_ERR_HIP, "upload genome words"));
  if (nmask_bytes > 0) {
    // Pad so the kernel's 5-byte reads at (p >> 3) stay inside
    // (cuda_search.py:365-381 pads to (total_bp >> 3) + 8).
    const int64_t need = std::max<int64_t>(nmask_bytes, (total_bp >> 3) + 16);
    if (hipMalloc(&s->d_nmask, need) != hipSuccess)
      return cleanup(fail(BGO_ERR_HIP, "hipMalloc N mask"));
    if (hipMemset(s->d_nmask, 0, need) != hipSuccess ||
        hipMemcpy(s->d_nmask, nmask_le, nmask_bytes, hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(BGO_ERR_HIP, "upload N mask"));
    s->has_nmask = 1;
  }
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
    return cleanup(fail(BGO_ERR_HIP, "stream/event create"));
  *out = s;
  return BGO_OK;
}

extern "C" int bgo_strand_destroy(bgo_strand* s) {
  if (!s) return BGO_OK;
  {
    // A scan running on another thread holds s->mu for its whole call
    // (uploads, kernels, copies back): the device memory is released only
    // after it returns.  Calling destroy while other threads may still start
    // scans on the handle is the caller's error; the Python driver releases a
    // strand only when no call holds a reference to it.
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->d_words) (void)hipFree(s->d_words);
    if (s->d_nmask) (void)hipFree(s->d_nmask);
    if (s->d_scratch) (void)hipFree(s->d_scratch);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    s->d_words = nullptr;
    s->d_nmask = nullptr;
    s->d_scratch = nullptr;
    s->ev0 = s->ev1 = nullptr;
    s->stream = nullptr;
  }
  delete s;
  return BGO_OK;
}

extern "C" int bgo_scan_plain_last_ms(float* ms_out) {
  if (!ms_out) return fail(BGO_ERR_ARG, "null pointer");
  *ms_out = plain_last().ms;
  return BGO_OK;
}

extern "C" int bgo_scan_plain(bgo_strand* s, const uint64_t* guides, const uint8_t* pam_masks,
                              const uint32_t* pats, const uint32_t* cmps, int32_t n_guides,
                              int32_t max_mismatches, int64_t chunk_size, int64_t rpg,
                              int32_t n_encoded, int32_t* counts, int64_t* positions,
                              int8_t* mismatches) {
  if (!s) return fail(BGO_ERR_ARG, "bgo_scan_plain: null strand");
  if (n_guides > BGO_PLAIN_MAX_GUIDES)
    return fail(BGO_ERR_ARG,
                "Too many guides (%d) for GPU search. Maximum supported is %d. Consider "
                "splitting the input into smaller batches.",
                n_guides, BGO_PLAIN_MAX_GUIDES);
  if (n_guides < 0 || rpg <= 0 || chunk_size <= 0 || n_encoded < 0 || n_encoded > 3)
    return fail(BGO_ERR_ARG, "bgo_scan_plain: bad arguments (n=%d rpg=%lld chunk=%lld enc=%d)",
                n_guides, (long long)rpg, (long long)chunk_size, n_encoded);
  if (n_guides > 0 && (!guides || !pam_masks || !pats || !cmps || !counts || !positions ||
                       !mismatches))
    return fail(BGO_ERR_ARG, "bgo_scan_plain: null buffer");
  BGO_TRY(injected());
  std::lock_guard<std::mutex> lock(s->mu);
  plain_last().ms = 0.f;
  if (n_guides == 0) return BGO_OK;
  BGO_HIP(hipSetDevice(s->device));

  const size_t n = static_cast<size_t>(n_guides);
  const size_t cells = n * static_cast<size_t>(rpg);
  // scratch layout
  size_t off = 0;
  const size_t o_guides = off; off = align_up(off + n * 8, 256);
  const size_t o_masks = off;  off = align_up(off + n * BGO_MAX_PAM_LEN, 256);
  const size_t o_pats = off;   off = align_up(off + n * 4, 256);
  const size_t o_cmps = off;   off = align_up(off + n * 4, 256);
  const size_t o_counts = off; off = align_up(off + n * 4, 256);
  const size_t o_pos = off;    off = align_up(off + cells * 8, 256);
  const size_t o_mm = off;     off = align_up(off + cells, 256);
  const size_t o_roff = off;   off = align_up(off + n * 8, 256);
  const size_t o_rlen = off;   off = align_up(off + n * 4, 256);
  const size_t fixed = off;
  if (s->scratch_bytes < fixed) {
    if (s->d_scratch) (void)hipFree(s->d_scratch);
    s->d_scratch = nullptr;
    s->scratch_bytes = 0;
    BGO_HIP(hipMalloc(&s->d_scratch, fixed));
    s->scratch_bytes = fixed;
  }
  char* base = static_cast<char*>(s->d_scratch);
  auto* d_guides = reinterpret_cast<uint64_t*>(base + o_guides);
  auto* d_masks = reinterpret_cast<uint8_t*>(base + o_masks);
  auto* d_pats = reinterpret_cast<uint32_t*>(base + o_pats);
  auto* d_cmps = reinterpret_cast<uint32_t*>(base + o_cmps);
  auto* d_counts = reinterpret_cast<int32_t*>(base + o_counts);
  auto* d_pos = reinterpret_cast<int64_t*>(base + o_pos);
  auto* d_mm = reinterpret_cast<int8_t*>(base + o_mm);
  auto* d_roff = reinterpret_cast<int64_t*>(base + o_roff);
  auto* d_rlen = reinterpret_cast<int32_t*>(base + o_rlen);

  hipStream_t st = s->stream;
  BGO_HIP(hipMemcpyAsync(d_guides, guides, n * 8, hipMemcpyHostToDevice, st));
  BGO_HIP(hipMemcpyAsync(d_masks, pam_masks, n * BGO_MAX_PAM_LEN, hipMemcpyHostToDevice, st));
  BGO_HIP(hipMemcpyAsync(d_pats, pats, n * 4, hipMemcpyHostToDevice, st));
  BGO_HIP(hipMemcpyAsync(d_cmps, cmps, n * 4, hipMemcpyHostToDevice, st));
  BGO_HIP(hipMemsetAsync(d_counts, 0, n * 4, st));

  const int64_t valid_end = s->total_bp - 26;
  BGO_HIP(hipEventRecord(s->ev0, st));
  if (valid_end > 0) {
    const int64_t n_chunks = (valid_end + chunk_size - 1) / chunk_size;
    for (int64_t c = 0; c < n_chunks; ++c) {
      const int64_t cs = c * chunk_size;
      const int64_t ce = std::min(cs + chunk_size, valid_end);
      const int64_t len = ce - cs;
      if (len <= 0) continue;
      const int64_t grid = (len + kPlainBlock - 1) / kPlainBlock;
      if (grid > 0x7FFFFFFFll) return fail(BGO_ERR_ARG, "chunk_size too large");
      hipLaunchKernelGGL(plain_scan_kernel, dim3(static_cast<uint32_t>(grid)), dim3(kPlainBlock),
                         0, st, s->d_words, static_cast<uint64_t>(s->n_words), s->d_nmask,
                         s->has_nmask, s->total_bp, d_guides, n_guides, d_masks, d_pats, d_cmps,
                         max_mismatches, cs, ce, rpg, static_cast<uint32_t>(n_encoded), d_counts,
                         d_pos, d_mm);
      BGO_HIP(hipGetLastError());
    }
  }
  BGO_HIP(hipEventRecord(s->ev1, st));
  BGO_HIP(hipMemcpyAsync(counts, d_counts, n * 4, hipMemcpyDeviceToHost, st));
  BGO_HIP(hipStreamSynchronize(st));
  float ms = 0.f;
  BGO_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  plain_last().ms = ms;

  // Copy back only the recorded prefix of each row.
  std::vector<int64_t> roff(n);
  std::vector<int32_t> rlen(n);
  int64_t total = 0;
  int32_t max_len = 0;
  for (size_t g = 0; g < n; ++g) {
    const int64_t c = counts[g];
    const int32_t len = static_cast<int32_t>(std::min<int64_t>(c, rpg));
    roff[g] = total;
    rlen[g] = len;
    total += len;
    max_len = std::max(max_len, len);
  }
  if (total == 0) return BGO_OK;
  // compact into the tail of the position buffer region is unsafe; use a
  // separate allocation sized to the recorded hits.
  int64_t* d_cpos = nullptr;
  int8_t* d_cmm = nullptr;
  BGO_HIP(hipMalloc(&d_cpos, total * 8));
  if (hipMalloc(&d_cmm, total) != hipSuccess) {
    (void)hipFree(d_cpos);
    return fail(BGO_ERR_HIP, "hipMalloc compact buffer");
  }
  int rc = BGO_OK;
  std::vector<int64_t> hpos(total);
  std::vector<int8_t> hmm(total);
  do {
    if (hipMemcpyAsync(d_roff, roff.data(), n * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_rlen, rlen.data(), n * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
      rc = fail(BGO_ERR_HIP, "upload row offsets");
      break;
    }
    const uint32_t gx = static_cast<uint32_t>(std::min<int64_t>((max_len + 255) / 256, 4096));
    hipLaunchKernelGGL(plain_gather_kernel, dim3(std::max(gx, 1u), static_cast<uint32_t>(n)),
                       dim3(256), 0, st, d_pos, d_mm, rpg, d_roff, d_rlen, n_guides, d_cpos,
                       d_cmm);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(hpos.data(), d_cpos, total * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(hmm.data(), d_cmm, total, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = fail(BGO_ERR_HIP, "gather/copy back hits");
      break;
    }
  } while (false);
  (void)hipFree(d_cpos);
  (void)hipFree(d_cmm);
  if (rc != BGO_OK) return rc;
  for (size_t g = 0; g < n; ++g) {
    if (rlen[g] == 0) continue;
    std::memcpy(positions + g * rpg, hpos.data() + roff[g], rlen[g] * sizeof(int64_t));
    std::memcpy(mismatches + g * rpg, hmm.data() + roff[g], rlen[g]);
  }
  return BGO_OK;
}

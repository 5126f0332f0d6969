// FASTA sequence bytes -> genome artefact arrays, on the GPU.
//
// Layout of beguidedoff/genome_converter.py:15-108 with encoding.py:6-11 and
// :36-46: bases are upper-cased, A C G T -> 0 1 2 3 and every other byte -> 0;
// 16 bases per uint32 word, base i at bits 2i, the last word zero-padded.
// The reverse-complement array holds, contig by contig at the same offsets,
// 3 - code in reverse order (an N, coded as A, becomes T).  The N mask is
// np.packbits of (upper-cased byte == 'N'): the first position of a byte in
// its bit 7, the last byte zero-padded.
//
// Both kernels are HBM-bound: one thread per output word.  The forward
// kernel reads its 16 sequence bytes in one load and writes the word and two
// mask bytes; the reverse-complement kernel builds its word from two forward
// words.
#include "bgo_common.h"

namespace bgo {
namespace enc {

// 2-bit code of a sequence byte (case-folded; non-ACGT -> 0).
__device__ inline uint32_t code_of(uint32_t c) {
  const uint32_t u = c & 0xDFu;  // clears the ASCII case bit only
  const bool acgt = u == 'A' || u == 'C' || u == 'G' || u == 'T';
  return acgt ? (((u >> 1) ^ (u >> 2)) & 3u) : 0u;  // A 0, C 1, G 2, T 3
}

__device__ inline bool is_n(uint32_t c) { return (c & 0xDFu) == 'N'; }

__global__ void fwd_kernel(const uint8_t* __restrict__ seq, uint64_t n, uint64_t n_words,
                           uint64_t n_mask_bytes, uint32_t* __restrict__ words,
                           uint8_t* __restrict__ nmask) {
  const uint64_t w = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (w >= n_words) return;
  const uint64_t p0 = w * 16;
  uint32_t b[4];
  if (p0 + 16 <= n) {
    const uint4 v = *reinterpret_cast<const uint4*>(seq + p0);
    b[0] = v.x; b[1] = v.y; b[2] = v.z; b[3] = v.w;
  } else {
    for (int i = 0; i < 4; ++i) b[i] = 0;
    for (uint64_t p = p0; p < n; ++p) b[(p - p0) >> 2] |= static_cast<uint32_t>(seq[p]) << (8 * ((p - p0) & 3));
  }
  uint32_t word = 0, mask = 0;  // mask bit 15 - j = position p0 + j is N
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const uint32_t c = (b[j >> 2] >> (8 * (j & 3))) & 0xFFu;
    word |= code_of(c) << (2 * j);
    mask |= static_cast<uint32_t>(is_n(c)) << (15 - j);
  }
  words[w] = word;
  if (nmask) {
    if (2 * w + 1 < n_mask_bytes) {  // bytes 2w, 2w + 1 as one little-endian store
      *reinterpret_cast<uint16_t*>(nmask + 2 * w) =
          static_cast<uint16_t>((mask >> 8) | ((mask & 0xFFu) << 8));
    } else if (2 * w < n_mask_bytes) {
      nmask[2 * w] = static_cast<uint8_t>(mask >> 8);
    }
  }
}

// Position q of the reverse-complement array in contig k (offset o, length
// len) holds 3 - (code of forward position 2o + len - 1 - q).  A word that
// lies inside one contig is the forward bases base - 15 .. base read
// backwards: a 32-bit window of the forward words, its 2-bit fields
// reversed (bit reverse, then swap the bits of each pair) and complemented
// (3 - c = c ^ 3).  Words that cross a contig end go base by base.
__global__ void rc_kernel(const uint32_t* __restrict__ fwd, uint64_t n, uint64_t n_words,
                          const int64_t* __restrict__ offsets, const int64_t* __restrict__ lengths,
                          int32_t n_contigs, uint32_t* __restrict__ words) {
  // One binary search per block: the last contig with offset <= the block's
  // first position (empty contigs share the next offset); each thread walks
  // forward from there.
  __shared__ int32_t k_block;
  const uint64_t w = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (threadIdx.x == 0) {
    const uint64_t qb = static_cast<uint64_t>(blockIdx.x) * blockDim.x * 16;
    int32_t lo = 0, hi = n_contigs;
    while (hi - lo > 1) {
      const int32_t mid = (lo + hi) >> 1;
      if (static_cast<uint64_t>(offsets[mid]) <= qb) lo = mid; else hi = mid;
    }
    k_block = lo;
  }
  __syncthreads();
  if (w >= n_words) return;
  const uint64_t q0 = w * 16;
  int32_t k = k_block;
  uint64_t off = static_cast<uint64_t>(offsets[k]), end = off + static_cast<uint64_t>(lengths[k]);
  while (q0 >= end && k + 1 < n_contigs) {
    ++k;
    off = static_cast<uint64_t>(offsets[k]);
    end = off + static_cast<uint64_t>(lengths[k]);
  }
  uint32_t word = 0;
  if (q0 + 16 <= end) {
    const uint64_t first = 2 * off + (end - off) - 1 - (q0 + 15);  // forward base of q0 + 15
    const uint64_t wi = first >> 4;
    const uint32_t sh = static_cast<uint32_t>(2 * (first & 15));
    const uint32_t w0 = fwd[wi], w1 = (sh && wi + 1 < n_words) ? fwd[wi + 1] : 0u;
    const uint32_t win = sh ? ((w0 >> sh) | (w1 << (32 - sh))) : w0;
    const uint32_t r = __builtin_bitreverse32(win);
    word = (((r >> 1) & 0x55555555u) | ((r & 0x55555555u) << 1)) ^ 0xFFFFFFFFu;
  } else {
    for (int j = 0; j < 16; ++j) {
      const uint64_t q = q0 + j;
      if (q >= n) break;
      while (q >= end) {  // the word runs into the next non-empty contig
        ++k;
        off = static_cast<uint64_t>(offsets[k]);
        end = off + static_cast<uint64_t>(lengths[k]);
      }
      const uint64_t s = 2 * off + (end - off) - 1 - q;
      word |= (3u - ((fwd[s >> 4] >> (2 * (s & 15))) & 3u)) << (2 * j);
    }
  }
  words[w] = word;
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace enc
}  // namespace bgo

extern "C" int bgo_encode_genome(int32_t device, const uint8_t* seq, int64_t n_bytes,
                                 const int64_t* contig_offsets, const int64_t* contig_lengths,
                                 int32_t n_contigs, uint32_t* words_fwd, uint32_t* words_rc,
                                 uint8_t* n_mask, float* kernel_ms) {
  using namespace bgo;
  BGO_TRY(injected());
  if (kernel_ms) *kernel_ms = 0.f;
  if (n_bytes < 0 || n_contigs < 0) return fail(BGO_ERR_ARG, "negative size");
  if (n_bytes > 0 && (!seq || !contig_offsets || !contig_lengths || n_contigs == 0))
    return fail(BGO_ERR_ARG, "null sequence or contig table");
  int64_t next = 0;
  for (int32_t k = 0; k < n_contigs; ++k) {
    if (contig_offsets[k] != next || contig_lengths[k] < 0)
      return fail(BGO_ERR_ARG, "contig %d: offset %lld / length %lld do not tile the sequence",
                  k, static_cast<long long>(contig_offsets[k]),
                  static_cast<long long>(contig_lengths[k]));
    next += contig_lengths[k];
  }
  if (next != n_bytes)
    return fail(BGO_ERR_ARG, "contig lengths sum to %lld, sequence has %lld bytes",
                static_cast<long long>(next), static_cast<long long>(n_bytes));
  if (n_bytes == 0) return BGO_OK;
  const uint64_t n = static_cast<uint64_t>(n_bytes);
  const uint64_t n_words = (n + 15) / 16, n_mask_bytes = (n + 7) / 8;
  if (n_words / 256 >= 0x7FFFFFFFull) return fail(BGO_ERR_ARG, "sequence too long");

  BGO_HIP(hipSetDevice(device));
  enc::DevBuf d_seq, d_fwd, d_rc, d_mask, d_tab;
  BGO_HIP(hipMalloc(&d_seq.p, n_words * 16));
  BGO_HIP(hipMalloc(&d_fwd.p, n_words * 4));  // the reverse complement reads it
  if (words_rc) BGO_HIP(hipMalloc(&d_rc.p, n_words * 4));
  if (n_mask) BGO_HIP(hipMalloc(&d_mask.p, n_mask_bytes));
  BGO_HIP(hipMalloc(&d_tab.p, 2 * static_cast<size_t>(n_contigs) * 8));
  auto* offs = static_cast<int64_t*>(d_tab.p);
  BGO_HIP(hipMemcpy(d_seq.p, seq, n, hipMemcpyHostToDevice));
  BGO_HIP(hipMemcpy(offs, contig_offsets, n_contigs * 8, hipMemcpyHostToDevice));
  BGO_HIP(hipMemcpy(offs + n_contigs, contig_lengths, n_contigs * 8, hipMemcpyHostToDevice));

  EventPair ev;
  BGO_HIP(ev.create());
  const hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
  const dim3 grid(static_cast<uint32_t>((n_words + 255) / 256)), block(256);
  const auto* s = static_cast<const uint8_t*>(d_seq.p);
  hipError_t err = hipEventRecord(e0, 0);
  if (err == hipSuccess) {
    hipLaunchKernelGGL(enc::fwd_kernel, grid, block, 0, 0, s, n, n_words, n_mask_bytes,
                       static_cast<uint32_t*>(d_fwd.p), static_cast<uint8_t*>(d_mask.p));
    err = hipGetLastError();
  }
  if (err == hipSuccess && words_rc) {
    hipLaunchKernelGGL(enc::rc_kernel, grid, block, 0, 0, static_cast<const uint32_t*>(d_fwd.p),
                       n, n_words, offs, offs + n_contigs, n_contigs,
                       static_cast<uint32_t*>(d_rc.p));
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipEventRecord(e1, 0);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  float ms = 0.f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  BGO_HIP(err);
  if (kernel_ms) *kernel_ms = ms;
  if (words_fwd) BGO_HIP(hipMemcpy(words_fwd, d_fwd.p, n_words * 4, hipMemcpyDeviceToHost));
  if (words_rc) BGO_HIP(hipMemcpy(words_rc, d_rc.p, n_words * 4, hipMemcpyDeviceToHost));
  if (n_mask) BGO_HIP(hipMemcpy(n_mask, d_mask.p, n_mask_bytes, hipMemcpyDeviceToHost));
  return BGO_OK;
}

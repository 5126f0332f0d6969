// The extensive finder's table assembly on the device: per-site epilogue,
// per-id order and top-N over a plan's site records in HBM
// (beguidedoff/finder.py:193-244, 405-472, 545-566).
//
//  1. emit (one thread per site): coordinates (the reverse-complement array
//     is reversed per contig), on-target exclusion, the fused mismatch mask
//     checked against the site's bases, and for each input mapped to the
//     site's key the editable-base test and the below-threshold count; a
//     (site, input) pair that is editable (or every pair, under return_all)
//     becomes a row candidate with its order keys.
//  2. order: four stable radix passes, least significant first: (reported
//     position, strand, pair order), chromosome-name rank, score, id; the
//     score is the record's f64, whose bit pattern orders positive values.
//  3. select: rank within the id, the top n_top kept (or all), gathered into
//     rows.
// Only the rows are copied to the host.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "bgo_common.h"

namespace bgo {
namespace as {

constexpr uint32_t kW20 = 0xFFFFFu;

struct Dev {  // device arrays of one call, freed at its end
  std::vector<void*> ptrs;
  template <class T>
  hipError_t alloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
  ~Dev() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

// Index of the contig holding x as numpy's searchsorted(offsets, x, 'right')
// - 1, with -1 wrapping to the last contig as Python indexing does
// (finder.py:36-40).
__device__ inline int32_t contig_of(const int64_t* __restrict__ offs, int32_t n, int64_t x) {
  int32_t lo = 0, hi = n;  // first offset > x
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (offs[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  const int32_t k = lo - 1;
  return k < 0 ? n - 1 : k;
}

struct EmitArgs {
  const bgo_hit* hits;
  int64_t n_hits;
  const bgo_asm_key* keys;
  int32_t n_keys;
  const bgo_asm_input* inputs;
  const int64_t* offs;
  const int64_t* lens;
  const uint32_t* name_rank;
  int32_t n_contigs;
  int32_t return_all;
  int32_t use_threshold;
  double threshold;
  // outputs
  unsigned long long* strand_sites;  // [n_keys * 2]
  unsigned long long* below;         // [n_ids]
  unsigned int* n_cand;              // [1]
  unsigned int* bad;                 // [1] fused mask disagreed with the site's bases
  uint32_t* c_hit;
  uint32_t* c_input;
  uint32_t* c_contig;
  uint32_t* c_local;
  uint64_t* c_lo;     // local << 32 | strand << 31 | pair order
  uint32_t* c_name;   // chromosome-name rank
  uint64_t* c_score;  // f64 bits
  uint32_t* c_id;
  uint64_t cap;
};

__global__ void emit_kernel(EmitArgs a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n_hits) return;
  const bgo_hit h = a.hits[i];
  const bgo_asm_key K = a.keys[h.guide];
  const int64_t pos = h.pos;
  const uint32_t s = h.strand;
  atomicAdd(&a.strand_sites[2u * h.guide + s], 1ull);
  const int64_t window = 20 + K.pam_len;
  int64_t lm = pos, report = pos;
  if (s == 1) {
    const int32_t k = contig_of(a.offs, a.n_contigs, pos);
    lm = a.offs[k] + (a.lens[k] - (pos - a.offs[k]) - window);
    report = lm + window - 1;
  }
  if (K.ot_contig >= 0) {
    const int32_t k = contig_of(a.offs, a.n_contigs, lm);
    if (k == K.ot_contig && lm - a.offs[k] == K.ot_local && static_cast<int32_t>(s) == K.ot_strand)
      return;
  }
  const uint32_t mm = ((h.x0 ^ K.g0) | (h.x1 ^ K.g1)) & kW20;
  if (mm != h.mm_mask) atomicOr(a.bad, 1u);
  const uint32_t is_a = ~(h.x0 | h.x1) & kW20;
  const uint32_t is_c = h.x0 & ~h.x1 & kW20;
  const int32_t rc = contig_of(a.offs, a.n_contigs, report);
  const uint32_t local = static_cast<uint32_t>(report - a.offs[rc]);
  for (uint32_t j = 0; j < K.n_inputs; ++j) {
    const uint32_t ii = K.first_input + j;
    const bgo_asm_input I = a.inputs[ii];
    const bool ed = ((I.target ? is_c : is_a) & I.window) != 0;
    if (a.use_threshold && ed && h.score <= a.threshold) atomicAdd(&a.below[I.id], 1ull);
    if (!(ed || a.return_all)) continue;
    const unsigned int slot = atomicAdd(a.n_cand, 1u);
    if (slot >= a.cap) continue;
    a.c_hit[slot] = static_cast<uint32_t>(i);
    a.c_input[slot] = ii;
    a.c_contig[slot] = static_cast<uint32_t>(rc);
    a.c_local[slot] = local;
    a.c_lo[slot] = (static_cast<uint64_t>(local) << 32) | (static_cast<uint64_t>(s) << 31) |
                   (I.order & 0x7FFFFFFFu);
    a.c_name[slot] = a.name_rank[rc];
    a.c_score[slot] = __double_as_longlong(h.score);
    a.c_id[slot] = I.id;
  }
}

template <class T>
__global__ void gather_kernel(const T* __restrict__ src, const uint32_t* __restrict__ perm,
                              T* __restrict__ dst, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

__global__ void seg_start_kernel(const uint32_t* __restrict__ ids, uint32_t n,
                                 uint32_t* __restrict__ start) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (i == 0 || ids[i] != ids[i - 1])) start[ids[i]] = i;
}

__global__ void keep_kernel(const uint32_t* __restrict__ ids, uint32_t n,
                            const uint32_t* __restrict__ start, int32_t n_top, int32_t all,
                            uint32_t* __restrict__ keep) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keep[i] = (all || i - start[ids[i]] < static_cast<uint32_t>(n_top)) ? 1u : 0u;
}

__global__ void rows_kernel(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ ids,
                            uint32_t n, const uint32_t* __restrict__ start,
                            const uint32_t* __restrict__ keep, const uint32_t* __restrict__ out_at,
                            const bgo_hit* __restrict__ hits, const uint32_t* __restrict__ c_hit,
                            const uint32_t* __restrict__ c_input,
                            const uint32_t* __restrict__ c_contig,
                            const uint32_t* __restrict__ c_local, bgo_asm_row* __restrict__ rows,
                            uint64_t rows_cap) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !keep[i] || out_at[i] >= rows_cap) return;
  const uint32_t c = perm[i];
  bgo_asm_row r;
  r.hit = hits[c_hit[c]];
  r.input = c_input[c];
  r.rank = i - start[ids[i]];
  r.contig = c_contig[c];
  r.local = c_local[c];
  rows[out_at[i]] = r;
}

inline uint32_t blocks(uint64_t n) { return static_cast<uint32_t>((n + 255) / 256); }

}  // namespace as
}  // namespace bgo

using namespace bgo;

extern "C" int bgo_plan_assemble(bgo_plan* P, const bgo_asm_key* keys, int32_t n_keys,
                                 const bgo_asm_input* inputs, int32_t n_inputs, int32_t n_ids,
                                 const int64_t* contig_offsets, const int64_t* contig_lengths,
                                 const uint32_t* contig_name_rank, int32_t n_contigs,
                                 int32_t n_top, int32_t return_all, double threshold,
                                 int32_t use_threshold, int64_t strand_ceiling, bgo_asm_row* rows,
                                 int64_t rows_cap, int64_t* n_rows_out, int64_t* below_out,
                                 int64_t* max_strand_sites_out) {
  if (!P || n_keys < 0 || (n_keys > 0 && !keys) || n_inputs < 0 || (n_inputs > 0 && !inputs) ||
      n_ids < 0 || n_contigs <= 0 || !contig_offsets || !contig_lengths || !contig_name_rank ||
      rows_cap < 0 || (rows_cap > 0 && !rows) || !n_rows_out || (n_ids > 0 && !below_out) ||
      !max_strand_sites_out)
    return fail(BGO_ERR_ARG, "bgo_plan_assemble: bad arguments");
  for (int32_t k = 0; k < n_keys; ++k) {
    if (static_cast<int64_t>(keys[k].first_input) + keys[k].n_inputs > n_inputs ||
        keys[k].pam_len < 1 || keys[k].pam_len > 6)
      return fail(BGO_ERR_ARG, "bgo_plan_assemble: key %d out of range", k);
  }
  uint32_t max_inputs = 0;
  for (int32_t k = 0; k < n_keys; ++k) max_inputs = std::max(max_inputs, keys[k].n_inputs);
  for (int32_t j = 0; j < n_inputs; ++j) {
    if (inputs[j].id >= static_cast<uint32_t>(n_ids) || inputs[j].order > 0x7FFFFFFFu)
      return fail(BGO_ERR_ARG, "bgo_plan_assemble: input %d out of range", j);
  }
  BGO_TRY(injected());
  const bgo_hit* d_hits = nullptr;
  int64_t stored = 0, total = 0;
  hipStream_t st = nullptr;
  int32_t plan_guides = 0;
  BGO_TRY(plan_device_hits(P, &d_hits, &stored, &total, &st, &plan_guides));
  if (plan_guides != n_keys)  // records index the keys by the plan's guide index
    return fail(BGO_ERR_ARG, "bgo_plan_assemble: %d keys for a plan of %d guides", n_keys,
                plan_guides);
  if (total > stored)
    return fail(BGO_ERR_ARG, "bgo_plan_assemble: %lld sites exceed the hit buffer (%lld stored)",
                (long long)total, (long long)stored);
  const uint64_t n_hits = static_cast<uint64_t>(stored);
  const uint64_t cap = n_hits * std::max<uint32_t>(max_inputs, 1);
  if (cap >= (uint64_t{1} << 31))
    return fail(BGO_ERR_ARG, "bgo_plan_assemble: %llu row candidates exceed 2^31",
                (unsigned long long)cap);
  as::Dev D;
  bgo_asm_key* d_keys;
  bgo_asm_input* d_inputs;
  int64_t *d_offs, *d_lens;
  uint32_t* d_name;
  unsigned long long *d_strand, *d_below;
  unsigned int *d_ncand, *d_bad;
  BGO_HIP(D.alloc(&d_keys, n_keys));
  BGO_HIP(D.alloc(&d_inputs, n_inputs));
  BGO_HIP(D.alloc(&d_offs, n_contigs));
  BGO_HIP(D.alloc(&d_lens, n_contigs));
  BGO_HIP(D.alloc(&d_name, n_contigs));
  BGO_HIP(D.alloc(&d_strand, 2 * static_cast<size_t>(n_keys)));
  BGO_HIP(D.alloc(&d_below, n_ids));
  BGO_HIP(D.alloc(&d_ncand, 2));
  d_bad = d_ncand + 1;
  as::EmitArgs a{};
  BGO_HIP(D.alloc(&a.c_hit, cap));
  BGO_HIP(D.alloc(&a.c_input, cap));
  BGO_HIP(D.alloc(&a.c_contig, cap));
  BGO_HIP(D.alloc(&a.c_local, cap));
  BGO_HIP(D.alloc(&a.c_lo, cap));
  BGO_HIP(D.alloc(&a.c_name, cap));
  BGO_HIP(D.alloc(&a.c_score, cap));
  BGO_HIP(D.alloc(&a.c_id, cap));
  if (n_keys) BGO_HIP(hipMemcpyAsync(d_keys, keys, n_keys * sizeof(bgo_asm_key), hipMemcpyHostToDevice, st));
  if (n_inputs)
    BGO_HIP(hipMemcpyAsync(d_inputs, inputs, n_inputs * sizeof(bgo_asm_input), hipMemcpyHostToDevice, st));
  BGO_HIP(hipMemcpyAsync(d_offs, contig_offsets, n_contigs * sizeof(int64_t), hipMemcpyHostToDevice, st));
  BGO_HIP(hipMemcpyAsync(d_lens, contig_lengths, n_contigs * sizeof(int64_t), hipMemcpyHostToDevice, st));
  BGO_HIP(hipMemcpyAsync(d_name, contig_name_rank, n_contigs * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  BGO_HIP(hipMemsetAsync(d_strand, 0, 2 * static_cast<size_t>(n_keys) * sizeof(unsigned long long), st));
  BGO_HIP(hipMemsetAsync(d_below, 0, std::max(n_ids, 1) * sizeof(unsigned long long), st));
  BGO_HIP(hipMemsetAsync(d_ncand, 0, 2 * sizeof(unsigned int), st));
  a.hits = d_hits;
  a.n_hits = static_cast<int64_t>(n_hits);
  a.keys = d_keys;
  a.n_keys = n_keys;
  a.inputs = d_inputs;
  a.offs = d_offs;
  a.lens = d_lens;
  a.name_rank = d_name;
  a.n_contigs = n_contigs;
  a.return_all = return_all;
  a.use_threshold = use_threshold;
  a.threshold = threshold;
  a.strand_sites = d_strand;
  a.below = d_below;
  a.n_cand = d_ncand;
  a.bad = d_bad;
  a.cap = cap;
  if (n_hits)
    hipLaunchKernelGGL(as::emit_kernel, dim3(as::blocks(n_hits)), dim3(256), 0, st, a);
  BGO_HIP(hipGetLastError());
  // the ceiling check: the most sites of one key on one strand
  std::vector<unsigned long long> ss(2 * static_cast<size_t>(n_keys) + 1, 0);
  unsigned int hc[2] = {0, 0};
  if (n_keys)
    BGO_HIP(hipMemcpyAsync(ss.data(), d_strand, 2 * n_keys * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, st));
  BGO_HIP(hipMemcpyAsync(hc, d_ncand, sizeof(hc), hipMemcpyDeviceToHost, st));
  BGO_HIP(hipStreamSynchronize(st));
  if (hc[1]) return fail(BGO_ERR_HIP, "fused epilogue mismatch mask disagrees with the site's bases");
  const int64_t mx = static_cast<int64_t>(*std::max_element(ss.begin(), ss.end()));
  *max_strand_sites_out = mx;
  *n_rows_out = 0;
  if (mx > strand_ceiling) return BGO_OK;
  const uint32_t n = hc[0];
  if (n_ids) {
    std::vector<unsigned long long> b(n_ids);
    BGO_HIP(hipMemcpyAsync(b.data(), d_below, n_ids * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, st));
    BGO_HIP(hipStreamSynchronize(st));
    for (int32_t i = 0; i < n_ids; ++i) below_out[i] = static_cast<int64_t>(b[i]);
  }
  if (n == 0) return BGO_OK;
  // order: stable LSD passes over the candidate permutation
  uint32_t *perm, *perm2, *k32, *k32s, *start, *keep, *out_at;
  uint64_t *k64, *k64s;
  BGO_HIP(D.alloc(&perm, n));
  BGO_HIP(D.alloc(&perm2, n));
  BGO_HIP(D.alloc(&k32, n));
  BGO_HIP(D.alloc(&k32s, n));
  BGO_HIP(D.alloc(&k64, n));
  BGO_HIP(D.alloc(&k64s, n));
  BGO_HIP(D.alloc(&start, std::max(n_ids, 1)));
  BGO_HIP(D.alloc(&keep, n));
  BGO_HIP(D.alloc(&out_at, n));
  size_t tb = 0, t1 = 0;
  BGO_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, k64, k64s, perm, perm2, static_cast<int>(n), 0, 64, st));
  tb = std::max(tb, t1);
  BGO_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, k32, k32s, perm, perm2, static_cast<int>(n), 0, 32, st));
  tb = std::max(tb, t1);
  BGO_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, keep, out_at, static_cast<int>(n), st));
  tb = std::max(tb, t1);
  void* temp;
  BGO_HIP(D.alloc(reinterpret_cast<char**>(&temp), tb));
  // pass 1: (local, strand, pair order); values: candidate index
  {
    // identity permutation
    std::vector<uint32_t> iota(n);
    for (uint32_t i = 0; i < n; ++i) iota[i] = i;
    BGO_HIP(hipMemcpyAsync(perm, iota.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    BGO_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tb, a.c_lo, k64s, perm, perm2,
                                               static_cast<int>(n), 0, 64, st));
    BGO_HIP(hipStreamSynchronize(st));  // iota stays alive until the copy is done
  }
  // pass 2: chromosome-name rank
  hipLaunchKernelGGL(as::gather_kernel<uint32_t>, dim3(as::blocks(n)), dim3(256), 0, st, a.c_name,
                     perm2, k32, n);
  BGO_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tb, k32, k32s, perm2, perm, static_cast<int>(n),
                                             0, 32, st));
  // pass 3: score
  hipLaunchKernelGGL(as::gather_kernel<uint64_t>, dim3(as::blocks(n)), dim3(256), 0, st, a.c_score,
                     perm, k64, n);
  BGO_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tb, k64, k64s, perm, perm2, static_cast<int>(n),
                                             0, 64, st));
  // pass 4: id
  hipLaunchKernelGGL(as::gather_kernel<uint32_t>, dim3(as::blocks(n)), dim3(256), 0, st, a.c_id,
                     perm2, k32, n);
  BGO_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tb, k32, k32s, perm2, perm, static_cast<int>(n),
                                             0, 32, st));
  // k32s: ids in final order, perm: candidate of each final position
  hipLaunchKernelGGL(as::seg_start_kernel, dim3(as::blocks(n)), dim3(256), 0, st, k32s, n, start);
  hipLaunchKernelGGL(as::keep_kernel, dim3(as::blocks(n)), dim3(256), 0, st, k32s, n, start,
                     n_top, return_all, keep);
  BGO_HIP(hipcub::DeviceScan::ExclusiveSum(temp, tb, keep, out_at, static_cast<int>(n), st));
  uint32_t last[2] = {0, 0};
  BGO_HIP(hipMemcpyAsync(&last[0], out_at + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  BGO_HIP(hipMemcpyAsync(&last[1], keep + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  BGO_HIP(hipStreamSynchronize(st));
  const int64_t n_rows = static_cast<int64_t>(last[0]) + last[1];
  *n_rows_out = n_rows;
  const int64_t m = std::min(n_rows, rows_cap);
  if (m <= 0) return BGO_OK;
  bgo_asm_row* d_rows;
  BGO_HIP(D.alloc(&d_rows, m));
  hipLaunchKernelGGL(as::rows_kernel, dim3(as::blocks(n)), dim3(256), 0, st, perm, k32s, n, start,
                     keep, out_at, d_hits, a.c_hit, a.c_input, a.c_contig, a.c_local, d_rows,
                     static_cast<uint64_t>(m));
  BGO_HIP(hipGetLastError());
  BGO_HIP(hipMemcpyAsync(rows, d_rows, m * sizeof(bgo_asm_row), hipMemcpyDeviceToHost, st));
  BGO_HIP(hipStreamSynchronize(st));
  return BGO_OK;
}

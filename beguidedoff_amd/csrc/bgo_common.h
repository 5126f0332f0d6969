// Shared host/device helpers of libbgo_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "bgo.h"

// A genome resident on one device: both strands as bit planes (block b: .x
// bit j = bit 0 of base 32b + j's code, .y = bit 1) and little-endian N words
// (one per block), zero padded past the genome.
struct bgo_genome {
  int device = 0;
  int64_t total_bp = 0;
  int64_t n_words = 0;
  uint64_t n_blocks = 0;  // 32-base blocks, including zero padding
  int n_strands = 0;
  uint2* d_planes[2] = {nullptr, nullptr};
  uint32_t* d_nwords[2] = {nullptr, nullptr};
  int64_t bytes = 0;
};

namespace bgo {

// ----------------------------------------------------------------- errors

inline std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}

inline int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  last_error() = buf;
  return code;
}

inline std::atomic<int>& fail_countdown() {
  static std::atomic<int> n{0};
  return n;
}

// Every entry point that touches the GPU calls this first; it turns an armed
// bgo_debug_fail_next() into an error before any device work happens.
inline int injected() {
  int v = fail_countdown().load();
  while (v > 0) {
    if (fail_countdown().compare_exchange_weak(v, v - 1)) {
      if (v == 1) return fail(BGO_ERR_INJECTED, "injected failure (bgo_debug_fail_next)");
      return BGO_OK;
    }
  }
  return BGO_OK;
}

#define BGO_HIP(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return ::bgo::fail(BGO_ERR_HIP, "%s failed: %s (%s:%d)", #expr,        \
                         hipGetErrorString(e_), __FILE__, __LINE__);         \
  } while (0)

#define BGO_TRY(expr)              \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != BGO_OK) return rc_; \
  } while (0)

// Two timing events, destroyed with the object (also on early returns).
struct EventPair {
  hipEvent_t e[2] = {nullptr, nullptr};
  hipError_t create() {
    hipError_t r = hipEventCreate(&e[0]);
    if (r == hipSuccess) r = hipEventCreate(&e[1]);
    return r;
  }
  ~EventPair() {
    for (hipEvent_t ev : e)
      if (ev) (void)hipEventDestroy(ev);
  }
};

// ------------------------------------------------------------ device math

// Gather bits 0, 2, 4, ..., 30 of x into bits 0..15.
__host__ __device__ inline uint32_t even_bits(uint32_t x) {
  x &= 0x55555555u;
  x = (x | (x >> 1)) & 0x33333333u;
  x = (x | (x >> 2)) & 0x0F0F0F0Fu;
  x = (x | (x >> 4)) & 0x00FF00FFu;
  x = (x | (x >> 8)) & 0x0000FFFFu;
  return x;
}

// Bits [sh, sh + 32) of the 64-bit value hi:lo (sh in 0..31).
__device__ inline uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t sh) {
  return __builtin_amdgcn_alignbit(hi, lo, sh);
}

// (a << sh) | b as one v_lshl_or_b32 (the compiler otherwise re-associates
// a constant-shifted OR into an OR followed by a shift).
__device__ inline uint32_t lshl_or(uint32_t a, uint32_t sh, uint32_t b) {
  uint32_t r;
  asm("v_lshl_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "i"(sh), "v"(b));
  return r;
}

__device__ inline uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of `mask` below this lane.
__device__ inline uint32_t lanes_below(uint64_t mask, uint32_t plus = 0) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), plus));
}

__device__ inline uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Per-device set-up of bgo_scan_batched_noalloc (bgo_batched.hip), run by
// bgo_init with the device current.
int noalloc_setup(int device);

// The last launch's site records of a plan, on its device (bgo_batched.hip).
int plan_device_hits(bgo_plan* P, const bgo_hit** hits, int64_t* stored, int64_t* total,
                     hipStream_t* stream, int32_t* n_guides);

}  // namespace bgo

// Library-level entry points of libbgo_hip.so: version, devices, errors.
#include "bgo_common.h"

#define BGO_ABI_VERSION 1

extern "C" int bgo_abi_version(void) { return BGO_ABI_VERSION; }

extern "C" const char* bgo_last_error(void) { return bgo::last_error().c_str(); }

extern "C" int bgo_debug_fail_next(int32_t n_calls) {
  if (n_calls < 0) return bgo::fail(BGO_ERR_ARG, "bgo_debug_fail_next: negative count");
  bgo::fail_countdown().store(n_calls);
  return BGO_OK;
}

extern "C" int bgo_device_count(int32_t* n_out) {
  if (!n_out) return bgo::fail(BGO_ERR_ARG, "null pointer");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *n_out = n;
  return BGO_OK;
}

extern "C" int bgo_init(int32_t device) {
  BGO_TRY(bgo::injected());
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return bgo::fail(BGO_ERR_NODEV, "no HIP device available");
  if (device < 0 || device >= n)
    return bgo::fail(BGO_ERR_NODEV, "device %d out of range (%d visible)", device, n);
  BGO_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  BGO_HIP(hipGetDeviceProperties(&prop, device));
  if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
    return bgo::fail(BGO_ERR_NODEV, "device %d is %s; this library is built for gfx950", device,
                     prop.gcnArchName);
  return bgo::noalloc_setup(device);
}

extern "C" int bgo_device_info(int32_t device, int32_t* cu_count, int32_t* clock_khz,
                               int64_t* hbm_bytes, char* arch, int32_t arch_len) {
  BGO_TRY(bgo::injected());
  hipDeviceProp_t prop;
  BGO_HIP(hipGetDeviceProperties(&prop, device));
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (clock_khz) *clock_khz = prop.clockRate;
  if (hbm_bytes) *hbm_bytes = static_cast<int64_t>(prop.totalGlobalMem);
  if (arch && arch_len > 0) {
    std::snprintf(arch, static_cast<size_t>(arch_len), "%s", prop.gcnArchName);
  }
  return BGO_OK;
}

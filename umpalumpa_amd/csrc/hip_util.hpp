// Host helpers shared by the C ABI's translation units: HIP status mapping, a scoped device
// switch, and the payload checks of the reference's Size / alignment rules.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

#include "umpalumpa_amd.h"

namespace umpa::detail {

inline int from_hip(hipError_t e)
{
  if (e == hipSuccess) return UMPA_OK;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return UMPA_ERR_OUT_OF_MEMORY;
  return UMPA_ERR_DEVICE;
}

// Makes `device` current for the lifetime of the guard.
struct DeviceGuard {
  int prev = -1;
  bool ok = false;
  explicit DeviceGuard(int device)
  {
    if (hipGetDevice(&prev) != hipSuccess) return;
    ok = (prev == device) || hipSetDevice(device) == hipSuccess;
  }
  ~DeviceGuard()
  {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
  DeviceGuard(const DeviceGuard &) = delete;
  DeviceGuard &operator=(const DeviceGuard &) = delete;
};

// Size::IsValid (data/size.hpp:43-57, 86): n >= 1 and a valid 1/2/3-D shape
inline bool size_valid(const umpa_size &s)
{
  if (s.n < 1) return false;
  const bool d3 = s.z >= 2 && s.y >= 2 && s.x >= 2;
  const bool d2 = s.z == 1 && s.y >= 2 && s.x >= 2;
  const bool d1 = s.z == 1 && s.y == 1 && s.x >= 1;
  return d3 || d2 || d1;
}

inline bool aligned(const void *p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

}  // namespace umpa::detail

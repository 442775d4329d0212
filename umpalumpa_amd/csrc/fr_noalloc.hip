// Device scratch of the allocation-free insertion entry point (umpa_fr_execute_noalloc).
//
// The operation handle keeps its work items and work counter in device memory it allocates at
// Init.  Callers that cannot allocate (code under stream capture, tight serving loops) use
// this file-scope array instead: a 256-byte header (the work counter at byte 0) followed by
// the work items, int4 each.  Every GPU holds its own copy until the process ends; host code
// reaches it only through the address hipGetSymbolAddress returns for the current device.
// All calls on one GPU share it: the schedule a call needs is uploaded only when it differs
// from the one the array holds, and the counter is cleared by a kernel before every launch.
// Calls that overlap on one GPU (other streams, other host threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "launch.hpp"

namespace umpa {

constexpr size_t kScratchHeader = 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

struct DeviceScratch {
  void *ptr = nullptr;  // device address of `scratch` on this GPU
  bool have = false;    // a schedule has been uploaded
  int vx = 0, vy = 0, vz = 0;
  bool split = false;
  int nItems = 0;
};

std::mutex g_mutex;
DeviceScratch g_devices[kNoAllocMaxDevices];

}  // namespace

hipError_t launch_fr_insert_noalloc(int device, const FrJob &job, const FrVariant &v,
                                    const int *items, int nItems, bool split, const float *table,
                                    int nTable, int numCUs, hipStream_t st)
{
  if (device < 0 || device >= kNoAllocMaxDevices || nItems < 0 ||
      (size_t)nItems > kNoAllocMaxItems)
    return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(g_mutex);
  DeviceScratch &d = g_devices[device];
  hipError_t e = hipSuccess;
  if (!d.ptr) {
    // once per device, with that device current (the caller's guard made it so)
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    d.ptr = p;
  }
  unsigned char *base = static_cast<unsigned char *>(d.ptr);
  if (!d.have || d.vx != job.vx || d.vy != job.vy || d.vz != job.vz || d.split != split ||
      d.nItems != nItems) {
    // earlier launches on this stream may still read the schedule being replaced
    e = hipStreamSynchronize(st);
    if (e == hipSuccess && nItems > 0)
      e = hipMemcpy(base + kScratchHeader, items, (size_t)nItems * 16, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      d.have = false;
      return e;
    }
    d.have = true;
    d.vx = job.vx;
    d.vy = job.vy;
    d.vz = job.vz;
    d.split = split;
    d.nItems = nItems;
  }
  FrLaunch l{};
  l.items = reinterpret_cast<const int *>(base + kScratchHeader);
  l.nItems = nItems;
  l.counter = reinterpret_cast<int *>(base);
  l.updates = reinterpret_cast<unsigned long long *>(base + 8);
  l.table = table;
  l.nTable = nTable;
  e = prepare_fr_insert(v, nTable);
  if (e == hipSuccess) e = launch_fr_insert(job, l, v, numCUs, st);
  return e;
}

}  // namespace umpa

// C ABI of umpalumpa_amd: operation handles, validation with the reference's Init/Execute
// semantics, and dispatch to the gfx950 kernels.  Nothing here throws across the ABI.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <string>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "fr_host.hpp"
#include "hip_util.hpp"
#include "launch.hpp"
#include "umpalumpa_amd.h"

namespace {

using umpa::detail::aligned;
using umpa::detail::DeviceGuard;
using umpa::detail::from_hip;
using umpa::detail::size_valid;

int query_cus(int device)
{
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      cus <= 0)
    cus = 256;
  return cus;
}

}  // namespace

struct umpa_fr_op {
  int device = 0;
  hipStream_t stream = nullptr;
  int numCUs = 256;
  bool initialized = false;
  umpa_fr_desc desc{};
  umpa_fr_settings settings{};
  umpa_fr_constants constants{};
  int strategy = UMPA_FR_EXACT;
  bool countUpdates = false;
  unsigned long long lastUpdates = 0;
  int *dBricks = nullptr;      // exact schedule: every brick owned by one wavefront
  int nBricks = 0;
  int *dBricksSplit = nullptr; // fast-math schedule: central bricks split over wavefronts
  int nBricksSplit = 0;
  int *dCounter = nullptr;
  unsigned long long *dUpdates = nullptr;
  int nb = 0;
  umpa::host::SchedParams params[2];  // schedule knobs: exact / fast-math schedules

  void release()
  {
    if (dBricks) (void)hipFree(dBricks);
    if (dBricksSplit) (void)hipFree(dBricksSplit);
    if (dCounter) (void)hipFree(dCounter);
    if (dUpdates) (void)hipFree(dUpdates);
    dBricks = nullptr;
    dBricksSplit = nullptr;
    nBricksSplit = 0;
    dCounter = nullptr;
    dUpdates = nullptr;
    nBricks = 0;
    initialized = false;
  }
};

// Work items as the kernel reads them (int4 per item) in a new device buffer.
static hipError_t upload_bricks(const std::vector<umpa::host::BrickOrigin> &bricks, int **dst)
{
  std::vector<int> flat(bricks.size() * 4);
  for (size_t i = 0; i < bricks.size(); ++i) {
    flat[4 * i + 0] = bricks[i].x;
    flat[4 * i + 1] = bricks[i].y;
    flat[4 * i + 2] = bricks[i].z;
    flat[4 * i + 3] = bricks[i].w;
  }
  hipError_t e = hipMalloc(dst, std::max<size_t>(flat.size(), 4) * sizeof(int));
  if (e == hipSuccess && !flat.empty())
    e = hipMemcpy(*dst, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice);
  return e;
}

struct umpa_fp_op {
  int device = 0;
  hipStream_t stream = nullptr;
  int numCUs = 256;
  bool initialized = false;
  umpa_fp_desc desc{};
  umpa_fp_settings settings{};
};

// IsValid (afr.hpp:47-54) and Strategy1::Init (fr_cpu.cpp:14-19): the checks of Init.
static int fr_check_init(const umpa_fr_desc &desc, const umpa_fr_settings &s)
{
  // odd volume x only
  if (!size_valid(desc.volume) || !size_valid(desc.fft)) return UMPA_ERR_INVALID_ARG;
  if (desc.volume.x % 2 == 0) return UMPA_ERR_INVALID_ARG;
  // the volume is indexed with one extent for y and z (fr_cpu_kernels.hpp:114-115)
  if (desc.volume.y != desc.volume.z || desc.volume.n != 1) return UMPA_ERR_INVALID_ARG;
  if (s.interpolation < 0 || s.interpolation > 1 || s.type < 0 || s.type > 1 || s.order < 0 ||
      s.order > 4 || !(s.blobRadius > 0.f) || !std::isfinite(s.alpha))
    return UMPA_ERR_INVALID_ARG;
  if (desc.volume.x > (uint64_t)INT_MAX || desc.volume.y > (uint64_t)INT_MAX ||
      desc.fft.x > (uint64_t)INT_MAX || desc.fft.y > (uint64_t)INT_MAX ||
      desc.fft.n > (uint64_t)INT_MAX || desc.blobTable > (uint64_t)INT_MAX)
    return UMPA_ERR_UNSUPPORTED;
  // a slice is read through a buffer resource with a 32-bit byte range (complex float pixels)
  if (desc.fft.x * desc.fft.y * 8 > (uint64_t)INT_MAX) return UMPA_ERR_UNSUPPORTED;
  return UMPA_OK;
}

// Widest blob footprint in pixels: floor(2r) + 1 (plus a margin for rounding at the edge),
// clamped to the unrolled widths 3..6; wider footprints take the kernel's generic path.
static int fr_footprint(const umpa_fr_settings &s) { return umpa::host::footprint(s); }

extern "C" {

int umpa_version(void) { return 101; }

const char *umpa_status_string(int status)
{
  switch (status) {
  case UMPA_OK: return "ok";
  case UMPA_ERR_INVALID_ARG: return "invalid argument";
  case UMPA_ERR_NOT_INITIALIZED: return "operation not initialized";
  case UMPA_ERR_NOT_EQUIVALENT: return "payload not equivalent to the initialized one";
  case UMPA_ERR_REJECTED: return "execution rejected (the reference would return false)";
  case UMPA_ERR_UNSUPPORTED: return "unsupported";
  case UMPA_ERR_DEVICE: return "HIP device error";
  case UMPA_ERR_OUT_OF_MEMORY: return "out of device memory";
  default: return "unknown status";
  }
}

// ---- FR ---------------------------------------------------------------------------------
int umpa_fr_create(umpa_fr_op **op, int device, void *stream)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  *op = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return UMPA_ERR_DEVICE;
  auto *h = new (std::nothrow) umpa_fr_op();
  if (!h) return UMPA_ERR_OUT_OF_MEMORY;
  h->device = device;
  h->stream = static_cast<hipStream_t>(stream);
  h->numCUs = query_cus(device);
  *op = h;
  return UMPA_OK;
}

int umpa_fr_init(umpa_fr_op *op, const umpa_fr_desc *desc, const umpa_fr_settings *s)
{
  if (!op || !desc || !s) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  // Init starts from a clean state (basic_operation.hpp:30-31)
  op->release();
  const int st = fr_check_init(*desc, *s);
  if (st != UMPA_OK) return st;
  try {
    op->desc = *desc;
    op->settings = *s;
    op->constants = umpa::host::create_constants(desc->volume.x, desc->volume.y,
                                                 desc->blobTable, *s);
    const int vx = (int)desc->volume.x, vy = (int)desc->volume.y, vz = (int)desc->volume.z;
    auto bricks = umpa::host::make_brick_schedule(vx, vy, vz, false, op->params[0]);
    auto split = umpa::host::make_brick_schedule(vx, vy, vz, true, op->params[1]);
    hipError_t e = upload_bricks(bricks, &op->dBricks);
    if (e == hipSuccess) e = upload_bricks(split, &op->dBricksSplit);
    if (e == hipSuccess) e = hipMalloc(&op->dCounter, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&op->dUpdates, sizeof(unsigned long long));
    if (e != hipSuccess) {
      op->release();
      return from_hip(e);
    }
    op->nBricksSplit = (int)split.size();
    op->nBricks = (int)bricks.size();
    op->nb = fr_footprint(*s);
    op->initialized = true;
    return UMPA_OK;
  } catch (const std::bad_alloc &) {
    op->release();
    return UMPA_ERR_OUT_OF_MEMORY;
  } catch (...) {
    op->release();
    return UMPA_ERR_DEVICE;
  }
}

// The reference's Execute preconditions on the payload of an operation initialized with
// (desc, s).
static int fr_check_payload(const umpa_fr_desc &desc, const umpa_fr_settings &s,
                            const void *fft, const float *scales, bool half, uint64_t fftCount,
                            const void *volume, const float *weights,
                            const umpa_traverse_space *spaces, uint64_t nSpaces,
                            const float *table)
{
  // BasicOperation::Execute equivalence (basic_operation.hpp:62-63): N may shrink
  if (fftCount > desc.fft.n) return UMPA_ERR_NOT_EQUIVALENT;
  // FRCUDA Strategy1::Execute rejects empty payloads (fr_cuda.cpp:165-167)
  if (!fft || !volume || !weights || fftCount == 0) return UMPA_ERR_REJECTED;
  // FRCPU rejects kLookup without a table (fr_cpu.cpp:27-29)
  if (s.interpolation == UMPA_FR_LOOKUP && (!table || desc.blobTable == 0))
    return UMPA_ERR_REJECTED;
  if (nSpaces > (uint64_t)INT_MAX) return UMPA_ERR_UNSUPPORTED;
  if (nSpaces > 0 && !spaces) return UMPA_ERR_INVALID_ARG;
  if (!aligned(volume, 8) || !aligned(weights, 4) || !aligned(fft, half ? 4 : 8) ||
      !aligned(spaces, 4) || (table && !aligned(table, 4)))
    return UMPA_ERR_INVALID_ARG;
  if (half && (!scales || !aligned(scales, 4))) return UMPA_ERR_INVALID_ARG;
  return UMPA_OK;
}

// Validation shared by execute and autotune: the reference's Execute preconditions.
static int fr_check_execute(umpa_fr_op *op, const void *fft, const float *scales, bool half,
                            uint64_t fftCount, const void *volume, const float *weights,
                            const umpa_traverse_space *spaces, uint64_t nSpaces,
                            const float *table)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  if (!op->initialized) return UMPA_ERR_NOT_INITIALIZED;
  return fr_check_payload(op->desc, op->settings, fft, scales, half, fftCount, volume, weights,
                          spaces, nSpaces, table);
}

static umpa::FrJob fr_job_record(const umpa_fr_desc &desc, const umpa_fr_settings &s,
                               const umpa_fr_constants &c, const void *fft,
                               const float *scales, uint64_t fftCount, void *volume,
                               float *weights, const umpa_traverse_space *spaces,
                               uint64_t nSpaces, bool boxFixed)
{
  umpa::FrJob a{};
  a.volume = volume;
  a.weights = weights;
  a.fft = fft;
  a.scales = scales;
  a.xSize = (int)desc.fft.x;
  a.ySize = (int)desc.fft.y;
  a.sliceElems = (int64_t)desc.fft.x * (int64_t)desc.fft.y;
  a.nProj = (int)fftCount;
  a.nSpaces = (int)nSpaces;
  a.spaceBegin = 0;
  a.spaceEnd = (int)nSpaces;
  a.spaces = spaces;
  a.c = c;
  a.vx = (int)desc.volume.x;
  a.vy = (int)desc.volume.y;
  a.vz = (int)desc.volume.z;
  const bool nearest = s.type == UMPA_FR_FAST;
  a.cullHalfWidth = nearest ? 1.6f : s.blobRadius;
  a.cullHalfSpace = nearest ? 0.f : s.blobRadius;
  a.shiftBox = !boxFixed && umpa::host::box_can_shift(s, desc.fft.x, desc.fft.y) ? 1 : 0;
  return a;
}

// The knobs' slot of an execution: 1 for the fast-math schedules (FP32 fast math, FP16)
static int fr_slot(const umpa_fr_op *op, bool half)
{
  return (half || op->strategy == UMPA_FR_FAST_MATH) ? 1 : 0;
}

static umpa::FrJob fr_make_job(const umpa_fr_op *op, const void *fft, const float *scales,
                               uint64_t fftCount, void *volume, float *weights,
                               const umpa_traverse_space *spaces, uint64_t nSpaces)
{
  return fr_job_record(op->desc, op->settings, op->constants, fft, scales, fftCount, volume,
                       weights, spaces, nSpaces, op->params[fr_slot(op, scales != nullptr)].boxFixed);
}

static umpa::FrVariant fr_variant_for(const umpa_fr_settings &s, int nb, int strategy,
                                      bool half, bool count,
                                      const umpa::host::SchedParams &knobs)
{
  umpa::FrVariant v{};
  if (s.type == UMPA_FR_FAST) {
    v.mode = umpa::kModeNearest;
  } else if (s.interpolation == UMPA_FR_LOOKUP) {
    v.mode = umpa::kModeLookup;
  } else {
    // FRCPU picks the rational I0 approximation for alpha <= 15 (fr_cpu.cpp:31)
    v.mode = s.alpha <= 15.f ? umpa::kModeFastKaiser : umpa::kModePreciseKaiser;
  }
  v.order = s.order;
  v.nb = nb;
  v.half = half;
  // bit-exact runs keep every brick on one wavefront; fast-math runs may split central bricks
  v.fastMath = half || strategy == UMPA_FR_FAST_MATH;
  v.count = count;
  v.batched = false;
  v.kernel = knobs.kernel;
  v.tableGlobal = knobs.tableGlobal;
  return v;
}

static umpa::FrVariant fr_make_variant(const umpa_fr_op *op, bool half)
{
  return fr_variant_for(op->settings, op->nb, op->strategy, half, op->countUpdates,
                        op->params[fr_slot(op, half)]);
}

static int fr_execute_impl(umpa_fr_op *op, const void *fft, const float *scales, bool half,
                           uint64_t fftCount, void *volume, float *weights,
                           const umpa_traverse_space *spaces, uint64_t nSpaces,
                           const float *table)
{
  int st = fr_check_execute(op, fft, scales, half, fftCount, volume, weights, spaces, nSpaces,
                            table);
  if (st != UMPA_OK) return st;
  op->lastUpdates = 0;
  if (nSpaces == 0) return UMPA_OK;
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;

  const umpa::FrJob a = fr_make_job(op, fft, scales, fftCount, volume, weights, spaces, nSpaces);
  const umpa::FrVariant v = fr_make_variant(op, half);
  if (half && v.mode != umpa::kModeLookup) return UMPA_ERR_UNSUPPORTED;
  umpa::FrLaunch l{};
  l.items = v.fastMath ? op->dBricksSplit : op->dBricks;
  l.nItems = v.fastMath ? op->nBricksSplit : op->nBricks;
  l.counter = op->dCounter;
  l.updates = op->dUpdates;
  l.table = table;
  l.nTable = (int)op->desc.blobTable;
  hipError_t e = hipSuccess;
  if (v.count) e = hipMemsetAsync(op->dUpdates, 0, sizeof(unsigned long long), op->stream);
  if (e == hipSuccess) e = umpa::prepare_fr_insert(v, l.nTable);
  if (e == hipSuccess) e = umpa::launch_fr_insert(a, l, v, op->numCUs, op->stream);
  if (e == hipSuccess && v.count) {
    e = hipMemcpyAsync(&op->lastUpdates, op->dUpdates, sizeof(unsigned long long),
                       hipMemcpyDeviceToHost, op->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(op->stream);
  }
  return from_hip(e);
}

int umpa_fr_execute(umpa_fr_op *op, const void *fft, uint64_t fftCount, void *volume,
                    float *weights, const umpa_traverse_space *spaces, uint64_t nSpaces,
                    const float *table)
{
  return fr_execute_impl(op, fft, nullptr, false, fftCount, volume, weights, spaces, nSpaces,
                         table);
}

int umpa_fr_execute_noalloc(int device, void *stream, const umpa_fr_desc *desc,
                            const umpa_fr_settings *s, int strategy, const void *fft,
                            const float *scales, uint64_t fftCount, void *volume,
                            float *weights, const umpa_traverse_space *spaces,
                            uint64_t nSpaces, const float *table)
{
  if (!desc || !s || (strategy != UMPA_FR_EXACT && strategy != UMPA_FR_FAST_MATH))
    return UMPA_ERR_INVALID_ARG;
  int st = fr_check_init(*desc, *s);
  if (st != UMPA_OK) return st;
  const bool half = scales != nullptr;
  st = fr_check_payload(*desc, *s, fft, scales, half, fftCount, volume, weights, spaces,
                        nSpaces, table);
  if (st != UMPA_OK) return st;
  const umpa::FrVariant v =
    fr_variant_for(*s, fr_footprint(*s), strategy, half, false, umpa::host::SchedParams());
  if (half && v.mode != umpa::kModeLookup) return UMPA_ERR_UNSUPPORTED;
  try {
    // the schedule a handle would build at Init (default knobs): it must fit the static
    // scratch, which holds volumes up to 513^3
    const auto bricks = umpa::host::make_brick_schedule(
      (int)desc->volume.x, (int)desc->volume.y, (int)desc->volume.z, v.fastMath,
      umpa::host::SchedParams());
    if (bricks.size() > umpa::kNoAllocMaxItems) return UMPA_ERR_UNSUPPORTED;
    if (nSpaces == 0) return UMPA_OK;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count ||
        device >= umpa::kNoAllocMaxDevices)
      return UMPA_ERR_DEVICE;
    DeviceGuard g(device);
    if (!g.ok) return UMPA_ERR_DEVICE;
    std::vector<int> items(bricks.size() * 4);
    for (size_t i = 0; i < bricks.size(); ++i) {
      items[4 * i + 0] = bricks[i].x;
      items[4 * i + 1] = bricks[i].y;
      items[4 * i + 2] = bricks[i].z;
      items[4 * i + 3] = bricks[i].w;
    }
    const umpa_fr_constants c =
      umpa::host::create_constants(desc->volume.x, desc->volume.y, desc->blobTable, *s);
    const umpa::FrJob a =
      fr_job_record(*desc, *s, c, fft, scales, fftCount, volume, weights, spaces, nSpaces, false);
    return from_hip(umpa::launch_fr_insert_noalloc(
      device, a, v, items.data(), (int)bricks.size(), v.fastMath, table,
      (int)desc->blobTable, query_cus(device), static_cast<hipStream_t>(stream)));
  } catch (const std::bad_alloc &) {
    return UMPA_ERR_OUT_OF_MEMORY;
  }
}

int umpa_fr_execute_f16(umpa_fr_op *op, const void *fft, const float *scales,
                        uint64_t fftCount, void *volume, float *weights,
                        const umpa_traverse_space *spaces, uint64_t nSpaces,
                        const float *table)
{
  return fr_execute_impl(op, fft, scales, true, fftCount, volume, weights, spaces, nSpaces,
                         table);
}

static int fr_autotune_impl(umpa_fr_op *op, const void *fft, const float *scales,
                            uint64_t fftCount, const umpa_traverse_space *spaces,
                            uint64_t nSpaces, const float *table, int useCache,
                            umpa_fr_sched_params *chosen)
{
  const bool half = scales != nullptr;
  // the checks of execute, with a placeholder for the output (scratch volumes are used)
  alignas(8) static const float kAligned[2] = { 0.f, 0.f };
  int st = fr_check_execute(op, fft, scales, half, fftCount, kAligned, kAligned, spaces, nSpaces,
                            table);
  if (st != UMPA_OK) return st;
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  umpa::FrVariant v = fr_make_variant(op, half);
  if (half && v.mode != umpa::kModeLookup) return UMPA_ERR_UNSUPPORTED;
  v.count = false;
  const int slot = v.fastMath ? 1 : 0;
  const int vx = (int)op->desc.volume.x, vy = (int)op->desc.volume.y, vz = (int)op->desc.volume.z;
  // spaces bucketed to a power of two
  uint64_t bucket = 1;
  while (bucket < nSpaces) bucket <<= 1;
  // "k2": the format of lines with kernel knobs
  const std::string key = "k2:op:cu" + std::to_string(op->numCUs) + ":m" +
                          std::to_string(v.mode) + ":f" + std::to_string(slot) + ":h" +
                          std::to_string(half ? 1 : 0) + ":nb" + std::to_string(v.nb) + ":" +
                          std::to_string(vx) + "x" + std::to_string(bucket);
  umpa::host::SchedParams best = op->params[slot];
  bool found = useCache && umpa::host::tuning_cache_lookup(key, best);
  hipError_t e = hipSuccess;
  double bestMs = 1e300;
  if (!found && nSpaces > 0) {
    const size_t n = (size_t)vx * vy * vz;
    void *vol = nullptr, *wts = nullptr;
    int *items = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    e = hipStreamSynchronize(op->stream);
    if (e == hipSuccess) e = hipMalloc(&vol, n * 8);
    if (e == hipSuccess) e = hipMalloc(&wts, n * 4);
    if (e == hipSuccess) e = hipMemsetAsync(vol, 0, n * 8, op->stream);
    if (e == hipSuccess) e = hipMemsetAsync(wts, 0, n * 4, op->stream);
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    // one candidate: its schedule, kernel and box, the best of two runs after a warm-up
    auto time_candidate = [&](const umpa::host::SchedParams &cand) {
      umpa::FrVariant vc = v;
      vc.kernel = cand.kernel;
      vc.tableGlobal = cand.tableGlobal;
      const umpa::FrJob a =
        fr_job_record(op->desc, op->settings, op->constants, fft, scales, fftCount, vol,
                      static_cast<float *>(wts), spaces, nSpaces, cand.boxFixed);
      const auto bricks = umpa::host::make_brick_schedule(vx, vy, vz, v.fastMath, cand);
      if (items) (void)hipFree(items);
      items = nullptr;
      e = upload_bricks(bricks, &items);
      if (e == hipSuccess) e = umpa::prepare_fr_insert(vc, (int)op->desc.blobTable);
      umpa::FrLaunch l{};
      l.items = items;
      l.nItems = (int)bricks.size();
      l.counter = op->dCounter;
      l.updates = op->dUpdates;
      l.table = table;
      l.nTable = (int)op->desc.blobTable;
      double ms = 1e300;
      for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {
        e = hipEventRecord(t0, op->stream);
        if (e == hipSuccess) e = umpa::launch_fr_insert(a, l, vc, op->numCUs, op->stream);
        if (e == hipSuccess) e = hipEventRecord(t1, op->stream);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, t0, t1);
        if (rep > 0) ms = std::min(ms, (double)t);  // the first run warms up
      }
      if (e == hipSuccess && ms < bestMs) {
        bestMs = ms;
        best = cand;
      }
    };
    try {
      // schedules (exact: with every kernel knob), then the kernel knobs over the best one
      for (const auto &cand : umpa::host::tuning_candidates(v.fastMath))
        if (e == hipSuccess) time_candidate(cand);
      const umpa::host::SchedParams sched = best;
      for (const auto &cand : umpa::host::kernel_candidates(sched, v.fastMath))
        if (e == hipSuccess) time_candidate(cand);
    } catch (const std::bad_alloc &) {
      e = hipErrorOutOfMemory;
    }
    (void)hipStreamSynchronize(op->stream);
    if (items) (void)hipFree(items);
    if (vol) (void)hipFree(vol);
    if (wts) (void)hipFree(wts);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (e != hipSuccess) return from_hip(e);
    umpa::host::tuning_cache_store(key, best, bestMs);
  }
  // rebuild the schedule of this strategy with the chosen knobs
  try {
    const auto bricks = umpa::host::make_brick_schedule(vx, vy, vz, v.fastMath, best);
    int *items = nullptr;
    e = upload_bricks(bricks, &items);
    if (e != hipSuccess) return from_hip(e);
    int *&dst = v.fastMath ? op->dBricksSplit : op->dBricks;
    // executes already enqueued on the stream may still read the old schedule
    e = hipStreamSynchronize(op->stream);
    if (e != hipSuccess) {
      (void)hipFree(items);
      return from_hip(e);
    }
    (void)hipFree(dst);
    dst = items;
    (v.fastMath ? op->nBricksSplit : op->nBricks) = (int)bricks.size();
  } catch (const std::bad_alloc &) {
    return UMPA_ERR_OUT_OF_MEMORY;
  }
  op->params[slot] = best;
  if (chosen) *chosen = umpa::host::to_c(best);
  return UMPA_OK;
}

int umpa_fr_autotune(umpa_fr_op *op, const void *fft, const float *scales, uint64_t fftCount,
                     const umpa_traverse_space *spaces, uint64_t nSpaces, const float *table,
                     int useCache, umpa_fr_sched_params *chosen)
{
  // the cache key and the tuning file are handled with std::string and streams: nothing they
  // throw crosses the C ABI (device buffers are released inside before any of them runs)
  try {
    return fr_autotune_impl(op, fft, scales, fftCount, spaces, nSpaces, table, useCache, chosen);
  } catch (const std::bad_alloc &) {
    return UMPA_ERR_OUT_OF_MEMORY;
  } catch (...) {
    return UMPA_ERR_DEVICE;
  }
}

int umpa_fr_get_sched_params(umpa_fr_op *op, umpa_fr_sched_params *out)
{
  if (!op || !out) return UMPA_ERR_INVALID_ARG;
  const int slot = op->strategy == UMPA_FR_FAST_MATH ? 1 : 0;
  *out = umpa::host::to_c(op->params[slot]);
  return UMPA_OK;
}

int umpa_fr_set_sched_params(umpa_fr_op *op, int strategy, const umpa_fr_sched_params *p)
{
  if (!op || !p || (strategy != UMPA_FR_EXACT && strategy != UMPA_FR_FAST_MATH) ||
      !umpa::host::sched_params_valid(*p))
    return UMPA_ERR_INVALID_ARG;
  const int slot = strategy == UMPA_FR_FAST_MATH ? 1 : 0;
  const umpa::host::SchedParams sp = umpa::host::from_c(*p);
  if (!op->initialized) {
    op->params[slot] = sp;  // used by the next Init
    return UMPA_OK;
  }
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  try {
    const auto bricks = umpa::host::make_brick_schedule(
      (int)op->desc.volume.x, (int)op->desc.volume.y, (int)op->desc.volume.z, slot == 1, sp);
    int *items = nullptr;
    hipError_t e = upload_bricks(bricks, &items);
    if (e != hipSuccess) return from_hip(e);
    // executes already enqueued on the stream may still read the old schedule
    e = hipStreamSynchronize(op->stream);
    if (e != hipSuccess) {
      (void)hipFree(items);
      return from_hip(e);
    }
    int *&dst = slot == 1 ? op->dBricksSplit : op->dBricks;
    (void)hipFree(dst);
    dst = items;
    (slot == 1 ? op->nBricksSplit : op->nBricks) = (int)bricks.size();
    op->params[slot] = sp;
    return UMPA_OK;
  } catch (const std::bad_alloc &) {
    return UMPA_ERR_OUT_OF_MEMORY;
  }
}

int umpa_fr_tuning_candidates(int strategy, umpa_fr_sched_params *out, uint64_t capacity,
                              uint64_t *count)
{
  if (!count || (strategy != UMPA_FR_EXACT && strategy != UMPA_FR_FAST_MATH))
    return UMPA_ERR_INVALID_ARG;
  try {
    const bool fast = strategy == UMPA_FR_FAST_MATH;
    auto all = umpa::host::tuning_candidates(fast);
    const auto second = umpa::host::kernel_candidates(umpa::host::SchedParams(), fast);
    all.insert(all.end(), second.begin(), second.end());
    *count = all.size();
    if (!out) return UMPA_OK;
    if (capacity < all.size()) return UMPA_ERR_INVALID_ARG;
    for (size_t i = 0; i < all.size(); ++i) out[i] = umpa::host::to_c(all[i]);
    return UMPA_OK;
  } catch (const std::bad_alloc &) {
    return UMPA_ERR_OUT_OF_MEMORY;
  }
}

int umpa_fr_brick_schedule(uint32_t vx, uint32_t vy, uint32_t vz, int splitHeavy,
                           const umpa_fr_sched_params *p, int32_t *items, uint64_t capacity,
                           uint64_t *count)
{
  if (!count || vx == 0 || vy == 0 || vz == 0 || vx > 65535 || vy > 65535 || vz > 65535)
    return UMPA_ERR_INVALID_ARG;
  umpa::host::SchedParams sp;
  if (p) {
    if (!umpa::host::sched_params_valid(*p)) return UMPA_ERR_INVALID_ARG;
    sp = umpa::host::from_c(*p);
  }
  try {
    const auto bricks =
      umpa::host::make_brick_schedule((int)vx, (int)vy, (int)vz, splitHeavy != 0, sp);
    *count = bricks.size();
    if (!items) return UMPA_OK;
    if (capacity < bricks.size()) return UMPA_ERR_INVALID_ARG;
    for (size_t i = 0; i < bricks.size(); ++i) {
      items[4 * i + 0] = bricks[i].x;
      items[4 * i + 1] = bricks[i].y;
      items[4 * i + 2] = bricks[i].z;
      items[4 * i + 3] = bricks[i].w;
    }
    return UMPA_OK;
  } catch (const std::bad_alloc &) {
    return UMPA_ERR_OUT_OF_MEMORY;
  }
}

int umpa_fr_set_strategy(umpa_fr_op *op, int strategy)
{
  if (!op || (strategy != UMPA_FR_EXACT && strategy != UMPA_FR_FAST_MATH))
    return UMPA_ERR_INVALID_ARG;
  op->strategy = strategy;
  return UMPA_OK;
}

int umpa_fr_set_count_updates(umpa_fr_op *op, int enable)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  op->countUpdates = enable != 0;
  return UMPA_OK;
}

int umpa_fr_last_updates(umpa_fr_op *op, uint64_t *updates)
{
  if (!op || !updates) return UMPA_ERR_INVALID_ARG;
  *updates = op->lastUpdates;
  return UMPA_OK;
}

int umpa_fr_synchronize(umpa_fr_op *op)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  return from_hip(hipStreamSynchronize(op->stream));
}

int umpa_fr_cleanup(umpa_fr_op *op)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(op->device);
  if (op->initialized) (void)hipStreamSynchronize(op->stream);
  op->release();
  return UMPA_OK;
}

void umpa_fr_destroy(umpa_fr_op *op)
{
  if (!op) return;
  umpa_fr_cleanup(op);
  delete op;
}

int umpa_fr_create_constants(const umpa_fr_desc *desc, const umpa_fr_settings *s,
                             umpa_fr_constants *out)
{
  if (!desc || !s || !out) return UMPA_ERR_INVALID_ARG;
  *out = umpa::host::create_constants(desc->volume.x, desc->volume.y, desc->blobTable, *s);
  return UMPA_OK;
}

int umpa_fr_fill_blob_table(float *table, uint64_t n, const umpa_fr_desc *desc,
                            const umpa_fr_settings *s)
{
  if (!table || !desc || !s || n == 0) return UMPA_ERR_INVALID_ARG;
  umpa::host::fill_blob_table(table, n, desc->volume.x, desc->volume.y, *s);
  return UMPA_OK;
}

int umpa_fr_compute_traverse_space(uint32_t imgSizeX, uint32_t imgSizeY,
                                   const float transform[9], umpa_traverse_space *space,
                                   uint32_t maxVolumeIndexX, uint32_t maxVolumeIndexYZ,
                                   int useFast, float blobRadius, float weight)
{
  if (!transform || !space) return UMPA_ERR_INVALID_ARG;
  umpa::host::compute_traverse_space(imgSizeX, imgSizeY, transform, *space, maxVolumeIndexX,
                                     maxVolumeIndexYZ, useFast != 0, blobRadius, weight);
  return UMPA_OK;
}

int umpa_fr_generate_traverse_spaces(const double *rotations, uint64_t n,
                                     const double *symmetries, uint32_t nSym,
                                     uint32_t paddedBox, const umpa_fr_settings *s, float weight,
                                     int32_t firstIndex, umpa_traverse_space *out)
{
  if ((!rotations && n) || !s || (!out && n) || paddedBox < 2) return UMPA_ERR_INVALID_ARG;
  if (symmetries && nSym == 0) return UMPA_ERR_INVALID_ARG;
  const double identity[9] = { 1, 0, 0, 0, 1, 0, 0, 0, 1 };
  const double *sym = symmetries ? symmetries : identity;
  const uint32_t k = symmetries ? nSym : 1;
  for (uint64_t i = 0; i < n; ++i) {
    const double *r = rotations + 9 * i;
    for (uint32_t q = 0; q < k; ++q) {
      const double *o = sym + 9 * q;
      // transform = R_i S_q in FP64, rounded once to float
      float t[9];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          t[3 * a + b] = (float)(r[3 * a] * o[b] + r[3 * a + 1] * o[3 + b] + r[3 * a + 2] * o[6 + b]);
      umpa_traverse_space &rec = out[i * k + q];
      umpa::host::compute_traverse_space(paddedBox / 2, paddedBox, t, rec, paddedBox, paddedBox,
                                         s->type == UMPA_FR_FAST, s->blobRadius, weight);
      rec.projectionIndex = firstIndex + (int32_t)i;
    }
  }
  return UMPA_OK;
}

int umpa_fr_symmetry_operators(const char *group, double *out, uint32_t capacity,
                               uint32_t *count)
{
  if (!group || !count) return UMPA_ERR_INVALID_ARG;
  const char f = group[0] == 'c' ? 'C' : group[0] == 'd' ? 'D' : group[0];
  char *end = nullptr;
  const long n = std::strtol(group + 1, &end, 10);
  if ((f != 'C' && f != 'D') || !end || *end != '\0' || end == group + 1 || n < 1 || n > 1000)
    return UMPA_ERR_INVALID_ARG;
  return umpa::host::symmetry_operators(f, (int)n, out, capacity, count) == 0
           ? UMPA_OK : UMPA_ERR_INVALID_ARG;
}

// ---- finalization -----------------------------------------------------------------------
int umpa_fr_finalize(const void *volume, const float *weights, uint64_t side, void *out,
                     int hermitian, void *stream)
{
  if (!volume || !weights || !out || side < 1 || side % 2 == 0 || side > 65535)
    return UMPA_ERR_INVALID_ARG;
  if (!aligned(volume, 8) || !aligned(weights, 4) || !aligned(out, 8))
    return UMPA_ERR_INVALID_ARG;
  return from_hip(umpa::launch_fr_finalize(volume, weights, (int)side, out, hermitian != 0,
                                           static_cast<hipStream_t>(stream)));
}

// ---- FP ---------------------------------------------------------------------------------
int umpa_fp_create(umpa_fp_op **op, int device, void *stream)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  *op = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return UMPA_ERR_DEVICE;
  auto *h = new (std::nothrow) umpa_fp_op();
  if (!h) return UMPA_ERR_OUT_OF_MEMORY;
  h->device = device;
  h->stream = static_cast<hipStream_t>(stream);
  h->numCUs = query_cus(device);
  *op = h;
  return UMPA_OK;
}

int umpa_fp_init(umpa_fp_op *op, const umpa_fp_desc *desc, const umpa_fp_settings *s)
{
  if (!op || !desc || !s) return UMPA_ERR_INVALID_ARG;
  op->initialized = false;
  // AFP::IsValid (afp.hpp:45-52) + Settings::SetMaxFreq's range (settings.hpp:38-43)
  if (!size_valid(desc->in) || !size_valid(desc->out) || !size_valid(desc->inSpatial))
    return UMPA_ERR_INVALID_ARG;
  if (desc->out.n < desc->in.n) return UMPA_ERR_INVALID_ARG;
  if (desc->out.x > desc->in.x || desc->out.y > desc->in.y) return UMPA_ERR_UNSUPPORTED;
  if (s->hasMaxFreq && !(s->maxFreq >= 0.f && s->maxFreq <= 0.5f)) return UMPA_ERR_INVALID_ARG;
  if (desc->inPaddedSpatialSingle == 0) return UMPA_ERR_INVALID_ARG;
  op->desc = *desc;
  op->settings = *s;
  op->initialized = true;
  return UMPA_OK;
}

int umpa_fp_execute(umpa_fp_op *op, const void *in, uint64_t inCount, void *out,
                    uint64_t outCount, const float *filter)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  if (!op->initialized) return UMPA_ERR_NOT_INITIALIZED;
  if (inCount > op->desc.in.n || outCount > op->desc.out.n) return UMPA_ERR_NOT_EQUIVALENT;
  // FPCPU Strategy1::Execute rejects empty payloads (fp_cpu.cpp:27-29)
  if (!in || !out || inCount == 0 || outCount == 0) return UMPA_ERR_REJECTED;
  if (outCount < inCount) return UMPA_ERR_INVALID_ARG;
  if (op->settings.applyFilter && !filter) return UMPA_ERR_INVALID_ARG;
  if (!aligned(in, 8) || !aligned(out, 8) || (filter && !aligned(filter, 4)))
    return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  umpa::FpArgs a{};
  a.in = in;
  a.out = out;
  a.filter = filter;
  a.inX = op->desc.in.x;
  a.inY = op->desc.in.y;
  a.inSingle = op->desc.in.x * op->desc.in.y * op->desc.in.z;
  a.spatialX = op->desc.inSpatial.x;
  a.spatialY = op->desc.inSpatial.y;
  a.outX = op->desc.out.x;
  a.outY = op->desc.out.y;
  a.outSingle = op->desc.out.x * op->desc.out.y * op->desc.out.z;
  a.n = inCount;
  a.normFactor = 1.f / static_cast<float>(op->desc.inPaddedSpatialSingle);
  a.maxFreqSquare = op->settings.hasMaxFreq ? op->settings.maxFreq : 0.f;
  a.applyFilter = op->settings.applyFilter != 0;
  a.normalize = op->settings.normalize != 0;
  a.center = op->settings.center != 0;
  a.cropFreq = op->settings.hasMaxFreq != 0;
  a.shift = op->settings.shift != 0;
  return from_hip(umpa::launch_fp_scale(a, op->numCUs, op->stream));
}

int umpa_fp_synchronize(umpa_fp_op *op)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  return from_hip(hipStreamSynchronize(op->stream));
}

int umpa_fp_cleanup(umpa_fp_op *op)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  op->initialized = false;
  return UMPA_OK;
}

void umpa_fp_destroy(umpa_fp_op *op) { delete op; }

}  // extern "C"

// Batched FR: several reconstructions inserted by one persistent launch per step.
//
// A step inserts a range of traverse spaces of every job it names.  Jobs that share a kernel
// variant (settings, blob table, slice precision) go into one launch whose work items are the
// bricks of all of them, heaviest first, so small and large volumes fill the GPU together.
// Voxels still receive each job's spaces in index order, so a job's result does not depend on
// the other jobs of the batch.
//
//   umpa_fr_batch_plan + execute: a static chunk loop (step c = spaces [c*C, (c+1)*C) of every
//     job), all device-side state built once, the loop captured as one HIP graph;
//   umpa_fr_batch_step: one step over explicit ranges, for streaming callers; job records and
//     work items travel through a ring of pinned staging slots recycled in stream order.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <string>
#include <tuple>
#include <utility>
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

// A validated job: its kernel variant and the record fields that do not depend on the step.
struct JobInfo {
  umpa::FrVariant v;
  umpa::FrJob rec;
  const float *table;
  int nTable;
};

// Checks one job like umpa_fr_init + umpa_fr_execute would and derives its kernel variant.
int check_job(const umpa_fr_job &j, int strategy, JobInfo &out)
{
  const umpa_fr_desc &d = j.desc;
  const umpa_fr_settings &s = j.settings;
  if (!size_valid(d.volume) || !size_valid(d.fft)) return UMPA_ERR_INVALID_ARG;
  if (d.volume.x % 2 == 0 || d.volume.y != d.volume.z || d.volume.n != 1)
    return UMPA_ERR_INVALID_ARG;
  if (s.interpolation < 0 || s.interpolation > 1 || s.type < 0 || s.type > 1 || s.order < 0 ||
      s.order > 4 || !(s.blobRadius > 0.f) || !std::isfinite(s.alpha))
    return UMPA_ERR_INVALID_ARG;
  // the work item packs z in 16 bits
  if (d.volume.x > 65535 || d.volume.y > 65535 || d.fft.x > (uint64_t)INT_MAX ||
      d.fft.y > (uint64_t)INT_MAX || d.fft.n > (uint64_t)INT_MAX ||
      d.traverseSpaces > (uint64_t)INT_MAX || d.blobTable > (uint64_t)INT_MAX)
    return UMPA_ERR_UNSUPPORTED;
  // a slice is read through a buffer resource with a 32-bit byte range
  if (d.fft.x * d.fft.y * 8 > (uint64_t)INT_MAX) return UMPA_ERR_UNSUPPORTED;
  const bool half = j.scales != nullptr;
  if (!j.fft || !j.volume || !j.weights) return UMPA_ERR_REJECTED;
  if (d.traverseSpaces > 0 && !j.spaces) return UMPA_ERR_INVALID_ARG;
  if (s.interpolation == UMPA_FR_LOOKUP && (!j.table || d.blobTable == 0)) return UMPA_ERR_REJECTED;
  if (!aligned(j.volume, 8) || !aligned(j.weights, 4) || !aligned(j.fft, half ? 4 : 8) ||
      !aligned(j.spaces, 4) || (j.table && !aligned(j.table, 4)) ||
      (j.scales && !aligned(j.scales, 4)))
    return UMPA_ERR_INVALID_ARG;
  if (strategy != UMPA_FR_EXACT && strategy != UMPA_FR_FAST_MATH) return UMPA_ERR_INVALID_ARG;

  umpa::FrVariant v{};
  const bool nearest = s.type == UMPA_FR_FAST;
  if (nearest) {
    v.mode = umpa::kModeNearest;
  } else if (s.interpolation == UMPA_FR_LOOKUP) {
    v.mode = umpa::kModeLookup;
  } else {
    v.mode = s.alpha <= 15.f ? umpa::kModeFastKaiser : umpa::kModePreciseKaiser;
  }
  if (half && v.mode != umpa::kModeLookup) return UMPA_ERR_UNSUPPORTED;
  if (v.mode == umpa::kModeLookup && d.blobTable > (uint64_t)umpa::kFrMaxLdsTable)
    return UMPA_ERR_UNSUPPORTED;
  v.order = s.order;
  v.nb = umpa::host::footprint(s);
  v.half = half;
  v.fastMath = half || strategy == UMPA_FR_FAST_MATH;
  v.count = false;
  v.batched = true;

  umpa::FrJob a{};
  a.volume = j.volume;
  a.weights = j.weights;
  a.fft = j.fft;
  a.scales = j.scales;
  a.spaces = j.spaces;
  a.xSize = (int)d.fft.x;
  a.ySize = (int)d.fft.y;
  a.sliceElems = (int64_t)d.fft.x * (int64_t)d.fft.y;
  a.nProj = (int)d.fft.n;
  a.nSpaces = (int)d.traverseSpaces;
  a.vx = (int)d.volume.x;
  a.vy = (int)d.volume.y;
  a.vz = (int)d.volume.z;
  a.c = umpa::host::create_constants(d.volume.x, d.volume.y, d.blobTable, s);
  a.cullHalfWidth = nearest ? 1.6f : s.blobRadius;
  a.cullHalfSpace = nearest ? 0.f : s.blobRadius;
  a.shiftBox = umpa::host::box_can_shift(s, d.fft.x, d.fft.y) ? 1 : 0;
  out.v = v;
  out.rec = a;
  out.table = v.mode == umpa::kModeLookup ? j.table : nullptr;
  out.nTable = v.mode == umpa::kModeLookup ? (int)d.blobTable : 0;
  return UMPA_OK;
}

// launches of one variant share a kernel and a blob table
using GroupKey = std::tuple<int, int, int, bool, bool, const float *, int>;
GroupKey group_key(const JobInfo &j)
{
  return GroupKey{ j.v.mode, j.v.order, j.v.nb, j.v.half, j.v.fastMath, j.table, j.nTable };
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Host image of one launch: its job records and work items, serialised for one upload.
struct LaunchImage {
  umpa::FrVariant v;
  const float *table;
  int nTable;
  size_t jobsOff, itemsOff, counterOff;
  int nItems;
};

}  // namespace

struct umpa_fr_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  bool ownStream = false;
  int numCUs = 256;
  bool count = false;
  unsigned long long lastUpdates = 0;
  unsigned long long *dUpdates = nullptr;
  std::map<std::tuple<int, int, int, float>, umpa::host::BrickSet> cands;
  std::vector<umpa::host::BrickOrigin> schedBuf;  // make_schedule's output, reused
  umpa::host::SchedParams params[2];  // schedule knobs for exact / fast-math launches
  int waveItems = 0;  // items per wave before it leaves (0: persistent workgroups)
  int maxWorkgroups = 0;  // persistent launches: at most this many workgroups (0: one per CU)

  // static plan
  std::vector<std::vector<LaunchImage>> steps;
  char *dPlan = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  // streaming steps: pinned staging slots
  struct Slot {
    char *host = nullptr;
    char *dev = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
  };
  Slot ring[4];
  int next = 0;

  std::deque<std::pair<uint64_t, hipEvent_t>> tickets;
  uint64_t nextTicket = 1;

  const umpa::host::BrickSet &candidates(int vx, int vy, int vz, float depth)
  {
    auto key = std::make_tuple(vx, vy, vz, depth);
    auto it = cands.find(key);
    if (it == cands.end())
      it = cands.emplace(key, umpa::host::make_brick_set(vx, vy, vz, depth)).first;
    return it->second;
  }

  void drop_plan()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (dPlan) (void)hipFree(dPlan);
    exec = nullptr;
    graph = nullptr;
    dPlan = nullptr;
    steps.clear();
  }
};

namespace {

// Builds the launches of one step: `ranges` = (job index into `infos`, begin, end).  Appends
// job records, work items and counters to `blob` (offsets relative to its start).
void build_step(umpa_fr_batch *b, const std::vector<JobInfo> &infos,
                const std::vector<std::tuple<int, int, int>> &ranges, std::vector<char> &blob,
                std::vector<LaunchImage> &out, const umpa::host::SchedParams *force = nullptr)
{
  std::map<GroupKey, std::vector<std::tuple<int, int, int>>> groups;
  for (const auto &r : ranges)
    if (std::get<2>(r) > std::get<1>(r)) groups[group_key(infos[std::get<0>(r)])].push_back(r);
  for (auto &g : groups) {
    const auto &members = g.second;
    const JobInfo &first = infos[std::get<0>(members[0])];
    const umpa::host::SchedParams &sp = force ? *force : b->params[first.v.fastMath ? 1 : 0];
    std::vector<umpa::FrJob> recs;
    std::vector<umpa::host::JobWork> work;
    for (size_t k = 0; k < members.size(); ++k) {
      const auto &[ji, s0, s1] = members[k];
      umpa::FrJob rec = infos[ji].rec;
      rec.spaceBegin = s0;
      rec.spaceEnd = s1;
      if (sp.boxFixed) rec.shiftBox = 0;
      recs.push_back(rec);
      work.push_back({ &b->candidates(rec.vx, rec.vy, rec.vz, sp.depthScale), (int)k,
                       (double)(s1 - s0) });
    }
    umpa::host::make_schedule(work, first.v.fastMath, sp, b->schedBuf);
    const auto &items = b->schedBuf;
    LaunchImage li{};
    li.v = first.v;
    // the carried-queue kernel reads a job's slices through one buffer resource: it runs only
    // when the knob names it and every job's slices fit one
    bool carryFits = true;
    for (const umpa::FrJob &rec : recs)
      carryFits = carryFits && (uint64_t)rec.nProj * (uint64_t)rec.sliceElems *
                                   (li.v.half ? 4u : 8u) <= 0x7fff0000ull;
    li.v.kernel = sp.kernel == UMPA_FR_KERNEL_CARRY && carryFits ? UMPA_FR_KERNEL_CARRY
                                                                 : UMPA_FR_KERNEL_PAIR;
    // the sliding-box instantiation only when every job of the launch slides
    li.v.shift = std::all_of(recs.begin(), recs.end(),
                             [](const umpa::FrJob &r) { return r.shiftBox != 0; });
    li.table = first.table;
    li.nTable = first.nTable;
    li.nItems = (int)items.size();
    li.jobsOff = align_up(blob.size(), 256);
    blob.resize(li.jobsOff + recs.size() * sizeof(umpa::FrJob));
    std::memcpy(blob.data() + li.jobsOff, recs.data(), recs.size() * sizeof(umpa::FrJob));
    li.itemsOff = align_up(blob.size(), 256);
    blob.resize(li.itemsOff + items.size() * 16);
    int *dst = reinterpret_cast<int *>(blob.data() + li.itemsOff);
    for (size_t i = 0; i < items.size(); ++i) {
      dst[4 * i + 0] = items[i].x;
      dst[4 * i + 1] = items[i].y;
      dst[4 * i + 2] = items[i].z;
      dst[4 * i + 3] = items[i].w;
    }
    li.counterOff = align_up(blob.size(), 256);
    blob.resize(li.counterOff + 256, 0);
    out.push_back(li);
  }
}

// Enqueues one launch.  Kernel attributes are set beforehand (umpa_fr_batch_plan) when the
// launch is being captured into a graph.
hipError_t enqueue(umpa_fr_batch *b, const LaunchImage &li, char *base, bool count,
                   bool capturing = false)
{
  umpa::FrLaunch l{};
  l.jobs = reinterpret_cast<const umpa::FrJob *>(base + li.jobsOff);
  l.items = reinterpret_cast<const int *>(base + li.itemsOff);
  l.nItems = li.nItems;
  l.counter = reinterpret_cast<int *>(base + li.counterOff);
  l.updates = b->dUpdates;
  l.table = li.table;
  l.nTable = li.nTable;
  l.waveItems = b->waveItems;
  l.maxBlocks = b->maxWorkgroups;
  umpa::FrVariant v = li.v;
  v.count = count;
  if (!capturing) {
    const hipError_t e = umpa::prepare_fr_insert(v, li.nTable);
    if (e != hipSuccess) return e;
  }
  umpa::FrJob unused{};
  return umpa::launch_fr_insert(unused, l, v, b->numCUs, b->stream);
}

// The volumes and weights the jobs accumulate into must be pairwise disjoint: a wavefront
// loads, sums and stores its brick without atomics, so two work items of one launch writing
// the same voxel would lose updates.
bool outputs_disjoint(const std::vector<JobInfo> &infos)
{
  std::vector<std::pair<uintptr_t, uintptr_t>> ranges;  // [begin, end) in bytes
  ranges.reserve(2 * infos.size());
  for (const JobInfo &j : infos) {
    const uint64_t n = (uint64_t)j.rec.vx * (uint64_t)j.rec.vy * (uint64_t)j.rec.vz;
    const uintptr_t v = reinterpret_cast<uintptr_t>(j.rec.volume);
    const uintptr_t w = reinterpret_cast<uintptr_t>(j.rec.weights);
    ranges.emplace_back(v, v + n * 8);
    ranges.emplace_back(w, w + n * 4);
  }
  std::sort(ranges.begin(), ranges.end());
  for (size_t i = 1; i < ranges.size(); ++i)
    if (ranges[i].first < ranges[i - 1].second) return false;
  return true;
}

int validate_all(const umpa_fr_job *jobs, uint32_t nJobs, int strategy, std::vector<JobInfo> &infos)
{
  if (!jobs || nJobs == 0 || nJobs > 32767) return UMPA_ERR_INVALID_ARG;
  infos.resize(nJobs);
  for (uint32_t i = 0; i < nJobs; ++i) {
    const int st = check_job(jobs[i], strategy, infos[i]);
    if (st != UMPA_OK) return st;
  }
  if (!outputs_disjoint(infos)) return UMPA_ERR_INVALID_ARG;
  return UMPA_OK;
}

}  // namespace

extern "C" {

int umpa_fr_batch_create(umpa_fr_batch **out, int device, void *stream)
{
  if (!out) return UMPA_ERR_INVALID_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return UMPA_ERR_DEVICE;
  DeviceGuard g(device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  auto *b = new (std::nothrow) umpa_fr_batch();
  if (!b) return UMPA_ERR_OUT_OF_MEMORY;
  b->device = device;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      cus > 0)
    b->numCUs = cus;
  hipError_t e = hipSuccess;
  if (stream) {
    b->stream = static_cast<hipStream_t>(stream);
  } else {
    e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    b->ownStream = e == hipSuccess;
  }
  if (e == hipSuccess) e = hipMalloc(&b->dUpdates, sizeof(unsigned long long));
  for (auto &s : b->ring)
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
  if (e != hipSuccess) {
    umpa_fr_batch_destroy(b);
    return from_hip(e);
  }
  *out = b;
  return UMPA_OK;
}

static int batch_plan_impl(umpa_fr_batch *b, const umpa_fr_job *jobs, uint32_t nJobs,
                           uint32_t chunkSpaces, int strategy, int useGraph)
{
  if (!b || chunkSpaces == 0) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(b->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  // the previous plan's launches may still be running
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return from_hip(e);
  b->drop_plan();
  std::vector<JobInfo> infos;
  int st = validate_all(jobs, nJobs, strategy, infos);
  if (st != UMPA_OK) return st;
  try {
    uint64_t maxSpaces = 0;
    for (const auto &j : infos) maxSpaces = std::max<uint64_t>(maxSpaces, (uint64_t)j.rec.nSpaces);
    const uint64_t nSteps = (maxSpaces + chunkSpaces - 1) / chunkSpaces;
    std::vector<char> blob;
    for (uint64_t c = 0; c < nSteps; ++c) {
      std::vector<std::tuple<int, int, int>> ranges;
      for (uint32_t i = 0; i < nJobs; ++i) {
        const uint64_t s0 = std::min<uint64_t>(infos[i].rec.nSpaces, c * chunkSpaces);
        const uint64_t s1 = std::min<uint64_t>(infos[i].rec.nSpaces, s0 + chunkSpaces);
        if (s1 > s0) ranges.emplace_back((int)i, (int)s0, (int)s1);
      }
      std::vector<LaunchImage> launches;
      build_step(b, infos, ranges, blob, launches);
      b->steps.push_back(std::move(launches));
    }
    e = hipMalloc(&b->dPlan, std::max<size_t>(blob.size(), 256));
    if (e == hipSuccess) e = hipMemcpy(b->dPlan, blob.data(), blob.size(), hipMemcpyHostToDevice);
    for (const auto &step : b->steps)
      for (const auto &li : step)
        if (e == hipSuccess) e = umpa::prepare_fr_insert(li.v, li.nTable);
    if (e == hipSuccess && useGraph) {
      // the chunk loop as one graph: a counter-reset kernel and the persistent launch per
      // (step, group), in stream order
      e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
      hipError_t le = hipSuccess;
      if (e == hipSuccess) {
        for (const auto &step : b->steps)
          for (const auto &li : step)
            if (le == hipSuccess) le = enqueue(b, li, b->dPlan, false, true);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(b->stream, &graph);
        b->graph = graph;
        if (e == hipSuccess) e = le;
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    }
    if (e != hipSuccess) {
      b->drop_plan();
      return from_hip(e);
    }
    return UMPA_OK;
  } catch (const std::bad_alloc &) {
    b->drop_plan();
    return UMPA_ERR_OUT_OF_MEMORY;
  }
}

int umpa_fr_batch_plan(umpa_fr_batch *b, const umpa_fr_job *jobs, uint32_t nJobs,
                       uint32_t chunkSpaces, int strategy, int useGraph)
{
  try {
    return batch_plan_impl(b, jobs, nJobs, chunkSpaces, strategy, useGraph);
  } catch (const std::bad_alloc &) {
    return UMPA_ERR_OUT_OF_MEMORY;
  } catch (...) {
    return UMPA_ERR_DEVICE;
  }
}

int umpa_fr_batch_execute(umpa_fr_batch *b)
{
  if (!b) return UMPA_ERR_INVALID_ARG;
  if (!b->dPlan) return UMPA_ERR_NOT_INITIALIZED;
  DeviceGuard g(b->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  hipError_t e = hipSuccess;
  b->lastUpdates = 0;
  if (b->count) {
    e = hipMemsetAsync(b->dUpdates, 0, sizeof(unsigned long long), b->stream);
    for (const auto &step : b->steps)
      for (const auto &li : step)
        if (e == hipSuccess) e = enqueue(b, li, b->dPlan, true);
    if (e == hipSuccess)
      e = hipMemcpyAsync(&b->lastUpdates, b->dUpdates, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return from_hip(e);
  }
  if (b->exec) return from_hip(hipGraphLaunch(b->exec, b->stream));
  for (const auto &step : b->steps)
    for (const auto &li : step)
      if (e == hipSuccess) e = enqueue(b, li, b->dPlan, false);
  return from_hip(e);
}

static int batch_step_impl(umpa_fr_batch *b, const umpa_fr_job *jobs, const uint64_t *begin,
                           const uint64_t *end, uint32_t nJobs, int strategy)
{
  if (!b || !begin || !end) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(b->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  std::vector<JobInfo> infos;
  int st = validate_all(jobs, nJobs, strategy, infos);
  if (st != UMPA_OK) return st;
  std::vector<std::tuple<int, int, int>> ranges;
  for (uint32_t i = 0; i < nJobs; ++i) {
    if (begin[i] > end[i] || end[i] > (uint64_t)infos[i].rec.nSpaces) return UMPA_ERR_INVALID_ARG;
    if (end[i] > begin[i]) ranges.emplace_back((int)i, (int)begin[i], (int)end[i]);
  }
  b->lastUpdates = 0;
  if (ranges.empty()) return UMPA_OK;
  try {
    std::vector<char> blob;
    std::vector<LaunchImage> launches;
    build_step(b, infos, ranges, blob, launches);
    umpa_fr_batch::Slot &slot = b->ring[b->next];
    b->next = (b->next + 1) % 4;
    // the slot's previous upload and launches must be done before it is overwritten
    hipError_t e = slot.pending ? hipEventSynchronize(slot.done) : hipSuccess;
    slot.pending = false;
    if (e == hipSuccess && slot.cap < blob.size()) {
      if (slot.host) (void)hipHostFree(slot.host);
      if (slot.dev) (void)hipFree(slot.dev);
      slot.host = nullptr;
      slot.dev = nullptr;
      slot.cap = 0;
      const size_t cap = align_up(blob.size() + blob.size() / 2, 1 << 16);
      e = hipHostMalloc((void **)&slot.host, cap, hipHostMallocDefault);
      if (e == hipSuccess) e = hipMalloc(&slot.dev, cap);
      if (e == hipSuccess) slot.cap = cap;
    }
    if (e == hipSuccess) {
      std::memcpy(slot.host, blob.data(), blob.size());
      e = hipMemcpyAsync(slot.dev, slot.host, blob.size(), hipMemcpyHostToDevice, b->stream);
    }
    if (e == hipSuccess && b->count)
      e = hipMemsetAsync(b->dUpdates, 0, sizeof(unsigned long long), b->stream);
    for (const auto &li : launches)
      if (e == hipSuccess) e = enqueue(b, li, slot.dev, b->count);
    if (e == hipSuccess) e = hipEventRecord(slot.done, b->stream);
    if (e == hipSuccess) slot.pending = true;
    if (e == hipSuccess && b->count) {
      e = hipMemcpyAsync(&b->lastUpdates, b->dUpdates, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, b->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    }
    return from_hip(e);
  } catch (const std::bad_alloc &) {
    return UMPA_ERR_OUT_OF_MEMORY;
  }
}

int umpa_fr_batch_step(umpa_fr_batch *b, const umpa_fr_job *jobs, const uint64_t *begin,
                       const uint64_t *end, uint32_t nJobs, int strategy)
{
  try {
    return batch_step_impl(b, jobs, begin, end, nJobs, strategy);
  } catch (const std::bad_alloc &) {
    return UMPA_ERR_OUT_OF_MEMORY;
  } catch (...) {
    return UMPA_ERR_DEVICE;
  }
}

static int batch_autotune_impl(umpa_fr_batch *b, const umpa_fr_job *jobs, uint32_t nJobs,
                               uint32_t chunkSpaces, int strategy, int useCache,
                               umpa_fr_sched_params *chosen)
{
  if (!b || chunkSpaces == 0) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(b->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  std::vector<JobInfo> infos;
  int st = validate_all(jobs, nJobs, strategy, infos);
  if (st != UMPA_OK) return st;
  const bool fastMath = infos[0].v.fastMath;
  for (const auto &j : infos)
    if (j.v.fastMath != fastMath) return UMPA_ERR_INVALID_ARG;
  // key: device, strategy, precision, footprint, and the first step's (volume, spaces) list
  std::vector<std::pair<int, int>> shape;
  std::vector<std::tuple<int, int, int>> ranges;
  for (uint32_t i = 0; i < nJobs; ++i) {
    const int s1 = std::min<int>(infos[i].rec.nSpaces, (int)chunkSpaces);
    if (s1 > 0) ranges.emplace_back((int)i, 0, s1);
    shape.emplace_back(infos[i].rec.vx, s1);
  }
  std::sort(shape.begin(), shape.end());
  std::string key = "k2:batch:cu" + std::to_string(b->numCUs) + ":m" +
                    std::to_string(infos[0].v.mode) + ":f" + std::to_string(fastMath ? 1 : 0) +
                    ":h" + std::to_string(infos[0].v.half ? 1 : 0) + ":nb" +
                    std::to_string(infos[0].v.nb) + ":";
  for (auto &p : shape) key += std::to_string(p.first) + "x" + std::to_string(p.second) + ",";
  umpa::host::SchedParams best;
  auto report = [&](const umpa::host::SchedParams &p) {
    b->params[fastMath ? 1 : 0] = p;
    if (chosen) *chosen = umpa::host::to_c(p);
    return UMPA_OK;
  };
  if (useCache && umpa::host::tuning_cache_lookup(key, best)) return report(best);
  if (ranges.empty()) return report(b->params[fastMath ? 1 : 0]);

  hipError_t e = hipStreamSynchronize(b->stream);
  // scratch volumes: the candidates accumulate into these, never into the caller's
  std::vector<void *> scratch;
  scratch.reserve(2 * infos.size());  // no reallocation between a hipMalloc and its record
  for (auto &j : infos) {
    const size_t n = (size_t)j.rec.vx * j.rec.vy * j.rec.vz;
    void *v = nullptr, *w = nullptr;
    if (e == hipSuccess) e = hipMalloc(&v, n * 8);
    if (e == hipSuccess) scratch.push_back(v);
    if (e == hipSuccess) e = hipMalloc(&w, n * 4);
    if (e == hipSuccess) scratch.push_back(w);
    if (e == hipSuccess) e = hipMemsetAsync(v, 0, n * 8, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(w, 0, n * 4, b->stream);
    j.rec.volume = v;
    j.rec.weights = static_cast<float *>(w);
  }
  char *dBlob = nullptr;
  size_t cap = 0;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  double bestMs = 1e300;
  // batches keep the table in LDS and run the per-pair kernel or (fast math) the carried
  // queues: those and boxFixed are tried over the fastest schedule
  std::vector<umpa::host::SchedParams> cands;
  for (auto c : umpa::host::tuning_candidates(fastMath)) {
    c.kernel = UMPA_FR_KERNEL_AUTO;
    c.tableGlobal = false;
    const bool dup = std::any_of(cands.begin(), cands.end(), [&](const auto &o) {
      return o.depthScale == c.depthScale && o.priority == c.priority &&
             o.maxParts == c.maxParts && o.splitAt == c.splitAt;
    });
    if (!dup) cands.push_back(c);
  }
  const size_t nSched = cands.size(), nCand = nSched + (fastMath ? 3 : 0);
  try {
    for (size_t ci = 0; ci < nCand; ++ci) {
      if (e != hipSuccess) break;
      if (ci == nSched) {  // the fastest schedule: fixed boxes, carried queues, both
        for (int k = 1; k <= 3; ++k) {
          umpa::host::SchedParams c = best;
          c.boxFixed = (k & 1) != 0;
          c.kernel = (k & 2) ? UMPA_FR_KERNEL_CARRY : UMPA_FR_KERNEL_AUTO;
          cands.push_back(c);
        }
      }
      const umpa::host::SchedParams cand = cands[ci];
      std::vector<char> blob;
      std::vector<LaunchImage> launches;
      build_step(b, infos, ranges, blob, launches, &cand);
      if (blob.size() > cap) {
        if (dBlob) (void)hipFree(dBlob);
        dBlob = nullptr;
        cap = align_up(blob.size() * 2, 1 << 16);
        e = hipMalloc(&dBlob, cap);
        if (e != hipSuccess) break;
      }
      e = hipMemcpy(dBlob, blob.data(), blob.size(), hipMemcpyHostToDevice);
      double ms = 1e300;
      for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {
        e = hipEventRecord(t0, b->stream);
        for (const auto &li : launches)
          if (e == hipSuccess) e = enqueue(b, li, dBlob, false);
        if (e == hipSuccess) e = hipEventRecord(t1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, t0, t1);
        if (rep > 0) ms = std::min(ms, (double)t);  // the first run warms up
      }
      if (e == hipSuccess && ms < bestMs) {
        bestMs = ms;
        best = cand;
      }
    }
  } catch (const std::bad_alloc &) {
    e = hipErrorOutOfMemory;
  }
  (void)hipStreamSynchronize(b->stream);
  if (dBlob) (void)hipFree(dBlob);
  for (void *p : scratch) (void)hipFree(p);
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  if (e != hipSuccess) return from_hip(e);
  umpa::host::tuning_cache_store(key, best, bestMs);
  return report(best);
}

int umpa_fr_batch_autotune(umpa_fr_batch *b, const umpa_fr_job *jobs, uint32_t nJobs,
                           uint32_t chunkSpaces, int strategy, int useCache,
                           umpa_fr_sched_params *chosen)
{
  try {
    return batch_autotune_impl(b, jobs, nJobs, chunkSpaces, strategy, useCache, chosen);
  } catch (const std::bad_alloc &) {
    return UMPA_ERR_OUT_OF_MEMORY;
  } catch (...) {
    return UMPA_ERR_DEVICE;
  }
}

int umpa_fr_batch_set_sched_params(umpa_fr_batch *b, int strategy, const umpa_fr_sched_params *p)
{
  if (!b || !p || (strategy != UMPA_FR_EXACT && strategy != UMPA_FR_FAST_MATH) ||
      !umpa::host::sched_params_valid(*p))
    return UMPA_ERR_INVALID_ARG;
  b->params[strategy == UMPA_FR_FAST_MATH ? 1 : 0] = umpa::host::from_c(*p);
  return UMPA_OK;
}

int umpa_fr_batch_set_wave_items(umpa_fr_batch *b, uint32_t itemsPerWave)
{
  if (!b || itemsPerWave > (1u << 20)) return UMPA_ERR_INVALID_ARG;
  b->waveItems = (int)itemsPerWave;
  return UMPA_OK;
}

int umpa_fr_batch_set_max_workgroups(umpa_fr_batch *b, uint32_t workgroups)
{
  if (!b || workgroups > 65535) return UMPA_ERR_INVALID_ARG;
  b->maxWorkgroups = (int)workgroups;
  return UMPA_OK;
}

int umpa_fr_batch_check_jobs(const umpa_fr_job *jobs, uint32_t nJobs, int strategy)
{
  try {
    std::vector<JobInfo> infos;
    return validate_all(jobs, nJobs, strategy, infos);
  } catch (const std::bad_alloc &) {
    return UMPA_ERR_OUT_OF_MEMORY;
  }
}

int umpa_fr_batch_num_steps(umpa_fr_batch *b, uint32_t *steps)
{
  if (!b || !steps) return UMPA_ERR_INVALID_ARG;
  *steps = (uint32_t)b->steps.size();
  return UMPA_OK;
}

int umpa_fr_batch_set_count_updates(umpa_fr_batch *b, int enable)
{
  if (!b) return UMPA_ERR_INVALID_ARG;
  b->count = enable != 0;
  return UMPA_OK;
}

int umpa_fr_batch_last_updates(umpa_fr_batch *b, uint64_t *updates)
{
  if (!b || !updates) return UMPA_ERR_INVALID_ARG;
  *updates = b->lastUpdates;
  return UMPA_OK;
}

int umpa_fr_batch_record(umpa_fr_batch *b, uint64_t *ticket)
{
  if (!b || !ticket) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(b->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  hipEvent_t ev = nullptr;
  hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(ev, b->stream);
  if (e != hipSuccess) {
    if (ev) (void)hipEventDestroy(ev);
    return from_hip(e);
  }
  *ticket = b->nextTicket++;
  b->tickets.emplace_back(*ticket, ev);
  return UMPA_OK;
}

int umpa_fr_batch_query(umpa_fr_batch *b, uint64_t ticket, int *done)
{
  if (!b || !done) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(b->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  // tickets complete in stream order: retire every finished one from the front
  while (!b->tickets.empty()) {
    auto &[t, ev] = b->tickets.front();
    const hipError_t e = hipEventQuery(ev);
    if (e == hipErrorNotReady) break;
    (void)hipEventDestroy(ev);
    b->tickets.pop_front();
    if (e != hipSuccess) return from_hip(e);
  }
  *done = (b->tickets.empty() || b->tickets.front().first > ticket) ? 1 : 0;
  if (ticket >= b->nextTicket) return UMPA_ERR_INVALID_ARG;
  return UMPA_OK;
}

int umpa_fr_batch_synchronize(umpa_fr_batch *b)
{
  if (!b) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(b->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  return from_hip(hipStreamSynchronize(b->stream));
}

void umpa_fr_batch_destroy(umpa_fr_batch *b)
{
  if (!b) return;
  DeviceGuard g(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->drop_plan();
  for (auto &s : b->ring) {
    if (s.host) (void)hipHostFree(s.host);
    if (s.dev) (void)hipFree(s.dev);
    if (s.done) (void)hipEventDestroy(s.done);
  }
  for (auto &t : b->tickets) (void)hipEventDestroy(t.second);
  if (b->dUpdates) (void)hipFree(b->dUpdates);
  if (b->ownStream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // extern "C"

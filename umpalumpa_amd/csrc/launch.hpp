// Launch interface between the host library (compiled by the host compiler) and the gfx950
// kernels (compiled by hipcc).  Plain types only, so both sides see the same layout.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "umpalumpa_amd.h"

namespace umpa {

// One reconstruction as the insertion kernel sees it.  A single Execute passes it by value;
// a batch keeps an array of them in device memory and every work item names its job.
struct FrJob {
  void *volume;         // [z][y][x] complex float (float2)
  float *weights;       // [z][y][x]
  const void *fft;      // slices [n][ySize][xSize]: float2, or __half2 with `scales`
  const float *scales;  // per-slice dequantisation factor (FP16 slices only)
  const umpa_traverse_space *spaces;
  int64_t sliceElems;   // xSize * ySize
  int xSize, ySize;
  int nProj;            // slices available in `fft`
  int nSpaces;
  int vx, vy, vz;       // volume extents
  int spaceBegin;       // this launch inserts spaces [spaceBegin, spaceEnd)
  int spaceEnd;
  int shiftBox;         // the candidate box may slide inside the slice (see gather_box)
  umpa_fr_constants c;
  float cullHalfWidth;  // slab half width used by the brick cull (blob radius, 1.6 for nearest)
  float cullHalfSpace;  // allowed q.x below zero (blob radius, 0 for nearest)
};
static_assert(sizeof(FrJob) == 128, "job records are 128 B on host and device");

// What one launch works on: work items (bricks of one or many jobs, heaviest first) and the
// blob table shared by all its jobs.
struct FrLaunch {
  const FrJob *jobs;    // device array of jobs (batched launches only)
  const int *items;     // int4 {x, y, z | job << 16, key} per work item
  int nItems;
  int *counter;         // work counter, zeroed before the launch
  unsigned long long *updates;  // voxel-update counter (counting variant only)
  const float *table;   // blob table (kLookup)
  int nTable;
  int waveItems;        // items per wave before it leaves (0: until the items run out)
  int maxBlocks;        // persistent launches: at most this many workgroups (0: one per CU)
};

// Largest blob table kept in LDS (floats); batches require it there.
constexpr int kFrMaxLdsTable = 11776;

enum : int { kModeLookup = 0, kModeFastKaiser = 1, kModePreciseKaiser = 2, kModeNearest = 3 };

struct FrVariant {
  int mode;       // kMode*
  int order;      // blob order (precise Kaiser only)
  int nb;         // unrolled footprint width, 0 = generic loops
  bool half;      // FP16 slices
  bool fastMath;  // fused multiply-add accumulation
  bool count;     // count voxel-updates
  bool batched;   // jobs from FrLaunch::jobs (else the FrJob argument)
  int kernel;     // kKernel*: which insertion kernel (kKernelAuto: the measured best)
  bool tableGlobal;  // read the blob table from global memory even where LDS holds it
  bool shift;        // candidate boxes slide inside the slice (FrJob::shiftBox of every job;
                     // a plain launch takes it from its job)
};

// Insertion kernels: one queue per (brick, space) pair; queues carried across spaces (fast
// math); a cross-space ring of staged entries (exact FP32 lookup, 12 wavefronts per CU).
// A kernel that does not cover a variant (mode, strategy, footprint, batching) or a chunk
// (slices beyond one 2 GiB buffer resource) is replaced by the per-pair kernel.
enum : int { kKernelAuto = 0, kKernelPair = 1, kKernelCarry = 2, kKernelStaged = 3 };

// Dynamic LDS and workgroup count of an insertion launch; `prepare` sets the kernel's LDS
// attribute once (outside any stream capture).
hipError_t prepare_fr_insert(const FrVariant &v, int nTable);
// Enqueues the counter reset and the insertion kernel on `st`.
hipError_t launch_fr_insert(const FrJob &job, const FrLaunch &l, const FrVariant &v, int numCUs,
                            hipStream_t st);

// Allocation-free insertion: the work counter and the work items live in a static device
// array of kNoAllocScratchBytes per GPU (a 256-byte header, then int4 items).  `items` is the
// host schedule of the job's volume (make_brick_schedule with default knobs, `split` for the
// fast-math schedules); it is uploaded only when the array holds another one.
constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;
constexpr size_t kNoAllocMaxItems = (kNoAllocScratchBytes - 256) / 16;
constexpr int kNoAllocMaxDevices = 64;
hipError_t launch_fr_insert_noalloc(int device, const FrJob &job, const FrVariant &v,
                                    const int *items, int nItems, bool split, const float *table,
                                    int nTable, int numCUs, hipStream_t st);

// Finalization of a (side^3) volume: optional Hermitian completion and the FP64 division by
// the weights, into `out` (complex float).
hipError_t launch_fr_finalize(const void *volume, const float *weights, int side, void *out,
                              bool hermitian, hipStream_t st);

struct FpArgs {
  const void *in;       // complex float
  void *out;            // complex float
  const float *filter;
  uint64_t inX, inY, inSingle;   // frequency-domain input size
  uint64_t spatialX, spatialY;   // spatial input size (for the frequency of an index)
  uint64_t outX, outY, outSingle;
  uint64_t n;                    // images
  float normFactor;
  float maxFreqSquare;
  int applyFilter, normalize, center, cropFreq, shift;
};

hipError_t launch_fp_scale(const FpArgs &a, int numCUs, hipStream_t st);

// Column pass of the fused FFT + FP step: `work` holds the row transforms [n][y][workX]
// (complex float, workX = spatial x / 2 + 1); y is a power of two from 4 to 1024.
struct FftFpArgs {
  const void *work;
  uint64_t workX;
  int y;                         // rows (= frequency rows of the input spectrum)
  uint64_t inY;
  void *out;                     // complex float [n][outY][outX]
  const float *filter;
  const void *twiddles;         // complex float exp(-2 pi i k / y), k < y
  uint64_t spatialX, spatialY;
  uint64_t outX, outY, outSingle;
  uint64_t n;
  float normFactor;
  float maxFreqSquare;
  int applyFilter, normalize, center, cropFreq, shift;
};

hipError_t prepare_fft_fp_columns(int y);
hipError_t launch_fft_fp_columns(const FftFpArgs &a, hipStream_t st);

}  // namespace umpa

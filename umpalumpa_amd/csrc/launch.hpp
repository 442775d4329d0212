// Launch interface between the host library (compiled by the host compiler) and the gfx950
// kernels (compiled by hipcc).  Plain types only, so both sides see the same layout.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "umpalumpa_amd.h"

namespace umpa {

// Argument block of one FR insertion launch.
struct FrArgs {
  void *volume;         // [z][y][x] complex float (float2)
  float *weights;       // [z][y][x]
  const void *fft;      // slices [n][ySize][xSize]: float2, or __half2 with `scales`
  const float *scales;  // per-slice dequantisation factor (FP16 slices only)
  int64_t sliceElems;   // xSize * ySize
  int xSize, ySize;
  int nProj;            // slices available in `fft`
  int nSpaces;
  const umpa_traverse_space *spaces;
  const float *table;
  int nTable;
  umpa_fr_constants c;
  int vx, vy, vz;       // volume extents
  const int *bricks;    // brick origins as int4 {x, y, z, key}, heaviest first
  int nBricks;
  int *counter;         // work counter, zeroed by the launch itself
  unsigned long long *updates;  // voxel-update counter (counting variant only)
  float cullHalfWidth;  // slab half width used by the brick cull (blob radius, 1.6 for nearest)
  float cullHalfSpace;  // allowed q.x below zero (blob radius, 0 for nearest)
};

enum : int { kModeLookup = 0, kModeFastKaiser = 1, kModePreciseKaiser = 2, kModeNearest = 3 };

struct FrVariant {
  int mode;       // kMode*
  int order;      // blob order (precise Kaiser only)
  int nb;         // unrolled footprint width, 0 = generic loops
  bool half;      // FP16 slices
  bool fastMath;  // fused multiply-add accumulation
  bool count;     // count voxel-updates
};

hipError_t launch_fr_insert(const FrArgs &a, const FrVariant &v, int numCUs, hipStream_t st);

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

}  // namespace umpa

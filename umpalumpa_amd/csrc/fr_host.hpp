// Host-side FR helpers of the product library (no GPU needed).
#pragma once

#include <cstdint>
#include <vector>

#include "umpalumpa_amd.h"

namespace umpa::host {

// AFR::CreateConstants (afr.cpp:6-19)
umpa_fr_constants create_constants(uint64_t volumeX, uint64_t volumeY, uint64_t tableElems,
                                   const umpa_fr_settings &s);

// AFR::FillBlobTable (afr.cpp:22-34)
void fill_blob_table(float *table, uint64_t n, uint64_t volumeX, uint64_t volumeY,
                     const umpa_fr_settings &s);

// ComputeTraverseSpace (traverse_space_generator.hpp:94-145)
void compute_traverse_space(uint32_t imgSizeX, uint32_t imgSizeY, const float transform[9],
                            umpa_traverse_space &space, uint32_t maxVolumeIndexX,
                            uint32_t maxVolumeIndexYZ, bool useFast, float blobRadius,
                            float weight);

// One work item of the insertion kernel: the origin of an 8 x 8 x D brick and
// w = D | heavy << 4 | parts << 8 | part << 16 (a brick split into `parts` slice ranges is
// accumulated by `parts` wavefronts, which merge their sums with atomics).
struct BrickOrigin {
  int x, y, z, w;
  float est;  // estimated cost, for ordering only
};

// Work items covering a vx*vy*vz volume, heaviest first.  `splitHeavy` splits the central
// bricks' slice ranges (changes the summation order: fast-math strategies only).
std::vector<BrickOrigin> make_brick_schedule(int vx, int vy, int vz, bool splitHeavy);

}  // namespace umpa::host

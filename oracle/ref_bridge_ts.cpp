// ORACLE build recipe helper -- test infrastructure only.
// Compiles the reference's traverse-space generator (traverse_space_generator.hpp) from the
// reference tree and exposes ComputeTraverseSpace with C linkage.  The header uses uint32_t
// without including <cstdint>, so the standard header is included first.
#include <cstdint>
#include <libumpalumpa/operations/fourier_reconstruction/traverse_space_generator.hpp>

extern "C" void ref_compute_traverse_space(uint32_t imgSizeX, uint32_t imgSizeY,
  const float *transform9, void *space, uint32_t maxVolumeIndexX, uint32_t maxVolumeIndexYZ,
  int useFast, float blobRadius, float weight)
{
  float t[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i][j] = transform9[i * 3 + j];
  umpalumpa::fourier_reconstruction::ComputeTraverseSpace(imgSizeX, imgSizeY, t,
    *reinterpret_cast<umpalumpa::fourier_reconstruction::TraverseSpace *>(space),
    maxVolumeIndexX, maxVolumeIndexYZ, useFast != 0, blobRadius, weight);
}

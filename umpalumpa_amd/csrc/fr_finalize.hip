// Finalization of an inserted volume: Hermitian completion and the weight normalisation in
// FP64.  The reference has no such step (FRCPU only accumulates, and skips the half-space
// q.x < -r, fr_cpu_kernels.hpp:93-95); it is defined here by its closed form.
//
// Every slice is the transform of a real image, so its values on the skipped half-plane are
// the complex conjugates of the stored ones at the mirrored frequency.  Inserting that virtual
// half would add, at voxel k, the conjugate of what the stored half adds at -k; so the full
// accumulation is V(k) + conj(V(-k)) with weight W(k) + W(-k), where -k is the voxel mirrored
// through the centre (side - 1) / 2.  The result is (V_full / W_full) with the sums and the
// division in double, rounded once to float; voxels with zero weight become 0.
#include <hip/hip_runtime.h>

#include "launch.hpp"

namespace umpa {

template <bool kHermitian>
__global__ __launch_bounds__(256) void fr_finalize_kernel(const float2 *__restrict__ volume,
                                                          const float *__restrict__ weights,
                                                          int side, float2 *__restrict__ out)
{
  const int64_t n = (int64_t)side * side * side;
  const int64_t plane = (int64_t)side * side;
  const int last = side - 1;  // mirror of index i is last - i (the centre is last / 2)
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = volume[i];
    double re = (double)v.x, im = (double)v.y, w = (double)weights[i];
    if constexpr (kHermitian) {
      const int z = (int)(i / plane);
      const int y = (int)((i - (int64_t)z * plane) / side);
      const int x = (int)(i - (int64_t)z * plane - (int64_t)y * side);
      const int64_t m = ((int64_t)(last - z) * side + (last - y)) * side + (last - x);
      const float2 vm = volume[m];
      re = re + (double)vm.x;
      im = im - (double)vm.y;
      w = w + (double)weights[m];
    }
    float2 r = make_float2(0.f, 0.f);
    if (w > 0.0) r = make_float2((float)(re / w), (float)(im / w));
    out[i] = r;
  }
}

hipError_t launch_fr_finalize(const void *volume, const float *weights, int side, void *out,
                              bool hermitian, hipStream_t st)
{
  const int64_t n = (int64_t)side * side * side;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 256 * 64) blocks = 256 * 64;  // grid-stride beyond 64 workgroups per CU
  if (hermitian) {
    hipLaunchKernelGGL(fr_finalize_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, st,
                       static_cast<const float2 *>(volume), weights, side,
                       static_cast<float2 *>(out));
  } else {
    hipLaunchKernelGGL(fr_finalize_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, st,
                       static_cast<const float2 *>(volume), weights, side,
                       static_cast<float2 *>(out));
  }
  return hipGetLastError();
}

}  // namespace umpa

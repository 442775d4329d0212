// FP crop / filter / normalize / center / shift kernel for gfx950.
//
// The reference CUDA kernel gives each thread one output pixel and loops over all images
// (fp_cuda_kernels.cu:9-61).  This is a pure HBM stream, so here every lane owns one complex
// output element, consecutive lanes take consecutive x of one row and the grid spans rows and
// images: loads of an input row and stores of an output row are both contiguous, and no index
// is divided out of a flat one.  Arithmetic and filter indexing follow the reference CPU path
// (ScaleFFT2DCPU, fp_cpu_kernels.hpp:8-41), compiled without contraction, so the output bits
// equal FPCPU's.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.hpp"

namespace umpa {


// Idx2Freq (fp_common_kernels.hpp:13-20): digital frequency of an FFT index
__device__ __forceinline__ float idx2freq(uint64_t idx, uint64_t size)
{
  if (size <= 1) return 0.f;
  float tmp = (idx <= (size / 2)) ? (float)idx : ((float)idx - (float)size);
  return tmp / (float)size;
}

// one lane per output element: x from the block's x (64 lanes of one row), the row from the
// block's y (4 rows), the image from the grid's z, so no index is divided out
__global__ __launch_bounds__(256) void fp_scale_kernel(FpArgs a)
{
  const uint64_t x = (uint64_t)blockIdx.x * 64 + threadIdx.x;
  const uint64_t y = (uint64_t)blockIdx.y * 4 + threadIdx.y;
  const uint64_t n = blockIdx.z;
  if (x >= a.outX || y >= a.outY) return;
  // low frequencies sit in the corners: keep the top outY/2+1 rows and the bottom ones
  const uint64_t origY = (y <= a.outY / 2) ? y : (a.inY - (a.outY - y));
  const uint64_t tmp = y + a.outY / 2;
  const uint64_t shiftedY = tmp >= a.outY ? tmp - a.outY : tmp;
  const uint64_t o = n * a.outSingle + (a.shift ? shiftedY : y) * a.outX + x;
  float re, im;
  bool zero = false;
  if (a.cropFreq) {
    // |f|^2 as std::norm(std::complex<float>) computes it: the square of hypotf
    const float fx = idx2freq(x, a.spatialX), fy = idx2freq(origY, a.spatialY);
    const float mag = (float)sqrt((double)fx * (double)fx + (double)fy * (double)fy);
    zero = (mag * mag) > a.maxFreqSquare;
  }
  if (zero) {
    re = 0.f;
    im = 0.f;
  } else {
    const float2 v = reinterpret_cast<const float2 *>(a.in)[n * a.inSingle + origY * a.inX + x];
    re = v.x;
    im = v.y;
  }
  if (a.applyFilter) {
    const float f = a.filter[y * a.outX + x];
    re *= f;
    im *= f;
  }
  if (a.normalize) {
    re *= a.normFactor;
    im *= a.normFactor;
  }
  if (a.center) {
    const float cc = (float)(1 - 2 * (((int)x + (int)y) & 1));
    re *= cc;
    im *= cc;
  }
  reinterpret_cast<float2 *>(a.out)[o] = make_float2(re, im);
}

hipError_t launch_fp_scale(const FpArgs &a, int numCUs, hipStream_t st)
{
  (void)numCUs;
  if (a.n == 0 || a.outX == 0 || a.outY == 0) return hipSuccess;
  const dim3 block(64, 4);
  const uint64_t gx = (a.outX + 63) / 64, gy = (a.outY + 3) / 4;
  if (gx > 0x7fffffffu || gy > 65535) return hipErrorInvalidValue;
  // images go in grid z, at most 65,535 per launch
  for (uint64_t n0 = 0; n0 < a.n; n0 += 65535) {
    FpArgs c = a;
    c.n = a.n - n0 < 65535 ? a.n - n0 : 65535;
    c.in = static_cast<const char *>(a.in) + n0 * a.inSingle * sizeof(float2);
    c.out = static_cast<char *>(a.out) + n0 * a.outSingle * sizeof(float2);
    hipLaunchKernelGGL(fp_scale_kernel, dim3((unsigned)gx, (unsigned)gy, (unsigned)c.n), block, 0,
                       st, c);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace umpa

// Column pass of the fused FFT + Fourier-processing step for gfx950.
//
// The reference harness transforms each padded particle image (FFTCUDA, fft_cuda.cpp:51-69)
// and then crops / filters / normalises / centres / shifts the spectrum in a second kernel
// (scaleFFT2DKernel, fp_cuda_kernels.cu:9-61), a full write and re-read of every spectrum in
// between.  Here the row transforms (real -> Hermitian half, batched, hipFFT) write a
// workspace, and this kernel does the rest in one pass: a workgroup takes consecutive columns
// of one image, every thread loads its share of a column at once (consecutive lanes on
// consecutive columns, so the loads are coalesced), the Y-point column FFT runs as
// R1 x R2 (four-step: an R1-point FFT in registers, twiddles, one transpose through LDS, an
// R2-point FFT in registers; twiddles from a table computed in double on the host), and the
// output rows the crop keeps are written with the FP arithmetic of fp_scale_kernel
// (ScaleFFT2DCPU, fp_cpu_kernels.hpp:8-41) applied on the way out.  Only the columns the crop
// keeps are ever transformed.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.hpp"

namespace umpa {

namespace {

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// -i * a
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }

// Idx2Freq (fp_common_kernels.hpp:13-20)
__device__ __forceinline__ float idx2freq(uint64_t idx, uint64_t size)
{
  if (size <= 1) return 0.f;
  float tmp = (idx <= (size / 2)) ? (float)idx : ((float)idx - (float)size);
  return tmp / (float)size;
}

}  // namespace

// In-register radix-2 decimation-in-time FFT of R points (R a power of two, indices known at
// compile time so v stays in registers); tw[k] = exp(-2 pi i k / Y), and the twiddle of a
// length-len stage is tw[k Y / len].
template <int R>
__device__ __forceinline__ void fft_regs(float2 (&v)[R], const float2 *tw, int Y)
{
  constexpr int kBits = __builtin_ctz(R);
#pragma unroll
  for (int i = 0; i < R; ++i) {
    int j = 0;
#pragma unroll
    for (int b = 0; b < kBits; ++b) j |= ((i >> b) & 1) << (kBits - 1 - b);
    if (j > i) {
      const float2 t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
    for (int i = 0; i < R; i += len) {
#pragma unroll
      for (int k = 0; k < len / 2; ++k) {
        const float2 u = v[i + k];
        const float2 t = k == 0 ? v[i + k + len / 2] : cmul(v[i + k + len / 2], tw[k * (Y / len)]);
        v[i + k] = cadd(u, t);
        v[i + k + len / 2] = csub(u, t);
      }
    }
  }
}

// Y = 2^kM point FFT down kCPW columns per workgroup of 256 threads, as Y = R1 x R2 (four-step):
// thread (c, n2) loads x[n2 + R2 n1] of column c for every n1 (coalesced across columns),
// transforms over n1 in registers, multiplies by exp(-2 pi i n2 k1 / Y), and writes through LDS
// transposed; thread (c, t) then transforms over n2 for k1 = t + R2 q and holds
// X[k1 + R1 k2].  Then the FP step on the way out.
template <int kM>
__global__ __launch_bounds__(256) void fft_fp_columns_kernel(FftFpArgs a)
{
  constexpr int Y = 1 << kM, R1 = 1 << ((kM + 1) / 2), R2 = 1 << (kM / 2);
  constexpr int kCPW = 256 / R2;
  constexpr int kStride = R2 + 1;                 // +1: transposed writes spread over banks
  extern __shared__ float2 s_lds[];
  const float2 *tw = reinterpret_cast<const float2 *>(a.twiddles);
  const uint64_t n = blockIdx.y;
  const int x0 = blockIdx.x * kCPW;
  const int c = threadIdx.x % kCPW, n2 = threadIdx.x / kCPW;
  const bool colOk = x0 + c < (int)a.outX;
  const float2 *src = reinterpret_cast<const float2 *>(a.work) + n * (uint64_t)Y * a.workX +
                      (uint64_t)(x0 + c);
  float2 v[R1];
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1)
    v[n1] = colOk ? src[(uint64_t)(n2 + R2 * n1) * a.workX] : make_float2(0.f, 0.f);
  fft_regs<R1>(v, tw, Y);
  float2 *col = s_lds + c * R1 * kStride;
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1)
    col[k1 * kStride + n2] = k1 == 0 ? v[k1] : cmul(v[k1], tw[(n2 * k1) & (Y - 1)]);
  __syncthreads();
  float2 *dst = reinterpret_cast<float2 *>(a.out) + n * a.outSingle;
  const uint64_t x = (uint64_t)(x0 + c);
  const uint64_t half = a.outY / 2, skip = a.inY - a.outY;
#pragma unroll
  for (int q = 0; q < R1 / R2; ++q) {
    const int k1 = n2 + R2 * q;
    float2 u[R2];
#pragma unroll
    for (int i = 0; i < R2; ++i) u[i] = col[k1 * kStride + i];
    fft_regs<R2>(u, tw, Y);
    if (!colOk) continue;
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      // frequency row f lands on output row y of the crop, if it is kept (fp_scale_kernel's
      // origY mapping read backwards)
      const uint64_t f = (uint64_t)(k1 + R1 * k2);
      uint64_t y;
      if (f <= half) {
        y = f;
      } else if (f > skip + half) {
        y = f - skip;
      } else {
        continue;
      }
      const uint64_t tmp = y + half;
      const uint64_t shiftedY = tmp >= a.outY ? tmp - a.outY : tmp;
      bool zero = false;
      if (a.cropFreq) {
        const float fx = idx2freq(x, a.spatialX), fy = idx2freq(f, a.spatialY);
        const float mag = (float)sqrt((double)fx * (double)fx + (double)fy * (double)fy);
        zero = (mag * mag) > a.maxFreqSquare;
      }
      float re = zero ? 0.f : u[k2].x, im = zero ? 0.f : u[k2].y;
      if (a.applyFilter) {
        const float fl = a.filter[y * a.outX + x];
        re *= fl;
        im *= fl;
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
      dst[(a.shift ? shiftedY : y) * a.outX + x] = make_float2(re, im);
    }
  }
}

namespace {

template <int kM>
constexpr size_t columns_lds()
{
  return (size_t)(256 / (1 << (kM / 2))) * (1 << ((kM + 1) / 2)) * ((1 << (kM / 2)) + 1) *
         sizeof(float2);
}

template <int kM>
hipError_t launch_m(const FftFpArgs &a, bool prepare, hipStream_t st)
{
  constexpr int kCPW = 256 / (1 << (kM / 2));
  const void *k = (const void *)fft_fp_columns_kernel<kM>;
  if (prepare)
    return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)columns_lds<kM>());
  hipLaunchKernelGGL(fft_fp_columns_kernel<kM>,
                     dim3((unsigned)((a.outX + kCPW - 1) / kCPW), (unsigned)a.n), dim3(256),
                     columns_lds<kM>(), st, a);
  return hipGetLastError();
}

hipError_t dispatch(int y, const FftFpArgs &a, bool prepare, hipStream_t st)
{
  switch (y) {
  case 4: return launch_m<2>(a, prepare, st);
  case 8: return launch_m<3>(a, prepare, st);
  case 16: return launch_m<4>(a, prepare, st);
  case 32: return launch_m<5>(a, prepare, st);
  case 64: return launch_m<6>(a, prepare, st);
  case 128: return launch_m<7>(a, prepare, st);
  case 256: return launch_m<8>(a, prepare, st);
  case 512: return launch_m<9>(a, prepare, st);
  case 1024: return launch_m<10>(a, prepare, st);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t prepare_fft_fp_columns(int y)
{
  FftFpArgs none{};
  return dispatch(y, none, true, nullptr);
}

hipError_t launch_fft_fp_columns(const FftFpArgs &a, hipStream_t st)
{
  if (a.n == 0 || a.outX == 0) return hipSuccess;
  // images go in grid y, at most 65,535 per launch
  for (uint64_t n0 = 0; n0 < a.n; n0 += 65535) {
    FftFpArgs c = a;
    c.n = a.n - n0 < 65535 ? a.n - n0 : 65535;
    c.work = static_cast<const char *>(a.work) + n0 * (uint64_t)a.y * a.workX * sizeof(float2);
    c.out = static_cast<char *>(a.out) + n0 * a.outSingle * sizeof(float2);
    const hipError_t e = dispatch(a.y, c, false, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace umpa

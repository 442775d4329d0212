// The AFFT operation (fourier_transformation/afft.hpp) on the device through hipFFT: batched
// out-of-place real-to-complex (forward) and complex-to-real (inverse, unnormalised) transforms
// of 1-, 2- or 3-D images, the upstream step of the FR harness (fr.cpp:134-164).  The plan
// mirrors FFTCUDA's (fft_cuda.cpp:16-47): contiguous images, rank from the image shape,
// distances of one padded spatial / frequency image.  Plans are kept per batch count, so an
// Execute over fewer images than at Init runs a plan of that many.
#include <hip/hip_runtime_api.h>
#include <hipfft/hipfft.h>

#include <climits>
#include <map>
#include <new>

#include "umpalumpa_amd.h"

namespace {

int from_hip(hipError_t e)
{
  if (e == hipSuccess) return UMPA_OK;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return UMPA_ERR_OUT_OF_MEMORY;
  return UMPA_ERR_DEVICE;
}

int from_fft(hipfftResult r)
{
  if (r == HIPFFT_SUCCESS) return UMPA_OK;
  if (r == HIPFFT_ALLOC_FAILED) return UMPA_ERR_OUT_OF_MEMORY;
  if (r == HIPFFT_INVALID_SIZE || r == HIPFFT_INVALID_VALUE) return UMPA_ERR_INVALID_ARG;
  return UMPA_ERR_DEVICE;
}

struct DeviceGuard {
  int prev = -1;
  bool ok = false;
  explicit DeviceGuard(int device)
  {
    if (hipGetDevice(&prev) != hipSuccess) return;
    ok = (prev == device) || hipSetDevice(device) == hipSuccess;
  }
  ~DeviceGuard()
  {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

}  // namespace

struct umpa_fft_op {
  int device = 0;
  hipStream_t stream = nullptr;
  bool initialized = false;
  umpa_fft_desc desc{};
  std::map<uint64_t, hipfftHandle> plans;  // by batch count

  void release()
  {
    for (auto &p : plans) (void)hipfftDestroy(p.second);
    plans.clear();
    initialized = false;
  }

  int plan(uint64_t count, hipfftHandle &out)
  {
    auto it = plans.find(count);
    if (it != plans.end()) {
      out = it->second;
      return UMPA_OK;
    }
    const umpa_size &s = desc.spatial;
    int n[3];
    int rank;
    if (s.z > 1) {
      rank = 3;
      n[0] = (int)s.z;
      n[1] = (int)s.y;
      n[2] = (int)s.x;
    } else if (s.y > 1) {
      rank = 2;
      n[0] = (int)s.y;
      n[1] = (int)s.x;
    } else {
      rank = 1;
      n[0] = (int)s.x;
    }
    const int spatial = (int)(s.x * s.y * s.z);
    const int freq = (int)((s.x / 2 + 1) * s.y * s.z);
    const bool fwd = desc.direction == UMPA_FFT_FORWARD;
    const bool dbl = desc.precision == UMPA_FFT_F64;
    const hipfftType type = fwd ? (dbl ? HIPFFT_D2Z : HIPFFT_R2C) : (dbl ? HIPFFT_Z2D : HIPFFT_C2R);
    hipfftHandle h = nullptr;
    hipfftResult r = hipfftCreate(&h);
    if (r == HIPFFT_SUCCESS) {
      size_t work = 0;
      r = hipfftMakePlanMany(h, rank, n, nullptr, 1, fwd ? spatial : freq, nullptr, 1,
                             fwd ? freq : spatial, type, (int)count, &work);
    }
    if (r == HIPFFT_SUCCESS) r = hipfftSetStream(h, stream);
    if (r != HIPFFT_SUCCESS) {
      if (h) (void)hipfftDestroy(h);
      return from_fft(r);
    }
    plans[count] = h;
    out = h;
    return UMPA_OK;
  }
};

extern "C" {

int umpa_fft_create(umpa_fft_op **op, int device, void *stream)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  *op = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return UMPA_ERR_DEVICE;
  auto *h = new (std::nothrow) umpa_fft_op();
  if (!h) return UMPA_ERR_OUT_OF_MEMORY;
  h->device = device;
  h->stream = static_cast<hipStream_t>(stream);
  *op = h;
  return UMPA_OK;
}

int umpa_fft_init(umpa_fft_op *op, const umpa_fft_desc *desc)
{
  if (!op || !desc) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  op->release();
  const umpa_size &s = desc->spatial;
  if (s.x < 1 || s.y < 1 || s.z < 1 || s.n < 1) return UMPA_ERR_INVALID_ARG;
  if ((s.z > 1 && s.y < 2) || (desc->direction != UMPA_FFT_FORWARD &&
                               desc->direction != UMPA_FFT_INVERSE) ||
      (desc->precision != UMPA_FFT_F32 && desc->precision != UMPA_FFT_F64))
    return UMPA_ERR_INVALID_ARG;
  // FFTCUDA::IsValid (fft_cuda.cpp:70-75): the element count must fit an int
  if (s.x * s.y * s.z * s.n > (uint64_t)INT_MAX) return UMPA_ERR_UNSUPPORTED;
  op->desc = *desc;
  hipfftHandle h = nullptr;
  const int st = op->plan(s.n, h);
  if (st != UMPA_OK) return st;
  op->initialized = true;
  return UMPA_OK;
}

int umpa_fft_execute(umpa_fft_op *op, const void *in, void *out, uint64_t count)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  if (!op->initialized) return UMPA_ERR_NOT_INITIALIZED;
  if (count > op->desc.spatial.n) return UMPA_ERR_NOT_EQUIVALENT;
  if (!in || !out || count == 0) return UMPA_ERR_REJECTED;
  if (in == out) return UMPA_ERR_UNSUPPORTED;  // in-place transforms are not offered
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  hipfftHandle h = nullptr;
  int st = op->plan(count, h);
  if (st != UMPA_OK) return st;
  void *i = const_cast<void *>(in);
  hipfftResult r;
  const bool dbl = op->desc.precision == UMPA_FFT_F64;
  if (op->desc.direction == UMPA_FFT_FORWARD) {
    r = dbl ? hipfftExecD2Z(h, static_cast<hipfftDoubleReal *>(i),
                            static_cast<hipfftDoubleComplex *>(out))
            : hipfftExecR2C(h, static_cast<hipfftReal *>(i), static_cast<hipfftComplex *>(out));
  } else {
    r = dbl ? hipfftExecZ2D(h, static_cast<hipfftDoubleComplex *>(i),
                            static_cast<hipfftDoubleReal *>(out))
            : hipfftExecC2R(h, static_cast<hipfftComplex *>(i), static_cast<hipfftReal *>(out));
  }
  return from_fft(r);
}

int umpa_fft_synchronize(umpa_fft_op *op)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  return from_hip(hipStreamSynchronize(op->stream));
}

int umpa_fft_cleanup(umpa_fft_op *op)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(op->device);
  op->release();
  return UMPA_OK;
}

void umpa_fft_destroy(umpa_fft_op *op)
{
  if (!op) return;
  {
    DeviceGuard g(op->device);
    (void)hipStreamSynchronize(op->stream);
    op->release();
  }
  delete op;
}

}  // extern "C"

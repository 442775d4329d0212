// The AFFT operation (fourier_transformation/afft.hpp) on the device through hipFFT: batched
// out-of-place real-to-complex (forward) and complex-to-real (inverse, unnormalised) transforms
// of 1-, 2- or 3-D images, the upstream step of the FR harness (fr.cpp:134-164).  The plan
// mirrors FFTCUDA's (fft_cuda.cpp:16-47): contiguous images, rank from the image shape,
// distances of one padded spatial / frequency image.  Plans are kept per batch count (at most
// eight), so an Execute over fewer images than at Init runs a plan of that many.
#include <hip/hip_runtime_api.h>
#include <hipfft/hipfft.h>

#include <climits>
#include <cmath>
#include <complex>
#include <map>
#include <new>
#include <vector>

#include "hip_util.hpp"
#include "launch.hpp"
#include "umpalumpa_amd.h"

namespace {

using umpa::detail::aligned;
using umpa::detail::DeviceGuard;
using umpa::detail::from_hip;

int from_fft(hipfftResult r)
{
  if (r == HIPFFT_SUCCESS) return UMPA_OK;
  if (r == HIPFFT_ALLOC_FAILED) return UMPA_ERR_OUT_OF_MEMORY;
  if (r == HIPFFT_INVALID_SIZE || r == HIPFFT_INVALID_VALUE) return UMPA_ERR_INVALID_ARG;
  return UMPA_ERR_DEVICE;
}

// Size::IsValid (data/size.hpp:43-57) for the 2-D images of the fused operation
bool valid_2d(const umpa_size &s) { return s.n >= 1 && s.z == 1 && s.y >= 2 && s.x >= 2; }

// Plans by image count.  A count not seen before gets a plan of its own (`make` creates the
// handle, plans it and sets the stream); past kMaxPlans the stream drains and every plan but
// the one for the Init count is destroyed, so a caller cycling through counts holds a bounded
// number of plans and their work areas.
constexpr size_t kMaxPlans = 8;

template <typename Make>
int cached_plan(std::map<uint64_t, hipfftHandle> &plans, uint64_t count, uint64_t initCount,
                hipStream_t stream, Make make, hipfftHandle &out)
{
  auto it = plans.find(count);
  if (it != plans.end()) {
    out = it->second;
    return UMPA_OK;
  }
  if (plans.size() >= kMaxPlans) {
    // enqueued transforms may still use the plans' work areas
    const hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return from_hip(e);
    for (auto p = plans.begin(); p != plans.end();) {
      if (p->first == initCount) {
        ++p;
        continue;
      }
      (void)hipfftDestroy(p->second);
      p = plans.erase(p);
    }
  }
  hipfftHandle h = nullptr;
  hipfftResult r = hipfftCreate(&h);
  if (r == HIPFFT_SUCCESS) r = make(h);
  if (r == HIPFFT_SUCCESS) r = hipfftSetStream(h, stream);
  if (r != HIPFFT_SUCCESS) {
    if (h) (void)hipfftDestroy(h);
    return from_fft(r);
  }
  try {
    plans.emplace(count, h);
  } catch (const std::bad_alloc &) {
    (void)hipfftDestroy(h);
    return UMPA_ERR_OUT_OF_MEMORY;
  }
  out = h;
  return UMPA_OK;
}

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
    auto make = [&](hipfftHandle h) {
      size_t work = 0;
      return hipfftMakePlanMany(h, rank, n, nullptr, 1, fwd ? spatial : freq, nullptr, 1,
                                fwd ? freq : spatial, type, (int)count, &work);
    };
    return cached_plan(plans, count, s.n, stream, make, out);
  }
};

struct umpa_fft_fp_op {
  int device = 0;
  hipStream_t stream = nullptr;
  bool initialized = false;
  umpa_fp_desc desc{};
  umpa_fp_settings settings{};
  std::map<uint64_t, hipfftHandle> plans;  // row transforms, by image count
  void *work = nullptr;                    // row spectra [n][y][x/2+1] complex float
  void *twiddles = nullptr;                // exp(-2 pi i k / y), k < y, complex float

  void release()
  {
    for (auto &p : plans) (void)hipfftDestroy(p.second);
    plans.clear();
    if (work) (void)hipFree(work);
    if (twiddles) (void)hipFree(twiddles);
    work = nullptr;
    twiddles = nullptr;
    initialized = false;
  }

  // batched 1-D R2C over every image row: count * y rows of x reals
  int plan(uint64_t count, hipfftHandle &out)
  {
    const umpa_size &s = desc.inSpatial;
    int n[1] = { (int)s.x };
    auto make = [&](hipfftHandle h) {
      size_t ws = 0;
      return hipfftMakePlanMany(h, 1, n, nullptr, 1, (int)s.x, nullptr, 1, (int)(s.x / 2 + 1),
                                HIPFFT_R2C, (int)(count * s.y), &ws);
    };
    return cached_plan(plans, count, s.n, stream, make, out);
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

// ---- fused FFT + FP ---------------------------------------------------------------------
int umpa_fft_fp_create(umpa_fft_fp_op **op, int device, void *stream)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  *op = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return UMPA_ERR_DEVICE;
  auto *h = new (std::nothrow) umpa_fft_fp_op();
  if (!h) return UMPA_ERR_OUT_OF_MEMORY;
  h->device = device;
  h->stream = static_cast<hipStream_t>(stream);
  *op = h;
  return UMPA_OK;
}

int umpa_fft_fp_init(umpa_fft_fp_op *op, const umpa_fp_desc *desc, const umpa_fp_settings *s)
{
  if (!op || !desc || !s) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  op->release();
  const umpa_size &sp = desc->inSpatial;
  // the FP checks of umpa_fp_init (afp.hpp:45-52, settings.hpp:38-43) on 2-D payloads
  if (!valid_2d(sp) || !valid_2d(desc->in) || !valid_2d(desc->out)) return UMPA_ERR_INVALID_ARG;
  if (desc->in.x != sp.x / 2 + 1 || desc->in.y != sp.y || desc->in.n != sp.n)
    return UMPA_ERR_INVALID_ARG;
  if (desc->out.n < desc->in.n) return UMPA_ERR_INVALID_ARG;
  if (desc->out.x > desc->in.x || desc->out.y > desc->in.y) return UMPA_ERR_UNSUPPORTED;
  if (s->hasMaxFreq && !(s->maxFreq >= 0.f && s->maxFreq <= 0.5f)) return UMPA_ERR_INVALID_ARG;
  if (desc->inPaddedSpatialSingle == 0) return UMPA_ERR_INVALID_ARG;
  // column transforms: power-of-two heights that fit the kernel's LDS tiles
  if (sp.y < 4 || sp.y > 1024 || (sp.y & (sp.y - 1)) != 0) return UMPA_ERR_UNSUPPORTED;
  if (sp.x > (uint64_t)INT_MAX || sp.n * sp.y > (uint64_t)INT_MAX ||
      sp.n * sp.y * sp.x > (uint64_t)INT_MAX)
    return UMPA_ERR_UNSUPPORTED;
  op->desc = *desc;
  op->settings = *s;
  try {
    std::vector<std::complex<float>> tw(sp.y);
    const double pi = 3.14159265358979323846;
    for (uint64_t k = 0; k < sp.y; ++k) {
      const double a = -2.0 * pi * (double)k / (double)sp.y;
      tw[k] = std::complex<float>((float)std::cos(a), (float)std::sin(a));
    }
    hipError_t e = hipMalloc(&op->twiddles, tw.size() * sizeof(tw[0]));
    if (e == hipSuccess)
      e = hipMemcpy(op->twiddles, tw.data(), tw.size() * sizeof(tw[0]), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&op->work, sp.n * sp.y * (sp.x / 2 + 1) * 8);
    if (e == hipSuccess) e = umpa::prepare_fft_fp_columns((int)sp.y);
    if (e != hipSuccess) {
      op->release();
      return from_hip(e);
    }
  } catch (const std::bad_alloc &) {
    op->release();
    return UMPA_ERR_OUT_OF_MEMORY;
  }
  hipfftHandle h = nullptr;
  const int st = op->plan(sp.n, h);
  if (st != UMPA_OK) {
    op->release();
    return st;
  }
  op->initialized = true;
  return UMPA_OK;
}

int umpa_fft_fp_execute(umpa_fft_fp_op *op, const float *images, uint64_t count, void *out,
                        const float *filter)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  if (!op->initialized) return UMPA_ERR_NOT_INITIALIZED;
  if (count > op->desc.inSpatial.n) return UMPA_ERR_NOT_EQUIVALENT;
  if (!images || !out || count == 0) return UMPA_ERR_REJECTED;
  if (op->settings.applyFilter && !filter) return UMPA_ERR_INVALID_ARG;
  if (!aligned(images, 4) || !aligned(out, 8) || (filter && !aligned(filter, 4)))
    return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  hipfftHandle h = nullptr;
  int st = op->plan(count, h);
  if (st != UMPA_OK) return st;
  hipfftResult r = hipfftExecR2C(h, const_cast<hipfftReal *>(images),
                                 static_cast<hipfftComplex *>(op->work));
  if (r != HIPFFT_SUCCESS) return from_fft(r);
  const umpa_fp_desc &d = op->desc;
  umpa::FftFpArgs a{};
  a.work = op->work;
  a.workX = d.inSpatial.x / 2 + 1;
  a.y = (int)d.inSpatial.y;
  a.inY = d.in.y;
  a.out = out;
  a.filter = filter;
  a.twiddles = op->twiddles;
  a.spatialX = d.inSpatial.x;
  a.spatialY = d.inSpatial.y;
  a.outX = d.out.x;
  a.outY = d.out.y;
  a.outSingle = d.out.x * d.out.y;
  a.n = count;
  a.normFactor = 1.f / static_cast<float>(d.inPaddedSpatialSingle);
  a.maxFreqSquare = op->settings.hasMaxFreq ? op->settings.maxFreq : 0.f;
  a.applyFilter = op->settings.applyFilter != 0;
  a.normalize = op->settings.normalize != 0;
  a.center = op->settings.center != 0;
  a.cropFreq = op->settings.hasMaxFreq != 0;
  a.shift = op->settings.shift != 0;
  return from_hip(umpa::launch_fft_fp_columns(a, op->stream));
}

int umpa_fft_fp_synchronize(umpa_fft_fp_op *op)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(op->device);
  if (!g.ok) return UMPA_ERR_DEVICE;
  return from_hip(hipStreamSynchronize(op->stream));
}

int umpa_fft_fp_cleanup(umpa_fft_fp_op *op)
{
  if (!op) return UMPA_ERR_INVALID_ARG;
  DeviceGuard g(op->device);
  if (op->initialized) (void)hipStreamSynchronize(op->stream);
  op->release();
  return UMPA_OK;
}

void umpa_fft_fp_destroy(umpa_fft_fp_op *op)
{
  if (!op) return;
  umpa_fft_fp_cleanup(op);
  delete op;
}

}  // extern "C"

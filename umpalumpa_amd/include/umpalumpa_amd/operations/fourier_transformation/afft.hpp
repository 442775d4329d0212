// Fourier transformation (FFT) operation: AFFT with its settings, and FFTHIP, the MI355X
// backend (batched out-of-place real <-> Hermitian-complex transforms through hipFFT).
//
// Same names and meaning as the reference's fourier_transformation::Settings
// (settings.hpp:6-45), AFFT (afft.hpp:29-58) and FFTCUDA (fft_cuda.hpp): the spatial payload is
// float (or double), the frequency payload complex float (or complex double) with x/2+1
// elements per row; the inverse is unnormalised.
#pragma once

#include <complex>
#include <stdexcept>

#include "../../data/payload.hpp"
#include "../basic_operation.hpp"
#include "locality.hpp"
#include "umpalumpa_amd.h"

namespace umpalumpa::fourier_transformation {

class Settings
{
public:
  Settings(Locality loc, Direction dir, unsigned thr = 1)
    : locality(loc), direction(dir), threads(thr)
  {}
  Locality GetLocality() const { return locality; }
  Direction GetDirection() const { return direction; }
  bool IsForward() const { return direction == Direction::kForward; }
  bool IsOutOfPlace() const { return locality == Locality::kOutOfPlace; }
  int GetVersion() const { return 1; }
  Settings CreateInverse() const
  {
    return Settings(locality, IsForward() ? Direction::kInverse : Direction::kForward, threads);
  }
  unsigned GetThreads() const { return threads; }  // CPU threads; no effect on the device

private:
  Locality locality;
  Direction direction;
  unsigned threads;
};

template<typename T> struct OutputDataWrapper : public data::PayloadWrapper<T>
{
  OutputDataWrapper(std::tuple<T> &t) : data::PayloadWrapper<T>(t) {}
  OutputDataWrapper(T &d) : data::PayloadWrapper<T>(d) {}
  const T &GetData() const { return std::get<0>(this->payloads); }
};

template<typename T> struct InputDataWrapper : public data::PayloadWrapper<T>
{
  InputDataWrapper(std::tuple<T> &t) : data::PayloadWrapper<T>(t) {}
  InputDataWrapper(T &d) : data::PayloadWrapper<T>(d) {}
  const T &GetData() const { return std::get<0>(this->payloads); }
};

class AFFT
  : public BasicOperation<OutputDataWrapper<data::Payload<data::FourierDescriptor>>,
      InputDataWrapper<data::Payload<data::FourierDescriptor>>,
      Settings>
{
public:
  static bool IsDouble(const OutputData &out, const InputData &in, Direction d)
  {
    if (d == Direction::kForward)
      return out.GetData().dataInfo.GetType().Is<std::complex<double>>() &&
             in.GetData().dataInfo.GetType().Is<double>();
    return out.GetData().dataInfo.GetType().Is<double>() &&
           in.GetData().dataInfo.GetType().Is<std::complex<double>>();
  }
  static bool IsFloat(const OutputData &out, const InputData &in, Direction d)
  {
    if (d == Direction::kForward)
      return out.GetData().dataInfo.GetType().Is<std::complex<float>>() &&
             in.GetData().dataInfo.GetType().Is<float>();
    return out.GetData().dataInfo.GetType().Is<float>() &&
           in.GetData().dataInfo.GetType().Is<std::complex<float>>();
  }

protected:
  bool IsValid(const OutputData &out, const InputData &in, const Settings &s) const override
  {
    return out.GetData().IsValid() && in.GetData().IsValid() &&
           (IsDouble(out, in, s.GetDirection()) || IsFloat(out, in, s.GetDirection()));
  }
};

// MI355X backend.  Payload pointers must be device accessible (device or managed memory).
class FFTHIP : public AFFT
{
public:
  explicit FFTHIP(int device = 0, void *stream = nullptr)
  {
    if (umpa_fft_create(&handle, device, stream) != UMPA_OK)
      throw std::runtime_error("FFTHIP: no usable HIP device");
  }
  ~FFTHIP() override { umpa_fft_destroy(handle); }
  FFTHIP(const FFTHIP &) = delete;
  FFTHIP &operator=(const FFTHIP &) = delete;

  void Synchronize() override { umpa_fft_synchronize(handle); }
  void Cleanup() override
  {
    AFFT::Cleanup();
    if (handle) umpa_fft_cleanup(handle);
  }
  int LastStatus() const { return lastStatus; }

protected:
  bool InitImpl(const OutputData &out, const InputData &in, const Settings &s) override
  {
    if (!s.IsOutOfPlace()) {
      lastStatus = UMPA_ERR_UNSUPPORTED;
      return false;
    }
    const auto &spatial = s.IsForward() ? in.GetData().info.GetPaddedSpatialSize()
                                        : out.GetData().info.GetPaddedSpatialSize();
    umpa_fft_desc d{ { spatial.x, spatial.y, spatial.z, spatial.n },
      s.IsForward() ? UMPA_FFT_FORWARD : UMPA_FFT_INVERSE,
      IsDouble(out, in, s.GetDirection()) ? UMPA_FFT_F64 : UMPA_FFT_F32 };
    lastStatus = umpa_fft_init(handle, &d);
    return lastStatus == UMPA_OK;
  }
  bool ExecuteImpl(const OutputData &out, const InputData &in) override
  {
    lastStatus = umpa_fft_execute(handle, in.GetData().GetPtr(), out.GetData().GetPtr(),
      in.GetData().info.GetSize().n);
    return lastStatus == UMPA_OK;
  }

private:
  umpa_fft_op *handle = nullptr;
  int lastStatus = UMPA_OK;
};

}  // namespace umpalumpa::fourier_transformation

// Fourier processing (FP) operation: AFP with its settings, and FPHIP, the MI355X backend
// (crop / filter / normalize / center / shift of Hermitian spectra).
//
// Same names and meaning as the reference's fourier_processing::Settings
// (fourier_processing/settings.hpp:10-84), AFP (afp.hpp:13-53) and FPCUDA (fp_cuda.hpp).
// The filter is indexed like the reference CPU path, filter[y * outX + x]
// (fp_cpu_kernels.hpp:31).
#pragma once

#include <optional>
#include <stdexcept>

#include "../../data/payload.hpp"
#include "../basic_operation.hpp"
#include "../fourier_transformation/locality.hpp"
#include "umpalumpa_amd.h"

namespace umpalumpa::fourier_processing {

class Settings
{
public:
  explicit Settings(fourier_transformation::Locality loc) : locality(loc) {}
  bool IsEquivalentTo(const Settings &r) const
  {
    return locality == r.locality && center == r.center && normalize == r.normalize &&
           applyFilter == r.applyFilter && shift == r.shift && maxFreq == r.maxFreq;
  }
  bool IsOutOfPlace() const { return locality == fourier_transformation::Locality::kOutOfPlace; }
  void SetCenter(bool v) { center = v; }
  void SetShift(bool v) { shift = v; }
  void SetNormalize(bool v) { normalize = v; }
  void SetApplyFilter(bool v) { applyFilter = v; }
  // max frequency in [0, 1/2]; compared with |f|^2 as in the reference (fp_cpu_kernels.hpp:30)
  void SetMaxFreq(float f) { maxFreq = f; }
  bool GetCenter() const { return center; }
  bool GetShift() const { return shift; }
  fourier_transformation::Locality GetLocality() const { return locality; }
  bool GetNormalize() const { return normalize; }
  bool GetApplyFilter() const { return applyFilter; }
  std::optional<float> GetMaxFreq() const { return maxFreq; }
  umpa_fp_settings ToC() const
  {
    return { center ? 1 : 0, normalize ? 1 : 0, applyFilter ? 1 : 0, shift ? 1 : 0,
      maxFreq.has_value() ? 1 : 0, maxFreq.value_or(0.f) };
  }

private:
  fourier_transformation::Locality locality;
  bool center = false;
  bool normalize = false;
  bool applyFilter = false;
  bool shift = false;
  std::optional<float> maxFreq;
};

template<typename T, typename U> struct InputDataWrapper : public data::PayloadWrapper<T, U>
{
  InputDataWrapper(std::tuple<T, U> &t) : data::PayloadWrapper<T, U>(t) {}
  InputDataWrapper(T &data, U &filter) : data::PayloadWrapper<T, U>(data, filter) {}
  const T &GetData() const { return std::get<0>(this->payloads); }
  const U &GetFilter() const { return std::get<1>(this->payloads); }
};

template<typename T> struct OutputDataWrapper : public data::PayloadWrapper<T>
{
  OutputDataWrapper(std::tuple<T> &t) : data::PayloadWrapper<T>(t) {}
  OutputDataWrapper(T &data) : data::PayloadWrapper<T>(data) {}
  const T &GetData() const { return std::get<0>(this->payloads); }
};

class AFP
  : public BasicOperation<OutputDataWrapper<data::Payload<data::FourierDescriptor>>,
      InputDataWrapper<data::Payload<data::FourierDescriptor>,
        data::Payload<data::LogicalDescriptor>>,
      Settings>
{
public:
  static bool IsFloat(const OutputData &out, const InputData &in)
  {
    return in.GetData().dataInfo.GetType().Is<std::complex<float>>() &&
           (in.GetFilter().dataInfo.GetType().Is<float>() ||
             in.GetFilter().dataInfo.GetType().Is<void>()) &&
           out.GetData().dataInfo.GetType().Is<std::complex<float>>();
  }

protected:
  bool IsValid(const OutputData &out, const InputData &in, const Settings &) const override
  {
    return out.IsValid() && in.IsValid();
  }
};

class FPHIP : public AFP
{
public:
  explicit FPHIP(int device = 0, void *stream = nullptr)
  {
    if (umpa_fp_create(&handle, device, stream) != UMPA_OK)
      throw std::runtime_error("FPHIP: no usable HIP device");
  }
  ~FPHIP() override { umpa_fp_destroy(handle); }
  FPHIP(const FPHIP &) = delete;
  FPHIP &operator=(const FPHIP &) = delete;
  void Synchronize() override { umpa_fp_synchronize(handle); }
  int LastStatus() const { return lastStatus; }

protected:
  bool InitImpl(const OutputData &out, const InputData &in, const Settings &s) override
  {
    // FPCPU/FPCUDA Strategy1::Init: complex float in/out, float or no filter; out-of-place
    if (!IsFloat(out, in) || !s.IsOutOfPlace()) return false;
    const auto &d = in.GetData().info;
    const auto &o = out.GetData().info;
    const umpa_fp_desc desc{ { d.GetSize().x, d.GetSize().y, d.GetSize().z, d.GetSize().n },
      { d.GetSpatialSize().x, d.GetSpatialSize().y, d.GetSpatialSize().z, d.GetSpatialSize().n },
      { o.GetSize().x, o.GetSize().y, o.GetSize().z, o.GetSize().n },
      d.GetPaddedSpatialSize().single };
    const umpa_fp_settings cs = s.ToC();
    lastStatus = umpa_fp_init(handle, &desc, &cs);
    return lastStatus == UMPA_OK;
  }
  bool ExecuteImpl(const OutputData &out, const InputData &in) override
  {
    lastStatus = umpa_fp_execute(handle, in.GetData().GetPtr(), in.GetData().info.GetSize().n,
      out.GetData().GetPtr(), out.GetData().info.GetSize().n,
      reinterpret_cast<const float *>(in.GetFilter().GetPtr()));
    return lastStatus == UMPA_OK;
  }

private:
  umpa_fp_op *handle = nullptr;
  int lastStatus = UMPA_OK;
};

}  // namespace umpalumpa::fourier_processing

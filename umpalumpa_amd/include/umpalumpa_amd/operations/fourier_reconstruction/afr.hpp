// Fourier reconstruction (FR) operation: AFR with its settings, and FRHIP, the MI355X backend
// that executes it with this library's gfx950 kernels through the C ABI (umpalumpa_amd.h).
//
// Same names and meaning as the reference's fourier_reconstruction::Settings
// (settings.hpp:8-67), BlobOrder (blob_order.hpp:5-23), AFR (afr.hpp:13-55) and FRCUDA
// (fr_cuda.hpp): a caller switches backend by constructing FRHIP instead of FRCUDA / FRCPU.
// FRHIP's Execute is asynchronous on its stream; Synchronize() waits (fr_cuda.cpp:243).
#pragma once

#include <stdexcept>

#include "../../data/payload.hpp"
#include "../basic_operation.hpp"
#include "umpalumpa_amd.h"

namespace umpalumpa::fourier_reconstruction {

enum class BlobOrder { k0, k1, k2, k3, k4 };

class Settings
{
public:
  enum class Interpolation { kDynamic, kLookup };
  enum class Type { kFast, kPrecise };

  Interpolation GetInterpolation() const { return interpolation; }
  void SetInterpolation(const Interpolation &i) { interpolation = i; }
  BlobOrder GetBlobOrder() const { return order; }
  void SetBlobOrder(const BlobOrder &o) { order = o; }
  float GetAlpha() const { return alpha; }
  void SetAlpha(float a) { alpha = a; }
  Type GetType() const { return type; }
  void SetType(const Type &t) { type = t; }
  float GetBlobRadius() const { return blobRadius; }
  void SetBlobRadius(float r) { blobRadius = r; }
  bool IsEquivalentTo(const Settings &o) const
  {
    return interpolation == o.interpolation && order == o.order && type == o.type &&
           alpha == o.alpha && blobRadius == o.blobRadius;
  }
  umpa_fr_settings ToC() const
  {
    return { static_cast<int32_t>(interpolation), static_cast<int32_t>(order),
      static_cast<int32_t>(type), alpha, blobRadius };
  }

private:
  Interpolation interpolation = Interpolation::kDynamic;
  BlobOrder order = BlobOrder::k0;
  Type type = Type::kFast;
  float alpha = 15.f;
  float blobRadius = 1.9f;
};

// 112-byte traverse-space record and 32-byte constants, byte compatible with the reference's
// TraverseSpace (traverse_space.hpp:14-48) and Constants (constants.hpp:6-16)
using TraverseSpace = umpa_traverse_space;
using Constants = umpa_fr_constants;

template<typename T, typename U> struct OutputDataWrapper : public data::PayloadWrapper<T, U>
{
  OutputDataWrapper(std::tuple<T, U> &t) : data::PayloadWrapper<T, U>(t) {}
  OutputDataWrapper(T &volume, U &weight) : data::PayloadWrapper<T, U>(volume, weight) {}
  const T &GetVolume() const { return std::get<0>(this->payloads); }
  const U &GetWeight() const { return std::get<1>(this->payloads); }
};

template<typename T, typename U>
struct InputDataWrapper : public data::PayloadWrapper<T, T, U, U, U>
{
  InputDataWrapper(std::tuple<T, T, U, U, U> &t) : data::PayloadWrapper<T, T, U, U, U>(t) {}
  InputDataWrapper(T &fft, T &volume, U &weight, U &traverseSpace, U &blobTable)
    : data::PayloadWrapper<T, T, U, U, U>(fft, volume, weight, traverseSpace, blobTable)
  {}
  const T &GetFFT() const { return std::get<0>(this->payloads); }
  const T &GetVolume() const { return std::get<1>(this->payloads); }
  const U &GetWeight() const { return std::get<2>(this->payloads); }
  const U &GetTraverseSpace() const { return std::get<3>(this->payloads); }
  const U &GetBlobTable() const { return std::get<4>(this->payloads); }
};

class AFR
  : public BasicOperation<OutputDataWrapper<data::Payload<data::FourierDescriptor>,
                            data::Payload<data::LogicalDescriptor>>,
      InputDataWrapper<data::Payload<data::FourierDescriptor>,
        data::Payload<data::LogicalDescriptor>>,
      Settings>
{
public:
  static umpa_fr_desc Describe(const InputData &in)
  {
    const auto &f = in.GetFFT().info.GetSize();
    const auto &v = in.GetVolume().info.GetSize();
    return { { f.x, f.y, f.z, f.n }, { v.x, v.y, v.z, v.n },
      in.GetTraverseSpace().info.GetSize().total, in.GetBlobTable().info.Elems() };
  }
  // AFR::CreateConstants (afr.cpp:6-19)
  static Constants CreateConstants(const InputData &in, const Settings &s)
  {
    Constants c{};
    const umpa_fr_desc d = Describe(in);
    const umpa_fr_settings cs = s.ToC();
    umpa_fr_create_constants(&d, &cs, &c);
    return c;
  }
  // AFR::FillBlobTable (afr.cpp:22-34): writes through the table payload's pointer, which must
  // be host accessible (host, pinned or managed memory)
  static void FillBlobTable(const InputData &in, const Settings &s)
  {
    const umpa_fr_desc d = Describe(in);
    const umpa_fr_settings cs = s.ToC();
    umpa_fr_fill_blob_table(reinterpret_cast<float *>(in.GetBlobTable().GetPtr()),
      in.GetBlobTable().info.Elems(), &d, &cs);
  }

protected:
  bool IsValid(const OutputData &out, const InputData &in, const Settings &) const override
  {
    auto validEqual = [](const auto &a, const auto &b) { return a.IsValid() && a == b; };
    return validEqual(out.GetVolume(), in.GetVolume()) &&
           validEqual(out.GetWeight(), in.GetWeight()) && in.GetFFT().IsValid() &&
           in.GetTraverseSpace().IsValid();
  }
};

// MI355X backend.  Payload pointers must be device accessible (device or managed memory).
class FRHIP : public AFR
{
public:
  explicit FRHIP(int device = 0, void *stream = nullptr)
  {
    if (umpa_fr_create(&handle, device, stream) != UMPA_OK)
      throw std::runtime_error("FRHIP: no usable HIP device");
  }
  ~FRHIP() override { umpa_fr_destroy(handle); }
  FRHIP(const FRHIP &) = delete;
  FRHIP &operator=(const FRHIP &) = delete;

  void Synchronize() override { umpa_fr_synchronize(handle); }
  void Cleanup() override
  {
    AFR::Cleanup();
    if (handle) umpa_fr_cleanup(handle);
  }
  // kernel strategy: UMPA_FR_EXACT (bit-identical to FRCPU) or UMPA_FR_FAST_MATH
  bool SetStrategy(int strategy) { return umpa_fr_set_strategy(handle, strategy) == UMPA_OK; }
  int LastStatus() const { return lastStatus; }
  umpa_fr_op *Handle() const { return handle; }

protected:
  bool InitImpl(const OutputData &, const InputData &in, const Settings &s) override
  {
    const umpa_fr_desc d = Describe(in);
    const umpa_fr_settings cs = s.ToC();
    lastStatus = umpa_fr_init(handle, &d, &cs);
    return lastStatus == UMPA_OK;
  }
  bool ExecuteImpl(const OutputData &, const InputData &in) override
  {
    lastStatus = umpa_fr_execute(handle, in.GetFFT().GetPtr(), in.GetFFT().info.GetSize().n,
      in.GetVolume().GetPtr(), reinterpret_cast<float *>(in.GetWeight().GetPtr()),
      reinterpret_cast<const TraverseSpace *>(in.GetTraverseSpace().GetPtr()),
      in.GetTraverseSpace().info.GetSize().total,
      reinterpret_cast<const float *>(in.GetBlobTable().GetPtr()));
    return lastStatus == UMPA_OK;
  }

private:
  umpa_fr_op *handle = nullptr;
  int lastStatus = UMPA_OK;
};

// ComputeTraverseSpace (traverse_space_generator.hpp:94-145)
inline void ComputeTraverseSpace(uint32_t imgSizeX, uint32_t imgSizeY, const float transform[3][3],
  TraverseSpace &space, uint32_t maxVolumeIndexX, uint32_t maxVolumeIndexYZ, bool useFast,
  float blobRadius, float weight)
{
  umpa_fr_compute_traverse_space(imgSizeX, imgSizeY, &transform[0][0], &space, maxVolumeIndexX,
    maxVolumeIndexYZ, useFast ? 1 : 0, blobRadius, weight);
}

}  // namespace umpalumpa::fourier_reconstruction

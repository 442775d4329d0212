// Logical, Fourier and physical descriptors of a payload: the semantics of
// data/logical_desriptor.hpp:10-92, data/fourier_descriptor.hpp:13-202,
// data/physical_desriptor.hpp:13-99, data/data_type.hpp and data/managed_by.hpp.
#pragma once

#include <complex>
#include <optional>
#include <typeinfo>
#include <utility>

#include "size.hpp"

namespace umpalumpa::data {

class LogicalDescriptor
{
public:
  explicit LogicalDescriptor(const Size &s) : size(s), paddedSize(s) {}
  LogicalDescriptor(const Size &s, const PaddingDescriptor &p)
    : size(s), paddedSize(p.Pad(s)), padding(p)
  {}
  bool IsValid() const { return size.IsValid() && paddedSize >= size; }
  const Size &GetSize() const { return size; }
  const Size &GetPaddedSize() const { return paddedSize; }
  const PaddingDescriptor &GetPadding() const { return padding; }
  size_t Elems() const { return paddedSize.total; }
  bool IsEquivalentTo(const LogicalDescriptor &ref) const
  {
    return size.IsEquivalentTo(ref.size) && padding == ref.padding;
  }
  bool operator==(const LogicalDescriptor &o) const
  {
    return size == o.size && paddedSize == o.paddedSize && padding == o.padding;
  }

private:
  Size size;
  Size paddedSize;
  PaddingDescriptor padding;
};

class FourierDescriptor
{
public:
  struct FourierSpaceDescriptor {
    bool isCentered = false;
    bool isShifted = false;
    bool isNormalized = false;
    bool hasSymetry = false;  // true: the whole x extent is stored, else the x/2+1 half
    bool operator==(const FourierSpaceDescriptor &o) const
    {
      return isCentered == o.isCentered && isShifted == o.isShifted &&
             isNormalized == o.isNormalized && hasSymetry == o.hasSymetry;
    }
  };

  // data still in the spatial domain
  explicit FourierDescriptor(const Size &s) : FourierDescriptor(s, PaddingDescriptor()) {}
  FourierDescriptor(const Size &s, const PaddingDescriptor &p)
    : size(s), paddedSize(p.Pad(s)), frequencySize(Frequency(s, {})), padding(p),
      isSpatial(true)
  {}
  // data already in the frequency domain; `s` is the spatial size it came from
  FourierDescriptor(const Size &s, const PaddingDescriptor &p, const FourierSpaceDescriptor &d)
    : size(s), paddedSize(p.Pad(s)), frequencySize(Frequency(s, d)), isSpatial(false), fsd(d)
  {}

  const Size &GetSpatialSize() const { return size; }
  const Size &GetPaddedSpatialSize() const { return paddedSize; }
  const Size &GetFrequencySize() const { return frequencySize; }
  const Size &GetSize() const { return isSpatial ? size : frequencySize; }
  const Size &GetPaddedSize() const { return isSpatial ? paddedSize : frequencySize; }
  const std::optional<FourierSpaceDescriptor> &GetFourierSpaceDescriptor() const { return fsd; }
  double GetNormFactor() const { return 1.0 / static_cast<double>(paddedSize.single); }
  bool IsValid() const { return size.IsValid() && paddedSize >= size; }
  size_t Elems() const { return GetPaddedSize().total; }
  bool IsEquivalentTo(const FourierDescriptor &ref) const
  {
    return size.IsEquivalentTo(ref.size) && padding == ref.padding &&
           isSpatial == ref.isSpatial && fsd == ref.fsd;
  }
  bool operator==(const FourierDescriptor &o) const
  {
    return size == o.size && paddedSize == o.paddedSize && frequencySize == o.frequencySize &&
           padding == o.padding && isSpatial == o.isSpatial && fsd == o.fsd;
  }

private:
  static Size Frequency(const Size &s, const FourierSpaceDescriptor &d)
  {
    return Size(d.hasSymetry ? s.x : s.x / 2 + 1, s.y, s.z, s.n);
  }
  Size size;
  Size paddedSize;
  Size frequencySize;
  PaddingDescriptor padding;
  bool isSpatial;
  std::optional<FourierSpaceDescriptor> fsd;
};

class DataType
{
public:
  DataType(const std::type_info &info, size_t bytes) : hash(info.hash_code()), size(bytes) {}
  template<typename T> static DataType Get() { return DataType(typeid(T), sizeof(T)); }
  size_t GetSize() const { return size; }
  template<typename T> bool Is() const { return hash == typeid(T).hash_code(); }
  bool operator==(const DataType &o) const { return hash == o.hash; }

private:
  size_t hash;
  size_t size;
};
template<> inline DataType DataType::Get<void>() { return DataType(typeid(void), 0); }

enum class ManagedBy { External, CUDA, Manually, StarPU, Unknown };

// Where the data lives: a raw pointer and its size; no ownership, move-only.
class PhysicalDescriptor
{
public:
  PhysicalDescriptor(void *data, size_t b, DataType t, ManagedBy m, void *h)
    : ptr(data), bytes(b), type(t), manager(m), handle(h)
  {}
  PhysicalDescriptor(PhysicalDescriptor &&) = default;
  PhysicalDescriptor &operator=(PhysicalDescriptor &&) = default;
  size_t GetBytes() const { return bytes; }
  DataType GetType() const { return type; }
  ManagedBy GetManager() const { return manager; }
  void *GetHandle() const { return handle; }
  void *GetPtr() const { return ptr; }
  bool IsEmpty() const { return bytes == 0 && ptr == nullptr; }
  bool IsValid() const { return IsEmpty() || (ptr != nullptr && bytes != 0); }
  PhysicalDescriptor CopyWithoutData() const
  {
    return PhysicalDescriptor(nullptr, 0, type, manager, nullptr);
  }
  PhysicalDescriptor CopyWithPtr(void *p) const
  {
    return PhysicalDescriptor(p, bytes, type, manager, handle);
  }
  bool operator==(const PhysicalDescriptor &o) const
  {
    return ptr == o.ptr && bytes == o.bytes && type == o.type && manager == o.manager &&
           handle == o.handle;
  }

private:
  PhysicalDescriptor(const PhysicalDescriptor &) = default;
  void *ptr;
  size_t bytes;
  DataType type;
  ManagedBy manager;
  void *handle;
};

}  // namespace umpalumpa::data

// Payload = logical descriptor + physical descriptor, and the wrapper of several payloads used
// as an operation's input / output: the semantics of data/payload.hpp:12-105 and
// data/payload_wrapper.hpp:54-172.  The framework never allocates or frees payload memory.
#pragma once

#include <string>
#include <tuple>
#include <utility>

#include "descriptors.hpp"

namespace umpalumpa::data {

template<typename T> class Payload
{
public:
  Payload(const T &ld, PhysicalDescriptor &&pd, const std::string & = "")
    : info(ld), dataInfo(std::move(pd))
  {}
  bool IsEmpty() const { return dataInfo.IsEmpty(); }
  bool IsValid() const
  {
    const bool ok = info.IsValid() && dataInfo.IsValid();
    return IsEmpty() ? ok : ok && dataInfo.GetBytes() >= GetRequiredBytes();
  }
  Payload CopyWithoutData() const { return Payload(info, dataInfo.CopyWithoutData()); }
  bool IsEquivalentTo(const Payload &ref) const
  {
    return info.IsEquivalentTo(ref.info) && dataInfo.GetType() == ref.dataInfo.GetType();
  }
  bool operator==(const Payload &o) const { return info == o.info && dataInfo == o.dataInfo; }
  size_t GetRequiredBytes() const { return info.Elems() * dataInfo.GetType().GetSize(); }
  void *GetPtr() const { return dataInfo.GetPtr(); }

  T info;
  PhysicalDescriptor dataInfo;
  typedef T LDType;
};

// References to the payloads of one side of an operation.
template<typename... Args> struct PayloadWrapper
{
  typedef std::tuple<Args...> PayloadCollection;

  bool IsValid() const
  {
    return std::apply([](const auto &...p) { return (p.IsValid() && ...); }, payloads);
  }
  // The reference's wrapper folds its per-payload results with a C varargs helper that reads
  // only the first two (payload_wrapper.hpp:160-162), so Execute tolerates, e.g., a traverse-
  // space payload with more records than at Init.  Kept for drop-in compatibility.
  bool IsEquivalentTo(const PayloadWrapper &ref) const
  {
    bool ok = std::get<0>(payloads).IsEquivalentTo(std::get<0>(ref.payloads));
    if constexpr (sizeof...(Args) > 1)
      ok = ok && std::get<1>(payloads).IsEquivalentTo(std::get<1>(ref.payloads));
    return ok;
  }
  auto CopyWithoutData() const
  {
    return std::apply([](const auto &...p) { return std::make_tuple(p.CopyWithoutData()...); },
                      payloads);
  }

protected:
  explicit PayloadWrapper(Args &...args) : payloads(args...) {}
  explicit PayloadWrapper(std::tuple<Args...> &t)
    : payloads(std::apply([](Args &...p) { return std::tie(p...); }, t))
  {}
  const std::tuple<Args &...> payloads;
};

}  // namespace umpalumpa::data

// Operation life cycle: the semantics of BasicOperation<O, I, S>
// (operations/basic_operation.hpp:7-181).  Init stores data-less copies of the payloads and
// settings, checks validity and picks a working implementation; Execute requires a successful
// Init and payloads equivalent to the Init ones (N may shrink).
#pragma once

#include <memory>

namespace umpalumpa {

template<typename O, typename I, typename S> class BasicOperation
{
public:
  typedef O OutputData;
  typedef I InputData;
  typedef S Settings;

  virtual ~BasicOperation() = default;

  [[nodiscard]] bool Init(const OutputData &out, const InputData &in, const Settings &s)
  {
    this->Cleanup();
    emptyOut = std::make_unique<typename O::PayloadCollection>(out.CopyWithoutData());
    outputRef = std::make_unique<OutputData>(*emptyOut);
    emptyIn = std::make_unique<typename I::PayloadCollection>(in.CopyWithoutData());
    inputRef = std::make_unique<InputData>(*emptyIn);
    settings = std::make_unique<Settings>(s);
    if (!this->IsValid(*outputRef, *inputRef, s)) return false;
    if (this->InitImpl(out, in, s)) {
      initialized = true;
      return true;
    }
    this->Cleanup();
    return false;
  }

  [[nodiscard]] bool Execute(const OutputData &out, const InputData &in)
  {
    if (!initialized || !out.IsEquivalentTo(*outputRef) || !in.IsEquivalentTo(*inputRef))
      return false;
    return this->ExecuteImpl(out, in);
  }

  virtual void Cleanup()
  {
    outputRef.reset();
    inputRef.reset();
    settings.reset();
    initialized = false;
  }

  virtual void Synchronize() = 0;

  bool IsInitialized() const { return initialized; }
  const OutputData &GetOutputRef() const { return *outputRef; }
  const InputData &GetInputRef() const { return *inputRef; }
  const Settings &GetSettings() const { return *settings; }
  virtual size_t GetUsedBytes() const { return 0; }

protected:
  virtual bool IsValid(const OutputData &out, const InputData &in, const Settings &s) const = 0;
  virtual bool InitImpl(const OutputData &out, const InputData &in, const Settings &s) = 0;
  virtual bool ExecuteImpl(const OutputData &out, const InputData &in) = 0;

private:
  std::unique_ptr<typename O::PayloadCollection> emptyOut;
  std::unique_ptr<typename I::PayloadCollection> emptyIn;
  std::unique_ptr<OutputData> outputRef;
  std::unique_ptr<InputData> inputRef;
  std::unique_ptr<Settings> settings;
  bool initialized = false;
};

}  // namespace umpalumpa

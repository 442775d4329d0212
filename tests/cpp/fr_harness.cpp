// The reference's Fourier-reconstruction harness (examples/fourier_reconstruction/fr.cpp:9-76,
// 134-312) on this library's drop-in classes: for every batch of images, GenerateData ->
// GenerateTraverseSpaces (slice orientations times symmetry operators) -> ConvertToFFT (FFTHIP)
// -> Crop (FPHIP: centre, normalise, shift) -> InsertToVolume (FRHIP, Init and FillBlobTable on
// the first batch).  Payloads live in managed memory, as in the reference's CUDA-managed
// variant of the harness.  The last batch holds the remaining images (fewer than at Init).
//
// Usage: fr_harness OUTDIR [X Y N BATCH GROUP]   (defaults 32 32 20 8 C2)
// Writes the inputs and results as raw little-endian arrays to OUTDIR for the comparison
// (tests/test_cpp_api.py): meta.txt, images.f32 [N][Y][X], spaces.bin (112-byte records, each
// batch's projection indices relative to the batch), crops.c64 [N][Y][X/2] (the FP output the
// insertion read), volume.c64 and weights.f32 [Y+1][Y+1][X+1].
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "umpalumpa_amd/operations/fourier_processing/afp.hpp"
#include "umpalumpa_amd/operations/fourier_reconstruction/afr.hpp"
#include "umpalumpa_amd/operations/fourier_transformation/afft.hpp"

using namespace umpalumpa;
using namespace umpalumpa::data;
namespace fr = umpalumpa::fourier_reconstruction;
namespace fp = umpalumpa::fourier_processing;
namespace ft = umpalumpa::fourier_transformation;

namespace {

[[noreturn]] void fail(const char *what)
{
  std::fprintf(stderr, "fr_harness: %s\n", what);
  std::exit(1);
}

// managed allocations, zero-initialised, freed at exit
struct Managed {
  std::vector<void *> blocks;
  PhysicalDescriptor Create(size_t bytes, DataType type)
  {
    void *p = nullptr;
    if (bytes) {
      if (hipMallocManaged(&p, bytes) != hipSuccess) fail("hipMallocManaged");
      std::memset(p, 0, bytes);
      blocks.push_back(p);
    }
    return PhysicalDescriptor(p, bytes, type, ManagedBy::Manually, nullptr);
  }
  ~Managed()
  {
    (void)hipDeviceSynchronize();
    for (void *p : blocks) (void)hipFree(p);
  }
};

void append(const std::string &path, const void *data, size_t bytes)
{
  FILE *f = std::fopen(path.c_str(), "ab");
  if (!f || std::fwrite(data, 1, bytes, f) != bytes) fail(path.c_str());
  std::fclose(f);
}

// uniform rotation from a unit quaternion of four Gaussians (FP64)
void random_rotation(std::mt19937_64 &g, double m[9])
{
  std::normal_distribution<double> n(0.0, 1.0);
  double q[4], s = 0;
  for (double &v : q) {
    v = n(g);
    s += v * v;
  }
  s = 1.0 / std::sqrt(s);
  const double w = q[0] * s, x = q[1] * s, y = q[2] * s, z = q[3] * s;
  const double r[9] = { 1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
    2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
    2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y) };
  std::memcpy(m, r, sizeof(r));
}

}  // namespace

int main(int argc, char **argv)
{
  if (argc < 2) fail("usage: fr_harness OUTDIR [X Y N BATCH GROUP]");
  const std::string dir = argv[1];
  const size_t X = argc > 2 ? std::strtoul(argv[2], nullptr, 10) : 32;
  const size_t Y = argc > 3 ? std::strtoul(argv[3], nullptr, 10) : 32;
  const size_t N = argc > 4 ? std::strtoul(argv[4], nullptr, 10) : 20;
  const size_t B = argc > 5 ? std::strtoul(argv[5], nullptr, 10) : 8;
  const std::string group = argc > 6 ? argv[6] : "C2";
  if (X % 2 || X < 4 || Y < 4 || N < 1 || B < 1) fail("bad sizes");  // fr.cpp:15
  for (const char *name : { "images.f32", "spaces.bin", "crops.c64", "volume.c64", "weights.f32" })
    std::remove((dir + "/" + name).c_str());

  // symmetry operators (the reference draws random matrices here, fr.cpp:81-88)
  uint32_t nSym = 0;
  if (umpa_fr_symmetry_operators(group.c_str(), nullptr, 0, &nSym) != UMPA_OK) fail("group");
  std::vector<double> sym(9 * nSym);
  if (umpa_fr_symmetry_operators(group.c_str(), sym.data(), nSym, &nSym) != UMPA_OK) fail("group");

  Managed mem;
  fr::Settings settings;  // fr.cpp:24-26
  settings.SetType(fr::Settings::Type::kPrecise);
  settings.SetInterpolation(fr::Settings::Interpolation::kLookup);
  const Size volumeSize(X + 1, Y + 1, Y + 1, 1);  // fr.cpp:19
  FourierDescriptor::FourierSpaceDescriptor volFsd;
  volFsd.hasSymetry = true;
  FourierDescriptor volDesc(volumeSize, PaddingDescriptor(), volFsd);
  Payload<FourierDescriptor> volume(volDesc,
    mem.Create(volDesc.Elems() * 8, DataType::Get<std::complex<float>>()));
  LogicalDescriptor wDesc(volumeSize);
  Payload<LogicalDescriptor> weight(wDesc, mem.Create(wDesc.Elems() * 4, DataType::Get<float>()));
  LogicalDescriptor tDesc(Size(10000, 1, 1, 1));
  Payload<LogicalDescriptor> table(tDesc, mem.Create(tDesc.Elems() * 4, DataType::Get<float>()));
  Payload<LogicalDescriptor> filter(LogicalDescriptor(Size(X / 2, Y, 1, 1)),
    mem.Create(0, DataType::Get<void>()));  // fr.cpp:257-264: an empty filter

  ft::FFTHIP fftOp;
  fp::FPHIP cropOp;
  fr::FRHIP frOp;
  std::mt19937_64 g(2024);
  std::normal_distribution<float> pixel(0.f, 1.f);
  std::vector<size_t> chunks;

  for (size_t i = 0; i < N; i += B) {
    const size_t m = std::min(B, N - i);
    // CreatePayloadImage + GenerateData (fr.cpp:230-276)
    FourierDescriptor imgDesc(Size(X, Y, 1, m));
    Payload<FourierDescriptor> img(imgDesc, mem.Create(imgDesc.Elems() * 4, DataType::Get<float>()));
    auto *px = reinterpret_cast<float *>(img.GetPtr());
    for (size_t k = 0; k < X * Y * m; ++k) px[k] = pixel(g);
    append(dir + "/images.f32", px, X * Y * m * 4);
    // GenerateTraverseSpaces (fr.cpp:278-312): image width X/2 as in the harness
    LogicalDescriptor sDesc(Size(1, 1, 1, m * nSym));
    Payload<LogicalDescriptor> space(sDesc,
      mem.Create(sDesc.Elems() * sizeof(fr::TraverseSpace), DataType::Get<fr::TraverseSpace>()));
    auto *sp = reinterpret_cast<fr::TraverseSpace *>(space.GetPtr());
    for (size_t k = 0, counter = 0; k < m; ++k) {
      double r[9];
      random_rotation(g, r);
      for (uint32_t o = 0; o < nSym; ++o, ++counter) {
        const double *s = &sym[9 * o];
        float t[3][3];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            t[a][b] = (float)(r[3 * a] * s[b] + r[3 * a + 1] * s[3 + b] + r[3 * a + 2] * s[6 + b]);
        fr::ComputeTraverseSpace((uint32_t)(X / 2), (uint32_t)Y, t, sp[counter],
          (uint32_t)(volumeSize.x - 1), (uint32_t)(volumeSize.y - 1), false,
          settings.GetBlobRadius(), 1.f);
        sp[counter].projectionIndex = (int)k;
      }
    }
    append(dir + "/spaces.bin", sp, m * nSym * sizeof(fr::TraverseSpace));
    // ConvertToFFT (fr.cpp:134-164)
    FourierDescriptor fDesc(Size(X, Y, 1, m), PaddingDescriptor(),
      FourierDescriptor::FourierSpaceDescriptor{});
    Payload<FourierDescriptor> fft(fDesc,
      mem.Create(fDesc.Elems() * 8, DataType::Get<std::complex<float>>()));
    {
      auto out = ft::AFFT::OutputData(fft);
      auto in = ft::AFFT::InputData(img);
      if (!fftOp.IsInitialized() &&
          !fftOp.Init(out, in, ft::Settings(ft::Locality::kOutOfPlace, ft::Direction::kForward)))
        fail("FFT Init");
      if (!fftOp.Execute(out, in)) fail("FFT Execute");
    }
    // Crop (fr.cpp:166-203): spatial (X - 2) x Y, i.e. X/2 frequency columns
    FourierDescriptor cDesc(Size(X - 2, Y, 1, m), PaddingDescriptor(),
      FourierDescriptor::FourierSpaceDescriptor{});
    Payload<FourierDescriptor> crop(cDesc,
      mem.Create(cDesc.Elems() * 8, DataType::Get<std::complex<float>>()));
    {
      auto out = fp::AFP::OutputData(crop);
      auto in = fp::AFP::InputData(fft, filter);
      if (!cropOp.IsInitialized()) {
        fp::Settings s(ft::Locality::kOutOfPlace);
        s.SetApplyFilter(false);
        s.SetCenter(true);
        s.SetNormalize(true);
        s.SetShift(true);
        if (!cropOp.Init(out, in, s)) fail("Crop Init");
      }
      if (!cropOp.Execute(out, in)) fail("Crop Execute");
    }
    // InsertToVolume (fr.cpp:205-228)
    {
      auto in = fr::AFR::InputData(crop, volume, weight, space, table);
      auto out = fr::AFR::OutputData(volume, weight);
      if (!frOp.IsInitialized()) {
        if (!frOp.Init(out, in, settings)) fail("FR Init");
        fr::AFR::FillBlobTable(in, settings);
      }
      if (!frOp.Execute(out, in)) fail("FR Execute");
    }
    frOp.Synchronize();  // the crop payload is read on the host below
    append(dir + "/crops.c64", crop.GetPtr(), cDesc.Elems() * 8);
    chunks.push_back(m);
  }
  frOp.Synchronize();
  append(dir + "/volume.c64", volume.GetPtr(), volDesc.Elems() * 8);
  append(dir + "/weights.f32", weight.GetPtr(), wDesc.Elems() * 4);
  FILE *f = std::fopen((dir + "/meta.txt").c_str(), "w");
  if (!f) fail("meta.txt");
  std::fprintf(f, "%zu %zu %zu %zu %u %s\n", X, Y, N, B, nSym, group.c_str());
  for (size_t c : chunks) std::fprintf(f, "%zu\n", c);
  std::fclose(f);
  std::printf("fr_harness: %zu images in %zu batches, %u symmetry operators\n", N, chunks.size(),
    nSym);
  return 0;
}

// ORACLE build recipe helper -- test infrastructure only.
// Compiles the reference's own CPU insertion (fr_cpu_kernels.hpp), FP crop (fp_cpu_kernels.hpp)
// and Bessel polynomials (math/bessier.hpp) straight from the reference tree (-I<reference>) and
// exposes them with C linkage, so tests can compare the oracle restatement against the
// reference's compiled code.  Nothing from the reference is copied here; the headers are read
// at build time only, and the resulting library lives in oracle/_ref/ (never committed).
#include <cstddef>
#include <cstdint>
#include <complex>
#include <libumpalumpa/operations/fourier_reconstruction/fr_cpu_kernels.hpp>
#include <libumpalumpa/operations/fourier_processing/fp_cpu_kernels.hpp>
#include <libumpalumpa/math/bessier.hpp>
#include <libumpalumpa/utils/bool_expand.hpp>

using umpalumpa::fourier_reconstruction::BlobOrder;
using umpalumpa::fourier_reconstruction::Constants;
using umpalumpa::fourier_reconstruction::FR;
using umpalumpa::fourier_reconstruction::TraverseSpace;

static_assert(sizeof(TraverseSpace) == 112, "TraverseSpace layout");
static_assert(sizeof(Constants) == 32, "Constants layout");

extern "C" {

// Same dispatch as FRCPU Strategy1::Execute (fr_cpu.cpp:24-44); constants are an input.
int ref_fr_execute(int type, int interpolation, int order, float alpha, float *volume,
  float *weights, int xSize, int ySize, const float *fft, const void *spaces, size_t nSpaces,
  const float *table, size_t tableElems, const void *constants)
{
  if (interpolation == 1 && (table == nullptr || tableElems == 0)) return -1;
  umpalumpa::utils::ExpandBools<FR>::Expand(type == 0,
    alpha <= 15.f,
    interpolation == 1,
    static_cast<BlobOrder>(order),
    reinterpret_cast<std::complex<float> *>(volume),
    weights,
    xSize,
    ySize,
    reinterpret_cast<const std::complex<float> *>(fft),
    reinterpret_cast<const TraverseSpace *>(spaces),
    nSpaces,
    table,
    *reinterpret_cast<const Constants *>(constants));
  return 0;
}

// Same dispatch as FPCPU Strategy1::Execute (fp_cpu.cpp:25-46); sizes are {x, y, z, n}.
void ref_fp_execute(const float *in, float *out, const size_t *inSize, const size_t *inSpatial,
  const size_t *outSize, const float *filter, float normFactor, float maxFreqSquare,
  int applyFilter, int normalize, int center, int cropFreq, int shift)
{
  using umpalumpa::data::Size;
  Size si(inSize[0], inSize[1], inSize[2], inSize[3]);
  Size ss(inSpatial[0], inSpatial[1], inSpatial[2], inSpatial[3]);
  Size so(outSize[0], outSize[1], outSize[2], outSize[3]);
  umpalumpa::utils::ExpandBools<umpalumpa::fourier_processing::ScaleFFT2DCPU>::Expand(
    applyFilter != 0, normalize != 0, center != 0, cropFreq != 0, shift != 0,
    reinterpret_cast<const std::complex<float> *>(in),
    reinterpret_cast<std::complex<float> *>(out), si, ss, so, filter, normFactor, maxFreqSquare);
}

double ref_bessi0(double x) { return umpalumpa::math::bessi0(x); }
double ref_bessi1(double x) { return umpalumpa::math::bessi1(x); }
double ref_bessi2(double x) { return umpalumpa::math::bessi2(x); }
double ref_bessi3(double x) { return umpalumpa::math::bessi3(x); }
double ref_bessi4(double x) { return umpalumpa::math::bessi4(x); }
float ref_bessi1_5f(float x) { return umpalumpa::math::bessi1_5<float>(x); }
float ref_get_bessi_order_alpha(int order, float alpha)
{
  return umpalumpa::math::getBessiOrderAlpha(static_cast<BlobOrder>(order), alpha);
}
}

// Locality and Direction of a Fourier transform (fourier_transformation/locality.hpp,
// direction.hpp): the same enumerators as the reference.
#pragma once

namespace umpalumpa::fourier_transformation {
enum class Locality { kInPlace, kOutOfPlace };
enum class Direction { kForward, kInverse };
}  // namespace umpalumpa::fourier_transformation

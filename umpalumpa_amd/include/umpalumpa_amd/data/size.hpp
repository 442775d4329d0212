// Size of up to three dimensions [x, y, z] plus a count n: the semantics of
// umpalumpa::data::Size (data/size.hpp:23-103), written anew for this library.
#pragma once

#include <cstddef>
#include <iostream>

namespace umpalumpa::data {

enum class Dimensionality { kInvalid = 0, k1Dim = 1, k2Dim = 2, k3Dim = 3 };

class Size
{
public:
  static Dimensionality GetDim(size_t x, size_t y, size_t z)
  {
    if (x >= 2 && y >= 2 && z >= 2) return Dimensionality::k3Dim;
    if (x >= 2 && y >= 2 && z == 1) return Dimensionality::k2Dim;
    if (x >= 1 && y == 1 && z == 1) return Dimensionality::k1Dim;
    return Dimensionality::kInvalid;
  }

  explicit Size(size_t xSize, size_t ySize, size_t zSize, size_t nSize)
    : x(xSize), y(ySize), z(zSize), n(nSize), dim(GetDim(xSize, ySize, zSize)),
      single(xSize * ySize * zSize), total(xSize * ySize * zSize * nSize)
  {}

  bool IsValid() const { return n >= 1 && dim != Dimensionality::kInvalid; }
  Size CopyFor(size_t newN) const { return Size(x, y, z, newN); }
  Dimensionality GetDim() const { return dim; }

  // same extents; N may be smaller than the reference's
  bool IsEquivalentTo(const Size &ref) const
  {
    return x == ref.x && y == ref.y && z == ref.z && n <= ref.n;
  }
  bool operator==(const Size &o) const
  {
    return x == o.x && y == o.y && z == o.z && n == o.n;
  }
  bool operator!=(const Size &o) const { return !(*this == o); }
  bool operator<(const Size &o) const { return total < o.total; }
  bool operator>=(const Size &o) const { return !(*this < o); }

  void Serialize(std::ostream &out) const { out << x << ' ' << y << ' ' << z << ' ' << n << '\n'; }

  size_t x, y, z, n;
  Dimensionality dim;
  size_t single, total;
};

// Padding at the beginning / end of each dimension (data/padding_descriptor.hpp:22-66)
class PaddingDescriptor
{
public:
  PaddingDescriptor() = default;
  PaddingDescriptor(size_t xB, size_t xE, size_t yB, size_t yE, size_t zB, size_t zE)
    : xBeg(xB), xEnd(xE), yBeg(yB), yEnd(yE), zBeg(zB), zEnd(zE)
  {}
  size_t GetXBeg() const { return xBeg; }
  size_t GetXEnd() const { return xEnd; }
  size_t GetYBeg() const { return yBeg; }
  size_t GetYEnd() const { return yEnd; }
  size_t GetZBeg() const { return zBeg; }
  size_t GetZEnd() const { return zEnd; }
  bool operator==(const PaddingDescriptor &o) const
  {
    return xBeg == o.xBeg && xEnd == o.xEnd && yBeg == o.yBeg && yEnd == o.yEnd &&
           zBeg == o.zBeg && zEnd == o.zEnd;
  }
  Size Pad(const Size &s) const
  {
    return Size(s.x + xBeg + xEnd, s.y + yBeg + yEnd, s.z + zBeg + zEnd, s.n);
  }

private:
  size_t xBeg = 0, xEnd = 0, yBeg = 0, yEnd = 0, zBeg = 0, zEnd = 0;
};

}  // namespace umpalumpa::data

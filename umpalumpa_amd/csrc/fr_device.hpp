// Device-side building blocks of the FR insertion kernels (gfx950).
//
// All arithmetic that decides which pixels reach a voxel, and with which weight, is written to
// give the same float results as the reference CPU path (fr_cpu_kernels.hpp:70-240, compiled
// without FMA contraction): this translation unit is compiled with -ffp-contract=off, every
// expression keeps the reference's evaluation order, and fused multiply-adds appear only where
// the reference calls fmaf itself (bessi0Fast) or where a kernel variant explicitly asks for
// them (the fast-math strategy).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "launch.hpp"

namespace umpa {

__device__ __forceinline__ float clampf_ref(float v, float lo, float hi)
{
  // clamp<T,U> (fr_common_kernels.hpp:13-17), NaN passes through unchanged
  float r = (v > hi) ? hi : v;
  return (r < lo) ? lo : r;
}

// getZ/getY/getX (math/geometry.hpp:11-33)
__device__ __forceinline__ float plane_z(float x, float y, const umpa_point3 &n, const umpa_point3 &p)
{
  return (-n.x * (x - p.x) - n.y * (y - p.y)) / n.z + p.z;
}
__device__ __forceinline__ float plane_y(float x, float z, const umpa_point3 &n, const umpa_point3 &p)
{
  return (-n.x * (x - p.x) - n.z * (z - p.z)) / n.y + p.y;
}
__device__ __forceinline__ float plane_x(float y, float z, const umpa_point3 &n, const umpa_point3 &p)
{
  return (-n.y * (y - p.y) - n.z * (z - p.z)) / n.x + p.x;
}

// Does the reference traversal of `s` (FR::Execute<order>, fr_cpu_kernels.hpp:166-239, kPrecise)
// visit voxel (x, y, z)?  The traversal walks the AABB's plane orthogonal to `dir` and, per
// column, the slab between the two bounding planes clamped to the volume.
__device__ inline bool visited_blob(const umpa_traverse_space &s, int x, int y, int z,
                                    int maxX, int maxYZ)
{
  int a, b, e;
  float h1, h2, lim;
  if (s.dir == 0) {
    if (x < s.minX || x > s.maxX || y < s.minY || y > s.maxY) return false;
    h1 = plane_z((float)x, (float)y, s.unitNormal, s.bottomOrigin);
    h2 = plane_z((float)x, (float)y, s.unitNormal, s.topOrigin);
    lim = (float)maxYZ;
    e = z;
  } else if (s.dir == 1) {
    if (x < s.minX || x > s.maxX || z < s.minZ || z > s.maxZ) return false;
    h1 = plane_y((float)x, (float)z, s.unitNormal, s.bottomOrigin);
    h2 = plane_y((float)x, (float)z, s.unitNormal, s.topOrigin);
    lim = (float)maxYZ;
    e = y;
  } else {
    if (y < s.minY || y > s.maxY || z < s.minZ || z > s.maxZ) return false;
    h1 = plane_x((float)y, (float)z, s.unitNormal, s.bottomOrigin);
    h2 = plane_x((float)y, (float)z, s.unitNormal, s.topOrigin);
    lim = (float)maxX;
    e = x;
  }
  h1 = clampf_ref(h1, 0.f, lim);
  h2 = clampf_ref(h2, 0.f, lim);
  a = (int)floorf(fminf(h1, h2));
  b = (int)ceilf(fmaxf(h1, h2));
  return e >= a && e <= b;
}

// Same for the nearest-neighbour traversal (Type::kFast): one rounded voxel per column.
__device__ inline bool visited_nearest(const umpa_traverse_space &s, int x, int y, int z)
{
  if (s.dir == 0) {
    if (x < s.minX || x > s.maxX || y < s.minY || y > s.maxY) return false;
    return z == (int)(plane_z((float)x, (float)y, s.unitNormal, s.bottomOrigin) + 0.5f);
  } else if (s.dir == 1) {
    if (x < s.minX || x > s.maxX || z < s.minZ || z > s.maxZ) return false;
    return y == (int)(plane_y((float)x, (float)z, s.unitNormal, s.bottomOrigin) + 0.5f);
  }
  if (y < s.minY || y > s.maxY || z < s.minZ || z > s.maxZ) return false;
  return x == (int)(plane_x((float)y, (float)z, s.unitNormal, s.bottomOrigin) + 0.5f);
}

// ---- Kaiser-Bessel blob in float (fr_common_kernels.hpp:30-170) ------------------------------
__device__ inline float bessi0_fast(float x)
{
  // Blair & Edwards rational minimax approximation of I0 (x <= 15)
  float x2 = x * x;
  float num = -0.8436825781374849e-19f;
  num = __builtin_fmaf(num, x2, -0.93466495199548700e-17f);
  num = __builtin_fmaf(num, x2, -0.15716375332511895e-13f);
  num = __builtin_fmaf(num, x2, -0.42520971595532318e-11f);
  num = __builtin_fmaf(num, x2, -0.13704363824102120e-8f);
  num = __builtin_fmaf(num, x2, -0.28508770483148419e-6f);
  num = __builtin_fmaf(num, x2, -0.44322160233346062e-4f);
  num = __builtin_fmaf(num, x2, -0.46703811755736946e-2f);
  num = __builtin_fmaf(num, x2, -0.31112484643702141e-0f);
  num = __builtin_fmaf(num, x2, -0.11512633616429962e+2f);
  num = __builtin_fmaf(num, x2, -0.18720283332732112e+3f);
  num = __builtin_fmaf(num, x2, -0.75281108169006924e+3f);
  float den = __builtin_fmaf(1.f, x2, -0.75281109410939403e+3f);
  return num / den;
}

__device__ inline float bessi0_f(float x)
{
  float ax = fabsf(x), y, r;
  if (ax < 3.75f) {
    y = x / 3.75f;
    y *= y;
    r = 1.f + y * (3.5156229f + y * (3.0899424f + y * (1.2067492f + y * (0.2659732f + y * (0.360768e-1f + y * 0.45813e-2f)))));
  } else {
    y = 3.75f / ax;
    r = (expf(ax) * (1 / sqrtf(ax)))
        * (0.39894228f
           + y * (0.1328592e-1f
                  + y * (0.225319e-2f
                         + y * (-0.157565e-2f
                                + y * (0.916281e-2f + y * (-0.2057706e-1f + y * (0.2635537e-1f + y * (-0.1647633e-1f + y * 0.392377e-2f))))))));
  }
  return r;
}

__device__ inline float bessi1_f(float x)
{
  float ax = fabsf(x), y, r;
  if (ax < 3.75f) {
    y = x / 3.75f;
    y *= y;
    r = ax * (0.5f + y * (0.87890594f + y * (0.51498869f + y * (0.15084934f + y * (0.2658733e-1f + y * (0.301532e-2f + y * 0.32411e-3f))))));
  } else {
    y = 3.75f / ax;
    r = 0.2282967e-1f + y * (-0.2895312e-1f + y * (0.1787654e-1f - y * 0.420059e-2f));
    r = 0.39894228f + y * (-0.3988024e-1f + y * (-0.362018e-2f + y * (0.163801e-2f + y * (-0.1031555e-1f + y * r))));
    r *= (expf(ax) * (1 / sqrtf(ax)));
  }
  return x < 0.f ? -r : r;
}
__device__ inline float bessi2_f(float x) { return (x == 0) ? 0 : bessi0_f(x) - ((2 * 1) / x) * bessi1_f(x); }
__device__ inline float bessi3_f(float x) { return (x == 0) ? 0 : bessi1_f(x) - ((2 * 2) / x) * bessi2_f(x); }
__device__ inline float bessi4_f(float x) { return (x == 0) ? 0 : bessi2_f(x) - ((2 * 3) / x) * bessi3_f(x); }

template <int kOrder>
__device__ inline float kaiser_value(float r, float a, const umpa_fr_constants &c)
{
  float rda = r / a;
  if (!(rda <= 1.f)) return 0.f;
  float rdas = rda * rda;
  float arg = c.cBlobAlpha * sqrtf(1.f - rdas);
  float w;
  if constexpr (kOrder == 0) {
    w = bessi0_f(arg) * c.cOneOverBessiOrderAlpha;
  } else if constexpr (kOrder == 1) {
    w = sqrtf(1.f - rdas);
    w *= bessi1_f(arg) * c.cOneOverBessiOrderAlpha;
  } else if constexpr (kOrder == 2) {
    w = sqrtf(1.f - rdas);
    w = w * w;
    w *= bessi2_f(arg) * c.cOneOverBessiOrderAlpha;
  } else if constexpr (kOrder == 3) {
    w = sqrtf(1.f - rdas);
    w = w * w * w;
    w *= bessi3_f(arg) * c.cOneOverBessiOrderAlpha;
  } else {
    w = sqrtf(1.f - rdas);
    w = w * w * w * w;
    w *= bessi4_f(arg) * c.cOneOverBessiOrderAlpha;
  }
  return w;
}

__device__ inline float kaiser_value_fast(float distSqr, const umpa_fr_constants &c)
{
  float arg = c.cBlobAlpha * sqrtf(1.f - (distSqr * c.cOneOverBlobRadiusSqr));
  return bessi0_fast(arg) * c.cOneOverBessiOrderAlpha * c.cIw0;
}

}  // namespace umpa

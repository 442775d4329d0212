/*
 * ORACLE -- test infrastructure only (see umpa_oracle.h for the pinning of each function).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this file.
 *
 * Build: gcc -std=c11 -O2 -ffp-contract=off (no -ffast-math, no FMA contraction), the same
 * arithmetic model as the reference's plain x86-64 build.  The reference operates on float with
 * left-to-right evaluation; every expression below keeps that order so the float results are
 * the same bits.
 */
#include "umpa_oracle.h"

#include <complex.h>
#include <float.h>
#include <math.h>
#include <string.h>

static const float kPiF = 3.14159265358979323846f;

/* ---- modified Bessel functions, Abramowitz & Stegun 9.8.1-9.8.4 polynomials -------------
 * as used in math/bessier.hpp:9-69 (double precision). */
double orc_bessi0(double x)
{
  double ax = fabs(x);
  if (ax < 3.75) {
    double t = x / 3.75;
    t *= t;
    return 1.0
           + t * (3.5156229
                  + t * (3.0899424 + t * (1.2067492 + t * (0.2659732 + t * (0.360768e-1 + t * 0.45813e-2)))));
  }
  double t = 3.75 / ax;
  return (exp(ax) / sqrt(ax))
         * (0.39894228
            + t * (0.1328592e-1
                   + t * (0.225319e-2
                          + t * (-0.157565e-2
                                 + t * (0.916281e-2
                                        + t * (-0.2057706e-1
                                               + t * (0.2635537e-1 + t * (-0.1647633e-1 + t * 0.392377e-2))))))));
}

double orc_bessi1(double x)
{
  double ax = fabs(x);
  double res;
  if (ax < 3.75) {
    double t = x / 3.75;
    t *= t;
    res = ax
          * (0.5
             + t * (0.87890594
                    + t * (0.51498869 + t * (0.15084934 + t * (0.2658733e-1 + t * (0.301532e-2 + t * 0.32411e-3))))));
  } else {
    double t = 3.75 / ax;
    double tail = 0.2282967e-1 + t * (-0.2895312e-1 + t * (0.1787654e-1 - t * 0.420059e-2));
    res = 0.39894228 + t * (-0.3988024e-1 + t * (-0.362018e-2 + t * (0.163801e-2 + t * (-0.1031555e-1 + t * tail))));
    res *= (exp(ax) / sqrt(ax));
  }
  return x < 0.0 ? -res : res;
}

/* upward recurrence I_{n+1} = I_{n-1} - (2n/x) I_n, bessier.hpp:80-84 */
double orc_bessi2(double x) { return (x == 0) ? 0 : orc_bessi0(x) - ((2 * 1) / x) * orc_bessi1(x); }
double orc_bessi3(double x) { return (x == 0) ? 0 : orc_bessi1(x) - ((2 * 2) / x) * orc_bessi2(x); }
double orc_bessi4(double x) { return (x == 0) ? 0 : orc_bessi2(x) - ((2 * 3) / x) * orc_bessi3(x); }

/* I_{3/2} closed form in float, bessier.hpp:72-76 instantiated with T = float */
float orc_bessi1_5f(float x)
{
  if (x == 0) return 0;
  return sqrtf(2.f / (kPiF * x)) * (coshf(x) - sinhf(x) / x);
}

/* getBessiOrderAlpha, bessier.hpp:90-107 */
float orc_get_bessi_order_alpha(int order, float alpha)
{
  switch (order) {
  case 0: return (float)orc_bessi0(alpha);
  case 1: return (float)orc_bessi1(alpha);
  case 2: return (float)orc_bessi2(alpha);
  case 3: return (float)orc_bessi3(alpha);
  case 4: return (float)orc_bessi4(alpha);
  default: return (float)NAN;
  }
}

/* KaiserValue<float>, math/kaiser.hpp:19-56 */
float orc_kaiser_value(float freq, float blobRadius, float alpha, int order)
{
  float w = 0;
  float rda = freq / blobRadius;
  if (rda <= 1.f) {
    float rdas = rda * rda;
    float arg = alpha * sqrtf(1 - rdas);
    switch (order) {
    case 0: w = (float)(orc_bessi0(arg) / orc_bessi0(alpha)); break;
    case 1:
      w = sqrtf(1 - rdas);
      if (alpha != 0) w *= (float)(orc_bessi1(arg) / orc_bessi1(alpha));
      break;
    case 2:
      w = sqrtf(1 - rdas);
      w = w * w;
      if (alpha != 0) w *= (float)(orc_bessi2(arg) / orc_bessi2(alpha));
      break;
    case 3:
      w = sqrtf(1 - rdas);
      w = w * w * w;
      if (alpha != 0) w *= (float)(orc_bessi3(arg) / orc_bessi3(alpha));
      break;
    case 4:
      w = sqrtf(1 - rdas);
      w = w * w * w * w;
      if (alpha != 0) w *= (float)(orc_bessi4(arg) / orc_bessi4(alpha));
      break;
    default: break;
    }
  }
  return w;
}

/* KaiserFourierValue<float>, math/kaiser.hpp:66-81 (order 0 only, NaN otherwise) */
float orc_kaiser_fourier_value(float freq, float blobRadius, float alpha, int order)
{
  float a = 2.f * kPiF * blobRadius * freq;
  float sigma = sqrtf(fabsf(alpha * alpha - a * a));
  if (order == 0 && (2 * kPiF * blobRadius * freq <= alpha)) {
    return powf(2.f * kPiF, 1.5f) * powf(blobRadius, 3.f) * orc_bessi1_5f(sigma)
           / ((float)orc_bessi0(alpha) * powf(sigma, 1.5f));
  }
  return (float)NAN;
}

/* AFR::CreateConstants, fourier_reconstruction/afr.cpp:6-19 */
orc_constants orc_create_constants(size_t volumeX, size_t volumeY, size_t tableElems,
                                   const orc_fr_settings *s)
{
  orc_constants c;
  memset(&c, 0, sizeof(c));
  c.cMaxVolumeIndexX = (int)(volumeX - 1);
  c.cMaxVolumeIndexYZ = (int)(volumeY - 1);
  c.cBlobRadius = s->blobRadius;
  c.cOneOverBlobRadiusSqr = 1.f / (s->blobRadius * s->blobRadius);
  c.cBlobAlpha = s->alpha;
  c.cIw0 = 1.f / orc_kaiser_fourier_value(0.f, s->blobRadius, s->alpha, s->order);
  /* size_t arithmetic: an empty table wraps to 2^64 - 1, as in the reference */
  c.cIDeltaSqrt = (float)(tableElems - 1) / (s->blobRadius * s->blobRadius);
  c.cOneOverBessiOrderAlpha = 1 / orc_get_bessi_order_alpha(s->order, s->alpha);
  return c;
}

/* AFR::FillBlobTable, fourier_reconstruction/afr.cpp:22-34: the table is indexed by squared
 * distance, entry i holds the blob value at distance r*sqrt(i/(N-1)), times cIw0. */
void orc_fill_blob_table(float *table, size_t n, size_t volumeX, size_t volumeY,
                         const orc_fr_settings *s)
{
  float freq = s->blobRadius * sqrtf(1.f / (float)(n - 1));
  orc_constants c = orc_create_constants(volumeX, volumeY, n, s);
  for (size_t i = 0; i < n; ++i) {
    table[i] = orc_kaiser_value(freq * (float)sqrt((double)i), s->blobRadius, s->alpha, s->order)
               * c.cIw0;
  }
}

/* ---- traverse space generation, traverse_space_generator.hpp:9-145 ------------------------ */
static void orc_mul3(const float t[3][3], orc_point3 *p)
{
  float a = t[0][0] * p->x + t[0][1] * p->y + t[0][2] * p->z;
  float b = t[1][0] * p->x + t[1][1] * p->y + t[1][2] * p->z;
  float c = t[2][0] * p->x + t[2][1] * p->y + t[2][2] * p->z;
  p->x = a;
  p->y = b;
  p->z = c;
}

void orc_compute_traverse_space(uint32_t imgSizeX, uint32_t imgSizeY, const float transform[9],
                                orc_traverse_space *space, uint32_t maxVolumeIndexX,
                                uint32_t maxVolumeIndexYZ, int useFast, float blobRadius,
                                float weight)
{
  float t[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i][j] = transform[i * 3 + j];
  float b = useFast ? 0.f : blobRadius;
  float sx = (float)imgSizeX;
  float halfY = (float)imgSizeY / 2.0f;
  /* cuboid corners (createProjectionCuboid, :30-40): index -> (x side, y side, z side) */
  orc_point3 cub[8];
  const float xlo = 0.f - b, xhi = sx + b, ylo = -(halfY + b), yhi = halfY + b, zlo = 0.f - b,
              zhi = 0.f + b;
  cub[0] = (orc_point3){ xhi, ylo, zhi };
  cub[1] = (orc_point3){ xhi, yhi, zhi };
  cub[2] = (orc_point3){ xlo, yhi, zhi };
  cub[3] = (orc_point3){ xlo, ylo, zhi };
  cub[4] = (orc_point3){ xhi, ylo, zlo };
  cub[5] = (orc_point3){ xhi, yhi, zlo };
  cub[6] = (orc_point3){ xlo, yhi, zlo };
  cub[7] = (orc_point3){ xlo, ylo, zlo };
  orc_point3 origin = { (float)maxVolumeIndexX / 2.f, (float)maxVolumeIndexYZ / 2.f,
                        (float)maxVolumeIndexYZ / 2.f };
  for (int i = 0; i < 8; ++i) {
    orc_mul3((const float(*)[3])t, &cub[i]);
    cub[i].x += origin.x;
    cub[i].y += origin.y;
    cub[i].z += origin.z;
  }
  /* computeAABB (:62-84); the running maximum starts at FLT_MIN, the smallest positive normal */
  orc_point3 lo = { FLT_MAX, FLT_MAX, FLT_MAX }, hi = { FLT_MIN, FLT_MIN, FLT_MIN };
  for (int i = 0; i < 8; ++i) {
    if (lo.x > cub[i].x) lo.x = cub[i].x;
    if (lo.y > cub[i].y) lo.y = cub[i].y;
    if (lo.z > cub[i].z) lo.z = cub[i].z;
    if (hi.x < cub[i].x) hi.x = cub[i].x;
    if (hi.y < cub[i].y) hi.y = cub[i].y;
    if (hi.z < cub[i].z) hi.z = cub[i].z;
  }
  if (lo.x < 0) lo.x = 0;
  if (lo.y < 0) lo.y = 0;
  if (lo.z < 0) lo.z = 0;
  if (hi.x > (float)maxVolumeIndexX) hi.x = (float)maxVolumeIndexX;
  if (hi.y > (float)maxVolumeIndexYZ) hi.y = (float)maxVolumeIndexYZ;
  if (hi.z > (float)maxVolumeIndexYZ) hi.z = (float)maxVolumeIndexYZ;
  space->minZ = (int)floorf(lo.z);
  space->minY = (int)floorf(lo.y);
  space->minX = (int)floorf(lo.x);
  space->maxZ = (int)ceilf(hi.z);
  space->maxY = (int)ceilf(hi.y);
  space->maxX = (int)ceilf(hi.x);
  space->topOrigin = cub[4];
  space->bottomOrigin = cub[0];
  space->maxDistanceSqr = ((float)imgSizeX + b) * ((float)imgSizeX + b);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) space->transformInv[i][j] = t[j][i];
  /* plane normal = T * z, dir = plane orthogonal to its largest component (:128-142) */
  space->unitNormal = (orc_point3){ 0.f, 0.f, 1.f };
  orc_mul3((const float(*)[3])t, &space->unitNormal);
  float nX = fabsf(space->unitNormal.x), nY = fabsf(space->unitNormal.y),
        nZ = fabsf(space->unitNormal.z);
  if (nX >= nY && nX >= nZ) {
    space->dir = 2;
  } else if (nY >= nX && nY >= nZ) {
    space->dir = 1;
  } else if (nZ >= nX && nZ >= nY) {
    space->dir = 0;
  }
  space->weight = weight;
}

/* ---- FR insertion (FRCPU), fr_cpu_kernels.hpp:15-278 ------------------------------------- */

/* device-side Kaiser helpers in float, fr_common_kernels.hpp:30-170 */
static float orc_bessi0_fast(float x)
{
  /* Blair & Edwards rational minimax approximation (fr_common_kernels.hpp:30-53) */
  float x2 = x * x;
  float num = -0.8436825781374849e-19f;
  num = fmaf(num, x2, -0.93466495199548700e-17f);
  num = fmaf(num, x2, -0.15716375332511895e-13f);
  num = fmaf(num, x2, -0.42520971595532318e-11f);
  num = fmaf(num, x2, -0.13704363824102120e-8f);
  num = fmaf(num, x2, -0.28508770483148419e-6f);
  num = fmaf(num, x2, -0.44322160233346062e-4f);
  num = fmaf(num, x2, -0.46703811755736946e-2f);
  num = fmaf(num, x2, -0.31112484643702141e-0f);
  num = fmaf(num, x2, -0.11512633616429962e+2f);
  num = fmaf(num, x2, -0.18720283332732112e+3f);
  num = fmaf(num, x2, -0.75281108169006924e+3f);
  float den = 1.f;
  den = fmaf(den, x2, -0.75281109410939403e+3f);
  return num / den;
}

static float orc_bessi0f(float x)
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

static float orc_bessi1f(float x)
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
static float orc_bessi2f(float x) { return (x == 0) ? 0 : orc_bessi0f(x) - ((2 * 1) / x) * orc_bessi1f(x); }
static float orc_bessi3f(float x) { return (x == 0) ? 0 : orc_bessi1f(x) - ((2 * 2) / x) * orc_bessi2f(x); }
static float orc_bessi4f(float x) { return (x == 0) ? 0 : orc_bessi2f(x) - ((2 * 3) / x) * orc_bessi3f(x); }

/* kaiserValue<order>, fr_common_kernels.hpp:129-161 */
static float orc_kaiser_value_dev(float r, float a, const orc_constants *c, int order)
{
  float w;
  float rda = r / a;
  if (rda <= 1.f) {
    float rdas = rda * rda;
    float arg = c->cBlobAlpha * sqrtf(1.f - rdas);
    if (order == 0) {
      w = orc_bessi0f(arg) * c->cOneOverBessiOrderAlpha;
    } else if (order == 1) {
      w = sqrtf(1.f - rdas);
      w *= orc_bessi1f(arg) * c->cOneOverBessiOrderAlpha;
    } else if (order == 2) {
      w = sqrtf(1.f - rdas);
      w = w * w;
      w *= orc_bessi2f(arg) * c->cOneOverBessiOrderAlpha;
    } else if (order == 3) {
      w = sqrtf(1.f - rdas);
      w = w * w * w;
      w *= orc_bessi3f(arg) * c->cOneOverBessiOrderAlpha;
    } else if (order == 4) {
      w = sqrtf(1.f - rdas);
      w = w * w * w * w;
      w *= orc_bessi4f(arg) * c->cOneOverBessiOrderAlpha;
    } else {
      w = 0.f;
    }
  } else {
    w = 0.f;
  }
  return w;
}

/* kaiserValueFast, fr_common_kernels.hpp:163-170 */
static float orc_kaiser_value_fast(float distSqr, const orc_constants *c)
{
  float arg = c->cBlobAlpha * sqrtf(1.f - (distSqr * c->cOneOverBlobRadiusSqr));
  return orc_bessi0_fast(arg) * c->cOneOverBessiOrderAlpha * c->cIw0;
}

typedef struct {
  int mode;  /* 0 lookup, 1 fast kaiser, 2 precise kaiser */
  int order;
  int nearest; /* Settings::Type::kFast */
  int xSize, ySize;
  const float *table;
  size_t tableElems;
  const orc_constants *c;
  size_t vx, vy, vz; /* volume extents (cMaxVolumeIndex + 1) */
  long long updates;
} orc_fr_ctx;

/* ProcessVoxel (nearest neighbour), fr_cpu_kernels.hpp:18-68 */
static void orc_process_voxel(orc_fr_ctx *k, float *vol, float *wts, int x, int y, int z,
                              const float *fft, const orc_traverse_space *s)
{
  const orc_constants *c = k->c;
  float dataWeight = s->weight;
  orc_point3 p;
  p.x = (float)(x - c->cMaxVolumeIndexX / 2);
  p.y = (float)(y - c->cMaxVolumeIndexYZ / 2);
  p.z = (float)(z - c->cMaxVolumeIndexYZ / 2);
  if (p.x * p.x + p.y * p.y + p.z * p.z > s->maxDistanceSqr) return;
  orc_mul3(s->transformInv, &p);
  if (p.x < 0.f) return;
  int imgX = (int)(p.x + 0.5f);
  imgX = imgX > k->xSize - 1 ? k->xSize - 1 : imgX;
  imgX = imgX < 0 ? 0 : imgX;
  int imgY = (int)(p.y + 0.5f + (float)(c->cMaxVolumeIndexYZ / 2));
  imgY = imgY > k->ySize - 1 ? k->ySize - 1 : imgY;
  imgY = imgY < 0 ? 0 : imgY;
  /* the reference indexes without a bounds check; voxels outside the volume are skipped */
  if (x < 0 || y < 0 || z < 0 || (size_t)x >= k->vx || (size_t)y >= k->vy || (size_t)z >= k->vz)
    return;
  size_t i3 = ((size_t)z * k->vy + (size_t)y) * k->vx + (size_t)x;
  size_t i2 = (size_t)imgY * (size_t)k->xSize + (size_t)imgX;
  float w = 1.f * dataWeight;
  vol[2 * i3 + 0] = vol[2 * i3 + 0] + fft[2 * i2 + 0] * w;
  vol[2 * i3 + 1] = vol[2 * i3 + 1] + fft[2 * i2 + 1] * w;
  wts[i3] = wts[i3] + w;
  k->updates++;
}

/* ProcessVoxelBlob, fr_cpu_kernels.hpp:70-152 */
static void orc_process_voxel_blob(orc_fr_ctx *k, float *vol, float *wts, int x, int y, int z,
                                   const float *fft, const orc_traverse_space *s)
{
  const orc_constants *c = k->c;
  orc_point3 p;
  p.x = (float)(x - c->cMaxVolumeIndexX / 2);
  p.y = (float)(y - c->cMaxVolumeIndexYZ / 2);
  p.z = (float)(z - c->cMaxVolumeIndexYZ / 2);
  if ((p.x * p.x + p.y * p.y + p.z * p.z) > s->maxDistanceSqr) return;
  orc_mul3(s->transformInv, &p);
  if (p.x < -c->cBlobRadius) return;
  p.y += (float)(c->cMaxVolumeIndexYZ / 2);
  float radiusSqr = c->cBlobRadius * c->cBlobRadius;
  float zSqr = p.z * p.z;
  if (zSqr > radiusSqr) return;
  int minX = (int)ceilf(p.x - c->cBlobRadius);
  int maxX = (int)floorf(p.x + c->cBlobRadius);
  int minY = (int)ceilf(p.y - c->cBlobRadius);
  int maxY = (int)floorf(p.y + c->cBlobRadius);
  if (minX < 0) minX = 0;
  if (minY < 0) minY = 0;
  if (maxX > k->xSize - 1) maxX = k->xSize - 1;
  if (maxY > k->ySize - 1) maxY = k->ySize - 1;
  float w = 0.f, vr = 0.f, vi = 0.f;
  float dataWeight = s->weight;
  for (int i = minY; i <= maxY; i++) {
    float ySqr = (p.y - (float)i) * (p.y - (float)i);
    float yzSqr = ySqr + zSqr;
    if (yzSqr > radiusSqr) continue;
    for (int j = minX; j <= maxX; j++) {
      float xD = p.x - (float)j;
      float distanceSqr = xD * xD + yzSqr;
      if (distanceSqr > radiusSqr) continue;
      size_t i2 = (size_t)i * (size_t)k->xSize + (size_t)j;
      float wBlob;
      if (k->mode == 0) {
        int aux = (int)((distanceSqr * c->cIDeltaSqrt + 0.5f));
        wBlob = k->table[aux];
      } else if (k->mode == 1) {
        wBlob = orc_kaiser_value_fast(distanceSqr, c);
      } else {
        wBlob = orc_kaiser_value_dev(sqrtf(distanceSqr), c->cBlobRadius, c, k->order) * c->cIw0;
      }
      float weight = wBlob * dataWeight;
      w += weight;
      vr += fft[2 * i2 + 0] * weight;
      vi += fft[2 * i2 + 1] * weight;
    }
  }
  size_t i3 = ((size_t)z * k->vy + (size_t)y) * k->vx + (size_t)x;
  vol[2 * i3 + 0] = vol[2 * i3 + 0] + vr;
  vol[2 * i3 + 1] = vol[2 * i3 + 1] + vi;
  wts[i3] = wts[i3] + w;
  k->updates++;
}

static float orc_clampf(float v, float lo, float hi)
{
  float r = (v > hi) ? hi : v;
  return (r < lo) ? lo : r;
}

/* getZ/getY/getX, math/geometry.hpp:11-33: plane through p0 with normal n */
static float orc_get_z(float x, float y, orc_point3 n, orc_point3 p0)
{
  return (-n.x * (x - p0.x) - n.y * (y - p0.y)) / n.z + p0.z;
}
static float orc_get_y(float x, float z, orc_point3 n, orc_point3 p0)
{
  return (-n.x * (x - p0.x) - n.z * (z - p0.z)) / n.y + p0.y;
}
static float orc_get_x(float y, float z, orc_point3 n, orc_point3 p0)
{
  return (-n.y * (y - p0.y) - n.z * (z - p0.z)) / n.x + p0.x;
}

/* FR::Execute<blobOrder> plane/slab traversal, fr_cpu_kernels.hpp:154-240 */
static void orc_traverse(orc_fr_ctx *k, float *vol, float *wts, const float *fft,
                         const orc_traverse_space *s)
{
  const orc_constants *c = k->c;
  switch (s->dir) {
  case 0:
    for (int idy = s->minY; idy <= s->maxY; idy++) {
      float fy = (float)idy;
      for (int idx = s->minX; idx <= s->maxX; idx++) {
        float fx = (float)idx;
        if (k->nearest) {
          float hit = orc_get_z(fx, fy, s->unitNormal, s->bottomOrigin);
          orc_process_voxel(k, vol, wts, idx, idy, (int)(hit + 0.5f), fft, s);
        } else {
          float a = orc_clampf(orc_get_z(fx, fy, s->unitNormal, s->bottomOrigin), 0.f,
                               (float)c->cMaxVolumeIndexYZ);
          float b = orc_clampf(orc_get_z(fx, fy, s->unitNormal, s->topOrigin), 0.f,
                               (float)c->cMaxVolumeIndexYZ);
          int lower = (int)floorf(fminf(a, b));
          int upper = (int)ceilf(fmaxf(a, b));
          for (int z = lower; z <= upper; z++) orc_process_voxel_blob(k, vol, wts, idx, idy, z, fft, s);
        }
      }
    }
    break;
  case 1:
    for (int idy = s->minZ; idy <= s->maxZ; idy++) {
      float fz = (float)idy;
      for (int idx = s->minX; idx <= s->maxX; idx++) {
        float fx = (float)idx;
        if (k->nearest) {
          float hit = orc_get_y(fx, fz, s->unitNormal, s->bottomOrigin);
          orc_process_voxel(k, vol, wts, idx, (int)(hit + 0.5f), idy, fft, s);
        } else {
          float a = orc_clampf(orc_get_y(fx, fz, s->unitNormal, s->bottomOrigin), 0.f,
                               (float)c->cMaxVolumeIndexYZ);
          float b = orc_clampf(orc_get_y(fx, fz, s->unitNormal, s->topOrigin), 0.f,
                               (float)c->cMaxVolumeIndexYZ);
          int lower = (int)floorf(fminf(a, b));
          int upper = (int)ceilf(fmaxf(a, b));
          for (int y = lower; y <= upper; y++) orc_process_voxel_blob(k, vol, wts, idx, y, idy, fft, s);
        }
      }
    }
    break;
  case 2:
    for (int idy = s->minZ; idy <= s->maxZ; idy++) {
      float fz = (float)idy;
      for (int idx = s->minY; idx <= s->maxY; idx++) {
        float fy = (float)idx;
        if (k->nearest) {
          float hit = orc_get_x(fy, fz, s->unitNormal, s->bottomOrigin);
          orc_process_voxel(k, vol, wts, (int)(hit + 0.5f), idx, idy, fft, s);
        } else {
          float a = orc_clampf(orc_get_x(fy, fz, s->unitNormal, s->bottomOrigin), 0.f,
                               (float)c->cMaxVolumeIndexX);
          float b = orc_clampf(orc_get_x(fy, fz, s->unitNormal, s->topOrigin), 0.f,
                               (float)c->cMaxVolumeIndexX);
          int lower = (int)floorf(fminf(a, b));
          int upper = (int)ceilf(fmaxf(a, b));
          for (int x = lower; x <= upper; x++) orc_process_voxel_blob(k, vol, wts, x, idx, idy, fft, s);
        }
      }
    }
    break;
  default: break;
  }
}

/* FRCPU Strategy1::Execute (fr_cpu.cpp:24-44) + FR::Execute (fr_cpu_kernels.hpp:242-277) */
int orc_fr_execute(const orc_fr_settings *s, float *volume, float *weights, int xSize, int ySize,
                   const float *fft, const orc_traverse_space *spaces, size_t nSpaces,
                   const float *table, size_t tableElems, const orc_constants *c,
                   long long *updates)
{
  if (s->interpolation == 1 && (table == NULL || tableElems == 0)) return -1;
  if (s->order < 0 || s->order > 4) return 0; /* unsupported order: the reference does nothing */
  orc_fr_ctx k;
  memset(&k, 0, sizeof(k));
  k.mode = (s->interpolation == 1) ? 0 : (s->alpha <= 15.f ? 1 : 2);
  k.order = s->order;
  k.nearest = (s->type == 0);
  k.xSize = xSize;
  k.ySize = ySize;
  k.table = table;
  k.tableElems = tableElems;
  k.c = c;
  k.vx = (size_t)c->cMaxVolumeIndexX + 1;
  k.vy = (size_t)c->cMaxVolumeIndexYZ + 1;
  k.vz = (size_t)c->cMaxVolumeIndexYZ + 1;
  for (size_t i = 0; i < nSpaces; ++i) {
    const orc_traverse_space *sp = &spaces[i];
    const float *f = fft + 2 * (size_t)xSize * (size_t)ySize * (size_t)sp->projectionIndex;
    orc_traverse(&k, volume, weights, f, sp);
  }
  if (updates) *updates = k.updates;
  return 0;
}

/* ---- FP64 host oracle: the same traversal and predicates with double intermediates -------- */
typedef struct { double x, y, z; } orc_d3;

static void orc_blob_f64(const orc_fr_ctx *k, double *vol, double *wts, int x, int y, int z,
                         const float *fft, const orc_traverse_space *s)
{
  const orc_constants *c = k->c;
  double px = (double)(x - c->cMaxVolumeIndexX / 2);
  double py = (double)(y - c->cMaxVolumeIndexYZ / 2);
  double pz = (double)(z - c->cMaxVolumeIndexYZ / 2);
  if (px * px + py * py + pz * pz > (double)s->maxDistanceSqr) return;
  const float(*t)[3] = s->transformInv;
  double qx = (double)t[0][0] * px + (double)t[0][1] * py + (double)t[0][2] * pz;
  double qy = (double)t[1][0] * px + (double)t[1][1] * py + (double)t[1][2] * pz;
  double qz = (double)t[2][0] * px + (double)t[2][1] * py + (double)t[2][2] * pz;
  double r = (double)c->cBlobRadius;
  if (qx < -r) return;
  qy += (double)(c->cMaxVolumeIndexYZ / 2);
  double r2 = r * r, zSqr = qz * qz;
  if (zSqr > r2) return;
  int minX = (int)ceil(qx - r), maxX = (int)floor(qx + r);
  int minY = (int)ceil(qy - r), maxY = (int)floor(qy + r);
  if (minX < 0) minX = 0;
  if (minY < 0) minY = 0;
  if (maxX > k->xSize - 1) maxX = k->xSize - 1;
  if (maxY > k->ySize - 1) maxY = k->ySize - 1;
  double w = 0, vr = 0, vi = 0, dw = (double)s->weight, cid = (double)c->cIDeltaSqrt;
  for (int i = minY; i <= maxY; i++) {
    double yz = (qy - i) * (qy - i) + zSqr;
    if (yz > r2) continue;
    for (int j = minX; j <= maxX; j++) {
      double d2 = (qx - j) * (qx - j) + yz;
      if (d2 > r2) continue;
      long aux = (long)(d2 * cid + 0.5);
      if (aux > (long)k->tableElems - 1) aux = (long)k->tableElems - 1;
      double wt = (double)k->table[aux] * dw;
      size_t i2 = (size_t)i * (size_t)k->xSize + (size_t)j;
      w += wt;
      vr += (double)fft[2 * i2] * wt;
      vi += (double)fft[2 * i2 + 1] * wt;
    }
  }
  size_t i3 = ((size_t)z * k->vy + (size_t)y) * k->vx + (size_t)x;
  vol[2 * i3] += vr;
  vol[2 * i3 + 1] += vi;
  wts[i3] += w;
}

int orc_fr_execute_f64(const orc_fr_settings *s, double *volume, double *weights, int xSize,
                       int ySize, const float *fft, const orc_traverse_space *spaces,
                       size_t nSpaces, const float *table, size_t tableElems,
                       const orc_constants *c)
{
  if (s->interpolation != 1 || s->type != 1 || table == NULL || tableElems == 0) return -1;
  orc_fr_ctx k;
  memset(&k, 0, sizeof(k));
  k.xSize = xSize;
  k.ySize = ySize;
  k.table = table;
  k.tableElems = tableElems;
  k.c = c;
  k.vx = (size_t)c->cMaxVolumeIndexX + 1;
  k.vy = (size_t)c->cMaxVolumeIndexYZ + 1;
  k.vz = (size_t)c->cMaxVolumeIndexYZ + 1;
  for (size_t n = 0; n < nSpaces; ++n) {
    const orc_traverse_space *sp = &spaces[n];
    const float *f = fft + 2 * (size_t)xSize * (size_t)ySize * (size_t)sp->projectionIndex;
    /* traverse in double: the column's slab bounds from the two bounding planes */
    orc_d3 nn = { sp->unitNormal.x, sp->unitNormal.y, sp->unitNormal.z };
    orc_d3 b0 = { sp->bottomOrigin.x, sp->bottomOrigin.y, sp->bottomOrigin.z };
    orc_d3 t0 = { sp->topOrigin.x, sp->topOrigin.y, sp->topOrigin.z };
    int d = sp->dir;
    int aLo = d == 0 ? sp->minY : sp->minZ, aHi = d == 0 ? sp->maxY : sp->maxZ;
    int bLo = d == 2 ? sp->minY : sp->minX, bHi = d == 2 ? sp->maxY : sp->maxX;
    double lim = (double)(d == 2 ? c->cMaxVolumeIndexX : c->cMaxVolumeIndexYZ);
    for (int a = aLo; a <= aHi; a++) {
      for (int bb = bLo; bb <= bHi; bb++) {
        double h1, h2;
        if (d == 0) {
          h1 = (-nn.x * (bb - b0.x) - nn.y * (a - b0.y)) / nn.z + b0.z;
          h2 = (-nn.x * (bb - t0.x) - nn.y * (a - t0.y)) / nn.z + t0.z;
        } else if (d == 1) {
          h1 = (-nn.x * (bb - b0.x) - nn.z * (a - b0.z)) / nn.y + b0.y;
          h2 = (-nn.x * (bb - t0.x) - nn.z * (a - t0.z)) / nn.y + t0.y;
        } else {
          h1 = (-nn.y * (bb - b0.y) - nn.z * (a - b0.z)) / nn.x + b0.x;
          h2 = (-nn.y * (bb - t0.y) - nn.z * (a - t0.z)) / nn.x + t0.x;
        }
        h1 = h1 < 0 ? 0 : (h1 > lim ? lim : h1);
        h2 = h2 < 0 ? 0 : (h2 > lim ? lim : h2);
        int lower = (int)floor(h1 < h2 ? h1 : h2), upper = (int)ceil(h1 < h2 ? h2 : h1);
        for (int e = lower; e <= upper; e++) {
          if (d == 0) orc_blob_f64(&k, volume, weights, bb, a, e, f, sp);
          else if (d == 1) orc_blob_f64(&k, volume, weights, bb, e, a, f, sp);
          else orc_blob_f64(&k, volume, weights, e, bb, a, f, sp);
        }
      }
    }
  }
  return 0;
}

/* ---- FP crop / normalize / center / shift, fp_cpu_kernels.hpp:8-41 ----------------------- */
float orc_idx2freq(size_t idx, size_t size)
{
  if (size <= 1) return 0;
  float tmp = (idx <= (size / 2)) ? (float)idx : ((float)idx - (float)size);
  return tmp / (float)size;
}

void orc_fp_execute(const float *in, float *out, const size_t inSize[4],
                    const size_t inSpatialSize[4], const size_t outSize[4], const float *filter,
                    float normFactor, float maxFreqSquare, int applyFilter, int normalize,
                    int center, int cropFreq, int shift)
{
  size_t inSingle = inSize[0] * inSize[1] * inSize[2];
  size_t outSingle = outSize[0] * outSize[1] * outSize[2];
  for (size_t n = 0; n < inSize[3]; ++n) {
    for (size_t y = 0; y < outSize[1]; ++y) {
      for (size_t x = 0; x < outSize[0]; ++x) {
        size_t origY = (y <= outSize[1] / 2) ? y : (inSize[1] - (outSize[1] - y));
        size_t iIndex = n * inSingle + origY * inSize[0] + x;
        size_t tmp = y + outSize[1] / 2;
        size_t shiftedY = tmp >= outSize[1] ? tmp - outSize[1] : tmp;
        size_t oIndex = n * outSingle + (shift ? shiftedY : y) * outSize[0] + x;
        float fx = orc_idx2freq(x, inSpatialSize[0]);
        float fy = orc_idx2freq(origY, inSpatialSize[1]);
        /* std::norm(std::complex<float>) in libstdc++ squares std::abs, i.e. cabsf */
        float a = cabsf(fx + fy * I);
        float nrm = a * a;
        float re, im;
        if (cropFreq && (nrm > maxFreqSquare)) {
          re = 0.f;
          im = 0.f;
        } else {
          re = in[2 * iIndex];
          im = in[2 * iIndex + 1];
        }
        if (applyFilter) {
          float f = filter[y * outSize[0] + x];
          re *= f;
          im *= f;
        }
        if (normalize) {
          re *= normFactor;
          im *= normFactor;
        }
        if (center) {
          float cc = (float)(1 - 2 * (((int)x + (int)y) & 1));
          re *= cc;
          im *= cc;
        }
        out[2 * oIndex] = re;
        out[2 * oIndex + 1] = im;
      }
    }
  }
}

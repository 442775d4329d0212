// FR insertion kernels for gfx950: voxel-owned gather.
//
// The reference scatters each rotated slice into the volume with three float atomics per voxel
// (fr_cuda_kernels.cu:266-365).  Here a wavefront owns a brick of 4 x 4 x 8 voxels (two per
// lane) for the whole launch: it keeps their running (complex, weight) sums in registers, culls
// the traverse spaces against the brick 64 at a time with one ballot, and walks the surviving
// spaces in index order.  Each voxel therefore receives its per-space contributions in exactly
// the order of the reference CPU loop over spaces (fr_cpu_kernels.hpp:254), with no atomics and
// no dependence on how bricks are scheduled.  Workgroups are persistent and take bricks from a
// counter, heaviest (most central) bricks first; the blob table is staged once per workgroup in
// LDS.
#include "fr_device.hpp"

namespace umpa {

constexpr int kMaxLdsTable = 16384;  // floats; larger tables are read from global memory
constexpr int kThreads = 256;

// Slice pixel fetch: FP32 complex, or FP16 complex (scaled once per voxel-update).
template <bool kHalf>
__device__ __forceinline__ float2 load_pixel(const void *slice, int64_t i)
{
  if constexpr (kHalf) {
    __half2 h = reinterpret_cast<const __half2 *>(slice)[i];
    return __half22float2(h);
  } else {
    return reinterpret_cast<const float2 *>(slice)[i];
  }
}

// One (voxel, traverse space) pair: the reference's ProcessVoxelBlob / ProcessVoxel
// (fr_cpu_kernels.hpp:18-152) preceded by the traversal-membership test, accumulated into the
// voxel's running sums.  Returns true when the pair reaches the accumulation step.
template <int kMode, int kOrder, int kNB, bool kHalf, bool kFastMath>
__device__ __forceinline__ bool insert_pair(const FrArgs &a, const umpa_traverse_space &s,
                                            const void *slice, float scale, const float *tbl,
                                            int x, int y, int z, float &accR, float &accI,
                                            float &accW)
{
  const umpa_fr_constants &c = a.c;
  float px = (float)(x - c.cMaxVolumeIndexX / 2);
  float py = (float)(y - c.cMaxVolumeIndexYZ / 2);
  float pz = (float)(z - c.cMaxVolumeIndexYZ / 2);
  if ((px * px + py * py + pz * pz) > s.maxDistanceSqr) return false;
  float qx = s.transformInv[0][0] * px + s.transformInv[0][1] * py + s.transformInv[0][2] * pz;
  float qy = s.transformInv[1][0] * px + s.transformInv[1][1] * py + s.transformInv[1][2] * pz;
  float qz = s.transformInv[2][0] * px + s.transformInv[2][1] * py + s.transformInv[2][2] * pz;

  if constexpr (kMode == kModeNearest) {
    if (qx < 0.f) return false;
    if (!visited_nearest(s, x, y, z)) return false;
    int ix = (int)(qx + 0.5f);
    ix = ix > a.xSize - 1 ? a.xSize - 1 : ix;
    ix = ix < 0 ? 0 : ix;
    int iy = (int)(qy + 0.5f + (float)(c.cMaxVolumeIndexYZ / 2));
    iy = iy > a.ySize - 1 ? a.ySize - 1 : iy;
    iy = iy < 0 ? 0 : iy;
    float w = 1.f * s.weight;
    float2 f = load_pixel<kHalf>(slice, (int64_t)iy * a.xSize + ix);
    if constexpr (kHalf) {
      f.x *= scale;
      f.y *= scale;
    }
    accR = accR + f.x * w;
    accI = accI + f.y * w;
    accW = accW + w;
    return true;
  } else {
    const float r = c.cBlobRadius;
    if (qx < -r) return false;
    qy += (float)(c.cMaxVolumeIndexYZ / 2);
    const float r2 = r * r;
    const float zSqr = qz * qz;
    if (zSqr > r2) return false;
    if (!visited_blob(s, x, y, z, c.cMaxVolumeIndexX, c.cMaxVolumeIndexYZ)) return false;

    int minX = (int)ceilf(qx - r), maxX = (int)floorf(qx + r);
    int minY = (int)ceilf(qy - r), maxY = (int)floorf(qy + r);
    minX = minX < 0 ? 0 : minX;
    minY = minY < 0 ? 0 : minY;
    maxX = maxX > a.xSize - 1 ? a.xSize - 1 : maxX;
    maxY = maxY > a.ySize - 1 ? a.ySize - 1 : maxY;
    const float dw = s.weight;
    float w = 0.f, vr = 0.f, vi = 0.f;

    auto pixel = [&](int64_t rowBase, float yz, int j) {
      float xD = qx - (float)j;
      float d2 = xD * xD + yz;
      if (d2 > r2) return;
      float wb;
      if constexpr (kMode == kModeLookup) {
        int aux = kFastMath ? (int)__builtin_fmaf(d2, c.cIDeltaSqrt, 0.5f)
                            : (int)(d2 * c.cIDeltaSqrt + 0.5f);
        aux = aux > a.nTable - 1 ? a.nTable - 1 : aux;
        wb = tbl[aux];
      } else if constexpr (kMode == kModeFastKaiser) {
        wb = kaiser_value_fast(d2, c);
      } else {
        wb = kaiser_value<kOrder>(sqrtf(d2), r, c) * c.cIw0;
      }
      float weight = wb * dw;
      float2 f = load_pixel<kHalf>(slice, rowBase + j);
      w += weight;
      if constexpr (kFastMath) {
        vr = __builtin_fmaf(f.x, weight, vr);
        vi = __builtin_fmaf(f.y, weight, vi);
      } else {
        vr += f.x * weight;
        vi += f.y * weight;
      }
    };
    auto row = [&](int i) {
      float dy = qy - (float)i;
      float yz = dy * dy + zSqr;
      if (yz > r2) return;
      int64_t rowBase = (int64_t)i * a.xSize;
      if constexpr (kNB > 0) {
#pragma unroll
        for (int jj = 0; jj < kNB; ++jj)
          if (minX + jj <= maxX) pixel(rowBase, yz, minX + jj);
        for (int j = minX + kNB; j <= maxX; ++j) pixel(rowBase, yz, j);
      } else {
        for (int j = minX; j <= maxX; ++j) pixel(rowBase, yz, j);
      }
    };
    if constexpr (kNB > 0) {
#pragma unroll
      for (int ii = 0; ii < kNB; ++ii)
        if (minY + ii <= maxY) row(minY + ii);
      for (int i = minY + kNB; i <= maxY; ++i) row(i);
    } else {
      for (int i = minY; i <= maxY; ++i) row(i);
    }
    if constexpr (kHalf) {
      vr *= scale;
      vi *= scale;
    }
    accR = accR + vr;
    accI = accI + vi;
    accW = accW + w;
    return true;
  }
}

template <int kMode, int kOrder, int kNB, bool kHalf, bool kFastMath, bool kCount, bool kTableLds>
__global__ __launch_bounds__(kThreads) void fr_insert_kernel(FrArgs a)
{
  extern __shared__ float s_table[];
  if constexpr (kMode == kModeLookup && kTableLds) {
    for (int i = threadIdx.x; i < a.nTable; i += kThreads) s_table[i] = a.table[i];
    __syncthreads();
  }
  const float *tbl = (kMode == kModeLookup && kTableLds) ? s_table : a.table;
  const int lane = threadIdx.x & 63;
  const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
  const int cxi = a.c.cMaxVolumeIndexX / 2, cyi = a.c.cMaxVolumeIndexYZ / 2;
  const float brickRadius = 4.0927375f;  // |(1.5, 1.5, 3.5)|, half diagonal of a brick

  for (;;) {
    int b = 0;
    if (lane == 0) b = atomicAdd(a.counter, 1);
    b = __builtin_amdgcn_readfirstlane(__shfl(b, 0));
    if (b >= a.nBricks) break;
    const int4 bo = reinterpret_cast<const int4 *>(a.bricks)[b];
    const int x = bo.x + lx, y = bo.y + ly;
    int zz[2] = { bo.z + lz, bo.z + lz + 4 };
    bool inside[2];
    int64_t vidx[2];
    float accR[2], accI[2], accW[2];
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      inside[v] = x < a.vx && y < a.vy && zz[v] < a.vz;
      vidx[v] = ((int64_t)zz[v] * a.vy + y) * a.vx + x;
      accR[v] = accI[v] = accW[v] = 0.f;
      if (inside[v]) {
        float2 t = reinterpret_cast<const float2 *>(a.volume)[vidx[v]];
        accR[v] = t.x;
        accI[v] = t.y;
        accW[v] = a.weights[vidx[v]];
      }
    }
    unsigned long long nupd = 0;
    // brick centre in centred volume coordinates
    const float bcx = (float)(bo.x - cxi) + 1.5f;
    const float bcy = (float)(bo.y - cyi) + 1.5f;
    const float bcz = (float)(bo.z - cyi) + 3.5f;
    const float bnorm = sqrtf(bcx * bcx + bcy * bcy + bcz * bcz);
    const float dmin = fmaxf(bnorm - brickRadius, 0.f);
    const float eps = 0.02f + 1e-5f * bnorm;

    for (int base = 0; base < a.nSpaces; base += 64) {
      const int si = base + lane;
      bool pass = false;
      if (si < a.nSpaces) {
        const umpa_traverse_space &sp = a.spaces[si];
        const float(*t)[3] = sp.transformInv;
        float qz = t[2][0] * bcx + t[2][1] * bcy + t[2][2] * bcz;
        float ez = 1.5f * (fabsf(t[2][0]) + fabsf(t[2][1])) + 3.5f * fabsf(t[2][2]);
        float qx = t[0][0] * bcx + t[0][1] * bcy + t[0][2] * bcz;
        float ex = 1.5f * (fabsf(t[0][0]) + fabsf(t[0][1])) + 3.5f * fabsf(t[0][2]);
        pass = fabsf(qz) <= a.cullHalfWidth + ez + eps && qx >= -a.cullHalfSpace - ex - eps &&
               dmin * dmin <= sp.maxDistanceSqr * 1.0001f + eps && sp.projectionIndex >= 0 &&
               sp.projectionIndex < a.nProj && sp.dir >= 0 && sp.dir <= 2;
      }
      unsigned long long mask = __ballot(pass);
      while (mask) {
        const int k = __builtin_ctzll(mask);
        mask &= mask - 1;
        const umpa_traverse_space sp = a.spaces[base + k];
        const int64_t off = (int64_t)sp.projectionIndex * a.sliceElems;
        const void *slice = kHalf ? (const void *)((const __half2 *)a.fft + off)
                                  : (const void *)((const float2 *)a.fft + off);
        const float scale = kHalf ? a.scales[sp.projectionIndex] : 1.f;
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          if (inside[v]) {
            bool hit = insert_pair<kMode, kOrder, kNB, kHalf, kFastMath>(
              a, sp, slice, scale, tbl, x, y, zz[v], accR[v], accI[v], accW[v]);
            if constexpr (kCount) nupd += hit ? 1 : 0;
          }
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      if (inside[v]) {
        reinterpret_cast<float2 *>(a.volume)[vidx[v]] = make_float2(accR[v], accI[v]);
        a.weights[vidx[v]] = accW[v];
      }
    }
    if constexpr (kCount) {
      if (nupd) atomicAdd(a.updates, nupd);
    }
  }
}

__global__ void fr_reset_counter(int *counter) { *counter = 0; }

// ---- host-side dispatch ------------------------------------------------------------------

template <int kMode, int kOrder, int kNB, bool kHalf, bool kFastMath, bool kCount>
static hipError_t launch_one(const FrArgs &a, int numCUs, hipStream_t st)
{
  const bool lds = kMode == kModeLookup && a.nTable <= kMaxLdsTable;
  const size_t shmem = lds ? (size_t)a.nTable * sizeof(float) : 0;
  const int wavesPerBlock = kThreads / 64;
  int blocks = (a.nBricks + wavesPerBlock - 1) / wavesPerBlock;
  int maxBlocks = numCUs * 4;
  if (blocks > maxBlocks) blocks = maxBlocks;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(fr_reset_counter, dim3(1), dim3(1), 0, st, a.counter);
  if (lds) {
    hipLaunchKernelGGL((fr_insert_kernel<kMode, kOrder, kNB, kHalf, kFastMath, kCount, true>),
                       dim3(blocks), dim3(kThreads), shmem, st, a);
  } else {
    hipLaunchKernelGGL((fr_insert_kernel<kMode, kOrder, kNB, kHalf, kFastMath, kCount, false>),
                       dim3(blocks), dim3(kThreads), 0, st, a);
  }
  return hipGetLastError();
}

template <int kMode, int kOrder, bool kHalf, bool kFastMath, bool kCount>
static hipError_t launch_nb(const FrArgs &a, int nb, int numCUs, hipStream_t st)
{
  switch (nb) {
  case 4: return launch_one<kMode, kOrder, 4, kHalf, kFastMath, kCount>(a, numCUs, st);
  case 5: return launch_one<kMode, kOrder, 5, kHalf, kFastMath, kCount>(a, numCUs, st);
  case 6: return launch_one<kMode, kOrder, 6, kHalf, kFastMath, kCount>(a, numCUs, st);
  default: return launch_one<kMode, kOrder, 0, kHalf, kFastMath, kCount>(a, numCUs, st);
  }
}

hipError_t launch_fr_insert(const FrArgs &a, const FrVariant &v, int numCUs, hipStream_t st)
{
  if (v.count) {
    // counting runs are for measurement only: generic footprint loops
    switch (v.mode) {
    case kModeLookup:
      if (v.half) return launch_one<kModeLookup, 0, 0, true, true, true>(a, numCUs, st);
      return v.fastMath ? launch_one<kModeLookup, 0, 0, false, true, true>(a, numCUs, st)
                        : launch_one<kModeLookup, 0, 0, false, false, true>(a, numCUs, st);
    case kModeFastKaiser: return launch_one<kModeFastKaiser, 0, 0, false, false, true>(a, numCUs, st);
    case kModeNearest: return launch_one<kModeNearest, 0, 0, false, false, true>(a, numCUs, st);
    default: return launch_one<kModePreciseKaiser, 0, 0, false, false, true>(a, numCUs, st);
    }
  }
  switch (v.mode) {
  case kModeLookup:
    if (v.half) return launch_nb<kModeLookup, 0, true, true, false>(a, v.nb, numCUs, st);
    if (v.fastMath) return launch_nb<kModeLookup, 0, false, true, false>(a, v.nb, numCUs, st);
    return launch_nb<kModeLookup, 0, false, false, false>(a, v.nb, numCUs, st);
  case kModeFastKaiser:
    return launch_one<kModeFastKaiser, 0, 0, false, false, false>(a, numCUs, st);
  case kModeNearest:
    return launch_one<kModeNearest, 0, 0, false, false, false>(a, numCUs, st);
  case kModePreciseKaiser:
    switch (v.order) {
    case 0: return launch_one<kModePreciseKaiser, 0, 0, false, false, false>(a, numCUs, st);
    case 1: return launch_one<kModePreciseKaiser, 1, 0, false, false, false>(a, numCUs, st);
    case 2: return launch_one<kModePreciseKaiser, 2, 0, false, false, false>(a, numCUs, st);
    case 3: return launch_one<kModePreciseKaiser, 3, 0, false, false, false>(a, numCUs, st);
    default: return launch_one<kModePreciseKaiser, 4, 0, false, false, false>(a, numCUs, st);
    }
  default: return hipErrorInvalidValue;
  }
}

}  // namespace umpa

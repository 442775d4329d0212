// FR insertion kernels for gfx950: voxel-owned gather.
//
// The reference scatters each rotated slice into the volume with three float atomics per voxel
// (fr_cuda_kernels.cu:266-365).  Here a wavefront owns a brick of 4 x 4 x 8 voxels (two per
// lane) for the whole launch and keeps their running (complex, weight) sums in registers.  It
// culls the traverse spaces against the brick 64 at a time with one ballot and walks the
// surviving spaces in index order, so every voxel receives its per-space contributions in the
// order of the reference CPU loop over spaces (fr_cpu_kernels.hpp:254): no atomics, and no
// dependence on how bricks are scheduled.
//
// Per (brick, space) pair the wave
//   1. evaluates the reference's per-voxel predicate for its 128 voxels,
//   2. compacts the voxels that pass (about a third) into a per-wave LDS queue, so the gather
//      runs with full lanes,
//   3. gathers the blob footprint of each queued voxel with a branch-free 4 x 4 (NB x NB)
//      candidate loop: a candidate outside the blob reads the zero entry appended to the LDS
//      blob table and an out-of-range buffer offset that the hardware returns as zero, which
//      adds exactly +0 to the sums, so the result is bit-identical to the reference's
//      skip-the-pixel loop while all table and pixel loads of an item issue back to back,
//   4. hands each contribution back to the owning lane, which adds it to its register sum.
// Workgroups are persistent and take bricks from a counter, most central (heaviest) first; the
// blob table is staged once per workgroup in LDS.
#include "fr_device.hpp"

namespace umpa {

constexpr int kMaxLdsTable = 16384;  // floats; larger tables are read from global memory
constexpr int kWaves = 8;            // wavefronts per workgroup (they share the LDS table)
constexpr int kThreads = 64 * kWaves;
constexpr int kQueue = 128;          // voxels per brick = queue slots per wave
constexpr uint32_t kOutOfRange = 0x7ffff000u;  // buffer offset past every slice: loads read 0
constexpr uint32_t kBufferConfig = 0x00020000u; // raw buffer descriptor word 3 on gfx9

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned lanes_below(unsigned long long mask)
{
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// Raw buffer loads through the LLVM intrinsics.  (The clang builtin for the 64-bit raw buffer
// load, __builtin_amdgcn_raw_buffer_load_b64, lowers to a single 32-bit load in this toolchain
// and returns that dword for both halves, so the vector form is declared directly.)
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ f32x2_t raw_buffer_load_f32x2(__amdgpu_buffer_rsrc_t rsrc, uint32_t voffset,
                                        uint32_t soffset,
                                        int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f32");
__device__ uint32_t raw_buffer_load_u32(__amdgpu_buffer_rsrc_t rsrc, uint32_t voffset,
                                        uint32_t soffset,
                                        int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.i32");

// Slice pixel through a bounds-checked buffer resource: offsets at or past the slice end
// return zero instead of faulting.
template <bool kHalf>
__device__ __forceinline__ float2 load_pixel(__amdgpu_buffer_rsrc_t rsrc, uint32_t byteOffset)
{
  if constexpr (kHalf) {
    const uint32_t v = raw_buffer_load_u32(rsrc, byteOffset, 0, 0);
    return __half22float2(__builtin_bit_cast(__half2, v));
  } else {
    const f32x2_t v = raw_buffer_load_f32x2(rsrc, byteOffset, 0, 0);
    return make_float2(v[0], v[1]);
  }
}

// Blob weight of one candidate pixel (kDynamic modes; kLookup reads the LDS table inline).
template <int kMode, int kOrder>
__device__ __forceinline__ float blob_weight(float d2, const umpa_fr_constants &c)
{
  if constexpr (kMode == kModeFastKaiser) {
    return kaiser_value_fast(d2, c);
  } else {
    return kaiser_value<kOrder>(sqrtf(d2), c.cBlobRadius, c) * c.cIw0;
  }
}

struct Contribution {
  float re, im, w;
};

// Reference-ordered gather of one voxel-update with generic loops (rows i ascending, columns j
// ascending, ProcessVoxelBlob fr_cpu_kernels.hpp:104-147).  Used for footprints wider than the
// unrolled NB x NB box, which only rounding at the blob edge can produce.
template <int kMode, int kOrder, bool kHalf, bool kFastMath>
__device__ Contribution gather_generic(const FrArgs &a, __amdgpu_buffer_rsrc_t rsrc,
                                       const float *tbl, float qx, float qy, float zSqr, float dw)
{
  const umpa_fr_constants &c = a.c;
  const float r = c.cBlobRadius, r2 = r * r;
  int minX = (int)ceilf(qx - r), maxX = (int)floorf(qx + r);
  int minY = (int)ceilf(qy - r), maxY = (int)floorf(qy + r);
  minX = minX < 0 ? 0 : minX;
  minY = minY < 0 ? 0 : minY;
  maxX = maxX > a.xSize - 1 ? a.xSize - 1 : maxX;
  maxY = maxY > a.ySize - 1 ? a.ySize - 1 : maxY;
  Contribution o{ 0.f, 0.f, 0.f };
  for (int i = minY; i <= maxY; ++i) {
    float dy = qy - (float)i;
    float yz = dy * dy + zSqr;
    if (yz > r2) continue;
    for (int j = minX; j <= maxX; ++j) {
      float xD = qx - (float)j;
      float d2 = xD * xD + yz;
      if (d2 > r2) continue;
      float wb;
      if constexpr (kMode == kModeLookup) {
        int aux = kFastMath ? (int)__builtin_fmaf(d2, c.cIDeltaSqrt, 0.5f)
                            : (int)(d2 * c.cIDeltaSqrt + 0.5f);
        aux = aux > a.nTable - 1 ? a.nTable - 1 : aux;
        wb = tbl[aux];
      } else {
        wb = blob_weight<kMode, kOrder>(d2, c);
      }
      float weight = wb * dw;
      float2 f = load_pixel<kHalf>(rsrc, (uint32_t)((i * a.xSize + j) * (kHalf ? 4 : 8)));
      o.w += weight;
      if constexpr (kFastMath) {
        o.re = __builtin_fmaf(f.x, weight, o.re);
        o.im = __builtin_fmaf(f.y, weight, o.im);
      } else {
        o.re += f.x * weight;
        o.im += f.y * weight;
      }
    }
  }
  return o;
}

// Branch-free gather of one voxel-update over the NB x NB candidate box around (qx, qy).
// Candidates the reference skips (outside the blob, outside the slice) read table[nTable] = 0
// and an out-of-range pixel (= 0), adding exactly +0; the running sums never hold -0, so the
// float results equal the reference's.  Returns false when the footprint is wider than NB.
template <int kMode, int kOrder, int kNB, bool kHalf, bool kFastMath, bool kUnitWeight, bool kTableLds>
__device__ __forceinline__ bool gather_box(const FrArgs &a, __amdgpu_buffer_rsrc_t rsrc,
                                           const float *tbl, float qx, float qy, float zSqr,
                                           float dw, Contribution &o)
{
  const umpa_fr_constants &c = a.c;
  const float r = c.cBlobRadius, r2 = r * r;
  const float inf = __builtin_inff();
  int minX = (int)ceilf(qx - r), maxX = (int)floorf(qx + r);
  int minY = (int)ceilf(qy - r), maxY = (int)floorf(qy + r);
  minX = minX < 0 ? 0 : minX;
  minY = minY < 0 ? 0 : minY;
  maxX = maxX > a.xSize - 1 ? a.xSize - 1 : maxX;
  maxY = maxY > a.ySize - 1 ? a.ySize - 1 : maxY;
  const bool fits = (maxX - minX < kNB) && (maxY - minY < kNB);
  constexpr int kPix = kHalf ? 4 : 8;

  float xd2[kNB], yz[kNB];
#pragma unroll
  for (int jj = 0; jj < kNB; ++jj) {
    float xD = qx - (float)(minX + jj);
    xd2[jj] = (minX + jj <= maxX) ? xD * xD : inf;
  }
#pragma unroll
  for (int ii = 0; ii < kNB; ++ii) {
    float dy = qy - (float)(minY + ii);
    yz[ii] = (minY + ii <= maxY) ? dy * dy + zSqr : inf;
  }
  const uint32_t rowStride = (uint32_t)(a.xSize * kPix);
  const uint32_t base = (uint32_t)(minY * a.xSize + minX) * kPix;
  float w = 0.f, vr = 0.f, vi = 0.f;
#pragma unroll
  for (int ii = 0; ii < kNB; ++ii) {
    const uint32_t rowOff = base + ii * rowStride;
#pragma unroll
    for (int jj = 0; jj < kNB; ++jj) {
      const float d2 = xd2[jj] + yz[ii];
      const bool in = d2 <= r2;
      float wb;
      if constexpr (kMode == kModeLookup) {
        int aux = kFastMath ? (int)__builtin_fmaf(d2, c.cIDeltaSqrt, 0.5f)
                            : (int)(d2 * c.cIDeltaSqrt + 0.5f);
        if constexpr (kTableLds) {
          wb = tbl[in ? aux : a.nTable];  // entry nTable of the LDS copy is 0
        } else {
          aux = aux > a.nTable - 1 ? a.nTable - 1 : aux;
          wb = in ? tbl[aux] : 0.f;
        }
      } else {
        wb = in ? blob_weight<kMode, kOrder>(d2, c) : 0.f;
      }
      const float weight = kUnitWeight ? wb : wb * dw;
      const float2 f = load_pixel<kHalf>(rsrc, (in ? rowOff : kOutOfRange) + jj * kPix);
      w += weight;
      if constexpr (kFastMath) {
        vr = __builtin_fmaf(f.x, weight, vr);
        vi = __builtin_fmaf(f.y, weight, vi);
      } else {
        vr += f.x * weight;
        vi += f.y * weight;
      }
    }
  }
  o.re = vr;
  o.im = vi;
  o.w = w;
  return fits;
}

// The reference's per-voxel predicate for one (voxel, space) pair in blob mode
// (ProcessVoxelBlob, fr_cpu_kernels.hpp:83-102) plus membership in the reference traversal of
// the space.  (qx, qy, zSqr) are returned for the gather, qy already shifted to slice rows.
__device__ __forceinline__ bool blob_predicate(const umpa_traverse_space &s, const umpa_fr_constants &c,
                                               int x, int y, int z, float px, float py, float pz,
                                               float dsq, float &qx, float &qy, float &zSqr)
{
  const float r = c.cBlobRadius;
  qx = s.transformInv[0][0] * px + s.transformInv[0][1] * py + s.transformInv[0][2] * pz;
  qy = s.transformInv[1][0] * px + s.transformInv[1][1] * py + s.transformInv[1][2] * pz;
  const float qz = s.transformInv[2][0] * px + s.transformInv[2][1] * py + s.transformInv[2][2] * pz;
  zSqr = qz * qz;
  bool pass = !(dsq > s.maxDistanceSqr) && !(qx < -r) && !(zSqr > r * r);
  qy += (float)(c.cMaxVolumeIndexYZ / 2);
  // traversal: the column must lie in the AABB's plane orthogonal to `dir`
  if (s.dir == 0) {
    pass = pass && x >= s.minX && x <= s.maxX && y >= s.minY && y <= s.maxY;
  } else if (s.dir == 1) {
    pass = pass && x >= s.minX && x <= s.maxX && z >= s.minZ && z <= s.maxZ;
  } else {
    pass = pass && y >= s.minY && y <= s.maxY && z >= s.minZ && z <= s.maxZ;
  }
  // ... and the voxel must lie in the column's slab between the two bounding planes.  A voxel
  // whose |qz| is at least 0.01 inside the blob radius is between the planes by at least that
  // much along any axis, far beyond the float error of the plane evaluation, so only voxels
  // near the slab faces need the reference's exact floor/ceil test.
  if (pass && !(fabsf(qz) <= r - 0.01f)) {
    pass = visited_blob(s, x, y, z, c.cMaxVolumeIndexX, c.cMaxVolumeIndexYZ);
  }
  return pass;
}

template <int kMode, int kOrder, int kNB, bool kHalf, bool kFastMath, bool kCount, bool kTableLds>
__global__ __launch_bounds__(kThreads) void fr_insert_kernel(FrArgs a)
{
  extern __shared__ float s_table[];
  __shared__ float4 s_queue[kWaves][kQueue];
  if constexpr (kMode == kModeLookup && kTableLds) {
    for (int i = threadIdx.x; i < a.nTable; i += kThreads) s_table[i] = a.table[i];
    if (threadIdx.x == 0) s_table[a.nTable] = 0.f;
    __syncthreads();
  }
  const float *tbl = (kMode == kModeLookup && kTableLds) ? s_table : a.table;
  const int lane = threadIdx.x & 63;
  float4 *queue = s_queue[threadIdx.x >> 6];
  const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
  const umpa_fr_constants &c = a.c;
  const int cxi = c.cMaxVolumeIndexX / 2, cyi = c.cMaxVolumeIndexYZ / 2;
  const float brickRadius = 4.0927375f;  // |(1.5, 1.5, 3.5)|, half diagonal of a brick
  constexpr int kPix = kHalf ? 4 : 8;

  for (;;) {
    int b = 0;
    if (lane == 0) b = atomicAdd(a.counter, 1);
    b = __builtin_amdgcn_readfirstlane(__shfl(b, 0));
    if (b >= a.nBricks) break;
    const int4 bo = reinterpret_cast<const int4 *>(a.bricks)[b];
    const int x = bo.x + lx, y = bo.y + ly;
    const int zz[2] = { bo.z + lz, bo.z + lz + 4 };
    bool inside[2];
    int64_t vidx[2];
    float accR[2], accI[2], accW[2], pz[2], dsq[2];
    const float px = (float)(x - cxi), py = (float)(y - cyi);
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      inside[v] = x < a.vx && y < a.vy && zz[v] < a.vz;
      vidx[v] = ((int64_t)zz[v] * a.vy + y) * a.vx + x;
      pz[v] = (float)(zz[v] - cyi);
      dsq[v] = px * px + py * py + pz[v] * pz[v];
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
        // the space index is wave-uniform: say so, so its record is read with scalar loads and
        // the slice's buffer descriptor lives in scalar registers
        const int si_u = __builtin_amdgcn_readfirstlane(base + k);
        const umpa_traverse_space sp = a.spaces[si_u];
        const int proj = __builtin_amdgcn_readfirstlane(sp.projectionIndex);
        const int64_t off = (int64_t)proj * a.sliceElems * kPix;
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
          (void *)((const char *)a.fft + off), 0, (int)(a.sliceElems * kPix), kBufferConfig);
        const float scale = kHalf ? a.scales[proj] : 1.f;

        // 1. predicate for the lane's two voxels
        bool pv[2];
        float qx[2], qy[2], zs[2];
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          if constexpr (kMode == kModeNearest) {
            qx[v] = sp.transformInv[0][0] * px + sp.transformInv[0][1] * py + sp.transformInv[0][2] * pz[v];
            qy[v] = sp.transformInv[1][0] * px + sp.transformInv[1][1] * py + sp.transformInv[1][2] * pz[v];
            zs[v] = 0.f;
            pv[v] = inside[v] && !(dsq[v] > sp.maxDistanceSqr) && !(qx[v] < 0.f) &&
                    visited_nearest(sp, x, y, zz[v]);
          } else {
            pv[v] = blob_predicate(sp, c, x, y, zz[v], px, py, pz[v], dsq[v], qx[v], qy[v], zs[v]);
            pv[v] = pv[v] && inside[v];
          }
        }
        // 2. compaction into the wave's queue
        const unsigned long long m0 = __ballot(pv[0]), m1 = __ballot(pv[1]);
        const int c0 = __popcll(m0);
        const int cnt = c0 + __popcll(m1);
        if (cnt == 0) continue;
        const int r0 = (int)lanes_below(m0), r1 = c0 + (int)lanes_below(m1);
        if (pv[0]) queue[r0] = make_float4(qx[0], qy[0], zs[0], 0.f);
        if (pv[1]) queue[r1] = make_float4(qx[1], qy[1], zs[1], 0.f);
        wave_sync();
        // 3. gather with full lanes
        const bool unit = sp.weight == 1.f;
        for (int t0 = 0; t0 < cnt; t0 += 64) {
          const int item = t0 + lane;
          const bool act = item < cnt;
          const float4 q = queue[act ? item : 0];
          Contribution o;
          if constexpr (kMode == kModeNearest) {
            // ProcessVoxel (fr_cpu_kernels.hpp:50-67): nearest pixel, weight = space weight
            int ix = (int)(q.x + 0.5f);
            ix = ix > a.xSize - 1 ? a.xSize - 1 : ix;
            ix = ix < 0 ? 0 : ix;
            int iy = (int)(q.y + 0.5f + (float)(c.cMaxVolumeIndexYZ / 2));
            iy = iy > a.ySize - 1 ? a.ySize - 1 : iy;
            iy = iy < 0 ? 0 : iy;
            const float wgt = 1.f * sp.weight;
            float2 f = load_pixel<kHalf>(rsrc, (uint32_t)((iy * a.xSize + ix) * kPix));
            o.re = f.x * wgt;
            o.im = f.y * wgt;
            o.w = wgt;
          } else {
            bool fits;
            if (unit) {
              fits = gather_box<kMode, kOrder, kNB, kHalf, kFastMath, true, kTableLds>(
                a, rsrc, tbl, q.x, q.y, q.z, sp.weight, o);
            } else {
              fits = gather_box<kMode, kOrder, kNB, kHalf, kFastMath, false, kTableLds>(
                a, rsrc, tbl, q.x, q.y, q.z, sp.weight, o);
            }
            if (!fits) o = gather_generic<kMode, kOrder, kHalf, kFastMath>(a, rsrc, tbl, q.x, q.y, q.z, sp.weight);
            if constexpr (kHalf) {
              o.re *= scale;
              o.im *= scale;
            }
          }
          wave_sync();
          if (act) queue[item] = make_float4(o.re, o.im, o.w, 0.f);
        }
        wave_sync();
        // 4. owners add their contributions in space order
        if (pv[0]) {
          const float4 r = queue[r0];
          accR[0] = accR[0] + r.x;
          accI[0] = accI[0] + r.y;
          accW[0] = accW[0] + r.z;
        }
        if (pv[1]) {
          const float4 r = queue[r1];
          accR[1] = accR[1] + r.x;
          accI[1] = accI[1] + r.y;
          accW[1] = accW[1] + r.z;
        }
        if constexpr (kCount) nupd += (unsigned long long)cnt;
        wave_sync();
      }
    }
    for (int v = 0; v < 2; ++v) {
      if (inside[v]) {
        reinterpret_cast<float2 *>(a.volume)[vidx[v]] = make_float2(accR[v], accI[v]);
        a.weights[vidx[v]] = accW[v];
      }
    }
    if constexpr (kCount) {
      if (lane == 0 && nupd) atomicAdd(a.updates, nupd);
    }
  }
}

__global__ void fr_reset_counter(int *counter) { *counter = 0; }

// ---- host-side dispatch ------------------------------------------------------------------
template <int kMode, int kOrder, int kNB, bool kHalf, bool kFastMath, bool kCount>
static hipError_t launch_one(const FrArgs &a, int numCUs, hipStream_t st)
{
  const bool lds = kMode == kModeLookup && a.nTable <= kMaxLdsTable;
  const size_t shmem = lds ? (size_t)(a.nTable + 1) * sizeof(float) : 0;
  int blocks = (a.nBricks + kWaves - 1) / kWaves;
  int maxBlocks = numCUs * 2;
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
  case 3: return launch_one<kMode, kOrder, 3, kHalf, kFastMath, kCount>(a, numCUs, st);
  case 4: return launch_one<kMode, kOrder, 4, kHalf, kFastMath, kCount>(a, numCUs, st);
  case 5: return launch_one<kMode, kOrder, 5, kHalf, kFastMath, kCount>(a, numCUs, st);
  default: return launch_one<kMode, kOrder, 6, kHalf, kFastMath, kCount>(a, numCUs, st);
  }
}

hipError_t launch_fr_insert(const FrArgs &a, const FrVariant &v, int numCUs, hipStream_t st)
{
  switch (v.mode) {
  case kModeLookup:
    if (v.count) {
      if (v.half) return launch_nb<kModeLookup, 0, true, true, true>(a, v.nb, numCUs, st);
      return v.fastMath ? launch_nb<kModeLookup, 0, false, true, true>(a, v.nb, numCUs, st)
                        : launch_nb<kModeLookup, 0, false, false, true>(a, v.nb, numCUs, st);
    }
    if (v.half) return launch_nb<kModeLookup, 0, true, true, false>(a, v.nb, numCUs, st);
    if (v.fastMath) return launch_nb<kModeLookup, 0, false, true, false>(a, v.nb, numCUs, st);
    return launch_nb<kModeLookup, 0, false, false, false>(a, v.nb, numCUs, st);
  case kModeFastKaiser:
    return v.count ? launch_one<kModeFastKaiser, 0, 4, false, false, true>(a, numCUs, st)
                   : launch_one<kModeFastKaiser, 0, 4, false, false, false>(a, numCUs, st);
  case kModeNearest:
    return v.count ? launch_one<kModeNearest, 0, 4, false, false, true>(a, numCUs, st)
                   : launch_one<kModeNearest, 0, 4, false, false, false>(a, numCUs, st);
  case kModePreciseKaiser:
    if (v.count) return launch_one<kModePreciseKaiser, 0, 4, false, false, true>(a, numCUs, st);
    switch (v.order) {
    case 0: return launch_one<kModePreciseKaiser, 0, 4, false, false, false>(a, numCUs, st);
    case 1: return launch_one<kModePreciseKaiser, 1, 4, false, false, false>(a, numCUs, st);
    case 2: return launch_one<kModePreciseKaiser, 2, 4, false, false, false>(a, numCUs, st);
    case 3: return launch_one<kModePreciseKaiser, 3, 4, false, false, false>(a, numCUs, st);
    default: return launch_one<kModePreciseKaiser, 4, 4, false, false, false>(a, numCUs, st);
    }
  default: return hipErrorInvalidValue;
  }
}

}  // namespace umpa

// FR insertion kernels for gfx950: voxel-owned gather.
//
// The reference scatters each rotated slice into the volume with three float atomics per voxel
// (fr_cuda_kernels.cu:266-365).  Here a wavefront owns a brick of 8 x 8 x D voxels (a z column
// of D per lane, D = 8, 4, 2 or 1 by distance from the centre) for the whole work item and
// keeps their running (complex, weight) sums in its LDS slot.  It culls the traverse spaces
// against the brick 64 at a time with one ballot and walks the surviving spaces in index
// order, so every voxel receives its per-space contributions in the order of the reference CPU
// loop over spaces (fr_cpu_kernels.hpp:254): no atomics, and no dependence on how bricks are
// scheduled.
//
// Per (brick, space) pair the wave
//   1. tests the reference's slab condition (q.z^2 <= r^2) for its voxels, 4 VALU each, and
//      compacts the voxels that pass into its LDS queue;
//   2. runs the rest of the reference predicate and the gather over the queue with full lanes:
//      a branch-free NB x NB candidate loop over whole pixel rows in which a candidate the
//      reference skips reads the zero entry appended to the LDS blob table, adding +-0 (finite
//      slice values), so the sums are bit-identical to the reference's skip-the-pixel loop
//      while all table and pixel loads of an item issue back to back; each voxel appears once
//      per space, so its lane adds the contribution to the voxel's LDS sum directly.
// Workgroups are persistent (one per CU: 16 or 12 wavefronts sharing the blob table in LDS)
// and take bricks from a counter, most central (heaviest) first.
//
// fr_insert_kernel covers every mode, precision and strategy; fr_insert_staged_kernel, for the
// exact FP32 lookup strategy, gathers the predicate's hits of several spaces together (see
// there).
#include "fr_device.hpp"

namespace umpa {

constexpr int kMaxLdsTable = kFrMaxLdsTable;  // floats next to the per-wave state; larger tables stay global
// Wavefronts per workgroup (they share the LDS table): 16 (4 per SIMD) in general; the exact
// FP32 lookup kernel runs faster with 12 (1.94 -> 1.86 ms per 1,000-slice chunk at box 256),
// the fast-math and FP16 kernels slower (1.46 -> 1.68 and 1.39 -> 1.60 ms).
constexpr int kMaxWaves = 16;
__host__ __device__ constexpr int fr_waves(int mode, bool half, bool fastMath)
{
  return (mode == kModeLookup && !half && !fastMath) ? 12 : kMaxWaves;
}
constexpr int kEdge = 8;             // brick edge: a wave owns 8 x 8 x 8 voxels, a z column per lane
constexpr int kVox = kEdge * kEdge * kEdge;
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
// and returns that dword for both halves, so the vector forms are declared directly.)  Every
// byte offset travels in the vector offset, which the hardware range-checks together with the
// instruction's immediate offset; the scalar offset is not range-checked, and loads that
// differed only in it were observed to be merged into one by this compiler, so it stays 0.
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
__device__ f32x2_t raw_buffer_load_f32x2(__amdgpu_buffer_rsrc_t rsrc, uint32_t voffset,
                                        uint32_t soffset,
                                        int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f32");
__device__ uint32_t raw_buffer_load_u32(__amdgpu_buffer_rsrc_t rsrc, uint32_t voffset,
                                        uint32_t soffset,
                                        int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.i32");
__device__ f32x4_t raw_buffer_load_f32x4(__amdgpu_buffer_rsrc_t rsrc, uint32_t voffset,
                                        uint32_t soffset,
                                        int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.v4f32");
__device__ u32x4_t raw_buffer_load_u32x4(__amdgpu_buffer_rsrc_t rsrc, uint32_t voffset,
                                        uint32_t soffset,
                                        int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.v4i32");

// Slice pixel at byte offset `byteOffset` through a bounds-checked buffer resource: offsets at
// or past the slice end return zero instead of faulting.
template <bool kHalf>
__device__ __forceinline__ float2 load_pixel(__amdgpu_buffer_rsrc_t rsrc, uint32_t byteOffset)
{
  if constexpr (kHalf) {
    const uint32_t v = raw_buffer_load_u32(rsrc, byteOffset, 0, 0);
    return __half22float2(__builtin_bit_cast(__half2, v));
  } else {
    const f32x2_t v = raw_buffer_load_f32x2(rsrc, byteOffset, 0, 0);
    return make_float2(v.x, v.y);
  }
}

// A slice pixel as loaded: complex float, or complex half packed in one dword (converted at
// the multiply-add, which the compiler fuses into a mixed-precision FMA).
template <bool kHalf> struct Pixel;
template <> struct Pixel<false> {
  float2 v;
  __device__ __forceinline__ float re() const { return v.x; }
  __device__ __forceinline__ float im() const { return v.y; }
};
template <> struct Pixel<true> {
  uint32_t v;
  __device__ __forceinline__ float re() const
  {
    return __half2float(__builtin_bit_cast(__half, (uint16_t)(v & 0xffffu)));
  }
  __device__ __forceinline__ float im() const
  {
    return __half2float(__builtin_bit_cast(__half, (uint16_t)(v >> 16)));
  }
};

template <bool kHalf>
__device__ __forceinline__ Pixel<kHalf> load_pixel_raw(__amdgpu_buffer_rsrc_t rsrc,
                                                       uint32_t byteOffset)
{
  Pixel<kHalf> p;
  if constexpr (kHalf) {
    p.v = raw_buffer_load_u32(rsrc, byteOffset, 0, 0);
  } else {
    const f32x2_t v = raw_buffer_load_f32x2(rsrc, byteOffset, 0, 0);
    p.v = make_float2(v.x, v.y);
  }
  return p;
}

// kNB consecutive pixels of a slice row starting at byte `voff`, with 16-byte loads (4 FP16 or
// 2 FP32 pixels each).  Dwords past the slice end read zero.  A load on the slice's last row
// can straddle the end; the raw buffer's range check then still returns its in-range dwords
// (the reference's 3 x 6 plane cases, where every last-row load straddles, are bit-exact).
template <bool kHalf, int kNB>
__device__ __forceinline__ void load_row(__amdgpu_buffer_rsrc_t rsrc, uint32_t voff,
                                         Pixel<kHalf> (&px)[kNB])
{
  constexpr int kPix = kHalf ? 4 : 8;
  constexpr int kPer = 16 / kPix;                  // pixels per 16-byte load
  constexpr int kLoads = (kNB + kPer - 1) / kPer;
  // (vector elements are named, never indexed by a loop variable: this compiler folded
  // v[e] in an unrolled loop over e to v[0] for every e, in the 16-byte loads here)
#pragma unroll
  for (int q = 0; q < kLoads; ++q) {
    if constexpr (kHalf) {
      const u32x4_t v = raw_buffer_load_u32x4(rsrc, voff + (uint32_t)(q * 16), 0, 0);
      const uint32_t w[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
      for (int e = 0; e < kPer; ++e) {
        const int j = q * kPer + e;
        if (j < kNB) px[j].v = w[e];
      }
    } else {
      const f32x4_t v = raw_buffer_load_f32x4(rsrc, voff + (uint32_t)(q * 16), 0, 0);
      const float w[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
      for (int e = 0; e < kPer; ++e) {
        const int j = q * kPer + e;
        if (j < kNB) px[j].v = make_float2(w[2 * e], w[2 * e + 1]);
      }
    }
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
__device__ Contribution gather_generic(const FrJob &a, int nTable, __amdgpu_buffer_rsrc_t rsrc,
                                       const float *tbl, float qx, float qy, float zSqr, float dw,
                                       uint32_t sliceOff = 0)
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
        aux = aux > nTable - 1 ? nTable - 1 : aux;
        wb = tbl[aux];
      } else {
        wb = blob_weight<kMode, kOrder>(d2, c);
      }
      float weight = wb * dw;
      float2 f = load_pixel<kHalf>(rsrc, sliceOff + (uint32_t)((i * a.xSize + j) * (kHalf ? 4 : 8)));
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

// load_row for the exact FP32 Kaiser-mode kernels: a row whose 16-byte loads straddle the slice
// end (byte `sliceBytes` of the resource) is read pixel by pixel.  This compiler merges those
// pixel loads back into the same 16-byte loads, so both sides load alike; the branch only
// splits the rows into scheduling regions, the form these kernels were measured and kept in.
template <bool kHalf, int kNB>
__device__ __forceinline__ void load_row_regions(__amdgpu_buffer_rsrc_t rsrc, uint32_t voff,
                                                 uint32_t sliceBytes, Pixel<kHalf> (&px)[kNB])
{
  constexpr int kPix = kHalf ? 4 : 8;
  constexpr int kLoads = (kNB * kPix + 15) / 16;
  if (voff < sliceBytes && voff + (uint32_t)(kLoads * 16) > sliceBytes) {
#pragma unroll
    for (int j = 0; j < kNB; ++j) px[j] = load_pixel_raw<kHalf>(rsrc, voff + (uint32_t)(j * kPix));
    return;
  }
  load_row<kHalf, kNB>(rsrc, voff, px);
}

// Blob weight of one candidate at squared distance d2, 0 for a candidate the reference skips
// (d2 > r^2, and d2 = +inf for box positions outside the slice).  Exact strategy: the
// reference's table position (int)(d2 * cIDeltaSqrt + 0.5); skipped candidates read the LDS
// copy's zero entry nTable.  Fast math with the table in LDS: a fused position, clamped into
// the zero entry, so +inf and every d2 past the blob edge weigh 0 without a test.
template <int kMode, int kOrder, bool kFastMath, bool kTableLds>
__device__ __forceinline__ float candidate_weight(float d2, float r2, int nTable, const float *tbl,
                                                  const umpa_fr_constants &c)
{
  if constexpr (kFastMath && kMode == kModeLookup && kTableLds) {
    // The reference's position (int)(d2 * cIDeltaSqrt + 0.5f): the product rounded to float,
    // then floor(t + 0.5), which v_cvt_rpi_i32_f32 computes in one instruction from the exact
    // sum; the reference's rounded sum truncates to the same integer for every t >= 0 outside
    // (0.5 - 2^-25, 0.5) (on the GPU: 2.16 M values, ties k + 0.5 and their neighbours).
    // Rounding the fused product to the nearest entry instead (one instruction less) moves
    // entries at rounding boundaries: 4.2e-5 rel-L2 from FRCPU at box 384 with 64 spaces,
    // whose central voxels carry most of the volume's norm.  Positions past the table and
    // +inf (saturated to INT_MAX) are clamped into the zero entry nTable, so candidates
    // outside the blob weigh 0 without a test.
    const float t = d2 * c.cIDeltaSqrt;
    int pos;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(pos) : "v"(t));
    const uint32_t aux = (uint32_t)pos;
    typedef __attribute__((address_space(3))) const float lds_float;
    return ((lds_float *)tbl)[aux < (uint32_t)nTable ? aux : (uint32_t)nTable];
  } else {
    const bool in = d2 <= r2;
    if constexpr (kMode == kModeLookup) {
      int aux = kFastMath ? (int)__builtin_fmaf(d2, c.cIDeltaSqrt, 0.5f)
                          : (int)(d2 * c.cIDeltaSqrt + 0.5f);
      if constexpr (kTableLds) {
        return tbl[in ? aux : nTable];  // entry nTable of the LDS copy is 0
      } else {
        aux = aux > nTable - 1 ? nTable - 1 : aux;
        return in ? tbl[aux] : 0.f;
      }
    } else {
      return in ? blob_weight<kMode, kOrder>(d2, c) : 0.f;
    }
  }
}

// Branch-free gather of one voxel-update over the NB x NB candidate box around (qx, qy).
//
// Exact strategy: candidates the reference skips (outside the blob, outside the slice) read
// table[nTable] = 0, so with finite slice values they add +-0; the running sums start at +0
// and never hold -0, so the float results equal the reference's (footprints wider than 4 read
// skipped pixels past the buffer end, as zeros).  Fast-math strategy (finite slice values):
// the table index is clamped into the zero entry, so out-of-blob candidates weigh 0 without a
// test.  Rows are read whole; pixels past the slice end read 0 through the bounds-checked
// buffer.  With kShift (the caller's instantiation for jobs with FrJob::shiftBox) the box slides
// inside the slice (see below).
// Returns false when the footprint is wider than NB.
template <int kMode, int kOrder, int kNB, bool kHalf, bool kFastMath, bool kUnitWeight, bool kTableLds,
          bool kShift>
__device__ __forceinline__ bool gather_box(const FrJob &a, int nTable,
                                           __amdgpu_buffer_rsrc_t rsrc, const float *tbl,
                                           float qx, float qy, float zSqr, float dw,
                                           Contribution &o, uint32_t sliceOff = 0)
{
  const umpa_fr_constants &c = a.c;
  const float r = c.cBlobRadius, r2 = r * r;
  const float inf = __builtin_inff();
  int minX = (int)ceilf(qx - r);
  int minY = (int)ceilf(qy - r);
  minX = minX < 0 ? 0 : minX;
  minY = minY < 0 ? 0 : minY;
  constexpr int kPix = kHalf ? 4 : 8;
  bool fits = true;
  float xd2[kNB], yz[kNB];
  if constexpr (kShift) {
    // The box slides inside the slice (slices of at least kNB x kNB pixels): it covers the
    // reference's bounding box [minX, maxX] x [minY, maxY] whenever that has at most kNB
    // columns and rows (the host sets shiftBox only where floor(2r) + 1 <= kNB), and every
    // other candidate in it lies farther than r from q, so it weighs 0 (the table's zero
    // entry; fast math: the clamp, or the last entry within 0.5 / cIDeltaSqrt of r^2), with no
    // per-candidate masking.
    minX = minX < a.xSize - kNB ? minX : a.xSize - kNB;
    minY = minY < a.ySize - kNB ? minY : a.ySize - kNB;
    const float fx0 = (float)minX, fy0 = (float)minY;
#pragma unroll
    for (int jj = 0; jj < kNB; ++jj) {
      const float xD = qx - (fx0 + (float)jj);
      xd2[jj] = xD * xD;
    }
#pragma unroll
    for (int ii = 0; ii < kNB; ++ii) {
      const float dy = qy - (fy0 + (float)ii);
      yz[ii] = dy * dy + zSqr;
    }
  } else {
    int maxX = (int)floorf(qx + r), maxY = (int)floorf(qy + r);
    maxX = maxX > a.xSize - 1 ? a.xSize - 1 : maxX;
    maxY = maxY > a.ySize - 1 ? a.ySize - 1 : maxY;
    const int nx = maxX - minX, ny = maxY - minY;
    fits = (nx < kNB) && (ny < kNB);
    // (float)j == (float)minX + jj exactly for slice indices
    const float fx0 = (float)minX, fy0 = (float)minY;
#pragma unroll
    for (int jj = 0; jj < kNB; ++jj) {
      const float xD = qx - (fx0 + (float)jj);
      xd2[jj] = (jj <= nx) ? xD * xD : inf;
    }
#pragma unroll
    for (int ii = 0; ii < kNB; ++ii) {
      const float dy = qy - (fy0 + (float)ii);
      yz[ii] = (ii <= ny) ? dy * dy + zSqr : inf;
    }
  }
  const uint32_t rowStride = (uint32_t)(a.xSize * kPix);
  // sliceOff: the slice's byte offset within the buffer resource (0 when the resource covers
  // the slice alone)
  const uint32_t base = sliceOff + (uint32_t)(minY * a.xSize + minX) * kPix;
  const uint32_t sliceEnd = sliceOff + (uint32_t)(a.sliceElems * kPix);
  float w = 0.f, vr = 0.f, vi = 0.f;
  if constexpr (kNB <= 4) {
    // Whole candidate rows, summed in the reference's order (rows, then columns).  Candidates
    // outside the blob or the slice weigh exactly 0 (the table's zero entry), so with finite
    // slice values they add +-0, which changes no running sum (the sums start at +0 and are
    // never -0): the result equals the reference's skip-the-pixel loop bit for bit.
    // The exact FP32 lookup kernel issues each row's table reads in a scheduling region of
    // its own, all four rows before the sums: left to itself this compiler waited on every
    // table read in turn there (2.70 ms per 1,000-slice chunk at box 256 instead of 1.94 ms).
    // The fast-math and FP16 kernels run faster with the compiler's own order (1.49 / 1.42 ms
    // instead of 1.53 / 1.46 ms with such regions).
    constexpr bool kRegions = !kHalf && !kFastMath && kMode == kModeLookup;
    float wr[kNB][kNB];
    if constexpr (kRegions) {
#pragma unroll
      for (int ii = 0; ii < kNB; ++ii) {
#pragma unroll
        for (int jj = 0; jj < kNB; ++jj)
          wr[ii][jj] = candidate_weight<kMode, kOrder, kFastMath, kTableLds>(
            xd2[jj] + yz[ii], r2, nTable, tbl, c);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int ii = 0; ii < kNB; ++ii) {
      Pixel<kHalf> px[kNB];
      // fast-math FP32: a row wholly outside the blob (every candidate weighs 0) reads zeros
      // from past the slice end instead of its pixels, a quarter of the rows at r = 1.9
      // (1.480 -> 1.458 ms per 1,000-slice chunk at box 256; the exact kernel's schedule
      // degrades with it, 1.94 -> 2.11 ms, and FP16 rows gain nothing)
      constexpr bool kSkipRows = kFastMath && !kHalf;
      const uint32_t rowOff =
        (!kSkipRows || yz[ii] <= r2) ? base + ii * rowStride : kOutOfRange;
      if constexpr (kRegions) {
        load_row<kHalf, kNB>(rsrc, rowOff, px);
      } else if constexpr (!kHalf && !kFastMath) {
        load_row_regions<kHalf, kNB>(rsrc, rowOff, sliceEnd, px);
      } else {
        load_row<kHalf, kNB>(rsrc, rowOff, px);
      }
#pragma unroll
      for (int jj = 0; jj < kNB; ++jj) {
        const float wb = kRegions ? wr[ii][jj]
                                  : candidate_weight<kMode, kOrder, kFastMath, kTableLds>(
                                      xd2[jj] + yz[ii], r2, nTable, tbl, c);
        const float weight = kUnitWeight ? wb : wb * dw;
        w += weight;
        if constexpr (kFastMath) {
          vr = __builtin_fmaf(px[jj].re(), weight, vr);
          vi = __builtin_fmaf(px[jj].im(), weight, vi);
        } else {
          vr += px[jj].re() * weight;
          vi += px[jj].im() * weight;
        }
      }
    }
  } else {
    // footprints wider than 4 keep one load per candidate (whole rows of a 5- or 6-wide box
    // would not fit the register budget of 16 wavefronts per CU); skipped candidates read past
    // the slice end (zero) and are masked, except under fast math with the table in LDS, where
    // their weight is the zero entry
#pragma unroll
    for (int ii = 0; ii < kNB; ++ii) {
#pragma unroll
      for (int jj = 0; jj < kNB; ++jj) {
        const float d2 = xd2[jj] + yz[ii];
        constexpr bool kZeroEntry = kFastMath && kMode == kModeLookup && kTableLds;
        const bool inb = kZeroEntry ? true : d2 <= r2;
        const Pixel<kHalf> pix = load_pixel_raw<kHalf>(
          rsrc, inb ? base + ii * rowStride + (uint32_t)(jj * kPix) : kOutOfRange);
        float fre = pix.re(), fim = pix.im();
        const float wb = candidate_weight<kMode, kOrder, kFastMath, kTableLds>(d2, r2, nTable,
                                                                                 tbl, c);
        if constexpr (!kZeroEntry) {
          const bool in = d2 <= r2;
          fre = in ? fre : 0.f;
          fim = in ? fim : 0.f;
        }
        const float weight = kUnitWeight ? wb : wb * dw;
        w += weight;
        if constexpr (kFastMath) {
          vr = __builtin_fmaf(fre, weight, vr);
          vi = __builtin_fmaf(fim, weight, vi);
        } else {
          vr += fre * weight;
          vi += fim * weight;
        }
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
// With kFma the rotation is evaluated with fused multiply-adds (fewer instructions; q moves by
// an ulp, which changes which nearly-zero-weight candidates at the blob edge are counted).  The
// kernels use the reference's separate products: fused, the weights of small volumes drift to
// about 1e-5 of FRCPU's, the edge of the fast-math strategy's tolerance.
template <bool kFma>
__device__ __forceinline__ bool blob_predicate(const umpa_traverse_space &s, const umpa_fr_constants &c,
                                               int x, int y, int z, float px, float py, float pz,
                                               bool brickInside, float &qx, float &qy,
                                               float &zSqr)
{
  const float r = c.cBlobRadius;
  const float(*t)[3] = s.transformInv;
  float qz;
  if constexpr (kFma) {
    qx = __builtin_fmaf(t[0][2], pz, __builtin_fmaf(t[0][1], py, t[0][0] * px));
    qy = __builtin_fmaf(t[1][2], pz, __builtin_fmaf(t[1][1], py, t[1][0] * px));
    qz = __builtin_fmaf(t[2][2], pz, __builtin_fmaf(t[2][1], py, t[2][0] * px));
  } else {
    qx = t[0][0] * px + t[0][1] * py + t[0][2] * pz;
    qy = t[1][0] * px + t[1][1] * py + t[1][2] * pz;
    qz = t[2][0] * px + t[2][1] * py + t[2][2] * pz;
  }
  zSqr = qz * qz;
  // & rather than &&: every test is evaluated, so the record's fields load together
  bool pass = !(qx < -r) & !(zSqr > r * r);
  qy += (float)(c.cMaxVolumeIndexYZ / 2);
  if (!brickInside) {
    // the sphere test, and the traversal: the column must lie in the AABB's plane orthogonal
    // to `dir` (both hold for every voxel of a brick that lies inside them: brick_inside)
    const float dsq = px * px + py * py + pz * pz;
    pass = pass & !(dsq > s.maxDistanceSqr);
    if (s.dir == 0) {
      pass = pass & (x >= s.minX) & (x <= s.maxX) & (y >= s.minY) & (y <= s.maxY);
    } else if (s.dir == 1) {
      pass = pass & (x >= s.minX) & (x <= s.maxX) & (z >= s.minZ) & (z <= s.maxZ);
    } else {
      pass = pass & (y >= s.minY) & (y <= s.maxY) & (z >= s.minZ) & (z <= s.maxZ);
    }
  }
  // ... and the voxel must lie in the column's slab between the two bounding planes.  A voxel
  // whose |qz| is at least 0.001 inside the blob radius is between the planes by at least that
  // much along the column (the axis of the normal's largest component), five times the float
  // error of the plane evaluation (about 2e-4 at coordinates up to 256), so the floor/ceil
  // range holds it; only voxels nearer the slab faces take the reference's exact test, whose
  // two IEEE divisions a wave runs whenever one lane needs it (a band of 0.01 sent about a
  // quarter of the batches there).
  if (pass && !(fabsf(qz) <= r - 0.001f)) {
    pass = visited_blob(s, x, y, z, c.cMaxVolumeIndexX, c.cMaxVolumeIndexYZ);
  }
  return pass;
}

// Per-wave LDS state: running sums of the wave's 8 x 8 x 8 brick and the queue of voxels that
// pass the slab test for the current space.
struct WaveScratch {
  float accR[kVox];
  float accI[kVox];
  float accW[kVox];
  uint16_t queue[kVox];
};

// Copies a wave-uniform value into scalar registers, dword by dword: a job record read from
// memory at a uniform address then lives in SGPRs like a kernel argument.
template <typename T>
__device__ __forceinline__ T make_uniform(const T &v)
{
  static_assert(sizeof(T) % 4 == 0, "dword-sized records only");
  int w[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}

// Element `index` (wave-uniform) of a buffer the kernels only read, through the constant
// address space: the compiler then reads it with scalar loads into SGPRs instead of 16-byte
// vector loads of one address.  The exact kernel with cross-space batches reads its space
// records this way (1.75 -> 1.62 ms per box-256 chunk).  In fr_insert_kernel the same change
// made the fast-math and FP16 kernels slower (1.46 -> 1.94 and 1.39 -> 1.87 ms; their SGPR
// count reached 100, with spills through VGPR lanes), so they keep the vector loads.
template <typename T>
__device__ __forceinline__ T uniform_load(const T *p, int index)
{
  static_assert(sizeof(T) % 4 == 0, "dword-sized records only");
  typedef __attribute__((address_space(4))) const int const_int;
  const const_int *src = (const const_int *)(p + index);
  int w[sizeof(T) / 4];
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = src[i];
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}

// ---- the brick's steps shared by the insertion kernels ----------------------------------------
constexpr float kHalfEdge = 0.5f * (kEdge - 1);

// Brick of the work item and the lane's column of it.
struct BrickCtx {
  int x0, y0, z0;     // brick origin
  int nz;             // levels inside the volume (wave-uniform)
  int x, y;           // the lane's column
  bool colIn;         // the column lies inside the volume
  float px, py;       // the column in centred coordinates
  float bcx, bcy, bcz, hz, dmin, eps;  // brick centre, half depth and the cull's margins
};

__device__ __forceinline__ BrickCtx make_brick(const FrJob &a, const int4 &bo, int lane)
{
  BrickCtx b;
  const umpa_fr_constants &c = a.c;
  const int cxi = c.cMaxVolumeIndexX / 2, cyi = c.cMaxVolumeIndexYZ / 2;
  const int depth = bo.w & 15;
  b.x0 = bo.x;
  b.y0 = bo.y;
  b.z0 = bo.z & 0xffff;
  b.x = bo.x + (lane & (kEdge - 1));
  b.y = bo.y + lane / kEdge;
  b.colIn = b.x < a.vx && b.y < a.vy;
  b.nz = min(depth, a.vz - b.z0);
  b.px = (float)(b.x - cxi);
  b.py = (float)(b.y - cyi);
  // brick centre in centred volume coordinates, half extents (3.5, 3.5, (depth - 1) / 2)
  b.hz = 0.5f * (float)(depth - 1);
  b.bcx = (float)(bo.x - cxi) + kHalfEdge;
  b.bcy = (float)(bo.y - cyi) + kHalfEdge;
  b.bcz = (float)(b.z0 - cyi) + b.hz;
  const float bnorm = sqrtf(b.bcx * b.bcx + b.bcy * b.bcy + b.bcz * b.bcz);
  b.dmin = fmaxf(bnorm - sqrtf(2.f * kHalfEdge * kHalfEdge + b.hz * b.hz), 0.f);
  b.eps = 0.02f + 1e-5f * bnorm;
  return b;
}

// Do all of the brick's voxels pass the reference's sphere test (|p|^2 <= maxDistanceSqr; the
// farthest corner's squared distance is an exact integer in float, as is every voxel's) and lie
// in the space's traversal plane (the AABB extents orthogonal to `dir`)?  Wave-uniform.
__device__ __forceinline__ bool brick_inside(const BrickCtx &b, const umpa_traverse_space &sp,
                                             int cxi, int cyi)
{
  const int x1 = b.x0 + kEdge - 1, y1 = b.y0 + kEdge - 1, z1 = b.z0 + b.nz - 1;
  const int mx = max(abs(b.x0 - cxi), abs(x1 - cxi));
  const int my = max(abs(b.y0 - cyi), abs(y1 - cyi));
  const int mz = max(abs(b.z0 - cyi), abs(z1 - cyi));
  const bool sphere = (float)(mx * mx + my * my + mz * mz) <= sp.maxDistanceSqr;
  const bool inX = b.x0 >= sp.minX && x1 <= sp.maxX;
  const bool inY = b.y0 >= sp.minY && y1 <= sp.maxY;
  const bool inZ = b.z0 >= sp.minZ && z1 <= sp.maxZ;
  const int dir = sp.dir;
  const bool box = dir == 0 ? (inX && inY) : dir == 1 ? (inX && inZ) : (inY && inZ);
  return __builtin_amdgcn_readfirstlane((int)(sphere && box)) != 0;
}

// The brick's running sums start from the volume (or from zero for a split part).
template <typename Scratch>
__device__ __forceinline__ void load_brick(Scratch &ws, const FrJob &a, const BrickCtx &b,
                                           int lane, bool fromZero)
{
  for (int k = 0; k < b.nz; ++k) {
    const int li = k * 64 + lane;
    float vr = 0.f, vi = 0.f, vw = 0.f;
    if (!fromZero && b.colIn) {
      const int64_t gi = ((int64_t)(b.z0 + k) * a.vy + b.y) * a.vx + b.x;
      const float2 t = reinterpret_cast<const float2 *>(a.volume)[gi];
      vr = t.x;
      vi = t.y;
      vw = a.weights[gi];
    }
    ws.accR[li] = vr;
    ws.accI[li] = vi;
    ws.accW[li] = vw;
  }
  wave_sync();
}

// The brick's sums back to the volume: stored, or added with atomics for a split part.
template <typename Scratch>
__device__ __forceinline__ void store_brick(const Scratch &ws, const FrJob &a,
                                            const BrickCtx &b, int lane, bool split)
{
  for (int k = 0; k < b.nz; ++k) {
    if (b.colIn) {
      const int li = k * 64 + lane;
      const int64_t gi = ((int64_t)(b.z0 + k) * a.vy + b.y) * a.vx + b.x;
      if (split) {
        float *vol = reinterpret_cast<float *>(a.volume);
        if (ws.accW[li] != 0.f || ws.accR[li] != 0.f || ws.accI[li] != 0.f) {
          atomicAdd(vol + 2 * gi, ws.accR[li]);
          atomicAdd(vol + 2 * gi + 1, ws.accI[li]);
          atomicAdd(a.weights + gi, ws.accW[li]);
        }
      } else {
        reinterpret_cast<float2 *>(a.volume)[gi] = make_float2(ws.accR[li], ws.accI[li]);
        a.weights[gi] = ws.accW[li];
      }
    }
  }
  wave_sync();
}

// Cull of space `si` (one per lane) against the brick: slab half-width, half-space and sphere
// with the brick's extents as margins; spaces naming a slice that does not exist never pass.
// Bricks this far from the centre (in voxels) exist only in volumes above box 256; there most
// spaces miss the brick's slab, so the cull rejects on the slab test before it loads the rest of
// the record (box 1024, D7: 141.5 -> 131.2 ms per 512-slice step), while closer bricks load every
// field at once (box 256: the early rejection would cost 1 to 3 %).
constexpr float kEarlyCullRadius = 128.f;
template <bool kEarly>
__device__ __forceinline__ bool cull_space(const FrJob &a, const BrickCtx &b, int si, int sEnd)
{
  if (si >= sEnd) return false;
  const umpa_traverse_space &sp = a.spaces[si];
  const float(*t)[3] = sp.transformInv;
  const float qz = t[2][0] * b.bcx + t[2][1] * b.bcy + t[2][2] * b.bcz;
  const float ez = kHalfEdge * (fabsf(t[2][0]) + fabsf(t[2][1])) + b.hz * fabsf(t[2][2]);
  const bool slab = fabsf(qz) <= a.cullHalfWidth + ez + b.eps;
  if (kEarly && !slab) return false;
  const float qx = t[0][0] * b.bcx + t[0][1] * b.bcy + t[0][2] * b.bcz;
  const float ex = kHalfEdge * (fabsf(t[0][0]) + fabsf(t[0][1])) + b.hz * fabsf(t[0][2]);
  // the (other) conditions are all evaluated (&, not &&), so their fields load together instead
  // of one after another behind the earlier tests
  const float mds = sp.maxDistanceSqr;
  const int proj = sp.projectionIndex, dir = sp.dir;
  return slab & (qx >= -a.cullHalfSpace - ex - b.eps) &
         (b.dmin * b.dmin <= mds * 1.0001f + b.eps) & (proj >= 0) & (proj < a.nProj) &
         (dir >= 0) & (dir <= 2);
}

// Phase 1: the slab test q.z^2 <= slab2 of the column's nz voxels; the voxels that pass go to
// the wave's queue (index k * 64 + lane), their number is returned.  qz = (t20 px + t21 py) +
// t22 pz is the reference's left-to-right sum, so the shared partial sum changes no bit.
__device__ __forceinline__ int slab_queue(WaveScratch &ws, const FrJob &a, const BrickCtx &b,
                                          const umpa_traverse_space &sp, float slab2, int lane)
{
  const int cyi = a.c.cMaxVolumeIndexYZ / 2;
  const float t20 = sp.transformInv[2][0], t21 = sp.transformInv[2][1],
              t22 = sp.transformInv[2][2];
  const float cz = t20 * b.px + t21 * b.py;
  // a column outside the volume compares against -1, which no square passes: the test is one
  // compare whose lane mask is the ballot (with the column flag and-ed in, this compiler turned
  // the mask into a 0/1 value and back, two more instructions per level: 1.227 -> 1.205 ms per
  // box-256 chunk under fast math)
  const float lim = b.colIn ? slab2 : -1.f;
  int cnt = 0;
  for (int k = 0; k < b.nz; ++k) {
    const float pz = (float)(b.z0 + k - cyi);
    const float qz = cz + t22 * pz;
    const bool ok = !(qz * qz > lim);
    const unsigned long long m = __ballot(ok);
    if (ok) ws.queue[cnt + (int)lanes_below(m)] = (uint16_t)(k * 64 + lane);
    cnt += __popcll(m);
  }
  return cnt;
}

template <int kMode, int kOrder, int kNB, bool kHalf, bool kFastMath, bool kCount, bool kTableLds,
          bool kBatched, bool kShift>
__global__ __launch_bounds__(64 * fr_waves(kMode, kHalf, kFastMath)) void fr_insert_kernel(FrJob job0, FrLaunch l)
{
  constexpr int kWaves = fr_waves(kMode, kHalf, kFastMath);
  constexpr int kThreads = 64 * kWaves;
  extern __shared__ float s_table[];
  __shared__ WaveScratch s_wave[kWaves];
  __shared__ umpa_traverse_space s_space[kWaves];  // the space the wave works on
  const int nTable = l.nTable;
  if constexpr (kMode == kModeLookup && kTableLds) {
    for (int i = threadIdx.x; i < nTable; i += kThreads) s_table[i] = l.table[i];
    if (threadIdx.x == 0) s_table[nTable] = 0.f;
    __syncthreads();
  }
  const float *tbl = (kMode == kModeLookup && kTableLds) ? s_table : l.table;
  const int lane = threadIdx.x & 63;
  WaveScratch &ws = s_wave[threadIdx.x >> 6];
  constexpr int kPix = kHalf ? 4 : 8;

  // bounded launches (batches): lane 0 claims at most l.waveItems items, then an index past
  // the items.  Only the batched instantiations carry the bound: in the others any bound on
  // this loop costs the compiler's schedule (per-pair FP16 kernel 84 -> 71 VGPRs, 1.22 ->
  // 1.75 ms per 1,000-slice chunk at box 256).
  int budget = l.waveItems > 0 ? l.waveItems : -1;
  for (;;) {
    int b = 0;
    if constexpr (kBatched) {
      if (lane == 0) b = budget-- == 0 ? 0x7fffffff : atomicAdd(l.counter, 1);
    } else {
      if (lane == 0) b = atomicAdd(l.counter, 1);
    }
    b = __builtin_amdgcn_readfirstlane(__shfl(b, 0));
    if (b >= l.nItems) break;
    const int4 bo = make_uniform(reinterpret_cast<const int4 *>(l.items)[b]);
    // work item: brick 8 x 8 x depth at (bo.x, bo.y, bo.z & 0xffff) of job bo.z >> 16; heavy
    // bricks run at raised wave priority; a brick split into `parts` slice ranges accumulates
    // its range from zero and merges the partial sums with atomics (fast-math schedules only)
    const FrJob a = kBatched ? make_uniform(l.jobs[bo.z >> 16]) : job0;
    const int s0 = a.spaceBegin, s1 = a.spaceEnd;
    if (s0 >= s1) continue;
    const bool heavy = (bo.w >> 4) & 1;
    const int parts = (bo.w >> 8) & 255, part = (bo.w >> 16) & 255;
    const bool split = parts > 1;
    const int sBegin = split ? s0 + (int)((int64_t)(s1 - s0) * part / parts) : s0;
    const int sEnd = split ? s0 + (int)((int64_t)(s1 - s0) * (part + 1) / parts) : s1;
    const umpa_fr_constants &c = a.c;
    const int cxi = c.cMaxVolumeIndexX / 2, cyi = c.cMaxVolumeIndexYZ / 2;
    const float r = c.cBlobRadius;
    const float r2 = r * r;
    // phase-1 filter on the squared slab coordinate: the reference's own test for blobs; for
    // the nearest-neighbour traversal the visited voxel lies within 1.5 of the plane
    const float slab2 = kMode == kModeNearest ? 2.56f : r2;
    if (heavy) __builtin_amdgcn_s_setprio(2);
    // the lane owns the z column (x, y, z0 .. z0 + depth - 1) of the brick
    const BrickCtx bk = make_brick(a, bo, lane);
    const int bz = bk.z0;
    load_brick(ws, a, bk, lane, split);
    // wave-uniform: the brick's distance from the centre
    const bool earlyCull = __builtin_amdgcn_readfirstlane((int)(bk.dmin > kEarlyCullRadius)) != 0;
    unsigned long long nupd = 0;
    for (int base = sBegin; base < sEnd; base += 64) {
      unsigned long long mask = earlyCull
                                  ? __ballot(cull_space<true>(a, bk, base + lane, sEnd))
                                  : __ballot(cull_space<false>(a, bk, base + lane, sEnd));
      // the space's record goes through the wave's LDS slot: one load of 28 dwords by 28
      // lanes, then broadcast reads, instead of seven 16-byte loads of one address by every
      // lane (its 28 SGPRs would not fit beside this kernel's other scalar state). The next
      // surviving space's dword is loaded while the current space is processed.
      constexpr int kRecDwords = (int)(sizeof(umpa_traverse_space) / 4);
      const bool recLane = lane < kRecDwords;
      int rec = 0;
      if (mask && recLane)
        rec = reinterpret_cast<const int *>(
          a.spaces + __builtin_amdgcn_readfirstlane(base + __builtin_ctzll(mask)))[lane];
      while (mask) {
        mask &= mask - 1;
        if (recLane) reinterpret_cast<int *>(&s_space[threadIdx.x >> 6])[lane] = rec;
        wave_sync();
        if (mask && recLane)
          rec = reinterpret_cast<const int *>(
            a.spaces + __builtin_amdgcn_readfirstlane(base + __builtin_ctzll(mask)))[lane];
        const umpa_traverse_space &sp = s_space[threadIdx.x >> 6];
        const int proj = __builtin_amdgcn_readfirstlane(sp.projectionIndex);
        const int64_t off = (int64_t)proj * a.sliceElems * kPix;
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
          (void *)((const char *)a.fft + off), 0, (int)(a.sliceElems * kPix), kBufferConfig);
        const float scale = kHalf ? a.scales[proj] : 1.f;

        // phase 1: slab test of the column's nz voxels (nz is wave-uniform: shallow bricks
        // run fewer levels)
        const int cnt = slab_queue(ws, a, bk, sp, slab2, lane);
        if (cnt == 0) continue;
        wave_sync();
        const bool unit = sp.weight == 1.f;
        const bool inside = brick_inside(bk, sp, cxi, cyi);

        // phase 2: the rest of the predicate and the gather, with full lanes
        for (int t0 = 0; t0 < cnt; t0 += 64) {
          const int item = t0 + lane;
          const bool act = item < cnt;
          const int v = act ? (int)ws.queue[item] : 0;
          const int vl = v & 63, vk = v >> 6;
          const int xv = bo.x + (vl & (kEdge - 1)), yv = bo.y + vl / kEdge, zv = bz + vk;
          const float pxv = (float)(xv - cxi), pyv = (float)(yv - cyi), pzv = (float)(zv - cyi);
          float qx, qy, zs;
          bool hit;
          if constexpr (kMode == kModeNearest) {
            const float dsq = pxv * pxv + pyv * pyv + pzv * pzv;
            qx = sp.transformInv[0][0] * pxv + sp.transformInv[0][1] * pyv + sp.transformInv[0][2] * pzv;
            qy = sp.transformInv[1][0] * pxv + sp.transformInv[1][1] * pyv + sp.transformInv[1][2] * pzv;
            zs = 0.f;
            hit = act && !(dsq > sp.maxDistanceSqr) && !(qx < 0.f) && visited_nearest(sp, xv, yv, zv);
          } else {
            hit = act && blob_predicate<false>(sp, c, xv, yv, zv, pxv, pyv, pzv, inside, qx, qy, zs);
          }
          if (hit) {
            Contribution o;
            if constexpr (kMode == kModeNearest) {
              // ProcessVoxel (fr_cpu_kernels.hpp:50-67): nearest pixel, weight = space weight
              int ix = (int)(qx + 0.5f);
              ix = ix > a.xSize - 1 ? a.xSize - 1 : ix;
              ix = ix < 0 ? 0 : ix;
              int iy = (int)(qy + 0.5f + (float)(c.cMaxVolumeIndexYZ / 2));
              iy = iy > a.ySize - 1 ? a.ySize - 1 : iy;
              iy = iy < 0 ? 0 : iy;
              const float wgt = 1.f * sp.weight;
              float2 f = load_pixel<kHalf>(rsrc, (uint32_t)((iy * a.xSize + ix) * kPix));
              if constexpr (kHalf) {
                f.x *= scale;
                f.y *= scale;
              }
              o.re = f.x * wgt;
              o.im = f.y * wgt;
              o.w = wgt;
            } else {
              bool fits;
              if (unit) {
                fits = gather_box<kMode, kOrder, kNB, kHalf, kFastMath, true, kTableLds, kShift>(
                  a, nTable, rsrc, tbl, qx, qy, zs, sp.weight, o);
              } else {
                fits = gather_box<kMode, kOrder, kNB, kHalf, kFastMath, false, kTableLds, kShift>(
                  a, nTable, rsrc, tbl, qx, qy, zs, sp.weight, o);
              }
              if (!fits)
                o = gather_generic<kMode, kOrder, kHalf, kFastMath>(a, nTable, rsrc, tbl, qx, qy, zs, sp.weight);
              if constexpr (kHalf) {
                o.re *= scale;
                o.im *= scale;
              }
            }
            // volume += contribution, in space order (one item per voxel per space)
            ws.accR[v] = ws.accR[v] + o.re;
            ws.accI[v] = ws.accI[v] + o.im;
            ws.accW[v] = ws.accW[v] + o.w;
          }
          if constexpr (kCount) nupd += (unsigned long long)__popcll(__ballot(hit));
        }
        wave_sync();
      }
    }
    store_brick(ws, a, bk, lane, split);
    if (heavy) __builtin_amdgcn_s_setprio(0);
    if constexpr (kCount) {
      if (lane == 0 && nupd) atomicAdd(l.updates, nupd);
    }
  }
}

// ---- fast-math lookup insertion with queues carried across spaces ---------------------------
//
// fr_insert_kernel gathers each (brick, space) queue 64 voxels at a time, so the last batch of
// every pair runs partly empty.  This variant (fast math, FP32 or FP16 slices, table in LDS,
// 16 wavefronts per CU) lets the slab voxels of consecutive spaces share one queue: a space's
// voxels are appended behind the ones left from the previous spaces, every full batch of 64 is
// gathered as soon as it forms, and a partial batch runs only at the end of a brick or when a
// third space would be pending.  An entry carries the slot (0 / 1) of its space's record, so a
// batch holds at most two spaces; it adds their contributions space by space, the earlier one
// first (no voxel appears twice within one space), so every voxel still receives its spaces in
// index order and the sums are the ones fr_insert_kernel forms.  All slices of the chunk are
// read through one buffer resource (chunks up to 2 GiB of slices).
constexpr int kCarryQueue = kVox + 64;  // leftovers (< 64) plus one space's slab voxels

struct CarrySlot {
  umpa_traverse_space rec;
  uint32_t sliceOff;  // byte offset of the space's slice in the chunk
  float scale;        // FP16 dequantisation factor of that slice
  int inside;         // brick_inside() of the brick and this space
  uint32_t pad;
};
static_assert(sizeof(CarrySlot) == 128, "two slots per wave");

struct CarryScratch {
  float accR[kVox];
  float accI[kVox];
  float accW[kVox];
  uint16_t queue[kCarryQueue];  // voxel index (bits 0-8) | slot (bit 9)
};

// Largest blob table the carry kernel keeps in LDS beside its per-wave state.
constexpr int kCarryMaxTable =
  (160 * 1024 - kMaxWaves * (int)(sizeof(CarryScratch) + 2 * sizeof(CarrySlot))) / 4 - 1;

// Slab test of the brick's voxels against `sp`, appended to the queue behind `pending`
// entries with the space's slot in bit 9; returns the number appended.
__device__ __forceinline__ int slab_queue_carry(CarryScratch &ws, const FrJob &a,
                                                const BrickCtx &b, const umpa_traverse_space &sp,
                                                float slab2, int lane, int pending, int slot)
{
  const int cyi = a.c.cMaxVolumeIndexYZ / 2;
  const float t20 = sp.transformInv[2][0], t21 = sp.transformInv[2][1],
              t22 = sp.transformInv[2][2];
  const float cz = t20 * b.px + t21 * b.py;
  const float lim = b.colIn ? slab2 : -1.f;  // as in slab_queue
  int cnt = 0;
  const uint32_t tag = (uint32_t)slot << 9;
  for (int k = 0; k < b.nz; ++k) {
    const float pz = (float)(b.z0 + k - cyi);
    const float qz = cz + t22 * pz;
    const bool ok = !(qz * qz > lim);
    const unsigned long long m = __ballot(ok);
    if (ok) ws.queue[pending + cnt + (int)lanes_below(m)] = (uint16_t)((uint32_t)(k * 64 + lane) | tag);
    cnt += __popcll(m);
  }
  return cnt;
}

// One batch of `n` (1..64) queue entries from `start`: the rest of the predicate, the gather and
// the accumulation, space by space.
template <int kNB, bool kHalf, bool kCount, bool kShift>
__device__ __forceinline__ void carry_batch(CarryScratch &ws, const CarrySlot *slots,
                                            const FrJob &a, const int4 &bo, int bz, int nTable,
                                            __amdgpu_buffer_rsrc_t rsrc, const float *tbl,
                                            int lane, int start, int n,
                                            unsigned long long &nupd)
{
  const umpa_fr_constants &c = a.c;
  const int cxi = c.cMaxVolumeIndexX / 2, cyi = c.cMaxVolumeIndexYZ / 2;
  const bool act = lane < n;
  const uint32_t e = act ? (uint32_t)ws.queue[start + lane] : 0u;
  const int v = (int)(e & 511u), tag = (int)(e >> 9);
  const CarrySlot &slt = slots[tag];
  const umpa_traverse_space &sp = slt.rec;
  const int vl = v & 63, vk = v >> 6;
  const int xv = bo.x + (vl & (kEdge - 1)), yv = bo.y + vl / kEdge, zv = bz + vk;
  const float pxv = (float)(xv - cxi), pyv = (float)(yv - cyi), pzv = (float)(zv - cyi);
  float qx, qy, zs;
  const bool hit =
    act && blob_predicate<false>(sp, c, xv, yv, zv, pxv, pyv, pzv, slt.inside != 0, qx, qy, zs);
  Contribution o{ 0.f, 0.f, 0.f };
  if (hit) {
    const uint32_t sliceOff = slt.sliceOff;
    const float dw = sp.weight;
    bool fits;
    if (__all(!hit || dw == 1.f)) {
      fits = gather_box<kModeLookup, 0, kNB, kHalf, true, true, true, kShift>(a, nTable, rsrc, tbl, qx,
                                                                       qy, zs, dw, o, sliceOff);
    } else {
      fits = gather_box<kModeLookup, 0, kNB, kHalf, true, false, true, kShift>(a, nTable, rsrc, tbl, qx,
                                                                        qy, zs, dw, o, sliceOff);
    }
    if (!fits)
      o = gather_generic<kModeLookup, 0, kHalf, true>(a, nTable, rsrc, tbl, qx, qy, zs, dw,
                                                     sliceOff);
    if constexpr (kHalf) {
      o.re *= slt.scale;
      o.im *= slt.scale;
    }
  }
  // lane 0 holds the batch's oldest entry: its space is added first
  const int first = __builtin_amdgcn_readfirstlane(tag);
  if (hit && tag == first) {
    ws.accR[v] = ws.accR[v] + o.re;
    ws.accI[v] = ws.accI[v] + o.im;
    ws.accW[v] = ws.accW[v] + o.w;
  }
  if (__any(hit && tag != first)) {
    wave_sync();
    if (hit && tag != first) {
      ws.accR[v] = ws.accR[v] + o.re;
      ws.accI[v] = ws.accI[v] + o.im;
      ws.accW[v] = ws.accW[v] + o.w;
    }
  }
  if constexpr (kCount) nupd += (unsigned long long)__popcll(__ballot(hit));
}

template <int kNB, bool kHalf, bool kCount, bool kBatched, bool kShift>
__global__ __launch_bounds__(64 * kMaxWaves) void fr_insert_carry_kernel(FrJob job0, FrLaunch l)
{
  constexpr int kWaves = kMaxWaves;
  constexpr int kThreads = 64 * kWaves;
  constexpr int kPix = kHalf ? 4 : 8;
  extern __shared__ float s_table[];
  __shared__ CarryScratch s_wave[kWaves];
  __shared__ CarrySlot s_slot[kWaves][2];
  const int nTable = l.nTable;
  for (int i = threadIdx.x; i < nTable; i += kThreads) s_table[i] = l.table[i];
  if (threadIdx.x == 0) s_table[nTable] = 0.f;
  __syncthreads();
  const float *tbl = s_table;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  CarryScratch &ws = s_wave[wv];
  CarrySlot *slots = s_slot[wv];

  // bounded launches (batches only, as in fr_insert_kernel)
  int budget = l.waveItems > 0 ? l.waveItems : -1;
  for (;;) {
    int b = 0;
    if constexpr (kBatched) {
      if (lane == 0) b = budget-- == 0 ? 0x7fffffff : atomicAdd(l.counter, 1);
    } else {
      if (lane == 0) b = atomicAdd(l.counter, 1);
    }
    b = __builtin_amdgcn_readfirstlane(__shfl(b, 0));
    if (b >= l.nItems) break;
    const int4 bo = make_uniform(reinterpret_cast<const int4 *>(l.items)[b]);
    const FrJob a = kBatched ? make_uniform(l.jobs[bo.z >> 16]) : job0;
    const int s0 = a.spaceBegin, s1 = a.spaceEnd;
    if (s0 >= s1) continue;
    const bool heavy = (bo.w >> 4) & 1;
    const int parts = (bo.w >> 8) & 255, part = (bo.w >> 16) & 255;
    const bool split = parts > 1;
    const int sBegin = split ? s0 + (int)((int64_t)(s1 - s0) * part / parts) : s0;
    const int sEnd = split ? s0 + (int)((int64_t)(s1 - s0) * (part + 1) / parts) : s1;
    const float r = a.c.cBlobRadius;
    const float slab2 = r * r;
    const int cxi = a.c.cMaxVolumeIndexX / 2, cyi = a.c.cMaxVolumeIndexYZ / 2;
    if (heavy) __builtin_amdgcn_s_setprio(2);
    const BrickCtx bk = make_brick(a, bo, lane);
    const int bz = bk.z0;
    load_brick(ws, a, bk, lane, split);
    const uint32_t sliceBytes = (uint32_t)(a.sliceElems * kPix);
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void *)a.fft, 0, (int)((uint32_t)a.nProj * sliceBytes), kBufferConfig);
    const bool earlyCull = __builtin_amdgcn_readfirstlane((int)(bk.dmin > kEarlyCullRadius)) != 0;
    unsigned long long nupd = 0;
    int pending = 0;     // queue entries waiting, queue[0 .. pending)
    int pendSpaces = 0;  // spaces they come from (at most 2)
    int slot = 0;        // record slot of the next space
    for (int base = sBegin; base < sEnd; base += 64) {
      unsigned long long mask = earlyCull
                                  ? __ballot(cull_space<true>(a, bk, base + lane, sEnd))
                                  : __ballot(cull_space<false>(a, bk, base + lane, sEnd));
      constexpr int kRecDwords = (int)(sizeof(umpa_traverse_space) / 4);
      const bool recLane = lane < kRecDwords;
      int rec = 0;
      if (mask && recLane)
        rec = reinterpret_cast<const int *>(
          a.spaces + __builtin_amdgcn_readfirstlane(base + __builtin_ctzll(mask)))[lane];
      while (mask) {
        mask &= mask - 1;
        if (pendSpaces == 2) {
          // the slot about to be rewritten still holds a pending space: gather everything
          wave_sync();
          carry_batch<kNB, kHalf, kCount, kShift>(ws, slots, a, bo, bz, nTable, rsrc, tbl, lane, 0,
                                          pending, nupd);
          pending = 0;
          pendSpaces = 0;
          wave_sync();
        }
        const int proj = __builtin_amdgcn_readlane(rec, 17);  // projectionIndex (byte 68)
        if (recLane) reinterpret_cast<int *>(&slots[slot])[lane] = rec;
        if (lane == kRecDwords) slots[slot].sliceOff = (uint32_t)proj * sliceBytes;
        if (lane == kRecDwords + 1) slots[slot].scale = kHalf ? a.scales[proj] : 1.f;
        wave_sync();
        if (mask && recLane)
          rec = reinterpret_cast<const int *>(
            a.spaces + __builtin_amdgcn_readfirstlane(base + __builtin_ctzll(mask)))[lane];
        const bool ins = brick_inside(bk, slots[slot].rec, cxi, cyi);
        if (lane == 0) slots[slot].inside = ins ? 1 : 0;
        const int cnt = slab_queue_carry(ws, a, bk, slots[slot].rec, slab2, lane, pending, slot);
        if (cnt == 0) continue;  // the slot stays free for the next space
        pending += cnt;
        ++pendSpaces;
        slot ^= 1;
        wave_sync();
        const int full = pending >> 6;
        if (full) {
          for (int t = 0; t < full; ++t)
            carry_batch<kNB, kHalf, kCount, kShift>(ws, slots, a, bo, bz, nTable, rsrc, tbl, lane,
                                            t * 64, 64, nupd);
          // the leftovers (< 64, all of the space just queued) move to the queue's front
          const int left = pending - (full << 6);
          const int x = lane < left ? (int)ws.queue[(full << 6) + lane] : 0;
          wave_sync();
          if (lane < left) ws.queue[lane] = (uint16_t)x;
          pending = left;
          pendSpaces = left ? 1 : 0;
        }
        wave_sync();
      }
    }
    if (pending) {
      carry_batch<kNB, kHalf, kCount, kShift>(ws, slots, a, bo, bz, nTable, rsrc, tbl, lane, 0, pending,
                                      nupd);
      wave_sync();
    }
    store_brick(ws, a, bk, lane, split);
    if (heavy) __builtin_amdgcn_s_setprio(0);
    if constexpr (kCount) {
      if (lane == 0 && nupd) atomicAdd(l.updates, nupd);
    }
  }
}

// ---- exact FP32 lookup insertion with cross-space gather batches -----------------------------
//
// In fr_insert_kernel a space's queue of slab voxels is gathered 64 at a time, so the last,
// partly filled batch of every (brick, space) pair leaves lanes idle: two thirds of the gather
// lanes do work at box 256.  This variant, for the exact strategy with the table in LDS (the
// one with LDS to spare at 12 wavefronts per CU), evaluates the predicate per space as before
// but appends the hits (voxel, space ordinal, q, slice offset, weight) to a per-wave ring in
// LDS and gathers whenever 64 are waiting, across spaces.  All slices are read through one
// buffer resource over the chunk.  A voxel still receives its contributions in space order:
// entries are appended in space order, and a batch adds its contributions space by space.
constexpr int kStagedWaves = 12;
constexpr int kStage = 128;  // ring entries per wave: a full batch plus one space's hits
// largest table the staged kernel's LDS holds beside its per-wave state (160 KiB in all)
constexpr int kStagedMaxTable =
  (160 * 1024 - kStagedWaves * (int)(sizeof(WaveScratch) + 6 * 4 * kStage)) / 4 - 1;

struct WaveStage {
  float qx[kStage], qy[kStage], zs[kStage], dw[kStage];
  uint32_t off[kStage];  // byte offset of the entry's slice in the chunk
  uint32_t key[kStage];  // voxel index in the brick (bits 0-15), space ordinal (bits 16-31)
};

template <int kNB, bool kShift>
__device__ __forceinline__ void staged_gather(const FrJob &a, int nTable,
                                              __amdgpu_buffer_rsrc_t rsrc, const float *tbl,
                                              WaveScratch &ws, const WaveStage &st, int lane,
                                              int head, int n)
{
  const int e = (head + lane) & (kStage - 1);
  const bool act = lane < n;
  const float qx = act ? st.qx[e] : 0.f, qy = act ? st.qy[e] : 0.f, zs = act ? st.zs[e] : 0.f;
  const float dw = act ? st.dw[e] : 1.f;
  const uint32_t off = act ? st.off[e] : 0u, key = act ? st.key[e] : 0u;
  Contribution o{ 0.f, 0.f, 0.f };
  if (act) {
    // a weight of 1 multiplies nothing; with another weight the product wb * dw is the
    // reference's, and wb * 1 = wb, so the choice per batch changes no bit
    bool fits;
    if (__all(dw == 1.f)) {
      fits = gather_box<kModeLookup, 0, kNB, false, false, true, true, kShift>(a, nTable, rsrc, tbl, qx,
                                                                       qy, zs, dw, o, off);
    } else {
      fits = gather_box<kModeLookup, 0, kNB, false, false, false, true, kShift>(a, nTable, rsrc, tbl, qx,
                                                                        qy, zs, dw, o, off);
    }
    if (!fits)
      o = gather_generic<kModeLookup, 0, false, false>(a, nTable, rsrc, tbl, qx, qy, zs, dw, off);
  }
  // add space by space (the lowest remaining lane holds the earliest space): within a space
  // every voxel appears once
  unsigned long long rem = __ballot(act);
  while (rem) {
    const int first = __builtin_ctzll(rem);
    const uint32_t ord = (uint32_t)__builtin_amdgcn_readlane((int)(key >> 16), first);
    const unsigned long long m = __ballot(act && (key >> 16) == ord) & rem;
    if ((m >> lane) & 1ull) {
      const int v = (int)(key & 0xffffu);
      ws.accR[v] = ws.accR[v] + o.re;
      ws.accI[v] = ws.accI[v] + o.im;
      ws.accW[v] = ws.accW[v] + o.w;
    }
    rem &= ~m;
  }
}

// kShift: the job's candidate boxes slide inside the slice (FrJob::shiftBox), a separate
// instantiation: decided per candidate at run time, the choice made the compiler reschedule the
// kernel (116 -> 83 VGPRs, 1.49 -> 1.92 ms per 1,000-slice chunk at box 256); as its own
// instantiation the sliding box takes the chunk from 1.502 to 1.445 ms, bit for bit the same.
template <int kNB, bool kShift>
__global__ __launch_bounds__(64 * kStagedWaves) void fr_insert_staged_kernel(FrJob a, FrLaunch l)
{
  constexpr int kWaves = kStagedWaves;
  constexpr int kThreads = 64 * kWaves;
  constexpr int kPix = 8;
  extern __shared__ float s_table[];
  __shared__ WaveScratch s_wave[kWaves];
  __shared__ WaveStage s_stage[kWaves];
  const int nTable = l.nTable;
  for (int i = threadIdx.x; i < nTable; i += kThreads) s_table[i] = l.table[i];
  if (threadIdx.x == 0) s_table[nTable] = 0.f;
  __syncthreads();
  const float *tbl = s_table;
  const int lane = threadIdx.x & 63;
  WaveScratch &ws = s_wave[threadIdx.x >> 6];
  WaveStage &st = s_stage[threadIdx.x >> 6];
  // one resource over the chunk's slices (the host keeps it under 2 GiB); a candidate row
  // read past its slice's end reads the next slice or zeros, always at weight 0
  const uint32_t sliceBytes = (uint32_t)(a.sliceElems * kPix);
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
    (void *)a.fft, 0, (int)((uint32_t)a.nProj * sliceBytes), kBufferConfig);
  const umpa_fr_constants &c = a.c;
  const int cxi = c.cMaxVolumeIndexX / 2, cyi = c.cMaxVolumeIndexYZ / 2;
  const float r2 = c.cBlobRadius * c.cBlobRadius;

  for (;;) {
    int b = 0;
    if (lane == 0) b = atomicAdd(l.counter, 1);
    b = __builtin_amdgcn_readfirstlane(__shfl(b, 0));
    if (b >= l.nItems) break;
    const int4 bo = uniform_load(reinterpret_cast<const int4 *>(l.items), b);
    // work item: brick 8 x 8 x depth at (bo.x, bo.y, bo.z); the exact schedules never split
    const int sBegin = a.spaceBegin, sEnd = a.spaceEnd;
    if (sBegin >= sEnd) continue;
    const bool heavy = (bo.w >> 4) & 1;
    if (heavy) __builtin_amdgcn_s_setprio(2);
    const BrickCtx bk = make_brick(a, bo, lane);
    const int bz = bk.z0;
    load_brick(ws, a, bk, lane, false);
    // wave-uniform: the brick's distance from the centre
    const bool earlyCull = __builtin_amdgcn_readfirstlane((int)(bk.dmin > kEarlyCullRadius)) != 0;
    int head = 0, waiting = 0;  // the ring's oldest entry and its number of entries
    uint32_t ordinal = 0;       // spaces taken so far by this work item (mod 2^16)
    for (int base = sBegin; base < sEnd; base += 64) {
      unsigned long long mask = earlyCull
                                  ? __ballot(cull_space<true>(a, bk, base + lane, sEnd))
                                  : __ballot(cull_space<false>(a, bk, base + lane, sEnd));
      while (mask) {
        const int k0 = __builtin_ctzll(mask);
        mask &= mask - 1;
        const umpa_traverse_space sp =
          uniform_load(a.spaces, __builtin_amdgcn_readfirstlane(base + k0));
        const uint32_t sliceOff =
          (uint32_t)__builtin_amdgcn_readfirstlane(sp.projectionIndex) * sliceBytes;
        // phase 1: the slab test, as in fr_insert_kernel
        const int cnt = slab_queue(ws, a, bk, sp, r2, lane);
        if (cnt == 0) continue;
        wave_sync();
        const bool inside = brick_inside(bk, sp, cxi, cyi);
        // phase 2: the rest of the predicate; hits go to the ring, gathered 64 at a time
        for (int t0 = 0; t0 < cnt; t0 += 64) {
          const int item = t0 + lane;
          const bool act = item < cnt;
          const int v = act ? (int)ws.queue[item] : 0;
          const int vl = v & 63, vk = v >> 6;
          const int xv = bo.x + (vl & (kEdge - 1)), yv = bo.y + vl / kEdge, zv = bz + vk;
          const float pxv = (float)(xv - cxi), pyv = (float)(yv - cyi), pzv = (float)(zv - cyi);
          float qx, qy, zs;
          const bool hit =
            act && blob_predicate<false>(sp, c, xv, yv, zv, pxv, pyv, pzv, inside, qx, qy, zs);
          const unsigned long long hm = __ballot(hit);
          if (hit) {
            const int e = (head + waiting + (int)lanes_below(hm)) & (kStage - 1);
            st.qx[e] = qx;
            st.qy[e] = qy;
            st.zs[e] = zs;
            st.dw[e] = sp.weight;
            st.off[e] = sliceOff;
            st.key[e] = (uint32_t)v | (ordinal << 16);
          }
          waiting += __popcll(hm);
          if (waiting >= 64) {
            wave_sync();
            staged_gather<kNB, kShift>(a, nTable, rsrc, tbl, ws, st, lane, head, 64);
            wave_sync();
            head = (head + 64) & (kStage - 1);
            waiting -= 64;
          }
        }
        ordinal = (ordinal + 1) & 0xffffu;
        wave_sync();
      }
    }
    if (waiting > 0) {
      wave_sync();
      staged_gather<kNB, kShift>(a, nTable, rsrc, tbl, ws, st, lane, head, waiting);
      wave_sync();
    }
    store_brick(ws, a, bk, lane, false);
    if (heavy) __builtin_amdgcn_s_setprio(0);
  }
}

__global__ void fr_reset_counter(int *counter) { *counter = 0; }

// ---- host-side dispatch ------------------------------------------------------------------
// The kernel is instantiated per (mode, order, footprint width, FP16, fast math, counting, LDS
// table, batched); `visit` calls `f` with a pointer to the instantiation a variant selects.
template <int kMode, int kOrder, int kNB, bool kHalf, bool kFastMath, bool kCount, bool kShift>
static const void *pick_lds(bool lds, bool batched)
{
  // batches with a lookup table keep it in LDS (tables up to kMaxLdsTable samples)
  if constexpr (kMode == kModeLookup) {
    if (batched && !lds) return nullptr;
    if (batched) return (const void *)fr_insert_kernel<kMode, kOrder, kNB, kHalf, kFastMath, kCount, true, true, kShift>;
  }
  if (lds) return (const void *)fr_insert_kernel<kMode, kOrder, kNB, kHalf, kFastMath, kCount, true, false, kShift>;
  return batched ? (const void *)fr_insert_kernel<kMode, kOrder, kNB, kHalf, kFastMath, kCount, false, true, kShift>
                 : (const void *)fr_insert_kernel<kMode, kOrder, kNB, kHalf, kFastMath, kCount, false, false, kShift>;
}

// The sliding box (FrJob::shiftBox, every job of the launch) is an instantiation of its own for
// the fast-math kernels: tested per gather at run time, it cost 2-3 % (fast math 1.229 against
// 1.251 ms, FP16 1.189 against 1.222 ms per 1,000-slice chunk at box 256).
template <int kMode, int kOrder, int kNB, bool kHalf, bool kFastMath, bool kCount>
static const void *pick_shift(bool lds, bool batched, bool shift)
{
  if constexpr (kFastMath) {
    if (shift) return pick_lds<kMode, kOrder, kNB, kHalf, kFastMath, kCount, true>(lds, batched);
  }
  return pick_lds<kMode, kOrder, kNB, kHalf, kFastMath, kCount, false>(lds, batched);
}

template <int kMode, int kOrder, bool kHalf, bool kFastMath, bool kCount>
static const void *pick_nb(int nb, bool lds, bool batched, bool shift)
{
  switch (nb) {
  case 3: return pick_shift<kMode, kOrder, 3, kHalf, kFastMath, kCount>(lds, batched, shift);
  case 4: return pick_shift<kMode, kOrder, 4, kHalf, kFastMath, kCount>(lds, batched, shift);
  case 5: return pick_shift<kMode, kOrder, 5, kHalf, kFastMath, kCount>(lds, batched, shift);
  default: return pick_shift<kMode, kOrder, 6, kHalf, kFastMath, kCount>(lds, batched, shift);
  }
}

// Precise Kaiser (kDynamic, alpha > 15): one instantiation per blob order
// (fr_common_kernels.hpp:129-161); orders outside 0..4 are rejected at Init.
template <bool kCount>
static const void *pick_order(int order, bool batched)
{
  switch (order) {
  case 0: return pick_lds<kModePreciseKaiser, 0, 4, false, false, kCount, false>(false, batched);
  case 1: return pick_lds<kModePreciseKaiser, 1, 4, false, false, kCount, false>(false, batched);
  case 2: return pick_lds<kModePreciseKaiser, 2, 4, false, false, kCount, false>(false, batched);
  case 3: return pick_lds<kModePreciseKaiser, 3, 4, false, false, kCount, false>(false, batched);
  case 4: return pick_lds<kModePreciseKaiser, 4, 4, false, false, kCount, false>(false, batched);
  default: return nullptr;
  }
}

static bool table_in_lds(const FrVariant &v, int nTable)
{
  return v.mode == kModeLookup && nTable <= kMaxLdsTable && !(v.tableGlobal && !v.batched);
}

static const void *pick_kernel(const FrVariant &v, int nTable)
{
  const bool lds = table_in_lds(v, nTable);
  const bool b = v.batched;
  switch (v.mode) {
  case kModeLookup:
    if (v.count) {
      if (v.half) return pick_nb<kModeLookup, 0, true, true, true>(v.nb, lds, b, v.shift);
      return v.fastMath ? pick_nb<kModeLookup, 0, false, true, true>(v.nb, lds, b, v.shift)
                        : pick_nb<kModeLookup, 0, false, false, true>(v.nb, lds, b, v.shift);
    }
    if (v.half) return pick_nb<kModeLookup, 0, true, true, false>(v.nb, lds, b, v.shift);
    if (v.fastMath) return pick_nb<kModeLookup, 0, false, true, false>(v.nb, lds, b, v.shift);
    return pick_nb<kModeLookup, 0, false, false, false>(v.nb, lds, b, v.shift);
  case kModeFastKaiser:
    return v.count ? pick_lds<kModeFastKaiser, 0, 4, false, false, true, false>(false, b)
                   : pick_lds<kModeFastKaiser, 0, 4, false, false, false, false>(false, b);
  case kModeNearest:
    return v.count ? pick_lds<kModeNearest, 0, 4, false, false, true, false>(false, b)
                   : pick_lds<kModeNearest, 0, 4, false, false, false, false>(false, b);
  case kModePreciseKaiser:
    // the counting variant evaluates the same blob order as the plain one
    return v.count ? pick_order<true>(v.order, b) : pick_order<false>(v.order, b);
  default: return nullptr;
  }
}

// The exact FP32 lookup kernel with cross-space gather batches, for the variants it covers
// (the launch also needs the chunk's slices under 2 GiB, one buffer resource).
static const void *pick_staged(const FrVariant &v, int nTable, bool shift)
{
  if (v.mode != kModeLookup || v.half || v.fastMath || v.count || v.batched) return nullptr;
  if (!table_in_lds(v, nTable) || nTable > kStagedMaxTable) return nullptr;
  if (v.nb == 3) return shift ? (const void *)fr_insert_staged_kernel<3, true>
                              : (const void *)fr_insert_staged_kernel<3, false>;
  if (v.nb == 4) return shift ? (const void *)fr_insert_staged_kernel<4, true>
                              : (const void *)fr_insert_staged_kernel<4, false>;
  return nullptr;
}

static bool staged_fits(const FrJob &job)
{
  return (uint64_t)job.nProj * (uint64_t)job.sliceElems * 8u <= 0x7fff0000ull;
}

// The fast-math kernel with queues carried across spaces, for the variants it covers (the
// launch also needs the chunk's slices under 2 GiB, one buffer resource).
static const void *pick_carry(const FrVariant &v, int nTable)
{
  if (v.mode != kModeLookup || !v.fastMath) return nullptr;
  if (!table_in_lds(v, nTable) || nTable > kCarryMaxTable) return nullptr;
  if (v.nb != 3 && v.nb != 4) return nullptr;
#define UMPA_CARRY(NB, H, B)                                                                 \
  (v.shift ? (v.count ? (const void *)fr_insert_carry_kernel<NB, H, true, B, true>           \
                      : (const void *)fr_insert_carry_kernel<NB, H, false, B, true>)         \
           : (v.count ? (const void *)fr_insert_carry_kernel<NB, H, true, B, false>          \
                      : (const void *)fr_insert_carry_kernel<NB, H, false, B, false>))
  if (v.batched) {
    if (v.half) return v.nb == 3 ? UMPA_CARRY(3, true, true) : UMPA_CARRY(4, true, true);
    return v.nb == 3 ? UMPA_CARRY(3, false, true) : UMPA_CARRY(4, false, true);
  }
  if (v.half) return v.nb == 3 ? UMPA_CARRY(3, true, false) : UMPA_CARRY(4, true, false);
  return v.nb == 3 ? UMPA_CARRY(3, false, false) : UMPA_CARRY(4, false, false);
#undef UMPA_CARRY
}

static bool carry_fits(const FrJob &job, bool half)
{
  return (uint64_t)job.nProj * (uint64_t)job.sliceElems * (half ? 4u : 8u) <= 0x7fff0000ull;
}

hipError_t prepare_fr_insert(const FrVariant &v, int nTable)
{
  if (!pick_kernel(v, nTable)) return hipErrorInvalidValue;
  if (!table_in_lds(v, nTable)) return hipSuccess;
  const int bytes = (int)((nTable + 1) * sizeof(float));
  hipError_t e = hipSuccess;
  // every kernel the launch may pick: either box form (a plain launch takes it from its job)
  for (bool shift : { false, true }) {
    FrVariant w = v;
    w.shift = shift;
    for (const void *k : { pick_kernel(w, nTable), pick_staged(w, nTable, shift),
                           pick_carry(w, nTable) })
      if (e == hipSuccess && k)
        e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  }
  return e;
}

hipError_t launch_fr_insert(const FrJob &job, const FrLaunch &l, const FrVariant &v0, int numCUs,
                            hipStream_t st)
{
  // a plain launch slides its boxes when its job does; a batch's host sets v.shift when every
  // job of the launch does
  FrVariant v = v0;
  if (!v.batched) v.shift = job.shiftBox != 0;
  const void *k = pick_kernel(v, l.nTable);
  if (!k) return hipErrorInvalidValue;
  const size_t shmem = table_in_lds(v, l.nTable) ? (size_t)(l.nTable + 1) * sizeof(float) : 0;
  // persistent workgroups: the per-wave LDS state allows one per CU
  int waves = fr_waves(v.mode, v.half, v.fastMath);
  // kKernelAuto: the staged kernel for the exact strategy, the carried queues for fast-math
  // FP32 slices (FP16 slices run faster on the per-pair kernel: 1.22 vs 1.26 ms per 1,000-slice
  // chunk at box 256); batches run the per-pair kernel unless they name the carried queues
  // (their host checks that every job's slices fit one buffer resource)
  const int want = v.kernel != kKernelAuto ? v.kernel
                   : v.batched            ? kKernelPair
                   : !v.fastMath          ? kKernelStaged
                   : v.half               ? kKernelPair
                                          : kKernelCarry;
  const void *staged = want == kKernelStaged ? pick_staged(v, l.nTable, v.shift) : nullptr;
  const void *carry = want == kKernelCarry ? pick_carry(v, l.nTable) : nullptr;
  if (staged && staged_fits(job)) {
    k = staged;
    waves = kStagedWaves;
  } else if (carry && carry_fits(job, v.half)) {
    k = carry;
  }
  // persistent workgroups, one per CU, until the items run out; with l.waveItems, every wave
  // leaves after that many items and the grid covers the items (CUs come free for other
  // streams' launches as workgroups end)
  int blocks;
  if (l.waveItems > 0) {
    const int64_t per = (int64_t)waves * l.waveItems;
    blocks = (int)((l.nItems + per - 1) / per);
  } else {
    blocks = (l.nItems + waves - 1) / waves;
    if (blocks > numCUs) blocks = numCUs;
    if (l.maxBlocks > 0 && blocks > l.maxBlocks) blocks = l.maxBlocks;
  }
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(fr_reset_counter, dim3(1), dim3(1), 0, st, l.counter);
  FrJob j = job;
  FrLaunch ll = l;
  void *args[] = { &j, &ll };
  return hipLaunchKernel(k, dim3(blocks), dim3(64 * waves), args, shmem, st);
}

}  // namespace umpa

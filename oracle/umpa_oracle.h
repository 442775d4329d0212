/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement of Umpalumpa's Fourier-reconstruction (FR) insertion path and of the
 * Fourier-processing (FP) crop step that feeds it.  It exists to check the HIP product path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as
 * the reference the HIP results are compared with or the reported CPU baseline, never as a
 * fallback of the product.
 *
 * Every function names the reference file:line it follows (paths relative to the reference
 * repository root).  Parity pins, per function:
 *   orc_bessi0/1/2/3/4          bit-exact vs the compiled reference bessier.hpp (oracle/_ref)
 *   orc_kaiser_value,
 *   orc_kaiser_fourier_value,
 *   orc_create_constants,
 *   orc_fill_blob_table         closed forms (I0(15), I_{3/2}(15), KB Fourier value at 0) to
 *                               1e-6 relative, table monotonicity and end points; orders 1-4
 *                               against s^k I_k(alpha s) / I_k(alpha) (mpmath), and the NaN
 *                               cIw0 / table that the order-0-only Fourier value gives them (the
 *                               reference afr.cpp / kaiser.hpp need spdlog, absent here: not
 *                               compiled)
 *   orc_compute_traverse_space  reference KATs (afr_common.hpp:136-254) + bit-exact vs the
 *                               compiled reference traverse_space_generator.hpp
 *   orc_fr_execute              bit-exact vs the compiled reference fr_cpu_kernels.hpp golden
 *                               vectors (tests/golden; every mode, blob orders 0-4), zero/non-
 *                               zero pattern of the reference tests, brute-force per-voxel
 *                               gather on tiny inputs
 *   orc_fr_execute_f64          exact agreement with orc_fr_execute where FP32 is exact (axis-
 *                               aligned rotations, integer slices, table and space weights),
 *                               agreement to FP32 accumulation error with random orientations
 *                               and non-unit space weights, exact scaling by a power-of-two
 *                               weight (tests/test_oracle.py)
 *   orc_fp_execute              bit-exact vs the compiled reference fp_cpu_kernels.hpp and the
 *                               reference value tests (fp_tests.hpp / afp_common.hpp:132-178)
 */
#ifndef UMPA_ORACLE_H
#define UMPA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 112-byte record, same layout as fourier_reconstruction/traverse_space.hpp:14-48 */
typedef struct orc_point3 { float x, y, z; } orc_point3;
typedef struct orc_traverse_space {
  int minY, minX, minZ;
  int maxY, maxX, maxZ;
  float maxDistanceSqr;
  int dir; /* 0 = XY, 1 = XZ, 2 = YZ (traverse_space.hpp:29-33) */
  orc_point3 unitNormal;
  orc_point3 topOrigin, bottomOrigin;
  int projectionIndex;
  float transformInv[3][3];
  float weight;
} orc_traverse_space;

/* 32-byte record, fourier_reconstruction/constants.hpp:6-16 */
typedef struct orc_constants {
  int cMaxVolumeIndexX;
  int cMaxVolumeIndexYZ;
  float cBlobRadius;
  float cOneOverBlobRadiusSqr;
  float cBlobAlpha;
  float cIw0;
  float cIDeltaSqrt;
  float cOneOverBessiOrderAlpha;
} orc_constants;

/* fourier_reconstruction/settings.hpp:8-67 (enum values as int) */
typedef struct orc_fr_settings {
  int interpolation; /* 0 kDynamic, 1 kLookup */
  int order;         /* 0..4 = BlobOrder k0..k4 */
  int type;          /* 0 kFast, 1 kPrecise */
  float alpha;
  float blobRadius;
} orc_fr_settings;

double orc_bessi0(double x);
double orc_bessi1(double x);
double orc_bessi2(double x);
double orc_bessi3(double x);
double orc_bessi4(double x);
float orc_bessi1_5f(float x);
float orc_get_bessi_order_alpha(int order, float alpha);
float orc_kaiser_value(float freq, float blobRadius, float alpha, int order);
float orc_kaiser_fourier_value(float freq, float blobRadius, float alpha, int order);

orc_constants orc_create_constants(size_t volumeX, size_t volumeY, size_t tableElems,
                                   const orc_fr_settings *s);
void orc_fill_blob_table(float *table, size_t n, size_t volumeX, size_t volumeY,
                         const orc_fr_settings *s);

void orc_compute_traverse_space(uint32_t imgSizeX, uint32_t imgSizeY, const float transform[9],
                                orc_traverse_space *space, uint32_t maxVolumeIndexX,
                                uint32_t maxVolumeIndexYZ, int useFast, float blobRadius,
                                float weight);

/* FRCPU insertion.  volume: (P+1)^3 interleaved complex float, weights: (P+1)^3 float,
 * fft: nProj x ySize x xSize interleaved complex float.  Returns 0 on success, -1 when the
 * reference FRCPU Execute returns false (kLookup with an empty table, fr_cpu.cpp:27-29).
 * updates (may be NULL) receives the number of (voxel, space) pairs reaching accumulation. */
int orc_fr_execute(const orc_fr_settings *s, float *volume, float *weights, int xSize, int ySize,
                   const float *fft, const orc_traverse_space *spaces, size_t nSpaces,
                   const float *table, size_t tableElems, const orc_constants *c,
                   long long *updates);

/* Same traversal and arithmetic with every intermediate in double precision (the FP64 host
 * oracle).  Supports kPrecise + kLookup only; returns -1 otherwise. */
int orc_fr_execute_f64(const orc_fr_settings *s, double *volume, double *weights, int xSize,
                       int ySize, const float *fft, const orc_traverse_space *spaces,
                       size_t nSpaces, const float *table, size_t tableElems,
                       const orc_constants *c);

/* ScaleFFT2DCPU (fp_cpu_kernels.hpp:8-41).  Sizes are {x, y, z, n}. */
void orc_fp_execute(const float *in, float *out, const size_t inSize[4],
                    const size_t inSpatialSize[4], const size_t outSize[4], const float *filter,
                    float normFactor, float maxFreqSquare, int applyFilter, int normalize,
                    int center, int cropFreq, int shift);

/* Idx2Freq (fp_common_kernels.hpp:13-20) */
float orc_idx2freq(size_t idx, size_t size);

#ifdef __cplusplus
}
#endif
#endif

/*
 * umpalumpa_amd -- C ABI of the MI355X-native Fourier-reconstruction (FR) insertion and
 * Fourier-processing (FP) crop operations.
 *
 * Every entry point names the reference interface it replaces (paths relative to the reference
 * repository root).  Conventions shared by all entry points:
 *   - status codes, never exceptions: 0 = OK, see UMPA_* below;
 *   - the caller owns every buffer ("We do not manage your memory", data/payload.hpp:7-11);
 *     data pointers passed to *_execute are device (or managed) pointers;
 *   - *_execute enqueues on the operation's HIP stream and returns; *_synchronize blocks
 *     (FRCUDA::Synchronize, fr_cuda.cpp:243).  Handles are not thread-safe individually;
 *     distinct handles are independent.
 */
#ifndef UMPALUMPA_AMD_H
#define UMPALUMPA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (BasicOperation::Init/Execute return bool, basic_operation.hpp:28-66) --- */
#define UMPA_OK 0
#define UMPA_ERR_INVALID_ARG 1     /* malformed descriptor / pointer / size: Init would be false */
#define UMPA_ERR_NOT_INITIALIZED 2 /* Execute before a successful Init (basic_operation.hpp:62) */
#define UMPA_ERR_NOT_EQUIVALENT 3  /* payload not equivalent to the Init one (:62-63) */
#define UMPA_ERR_REJECTED 4        /* the reference Execute returns false (fr_cpu.cpp:27-29) */
#define UMPA_ERR_UNSUPPORTED 5     /* valid in the reference, not offered here */
#define UMPA_ERR_DEVICE 6          /* HIP runtime error */
#define UMPA_ERR_OUT_OF_MEMORY 7

/* Size (data/size.hpp:23-103): x, y, z, n (n = number of objects) */
typedef struct umpa_size {
  uint64_t x, y, z, n;
} umpa_size;

/* TraverseSpace (fourier_reconstruction/traverse_space.hpp:14-48), byte compatible: 112 B */
typedef struct umpa_point3 {
  float x, y, z;
} umpa_point3;
typedef struct umpa_traverse_space {
  int32_t minY, minX, minZ;
  int32_t maxY, maxX, maxZ;
  float maxDistanceSqr;
  int32_t dir; /* 0 = XY, 1 = XZ, 2 = YZ */
  umpa_point3 unitNormal;
  umpa_point3 topOrigin, bottomOrigin;
  int32_t projectionIndex;
  float transformInv[3][3];
  float weight;
} umpa_traverse_space;

/* Constants (fourier_reconstruction/constants.hpp:6-16), 32 B */
typedef struct umpa_fr_constants {
  int32_t cMaxVolumeIndexX;
  int32_t cMaxVolumeIndexYZ;
  float cBlobRadius;
  float cOneOverBlobRadiusSqr;
  float cBlobAlpha;
  float cIw0;
  float cIDeltaSqrt;
  float cOneOverBessiOrderAlpha;
} umpa_fr_constants;

/* fourier_reconstruction::Settings (settings.hpp:8-67) */
#define UMPA_FR_DYNAMIC 0 /* Interpolation::kDynamic */
#define UMPA_FR_LOOKUP 1  /* Interpolation::kLookup */
#define UMPA_FR_FAST 0    /* Type::kFast (nearest neighbour) */
#define UMPA_FR_PRECISE 1 /* Type::kPrecise (blob) */
typedef struct umpa_fr_settings {
  int32_t interpolation;
  int32_t order; /* BlobOrder k0..k4 (blob_order.hpp:5-23) */
  int32_t type;
  float alpha;       /* default 15 */
  float blobRadius;  /* default 1.9 */
} umpa_fr_settings;

/* Shapes of AFR's payloads (afr.hpp:13-55, fr.cpp:240-249).
 *   fft:    frequency-domain size of the slices, x = Hx, y = Hy, z = 1, n = slice count
 *   volume: frequency-domain size of the volume (hasSymetry: full x stored), z = y, n = 1
 *   traverseSpaces: capacity (Size(1,1,1,n)); blobTable: element count (0 for kDynamic) */
typedef struct umpa_fr_desc {
  umpa_size fft;
  umpa_size volume;
  uint64_t traverseSpaces;
  uint64_t blobTable;
} umpa_fr_desc;

/* Precision of the slice data handed to execute */
#define UMPA_SLICE_F32 0 /* interleaved complex float (float2), 8 B per pixel */
#define UMPA_SLICE_F16 1 /* interleaved complex half + one float scale per slice, 4 B per pixel */

/* Kernel strategy of an FR operation (the counterpart of FRCUDA's tuned kernel variants) */
#define UMPA_FR_EXACT 0 /* FRCPU arithmetic and order: bit-identical to the reference CPU path */
#define UMPA_FR_FAST_MATH 1 /* fused multiply-add, same order; within 1e-5 rel-L2 of FRCPU */

typedef struct umpa_fr_op umpa_fr_op;

/* ---- FR: the AFR operation, executed by hand-written gfx950 kernels (drop-in for FRCUDA,
 *      fr_cuda.cpp:12-250, beside FRCPU, fr_cpu.cpp:9-54) ----------------------------------- */

/* New operation bound to `device` and HIP stream `stream` (NULL = the null stream).
 * Replaces constructing FRCUDA/FRCPU (fr_cuda.hpp, fr_cpu.hpp). */
int umpa_fr_create(umpa_fr_op **op, int device, void *stream);

/* AFR::Init (basic_operation.hpp:28-48 + Strategy1::Init, fr_cpu.cpp:14-19,
 * fr_cuda.cpp:71-75): false (UMPA_ERR_INVALID_ARG) for an even volume x, invalid sizes or a
 * volume whose y != z.  Allocates the operation's own brick schedule on the device. */
int umpa_fr_init(umpa_fr_op *op, const umpa_fr_desc *desc, const umpa_fr_settings *settings);

/* AFR::Execute (basic_operation.hpp:60-66 + FRCPU Strategy1::Execute, fr_cpu.cpp:21-44).
 * Accumulates every traverse space into `volume` (complex float, [z][y][x]) and `weights`
 * (float, [z][y][x]) in place; the caller zeroes them first (fr.cpp:46).  `fft` holds
 * fftCount slices (n of the Init descriptor or fewer), `spaces` nSpaces records (device
 * memory), `table` the blob table (device memory, blobTable elements, kLookup only).
 * Returns UMPA_ERR_REJECTED where FRCPU returns false (kLookup without a table). */
int umpa_fr_execute(umpa_fr_op *op, const void *fft, uint64_t fftCount, void *volume,
                    float *weights, const umpa_traverse_space *spaces, uint64_t nSpaces,
                    const float *table);

/* Same as umpa_fr_execute with FP16 slices: `fft` is complex half [n][Hy][Hx] and `scales`
 * one float per slice (pixel = half value * scale); accumulation stays FP32.  North-star
 * precision mode; no counterpart in the reference. */
int umpa_fr_execute_f16(umpa_fr_op *op, const void *fft, const float *scales,
                        uint64_t fftCount, void *volume, float *weights,
                        const umpa_traverse_space *spaces, uint64_t nSpaces,
                        const float *table);

/* Init + Execute in one call that allocates no memory: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under stream capture, tight serving
 * loops).  Arguments as umpa_fr_init + umpa_fr_execute (umpa_fr_execute_f16 when `scales` is
 * not NULL) of an operation on `device` and `stream` with `strategy`; same checks, status
 * codes and results as that operation with its default schedule.  The work counter and the
 * work items live in a static device array of 8 MiB per GPU (held until the process ends),
 * which takes volumes up to 513^3: larger ones give UMPA_ERR_UNSUPPORTED.  A call whose
 * volume shape (or strategy class) differs from the previous call's on that GPU first waits
 * for `stream` and copies the new work items in (a blocking copy, so such a call is not
 * capturable); other calls only enqueue kernels.  All calls on one GPU share the array: calls
 * that overlap on one GPU (other streams or host threads) race. */
int umpa_fr_execute_noalloc(int device, void *stream, const umpa_fr_desc *desc,
                            const umpa_fr_settings *settings, int strategy, const void *fft,
                            const float *scales, uint64_t fftCount, void *volume,
                            float *weights, const umpa_traverse_space *spaces,
                            uint64_t nSpaces, const float *table);

/* Select the kernel strategy (UMPA_FR_EXACT default).  No counterpart in the reference API:
 * FRCUDA picks kernel variants through KTT (fr_cuda.cpp:100-157). */
int umpa_fr_set_strategy(umpa_fr_op *op, int strategy);

/* FRCUDA::Synchronize (fr_cuda.cpp:243): wait for every enqueued execute */
int umpa_fr_synchronize(umpa_fr_op *op);

/* BasicOperation::Cleanup (basic_operation.hpp:72-79): drop Init state and device buffers */
int umpa_fr_cleanup(umpa_fr_op *op);

/* destroy the handle (the operation's destructor, basic_operation.hpp:14) */
void umpa_fr_destroy(umpa_fr_op *op);

/* voxel-updates counted by the last execute when counting is enabled (0 otherwise) */
int umpa_fr_set_count_updates(umpa_fr_op *op, int enable);
int umpa_fr_last_updates(umpa_fr_op *op, uint64_t *updates);

/* ---- Batched FR: many reconstructions through one persistent kernel ---------------------
 * No counterpart in the reference API, which runs one AFR per volume (fr.cpp:9-76); a batch
 * is the same AFR::Execute applied to several (slices, volume) jobs, each job's voxels
 * receiving its spaces in index order as FRCPU would (results identical to running the jobs
 * one by one with the same strategy; bit-identical to FRCPU under UMPA_FR_EXACT). */

/* One reconstruction of a batch.  desc/settings as for umpa_fr_init (desc.fft.n = slices in
 * `fft`); spaces: desc.traverseSpaces records, projectionIndex < desc.fft.n; fft complex
 * float, or complex half when `scales` (one float per slice) is not NULL; all pointers device
 * memory; table: device blob table of desc.blobTable floats (kLookup).  The volumes and
 * weights of the jobs passed to one plan / step must not overlap one another (each job owns
 * its outputs; a voxel is summed by one wavefront without atomics): overlapping outputs give
 * UMPA_ERR_INVALID_ARG.  To insert two slice sets into one volume, make them one job or run
 * them in successive steps. */
typedef struct umpa_fr_job {
  umpa_fr_desc desc;
  umpa_fr_settings settings;
  const void *fft;
  const float *scales;
  void *volume;
  float *weights;
  const umpa_traverse_space *spaces;
  const float *table;
} umpa_fr_job;

typedef struct umpa_fr_batch umpa_fr_batch;

/* new batch runner on `device` and `stream` (NULL: the runner creates its own stream, which a
 * graph capture needs) */
int umpa_fr_batch_create(umpa_fr_batch **b, int device, void *stream);

/* Plans the chunk loop of `nJobs` jobs: step c inserts spaces [c * chunkSpaces,
 * (c + 1) * chunkSpaces) of every job that has them, one persistent launch per group of jobs
 * sharing settings, table and slice precision; work items (bricks of all jobs of the step)
 * are ordered heaviest first.  With `useGraph`, the whole loop is captured once as a HIP graph
 * and replayed by each umpa_fr_batch_execute.  The job array is copied; the buffers it names
 * must stay valid while the plan is used.  kLookup jobs keep their blob table in LDS: tables
 * of more than 11,776 samples give UMPA_ERR_UNSUPPORTED. */
int umpa_fr_batch_plan(umpa_fr_batch *b, const umpa_fr_job *jobs, uint32_t nJobs,
                       uint32_t chunkSpaces, int strategy, int useGraph);

/* Enqueues every step of the plan (asynchronous; volumes accumulate in place). */
int umpa_fr_batch_execute(umpa_fr_batch *b);

/* Enqueues one step over explicit space ranges: job i inserts spaces [begin[i], end[i]).  For
 * streaming callers whose jobs advance independently (continuous batching); staging buffers
 * are pinned and recycled by stream order, so successive calls do not block. */
int umpa_fr_batch_step(umpa_fr_batch *b, const umpa_fr_job *jobs, const uint64_t *begin,
                       const uint64_t *end, uint32_t nJobs, int strategy);

/* The checks umpa_fr_batch_plan and umpa_fr_batch_step apply to a job list (each job as
 * umpa_fr_init + umpa_fr_execute would check it, and outputs that do not overlap), on the
 * host: no device memory is read and no device is needed.  Same status codes. */
int umpa_fr_batch_check_jobs(const umpa_fr_job *jobs, uint32_t nJobs, int strategy);

/* number of steps of the current plan */
int umpa_fr_batch_num_steps(umpa_fr_batch *b, uint32_t *steps);
/* counting variant for the next execute / step (bypasses the graph); result after it */
int umpa_fr_batch_set_count_updates(umpa_fr_batch *b, int enable);
int umpa_fr_batch_last_updates(umpa_fr_batch *b, uint64_t *updates);
/* Bounded launches: with itemsPerWave > 0 every wavefront of the batch's later launches (and
 * plans) leaves after that many work items, and the grid covers all items, so CUs come free
 * while a launch runs and another stream's launches (a second batch) start at once instead of
 * after it; 0 (the default) keeps one persistent workgroup per CU.  Results do not change. */
int umpa_fr_batch_set_wave_items(umpa_fr_batch *b, uint32_t itemsPerWave);
/* Persistent launches (itemsPerWave 0) of the batch's later steps (and plans) use at most this
 * many workgroups, one per CU, leaving the other CUs to another stream's launches; 0 (the
 * default): one per CU.  Results do not change. */
int umpa_fr_batch_set_max_workgroups(umpa_fr_batch *b, uint32_t workgroups);

/* records an event on the batch stream after the last enqueued work; query / wait on it */
int umpa_fr_batch_record(umpa_fr_batch *b, uint64_t *ticket);
int umpa_fr_batch_query(umpa_fr_batch *b, uint64_t ticket, int *done);
int umpa_fr_batch_synchronize(umpa_fr_batch *b);
void umpa_fr_batch_destroy(umpa_fr_batch *b);

/* ---- Autotuning -------------------------------------------------------------------------
 * The project's own tuner, in place of KTT (fr_cuda.cpp:100-157) and of its tuning data kept
 * beside the executable (strategy_manager.cpp:32-45, utils/system.cpp:39-45).  It times every
 * candidate schedule on scratch volumes (the caller's volumes are not touched) with the
 * caller's slices, spaces and table, keeps the fastest for the handle, and records it in
 * $UMPA_TUNING_DIR/fr_tuning.txt (default $HOME/.cache/umpalumpa_amd) under a key of the
 * device, strategy, precision, footprint and problem shape; with useCache a recorded result is
 * reused without timing.  Results are the same, bit for bit, under every schedule and kernel
 * knob for UMPA_FR_EXACT. */
typedef struct umpa_fr_sched_params {
  float depthScale;  /* bricks 8x8x{8,4,2,1} beyond radius {48,24,12} x depthScale */
  int32_t priority;  /* heavy bricks at raised wave priority */
  int32_t maxParts;  /* split heavy bricks into at most this many slice ranges (fast math) */
  float splitAt;     /* ... above splitAt x one wavefront's fair share of the launch */
  /* Kernel knobs; 0 selects the default in each.  Batches keep the table in LDS and run the
   * per-pair kernel, or the carried queues when `kernel` names them (fast math) and every
   * job's slices fit one 2 GiB buffer resource. */
  int32_t kernel;      /* UMPA_FR_KERNEL_*: insertion kernel; one that does not cover the
                          variant (mode, strategy, footprint) runs as UMPA_FR_KERNEL_PAIR */
  int32_t tableGlobal; /* 1: blob table read from global memory instead of LDS */
  int32_t boxFixed;    /* 1: candidate boxes stay at the blob's bounding box instead of
                          sliding inside the slice (fast-math kernels and the staged exact
                          kernel; sliding needs no masking of out-of-box candidates) */
} umpa_fr_sched_params;

enum {
  UMPA_FR_KERNEL_AUTO = 0,   /* staged for UMPA_FR_EXACT, carried for fast-math FP32, else pair */
  UMPA_FR_KERNEL_PAIR = 1,   /* one queue of slab voxels per (brick, space), 16 or 12 waves */
  UMPA_FR_KERNEL_CARRY = 2,  /* queues carried across spaces (fast math, footprint 3-4) */
  UMPA_FR_KERNEL_STAGED = 3  /* cross-space ring of entries (exact FP32, footprint 3-4) */
};

/* The candidates umpa_fr_autotune times under `strategy` (UMPA_FR_EXACT: every schedule with
 * every kernel knob; UMPA_FR_FAST_MATH: the schedules with default kernel knobs, then the
 * kernel knobs over the default schedule, which the tuner applies to the fastest schedule).
 * out may be NULL to query *count, else capacity must hold them. */
int umpa_fr_tuning_candidates(int strategy, umpa_fr_sched_params *out, uint64_t capacity,
                              uint64_t *count);

/* Tunes the schedule the op uses under its current strategy for inputs like these (as for
 * umpa_fr_execute; scales != NULL for FP16 slices).  Not under graph capture (allocates). */
int umpa_fr_autotune(umpa_fr_op *op, const void *fft, const float *scales, uint64_t fftCount,
                     const umpa_traverse_space *spaces, uint64_t nSpaces, const float *table,
                     int useCache, umpa_fr_sched_params *chosen);
int umpa_fr_get_sched_params(umpa_fr_op *op, umpa_fr_sched_params *out);
/* Sets the schedule knobs the op uses under `strategy` (its work items are rebuilt; waits for
 * the op's stream first).  UMPA_FR_EXACT ignores maxParts / splitAt (it never splits). */
int umpa_fr_set_sched_params(umpa_fr_op *op, int strategy, const umpa_fr_sched_params *p);

/* Tunes the schedule of a batch's later plans / steps under `strategy` on the first step of
 * the chunk loop these jobs would form. */
int umpa_fr_batch_autotune(umpa_fr_batch *b, const umpa_fr_job *jobs, uint32_t nJobs,
                           uint32_t chunkSpaces, int strategy, int useCache,
                           umpa_fr_sched_params *chosen);
int umpa_fr_batch_set_sched_params(umpa_fr_batch *b, int strategy,
                                   const umpa_fr_sched_params *p);

/* ---- FR host-side helpers (no GPU needed) ----------------------------------------------- */

/* The insertion's work items for a vx * vy * vz volume under the schedule knobs `p`, heaviest
 * first, as int32 quadruples {x, y, z, flags}: the brick at (x, y, z) is 8 x 8 x D voxels with
 * D = flags & 15, raised priority when flags bit 4 is set, and a brick split over P = (flags
 * >> 8) & 255 wavefronts (P > 1 only with splitHeavy, the fast-math schedules) appears P times
 * with part (flags >> 16) & 255.  *count receives the number of items; items may be NULL to
 * query it, else capacity must hold them. */
int umpa_fr_brick_schedule(uint32_t vx, uint32_t vy, uint32_t vz, int splitHeavy,
                           const umpa_fr_sched_params *p, int32_t *items, uint64_t capacity,
                           uint64_t *count);

/* AFR::CreateConstants (afr.cpp:6-19) */
int umpa_fr_create_constants(const umpa_fr_desc *desc, const umpa_fr_settings *settings,
                             umpa_fr_constants *out);

/* AFR::FillBlobTable (afr.cpp:22-34): host memory, n = blobTable elements */
int umpa_fr_fill_blob_table(float *table, uint64_t n, const umpa_fr_desc *desc,
                            const umpa_fr_settings *settings);

/* ComputeTraverseSpace (traverse_space_generator.hpp:94-145); transform is row-major 3x3 */
int umpa_fr_compute_traverse_space(uint32_t imgSizeX, uint32_t imgSizeY,
                                   const float transform[9], umpa_traverse_space *space,
                                   uint32_t maxVolumeIndexX, uint32_t maxVolumeIndexYZ,
                                   int useFast, float blobRadius, float weight);

/* GenerateTraverseSpaces (examples/fourier_reconstruction/fr.cpp:278-312) for n slices of a
 * padded box P: record i * nSym + q gets transform R_i S_q (rotations: n row-major 3x3
 * doubles, symmetries: nSym row-major 3x3 doubles or NULL for the identity; the product is
 * formed in FP64 and rounded once to float), imgSizeX = P / 2, imgSizeY = P, maximum volume
 * index P (afr_common.hpp:50-68) and projectionIndex = firstIndex + i. */
int umpa_fr_generate_traverse_spaces(const double *rotations, uint64_t n,
                                     const double *symmetries, uint32_t nSym,
                                     uint32_t paddedBox, const umpa_fr_settings *settings,
                                     float weight, int32_t firstIndex,
                                     umpa_traverse_space *out);

/* Rotation operators of a point group, for the `symmetries` of
 * umpa_fr_generate_traverse_spaces: "C<n>" (n rotations about z) or "D<n>" (those and n
 * two-fold axes in the xy plane), row-major 3x3 doubles.  *count receives n or 2n; out may be
 * NULL to query it.  The reference's harness draws random matrices instead (fr.cpp:296,
 * fr.hpp:123-130); point groups are what a reconstruction with symmetry applies. */
int umpa_fr_symmetry_operators(const char *group, double *out, uint32_t capacity,
                               uint32_t *count);

/* ---- Finalization (no counterpart in the reference, which stops after accumulation) ------
 * out[k] = (V(k) + conj(V(-k))) / (W(k) + W(-k)) with `hermitian`, else V(k) / W(k): sums
 * and division in FP64, rounded once to complex float; 0 where the weight sum is 0.  -k is
 * the voxel mirrored through the centre (side - 1) / 2 of the side^3 volume (side odd).
 * Enqueued on `stream` (device pointers; `out` must not alias `volume`). */
int umpa_fr_finalize(const void *volume, const float *weights, uint64_t side, void *out,
                     int hermitian, void *stream);

/* ---- FP: the AFP operation (afp.hpp:13-53), crop/filter/normalize/center/shift ----------- */

/* fourier_processing::Settings (fourier_processing/settings.hpp:10-84), out-of-place only */
typedef struct umpa_fp_settings {
  int32_t center;
  int32_t normalize;
  int32_t applyFilter;
  int32_t shift;
  int32_t hasMaxFreq;
  float maxFreq;
} umpa_fp_settings;

/* Payload shapes: in/out are frequency-domain sizes (x = X/2+1), inSpatial/outSpatial the
 * spatial sizes they came from; inPaddedSpatialSingle = elements of one padded spatial
 * image (the normalization is 1/that, fp_cpu.cpp:42). */
typedef struct umpa_fp_desc {
  umpa_size in;
  umpa_size inSpatial;
  umpa_size out;
  uint64_t inPaddedSpatialSingle;
} umpa_fp_desc;

typedef struct umpa_fp_op umpa_fp_op;

/* new FP operation (FPCUDA / FPCPU) on `device` and `stream` */
int umpa_fp_create(umpa_fp_op **op, int device, void *stream);
/* AFP::Init (basic_operation.hpp:28-48, fp_cpu.cpp:14-21: complex float in/out) */
int umpa_fp_init(umpa_fp_op *op, const umpa_fp_desc *desc, const umpa_fp_settings *settings);
/* AFP::Execute (fp_cpu.cpp:25-46 / fp_cuda.cpp:144-198): in/out complex float device
 * memory, filter float [out.y][out.x] (NULL unless applyFilter) */
int umpa_fp_execute(umpa_fp_op *op, const void *in, uint64_t inCount, void *out,
                    uint64_t outCount, const float *filter);
int umpa_fp_synchronize(umpa_fp_op *op);
int umpa_fp_cleanup(umpa_fp_op *op);
void umpa_fp_destroy(umpa_fp_op *op);

/* ---- FFT: the AFFT operation (fourier_transformation/afft.hpp), out-of-place, via hipFFT -- */

#define UMPA_FFT_FORWARD 0 /* Direction::kForward: real -> Hermitian half (x/2+1) complex */
#define UMPA_FFT_INVERSE 1 /* Direction::kInverse: complex -> real, unnormalised (may
                              overwrite its input, as cuFFT's C2R does) */
#define UMPA_FFT_F32 0     /* float / complex float */
#define UMPA_FFT_F64 1     /* double / complex double */

/* spatial: image size (x, y, z; n images), contiguous; the frequency side holds x/2+1 */
typedef struct umpa_fft_desc {
  umpa_size spatial;
  int32_t direction;
  int32_t precision;
} umpa_fft_desc;

typedef struct umpa_fft_op umpa_fft_op;

/* new FFT operation (FFTCUDA, fft_cuda.cpp:107-116) on `device` and `stream` */
int umpa_fft_create(umpa_fft_op **op, int device, void *stream);
/* AFFT::Init (afft.hpp:40-57, fft_cuda.cpp:51-57): plans the batched transform; element count
 * must fit an int (fft_cuda.cpp:70-75) */
int umpa_fft_init(umpa_fft_op *op, const umpa_fft_desc *desc);
/* AFFT::Execute (fft_cuda.cpp:59-68): `count` images (at most n of Init), device memory;
 * in-place transforms (in == out) are not offered */
int umpa_fft_execute(umpa_fft_op *op, const void *in, void *out, uint64_t count);
int umpa_fft_synchronize(umpa_fft_op *op);
int umpa_fft_cleanup(umpa_fft_op *op);
void umpa_fft_destroy(umpa_fft_op *op);

/* ---- fused FFT + FP: the harness's ConvertToFFT + Crop (fr.cpp:134-203) in one operation --
 * No single counterpart in the reference: it replaces the pair FFTCUDA (fft_cuda.cpp:51-69)
 * then FPCUDA (fp_cuda.cpp:144-198) for 2-D images, with the FP semantics of umpa_fp_*
 * (crop, filter, normalise, centre, max frequency, shift) applied to the forward R2C
 * transform of the images.  Rows are transformed by hipFFT into a workspace; the column
 * transforms and the FP step run in one gfx950 kernel.  Results equal the two operations in
 * sequence up to the rounding of the transform. */
typedef struct umpa_fft_fp_op umpa_fft_fp_op;

int umpa_fft_fp_create(umpa_fft_fp_op **op, int device, void *stream);
/* desc as for umpa_fp_init, with in = the R2C size of inSpatial (in.x = inSpatial.x / 2 + 1,
 * in.y = inSpatial.y), 2-D images, inSpatial.y a power of two from 4 to 1024 (else
 * UMPA_ERR_UNSUPPORTED); allocates the workspace for desc->inSpatial.n images */
int umpa_fft_fp_init(umpa_fft_fp_op *op, const umpa_fp_desc *desc, const umpa_fp_settings *s);
/* images: float [count][inSpatial.y][inSpatial.x] device memory; out: complex float
 * [count][out.y][out.x]; filter float [out.y][out.x] (NULL unless applyFilter) */
int umpa_fft_fp_execute(umpa_fft_fp_op *op, const float *images, uint64_t count, void *out,
                        const float *filter);
int umpa_fft_fp_synchronize(umpa_fft_fp_op *op);
int umpa_fft_fp_cleanup(umpa_fft_fp_op *op);
void umpa_fft_fp_destroy(umpa_fft_fp_op *op);

/* ---- misc --------------------------------------------------------------------------------- */
const char *umpa_status_string(int status);
/* library version, for the loader's sanity check */
int umpa_version(void);

#ifdef __cplusplus
}
#endif
#endif

// Host-side FR helpers of the product library (no GPU needed).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "umpalumpa_amd.h"

namespace umpa::host {

// AFR::CreateConstants (afr.cpp:6-19)
umpa_fr_constants create_constants(uint64_t volumeX, uint64_t volumeY, uint64_t tableElems,
                                   const umpa_fr_settings &s);

// AFR::FillBlobTable (afr.cpp:22-34)
void fill_blob_table(float *table, uint64_t n, uint64_t volumeX, uint64_t volumeY,
                     const umpa_fr_settings &s);

// ComputeTraverseSpace (traverse_space_generator.hpp:94-145)
void compute_traverse_space(uint32_t imgSizeX, uint32_t imgSizeY, const float transform[9],
                            umpa_traverse_space &space, uint32_t maxVolumeIndexX,
                            uint32_t maxVolumeIndexYZ, bool useFast, float blobRadius,
                            float weight);

// Width of the unrolled candidate box for blob radius r: floor(2r) + 1 columns (with a margin
// for rounding at the blob edge), clamped to the kernels' widths 3..6; wider footprints take
// the generic gather.
int footprint(const umpa_fr_settings &s);

// May the candidate box slide inside the slice (no per-candidate masking)?  Slices of at least
// footprint(s) pixels each way, and a footprint that holds the blob's whole bounding box.
bool box_can_shift(const umpa_fr_settings &s, uint64_t sliceX, uint64_t sliceY);

// One work item of the insertion kernel: the origin of an 8 x 8 x D brick of job `job`
// (z | job << 16) and w = D | heavy << 4 | parts << 8 | part << 16 (a brick split into `parts`
// slice ranges is accumulated by `parts` wavefronts, which merge their sums with atomics).
struct BrickOrigin {
  int x, y, z, w;
  float est;  // estimated cost, for ordering only
};

// Tunable knobs of the work-item schedule and of the kernel (the autotuner's search space).
struct SchedParams {
  float depthScale = 1.f;  // brick depth 8 / 4 / 2 / 1 beyond radius 48 / 24 / 12 times this
  bool priority = true;    // run bricks above a quarter of a wave's fair share at priority 2
  int maxParts = 16;       // split heavy bricks into at most this many slice ranges (fast math)
  float splitAt = 0.5f;    // ... when their cost exceeds splitAt times a wave's fair share
  int kernel = 0;          // UMPA_FR_KERNEL_*
  bool tableGlobal = false;  // blob table from global memory
  bool boxFixed = false;     // candidate box kept at the blob's bounding box
};

// The C struct's knobs in range (finite, positive scales; flags 0 / 1)?
bool sched_params_valid(const umpa_fr_sched_params &p);
SchedParams from_c(const umpa_fr_sched_params &p);
umpa_fr_sched_params to_c(const SchedParams &p);

// The bricks covering a vx*vy*vz volume with their estimated cost per inserted slice.
struct BrickCand {
  int x, y, z, d;
  float est;
};
std::vector<BrickCand> brick_candidates(int vx, int vy, int vz, float depthScale = 1.f);

// A volume's bricks sorted by estimated cost (heaviest first, ties in brick order) and the
// costs' sum: what make_schedule needs of a volume, built once per volume shape.
struct BrickSet {
  std::vector<BrickCand> sorted;
  double estSum = 0;
};
BrickSet make_brick_set(int vx, int vy, int vz, float depthScale = 1.f);

// The bricks of one job in a launch, `scale` = slices the launch inserts for it.
struct JobWork {
  const BrickSet *set;
  int job;
  double scale;
};

// Work items of one launch over several jobs, heaviest first (equal costs in job order, then
// brick order).  `splitHeavy` splits the heaviest bricks' slice ranges (changes the summation
// order: fast-math strategies only).
std::vector<BrickOrigin> make_schedule(const std::vector<JobWork> &jobs, bool splitHeavy,
                                       const SchedParams &sp = SchedParams());
// The same into `out` (its storage reused from call to call).
void make_schedule(const std::vector<JobWork> &jobs, bool splitHeavy, const SchedParams &sp,
                   std::vector<BrickOrigin> &out);

// make_schedule for a single job
std::vector<BrickOrigin> make_brick_schedule(int vx, int vy, int vz, bool splitHeavy,
                                             const SchedParams &sp = SchedParams());

// Persisted tuning results: one "key depthScale priority maxParts splitAt kernel tableGlobal
// boxFixed ms" line per key in $UMPA_TUNING_DIR/fr_tuning.txt (default
// $HOME/.cache/umpalumpa_amd); keys carry a format tag, so lines of another format are skipped.
bool tuning_cache_lookup(const std::string &key, SchedParams &out);
void tuning_cache_store(const std::string &key, const SchedParams &p, double ms);

// Rotation matrices (row-major 3x3, FP64) of the point group C_n or D_n ('C' / 'D'); the
// number is written to *count (n or 2n); out == nullptr only queries it.  0 on success.
int symmetry_operators(char family, int n, double *out, uint32_t cap, uint32_t *count);

// The autotuner's first-stage candidates for a strategy: exact, every schedule (no splitting)
// with every kernel knob; fast math, the schedules with default kernel knobs.
std::vector<SchedParams> tuning_candidates(bool fastMath);
// Second stage (fast math only, else empty): the kernel knobs over the schedule `base`.
std::vector<SchedParams> kernel_candidates(const SchedParams &base, bool fastMath);

}  // namespace umpa::host

// Host-side FR helpers of the product library: the AFR statics (CreateConstants,
// FillBlobTable), traverse-space generation and the brick schedule of the insertion kernel.
//
// Compiled with g++ -ffp-contract=off like the reference's plain x86-64 build, so the blob
// table and traverse spaces come out with the same bits as the reference's host code.
#include "fr_host.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <fstream>
#include <limits>
#include <map>
#include <mutex>
#include <sstream>
#include <sys/stat.h>

namespace umpa::host {

namespace {

constexpr float kPi = 3.14159265358979323846f;

// Modified Bessel functions I0, I1 (Abramowitz & Stegun 9.8.1-9.8.4), the polynomials of
// math/bessier.hpp:9-69, in double precision.
double i0(double x)
{
  const double ax = std::fabs(x);
  if (ax < 3.75) {
    double t = x / 3.75;
    t *= t;
    return 1.0 + t * (3.5156229 + t * (3.0899424 + t * (1.2067492 + t * (0.2659732 + t * (0.360768e-1 + t * 0.45813e-2)))));
  }
  const double t = 3.75 / ax;
  return (std::exp(ax) / std::sqrt(ax))
         * (0.39894228
            + t * (0.1328592e-1
                   + t * (0.225319e-2
                          + t * (-0.157565e-2
                                 + t * (0.916281e-2 + t * (-0.2057706e-1 + t * (0.2635537e-1 + t * (-0.1647633e-1 + t * 0.392377e-2))))))));
}

double i1(double x)
{
  const double ax = std::fabs(x);
  double v;
  if (ax < 3.75) {
    double t = x / 3.75;
    t *= t;
    v = ax * (0.5 + t * (0.87890594 + t * (0.51498869 + t * (0.15084934 + t * (0.2658733e-1 + t * (0.301532e-2 + t * 0.32411e-3))))));
  } else {
    const double t = 3.75 / ax;
    double tail = 0.2282967e-1 + t * (-0.2895312e-1 + t * (0.1787654e-1 - t * 0.420059e-2));
    v = 0.39894228 + t * (-0.3988024e-1 + t * (-0.362018e-2 + t * (0.163801e-2 + t * (-0.1031555e-1 + t * tail))));
    v *= (std::exp(ax) / std::sqrt(ax));
  }
  return x < 0.0 ? -v : v;
}

// I_{n+1}(x) = I_{n-1}(x) - (2n/x) I_n(x)   (bessier.hpp:80-84)
double iN(int n, double x)
{
  if (n == 0) return i0(x);
  if (n == 1) return i1(x);
  if (x == 0) return 0;
  double a = i0(x), b = i1(x);
  for (int k = 1; k < n; ++k) {
    double c = a - ((2 * k) / x) * b;
    a = b;
    b = c;
  }
  return b;
}

// I_{3/2} in float (bessier.hpp:72-76 with T = float)
float i1_5f(float x)
{
  if (x == 0) return 0;
  return std::sqrt(2.f / (kPi * x)) * (std::cosh(x) - std::sinh(x) / x);
}

// KaiserValue<float> (math/kaiser.hpp:19-56): blob profile at distance `freq`
float kaiser_value(float freq, float radius, float alpha, int order)
{
  float w = 0;
  const float rda = freq / radius;
  if (rda <= 1.f) {
    const float rdas = rda * rda;
    const float arg = alpha * std::sqrt(1 - rdas);
    if (order == 0) {
      w = static_cast<float>(i0(arg) / i0(alpha));
    } else if (order >= 1 && order <= 4) {
      const float s = std::sqrt(1 - rdas);
      w = s;
      for (int k = 1; k < order; ++k) w = w * s;
      if (alpha != 0) w *= static_cast<float>(iN(order, arg) / iN(order, alpha));
    }
  }
  return w;
}

// KaiserFourierValue<float> (math/kaiser.hpp:66-81): order 0 only, NaN otherwise
float kaiser_fourier_value(float freq, float radius, float alpha, int order)
{
  const float a = 2.f * kPi * radius * freq;
  const float sigma = std::sqrt(std::fabs(alpha * alpha - a * a));
  if (order == 0 && (2 * kPi * radius * freq <= alpha)) {
    return std::pow(2.f * kPi, 1.5f) * std::pow(radius, 3.f) * i1_5f(sigma)
           / (static_cast<float>(i0(alpha)) * std::pow(sigma, 1.5f));
  }
  return std::numeric_limits<float>::quiet_NaN();
}

float bessi_order_alpha(int order, float alpha)
{
  if (order < 0 || order > 4) return std::numeric_limits<float>::quiet_NaN();
  return static_cast<float>(iN(order, alpha));
}

struct P3 {
  float x, y, z;
};

P3 rotate(const float t[3][3], P3 p)
{
  return { t[0][0] * p.x + t[0][1] * p.y + t[0][2] * p.z,
    t[1][0] * p.x + t[1][1] * p.y + t[1][2] * p.z,
    t[2][0] * p.x + t[2][1] * p.y + t[2][2] * p.z };
}

}  // namespace

umpa_fr_constants create_constants(uint64_t volumeX, uint64_t volumeY, uint64_t tableElems,
                                   const umpa_fr_settings &s)
{
  umpa_fr_constants c{};
  c.cMaxVolumeIndexX = static_cast<int>(volumeX - 1);
  c.cMaxVolumeIndexYZ = static_cast<int>(volumeY - 1);
  c.cBlobRadius = s.blobRadius;
  c.cOneOverBlobRadiusSqr = 1.f / (s.blobRadius * s.blobRadius);
  c.cBlobAlpha = s.alpha;
  c.cIw0 = 1.f / kaiser_fourier_value(0.f, s.blobRadius, s.alpha, s.order);
  // unsigned arithmetic: an empty (kDynamic) table wraps to 2^64 - 1, as in afr.cpp:16
  c.cIDeltaSqrt = static_cast<float>(tableElems - 1) / (s.blobRadius * s.blobRadius);
  c.cOneOverBessiOrderAlpha = 1 / bessi_order_alpha(s.order, s.alpha);
  return c;
}

void fill_blob_table(float *table, uint64_t n, uint64_t volumeX, uint64_t volumeY,
                     const umpa_fr_settings &s)
{
  // entry i: blob value at squared distance i * r^2 / (n - 1), times 1/KaiserFourierValue(0)
  const float step = s.blobRadius * std::sqrt(1.f / static_cast<float>(n - 1));
  const umpa_fr_constants c = create_constants(volumeX, volumeY, n, s);
  for (uint64_t i = 0; i < n; ++i) {
    const float d = step * static_cast<float>(std::sqrt(static_cast<double>(i)));
    table[i] = kaiser_value(d, s.blobRadius, s.alpha, s.order) * c.cIw0;
  }
}

void compute_traverse_space(uint32_t imgSizeX, uint32_t imgSizeY, const float transform[9],
                            umpa_traverse_space &space, uint32_t maxVolumeIndexX,
                            uint32_t maxVolumeIndexYZ, bool useFast, float blobRadius,
                            float weight)
{
  float t[3][3];
  for (int i = 0; i < 9; ++i) t[i / 3][i % 3] = transform[i];
  const float pad = useFast ? 0.f : blobRadius;
  const float halfY = static_cast<float>(imgSizeY) / 2.0f;
  const float x0 = 0.f - pad, x1 = static_cast<float>(imgSizeX) + pad;
  const float y0 = -(halfY + pad), y1 = halfY + pad;
  const float z0 = 0.f - pad, z1 = 0.f + pad;
  // slab around the slice's half plane, corners in the reference's numbering
  // (traverse_space_generator.hpp:19-40): 0..3 on the +pad side, 4..7 on the -pad side
  P3 corner[8] = { { x1, y0, z1 }, { x1, y1, z1 }, { x0, y1, z1 }, { x0, y0, z1 },
    { x1, y0, z0 }, { x1, y1, z0 }, { x0, y1, z0 }, { x0, y0, z0 } };
  const P3 centre = { static_cast<float>(maxVolumeIndexX) / 2.f,
    static_cast<float>(maxVolumeIndexYZ) / 2.f, static_cast<float>(maxVolumeIndexYZ) / 2.f };
  P3 lo = { FLT_MAX, FLT_MAX, FLT_MAX };
  P3 hi = { FLT_MIN, FLT_MIN, FLT_MIN };  // the reference's running maximum starts at FLT_MIN
  for (auto &p : corner) {
    p = rotate(t, p);
    p.x += centre.x;
    p.y += centre.y;
    p.z += centre.z;
    lo.x = lo.x > p.x ? p.x : lo.x;
    lo.y = lo.y > p.y ? p.y : lo.y;
    lo.z = lo.z > p.z ? p.z : lo.z;
    hi.x = hi.x < p.x ? p.x : hi.x;
    hi.y = hi.y < p.y ? p.y : hi.y;
    hi.z = hi.z < p.z ? p.z : hi.z;
  }
  lo.x = lo.x < 0.f ? 0.f : lo.x;
  lo.y = lo.y < 0.f ? 0.f : lo.y;
  lo.z = lo.z < 0.f ? 0.f : lo.z;
  hi.x = hi.x > static_cast<float>(maxVolumeIndexX) ? static_cast<float>(maxVolumeIndexX) : hi.x;
  hi.y = hi.y > static_cast<float>(maxVolumeIndexYZ) ? static_cast<float>(maxVolumeIndexYZ) : hi.y;
  hi.z = hi.z > static_cast<float>(maxVolumeIndexYZ) ? static_cast<float>(maxVolumeIndexYZ) : hi.z;
  space.minX = static_cast<int>(std::floor(lo.x));
  space.minY = static_cast<int>(std::floor(lo.y));
  space.minZ = static_cast<int>(std::floor(lo.z));
  space.maxX = static_cast<int>(std::ceil(hi.x));
  space.maxY = static_cast<int>(std::ceil(hi.y));
  space.maxZ = static_cast<int>(std::ceil(hi.z));
  space.topOrigin = { corner[4].x, corner[4].y, corner[4].z };
  space.bottomOrigin = { corner[0].x, corner[0].y, corner[0].z };
  const float reach = static_cast<float>(imgSizeX) + pad;
  space.maxDistanceSqr = reach * reach;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) space.transformInv[i][j] = t[j][i];
  const P3 n = rotate(t, P3{ 0.f, 0.f, 1.f });
  space.unitNormal = { n.x, n.y, n.z };
  const float ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
  // iterate the plane orthogonal to the dominant normal component
  if (ax >= ay && ax >= az) {
    space.dir = 2;
  } else if (ay >= ax && ay >= az) {
    space.dir = 1;
  } else if (az >= ax && az >= ay) {
    space.dir = 0;
  }
  space.weight = weight;
}

int footprint(const umpa_fr_settings &s)
{
  const int nb = (int)std::floor(2.0 * (double)s.blobRadius + 1e-3) + 1;
  return nb < 3 ? 3 : (nb > 6 ? 6 : nb);
}

bool box_can_shift(const umpa_fr_settings &s, uint64_t sliceX, uint64_t sliceY)
{
  const int need = (int)std::floor(2.0 * (double)s.blobRadius + 1e-3) + 1;
  // the box width the kernel runs with: the footprint for table lookups, 4 for the other
  // blob modes (which fall back to the general gather past it)
  const bool lookup = s.type != UMPA_FR_FAST && s.interpolation == UMPA_FR_LOOKUP;
  const int nb = lookup ? footprint(s) : 4;
  // positions below 8192 round qx - r and qx + r by at most 2^-11 together, inside the
  // 1e-3 margin of `need`, so the bounding box never has more than `need` columns or rows
  return need <= nb && sliceX >= (uint64_t)nb && sliceY >= (uint64_t)nb && sliceX <= 8192 &&
         sliceY <= 8192;
}

std::vector<BrickCand> brick_candidates(int vx, int vy, int vz, float depthScale)
{
  // Bricks are 8 x 8 x D columns, D in {8, 4, 2, 1}.  A voxel at distance R from the centre
  // is crossed by a fraction ~ r/R of random slices, so the work of a brick grows towards the
  // centre; D shrinks with R to even the work per brick out.
  std::vector<BrickCand> cand;
  const float cx = static_cast<float>((vx - 1) / 2), cy = static_cast<float>((vy - 1) / 2);
  auto radius = [&](int x, int y, int z, int d) {
    const float dx = x + 3.5f - cx, dy = y + 3.5f - cy, dz = z + 0.5f * (d - 1) - cy;
    return std::sqrt(dx * dx + dy * dy + dz * dz);
  };
  const float r8 = 48.f * depthScale, r4 = 24.f * depthScale, r2 = 12.f * depthScale;
  auto depth_for = [&](float R) { return R > r8 ? 8 : R > r4 ? 4 : R > r2 ? 2 : 1; };
  for (int y = 0; y < vy; y += 8)
    for (int x = 0; x < vx; x += 8)
      for (int z = 0; z < vz;) {
        int d = 8;
        while (d > 1 && depth_for(radius(x, y, z, d)) < d) d /= 2;
        const float R = radius(x, y, z, d);
        // expected cost per slice: P(active) times the cost of an active (brick, slice) pair.
        // A random plane passes within w of a point at distance R with probability w/R; away
        // from the centre the half-space test removes half of those.
        const float w = 1.9f + 3.5f * 1.5f + 0.25f * (float)d;
        const float pslab = w / std::max(R, w);
        const float p = pslab * (0.5f + 0.5f * pslab);
        const float passes = std::max(1.f, 0.3f * (float)d);  // ~30% of an active brick is in the slab
        const float est = p * (68.f + 8.f * d + 300.f * passes) + 2.f;
        cand.push_back({ x, y, z, d, est });
        z += d;
      }
  return cand;
}

BrickSet make_brick_set(int vx, int vy, int vz, float depthScale)
{
  BrickSet set;
  set.sorted = brick_candidates(vx, vy, vz, depthScale);
  std::stable_sort(set.sorted.begin(), set.sorted.end(),
                   [](const BrickCand &a, const BrickCand &b) { return a.est > b.est; });
  for (const BrickCand &b : set.sorted) set.estSum += b.est;
  return set;
}

void make_schedule(const std::vector<JobWork> &jobs, bool splitHeavy, const SchedParams &sp,
                   std::vector<BrickOrigin> &bricks)
{
  double total = 0;
  size_t count = 0;
  for (const JobWork &j : jobs) {
    total += j.set->estSum * j.scale;
    count += j.set->sorted.size();
  }
  // a wavefront's share of the work on a full MI355X (16 wavefronts on each of 256 CUs); only
  // bricks well above both that share and the typical brick are prioritised or split
  const double fair = total / 4096.0;
  const double mean = total / (double)std::max<size_t>(count, 1);
  const double at = fair * (double)sp.splitAt;
  // An item's sort key: its float cost, then (job position, rank of the brick in the job's
  // sorted bricks, part).  A job's bricks come sorted by cost and all scale by the same
  // factor, so each job is a sorted stream; heavy bricks, the only ones prioritised or split,
  // form a prefix of it.  Their parts (cost / parts) are sorted apart, then the streams are
  // merged: a stable sort by cost of the jobs' sorted bricks in turn gives the same sequence.
  struct Key {
    float est;
    int job;
    uint32_t rank;
    int part;
  };
  auto before = [](const Key &a, const Key &b) {
    if (a.est != b.est) return a.est > b.est;
    if (a.job != b.job) return a.job < b.job;
    if (a.rank != b.rank) return a.rank < b.rank;
    return a.part < b.part;
  };
  struct Item {
    Key key;
    BrickOrigin b;
  };
  std::vector<Item> heavy;
  std::vector<size_t> pos(jobs.size(), 0);
  for (size_t jj = 0; jj < jobs.size(); ++jj) {
    const JobWork &j = jobs[jj];
    const auto &sorted = j.set->sorted;
    size_t p = 0;
    for (; p < sorted.size(); ++p) {
      const BrickCand &c = sorted[p];
      const double est = c.est * j.scale;
      if (!(est > 4 * mean && (est > fair / 4 || est > at))) break;  // the rest is plain
      int flags = c.d;
      if (sp.priority && est > fair / 4) flags |= 1 << 4;  // raised priority
      int parts = 1;
      if (splitHeavy && sp.maxParts > 1 && est > at)
        parts = (int)std::min<double>(sp.maxParts, std::ceil(est / at));
      for (int k = 0; k < parts; ++k) {
        BrickOrigin b{ c.x, c.y, c.z | (j.job << 16), flags | (parts << 8) | (k << 16),
                       (float)(est / parts) };
        heavy.push_back({ { b.est, (int)jj, (uint32_t)p, k }, b });
      }
    }
    pos[jj] = p;
  }
  std::sort(heavy.begin(), heavy.end(),
            [&](const Item &a, const Item &b) { return before(a.key, b.key); });
  bricks.clear();
  bricks.reserve(count + heavy.size());
  size_t h = 0;
  auto emit_plain = [&](size_t jj, size_t p) {
    const JobWork &j = jobs[jj];
    const BrickCand &c = j.set->sorted[p];
    bricks.push_back({ c.x, c.y, c.z | (j.job << 16), c.d | (1 << 8),
                       (float)(c.est * j.scale) });
  };
  if (jobs.size() == 1) {
    // one stream: merge the heavy parts into it
    const size_t n = jobs[0].set->sorted.size();
    const double scale = jobs[0].scale;
    for (size_t p = pos[0]; p < n; ++p) {
      const Key k{ (float)(jobs[0].set->sorted[p].est * scale), 0, (uint32_t)p, 0 };
      while (h < heavy.size() && before(heavy[h].key, k)) bricks.push_back(heavy[h++].b);
      emit_plain(0, p);
    }
    while (h < heavy.size()) bricks.push_back(heavy[h++].b);
    return;
  }
  // k-way merge of the heavy items and the jobs' remaining streams.  A stream head's key packs
  // into 64 bits, smallest first: the cost's float bits inverted (costs are positive, so their
  // bits grow with them), the job position (8 bits) and the rank (24 bits).
  auto head = [&](size_t jj) {
    const JobWork &j = jobs[jj];
    const float e = (float)(j.set->sorted[pos[jj]].est * j.scale);
    uint32_t eb;
    std::memcpy(&eb, &e, 4);
    return ((uint64_t)~eb << 32) | ((uint64_t)jj << 24) | (uint64_t)pos[jj];
  };
  bool packable = jobs.size() <= 256;
  for (const JobWork &j : jobs) packable = packable && j.set->sorted.size() < (1u << 24);
  if (!packable) {
    // beyond the packed key's ranges: one sort of every item by the full key
    std::vector<Item> all;
    all.reserve(count - pos.size() + heavy.size());
    for (size_t jj = 0; jj < jobs.size(); ++jj)
      for (size_t p = pos[jj]; p < jobs[jj].set->sorted.size(); ++p) {
        const BrickCand &c = jobs[jj].set->sorted[p];
        const float e = (float)(c.est * jobs[jj].scale);
        all.push_back({ { e, (int)jj, (uint32_t)p, 0 },
                        { c.x, c.y, c.z | (jobs[jj].job << 16), c.d | (1 << 8), e } });
      }
    all.insert(all.end(), heavy.begin(), heavy.end());
    std::sort(all.begin(), all.end(),
              [&](const Item &a, const Item &b) { return before(a.key, b.key); });
    for (const Item &it : all) bricks.push_back(it.b);
    return;
  }
  std::vector<uint64_t> q;
  for (size_t jj = 0; jj < jobs.size(); ++jj)
    if (pos[jj] < jobs[jj].set->sorted.size()) q.push_back(head(jj));
  std::make_heap(q.begin(), q.end(), std::greater<uint64_t>());
  while (!q.empty()) {
    const uint64_t top = q.front();
    const size_t jj = (size_t)((top >> 24) & 255u);
    if (h < heavy.size()) {
      uint32_t eb = ~(uint32_t)(top >> 32);
      float e;
      std::memcpy(&e, &eb, 4);
      const Key k{ e, (int)jj, (uint32_t)(top & 0xffffffu), 0 };
      while (h < heavy.size() && before(heavy[h].key, k)) bricks.push_back(heavy[h++].b);
    }
    std::pop_heap(q.begin(), q.end(), std::greater<uint64_t>());
    emit_plain(jj, pos[jj]);
    if (++pos[jj] < jobs[jj].set->sorted.size()) {
      q.back() = head(jj);
      std::push_heap(q.begin(), q.end(), std::greater<uint64_t>());
    } else {
      q.pop_back();
    }
  }
  while (h < heavy.size()) bricks.push_back(heavy[h++].b);
}

std::vector<BrickOrigin> make_schedule(const std::vector<JobWork> &jobs, bool splitHeavy,
                                       const SchedParams &sp)
{
  std::vector<BrickOrigin> out;
  make_schedule(jobs, splitHeavy, sp, out);
  return out;
}

std::vector<BrickOrigin> make_brick_schedule(int vx, int vy, int vz, bool splitHeavy,
                                             const SchedParams &sp)
{
  const BrickSet set = make_brick_set(vx, vy, vz, sp.depthScale);
  return make_schedule({ JobWork{ &set, 0, 1.0 } }, splitHeavy, sp);
}

int symmetry_operators(char family, int n, double *out, uint32_t cap, uint32_t *count)
{
  // C_n: rotations by 2 pi k / n about z.  D_n: C_n and the n two-fold rotations about the
  // axes at angles pi k / n in the xy plane (z -> -z).
  if (n < 1 || !count || (family != 'C' && family != 'D')) return -1;
  const uint32_t total = (uint32_t)(family == 'C' ? n : 2 * n);
  *count = total;
  if (!out) return 0;
  if (cap < total) return -1;
  const double pi = 3.14159265358979323846;
  uint32_t k = 0;
  for (int i = 0; i < n; ++i, ++k) {
    const double a = 2.0 * pi * i / n, c = std::cos(a), s = std::sin(a);
    const double m[9] = { c, -s, 0, s, c, 0, 0, 0, 1 };
    std::copy(m, m + 9, out + 9 * k);
  }
  if (family == 'D') {
    for (int i = 0; i < n; ++i, ++k) {
      const double t = pi * i / n, c = std::cos(2 * t), s = std::sin(2 * t);
      const double m[9] = { c, s, 0, s, -c, 0, 0, 0, -1 };
      std::copy(m, m + 9, out + 9 * k);
    }
  }
  return 0;
}

bool sched_params_valid(const umpa_fr_sched_params &p)
{
  return std::isfinite(p.depthScale) && p.depthScale > 0.f && p.depthScale <= 64.f &&
         p.maxParts >= 1 && p.maxParts <= 255 && std::isfinite(p.splitAt) && p.splitAt > 0.f &&
         p.splitAt <= 1024.f && p.kernel >= UMPA_FR_KERNEL_AUTO &&
         p.kernel <= UMPA_FR_KERNEL_STAGED && (p.tableGlobal == 0 || p.tableGlobal == 1) &&
         (p.boxFixed == 0 || p.boxFixed == 1);
}

SchedParams from_c(const umpa_fr_sched_params &p)
{
  SchedParams s;
  s.depthScale = p.depthScale;
  s.priority = p.priority != 0;
  s.maxParts = p.maxParts;
  s.splitAt = p.splitAt;
  s.kernel = p.kernel;
  s.tableGlobal = p.tableGlobal != 0;
  s.boxFixed = p.boxFixed != 0;
  return s;
}

umpa_fr_sched_params to_c(const SchedParams &p)
{
  umpa_fr_sched_params c{};
  c.depthScale = p.depthScale;
  c.priority = p.priority ? 1 : 0;
  c.maxParts = p.maxParts;
  c.splitAt = p.splitAt;
  c.kernel = p.kernel;
  c.tableGlobal = p.tableGlobal ? 1 : 0;
  c.boxFixed = p.boxFixed ? 1 : 0;
  return c;
}

std::vector<SchedParams> tuning_candidates(bool fastMath)
{
  std::vector<SchedParams> out;
  for (float depth : { 0.5f, 1.f, 2.f })
    for (bool prio : { true, false }) {
      SchedParams p;
      p.depthScale = depth;
      p.priority = prio;
      if (!fastMath) {
        // exact: the two kernels at 12 waves per CU, the table in LDS or global memory
        p.maxParts = 1;
        for (int kernel : { UMPA_FR_KERNEL_STAGED, UMPA_FR_KERNEL_PAIR })
          for (bool global : { false, true }) {
            p.kernel = kernel;
            p.tableGlobal = global;
            out.push_back(p);
          }
        continue;
      }
      for (auto [parts, at] : { std::pair<int, float>{ 1, 0.5f }, { 4, 0.5f }, { 16, 0.5f },
                                { 16, 0.25f } }) {
        p.maxParts = parts;
        p.splitAt = at;
        out.push_back(p);
      }
    }
  return out;
}

std::vector<SchedParams> kernel_candidates(const SchedParams &base, bool fastMath)
{
  std::vector<SchedParams> out;
  if (!fastMath) return out;
  // fast math: per-pair (16 waves) or carried queues, sliding or fixed boxes, table placement
  for (int kernel : { UMPA_FR_KERNEL_CARRY, UMPA_FR_KERNEL_PAIR })
    for (bool fixed : { false, true })
      for (bool global : { false, true }) {
        SchedParams p = base;
        p.kernel = kernel;
        p.boxFixed = fixed;
        p.tableGlobal = global;
        out.push_back(p);
      }
  return out;
}

namespace {
std::mutex g_cacheMutex;

std::string tuning_file()
{
  const char *dir = std::getenv("UMPA_TUNING_DIR");
  std::string d;
  if (dir && *dir) {
    d = dir;
  } else {
    const char *home = std::getenv("HOME");
    d = std::string(home && *home ? home : "/tmp") + "/.cache";
    ::mkdir(d.c_str(), 0755);
    d += "/umpalumpa_amd";
  }
  ::mkdir(d.c_str(), 0755);
  return d + "/fr_tuning.txt";
}
}  // namespace

bool tuning_cache_lookup(const std::string &key, SchedParams &out)
{
  std::lock_guard<std::mutex> lock(g_cacheMutex);
  std::ifstream f(tuning_file());
  std::string line;
  bool found = false;
  // the last line of a key wins
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string k;
    umpa_fr_sched_params c{};
    double ms = 0;
    // entries outside the knobs' valid ranges (a damaged or hand-edited file) are ignored
    if (is >> k >> c.depthScale >> c.priority >> c.maxParts >> c.splitAt >> c.kernel >>
          c.tableGlobal >> c.boxFixed >> ms &&
        k == key && (c.priority == 0 || c.priority == 1) && sched_params_valid(c)) {
      out = from_c(c);
      found = true;
    }
  }
  return found;
}

void tuning_cache_store(const std::string &key, const SchedParams &p, double ms)
{
  std::lock_guard<std::mutex> lock(g_cacheMutex);
  std::ofstream f(tuning_file(), std::ios::app);
  if (f) f << key << ' ' << p.depthScale << ' ' << (p.priority ? 1 : 0) << ' ' << p.maxParts
           << ' ' << p.splitAt << ' ' << p.kernel << ' ' << (p.tableGlobal ? 1 : 0) << ' '
           << (p.boxFixed ? 1 : 0) << ' ' << ms << '\n';
}

}  // namespace umpa::host

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

std::vector<BrickOrigin> make_schedule(const std::vector<JobWork> &jobs, bool splitHeavy,
                                       const SchedParams &sp)
{
  double total = 0;
  size_t count = 0;
  for (const JobWork &j : jobs) {
    double sum = 0;
    for (const BrickCand &c : *j.cands) sum += c.est;
    total += sum * j.scale;
    count += j.cands->size();
  }
  // a wavefront's share of the work on a full MI355X (16 wavefronts on each of 256 CUs); only
  // bricks well above both that share and the typical brick are prioritised or split
  const double fair = total / 4096.0;
  const double mean = total / (double)std::max<size_t>(count, 1);
  std::vector<BrickOrigin> bricks;
  bricks.reserve(count + count / 8);
  for (const JobWork &j : jobs) {
    for (const BrickCand &c : *j.cands) {
      const double est = c.est * j.scale;
      int flags = c.d;
      if (sp.priority && est > fair / 4 && est > 4 * mean) flags |= 1 << 4;  // raised priority
      int parts = 1;
      const double at = fair * (double)sp.splitAt;
      if (splitHeavy && sp.maxParts > 1 && est > at && est > 4 * mean)
        parts = (int)std::min<double>(sp.maxParts, std::ceil(est / at));
      for (int k = 0; k < parts; ++k) {
        BrickOrigin b{ c.x, c.y, c.z | (j.job << 16), flags | (parts << 8) | (k << 16), 0.f };
        b.est = (float)(est / parts);
        bricks.push_back(b);
      }
    }
  }
  // heaviest first, so the long central bricks start at once
  std::stable_sort(bricks.begin(), bricks.end(),
    [](const BrickOrigin &a, const BrickOrigin &b) { return a.est > b.est; });
  return bricks;
}

std::vector<BrickOrigin> make_brick_schedule(int vx, int vy, int vz, bool splitHeavy,
                                             const SchedParams &sp)
{
  const std::vector<BrickCand> cands = brick_candidates(vx, vy, vz, sp.depthScale);
  return make_schedule({ JobWork{ &cands, 0, 1.0 } }, splitHeavy, sp);
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

std::vector<SchedParams> tuning_candidates(bool fastMath)
{
  std::vector<SchedParams> out;
  for (float depth : { 0.5f, 1.f, 2.f })
    for (bool prio : { true, false }) {
      if (!fastMath) {
        out.push_back({ depth, prio, 1, 0.5f });
        continue;
      }
      for (auto [parts, at] : { std::pair<int, float>{ 1, 0.5f }, { 4, 0.5f }, { 16, 0.5f },
                                { 16, 0.25f } })
        out.push_back({ depth, prio, parts, at });
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
    SchedParams p;
    int prio = 1;
    double ms = 0;
    // entries outside the knobs' valid ranges (a damaged or hand-edited file) are ignored
    if (is >> k >> p.depthScale >> prio >> p.maxParts >> p.splitAt >> ms && k == key &&
        std::isfinite(p.depthScale) && p.depthScale > 0.f && p.depthScale <= 64.f &&
        p.maxParts >= 1 && p.maxParts <= 255 && std::isfinite(p.splitAt) && p.splitAt > 0.f &&
        p.splitAt <= 1024.f) {
      p.priority = prio != 0;
      out = p;
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
           << ' ' << p.splitAt << ' ' << ms << '\n';
}

}  // namespace umpa::host

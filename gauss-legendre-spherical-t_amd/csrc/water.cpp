// Synthetic rigid-water boxes with a generator defined here (SplitMix64
// seeding xoshiro256**), so every host produces bit-identical inputs.
#include "glst_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>

namespace {

struct Xoshiro256ss {
  uint64_t s[4];
  explicit Xoshiro256ss(uint64_t seed) {
    uint64_t x = seed;
    for (auto &v : s) {  // SplitMix64
      x += 0x9E3779B97F4A7C15ull;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      v = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform in [0, 1) with 53 random bits
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

constexpr double kPi = 3.14159265358979323846;

}  // namespace

extern "C" glst_status_t glst_make_water_box(uint32_t nwater, uint32_t nion,
                                             double L, uint64_t seed,
                                             double *rx, double *ry,
                                             double *rz, double *qc) {
  const size_t n = 3ull * nwater + 2ull * nion;
  if (n > 0 && (!rx || !ry || !rz || !qc)) return GLST_ERR_INVALID;
  if (!(L > 0)) return GLST_ERR_INVALID;
  Xoshiro256ss rng(seed);
  // H at O + R (+-sin(a/2), 0, cos(a/2)), a = 109.47 deg, |O-H| = 1.0 A
  const double half = 0.5 * 109.47 * kPi / 180.0;
  const double hx = std::sin(half), hz = std::cos(half);
  size_t a = 0;
  for (uint32_t w = 0; w < nwater; ++w) {
    const double ox = L * rng.uniform(), oy = L * rng.uniform(),
                 oz = L * rng.uniform();
    // uniform rotation from a uniform unit quaternion (Shoemake)
    const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
    const double s1 = std::sqrt(1.0 - u1), s2 = std::sqrt(u1);
    const double qx = s1 * std::sin(2.0 * kPi * u2);
    const double qy = s1 * std::cos(2.0 * kPi * u2);
    const double qz = s2 * std::sin(2.0 * kPi * u3);
    const double qw = s2 * std::cos(2.0 * kPi * u3);
    const double r00 = 1 - 2 * (qy * qy + qz * qz), r02 = 2 * (qx * qz + qy * qw);
    const double r10 = 2 * (qx * qy + qz * qw), r12 = 2 * (qy * qz - qx * qw);
    const double r20 = 2 * (qx * qz - qy * qw), r22 = 1 - 2 * (qx * qx + qy * qy);
    rx[a] = ox;
    ry[a] = oy;
    rz[a] = oz;
    qc[a] = -0.8476;
    ++a;
    for (int sgn = 1; sgn >= -1; sgn -= 2) {
      const double vx = sgn * hx, vz = hz;  // body frame (vx, 0, vz)
      rx[a] = ox + r00 * vx + r02 * vz;
      ry[a] = oy + r10 * vx + r12 * vz;
      rz[a] = oz + r20 * vx + r22 * vz;
      qc[a] = 0.4238;
      ++a;
    }
  }
  for (int sgn = 1; sgn >= -1; sgn -= 2)
    for (uint32_t i = 0; i < nion; ++i) {
      rx[a] = L * rng.uniform();
      ry[a] = L * rng.uniform();
      rz[a] = L * rng.uniform();
      qc[a] = static_cast<double>(sgn);
      ++a;
    }
  // recenter: minimum of each coordinate at 0 (utils/utils.cpp:35-58)
  double mn[3] = {std::numeric_limits<double>::max(),
                  std::numeric_limits<double>::max(),
                  std::numeric_limits<double>::max()};
  for (size_t i = 0; i < n; ++i) {
    mn[0] = std::min(mn[0], rx[i]);
    mn[1] = std::min(mn[1], ry[i]);
    mn[2] = std::min(mn[2], rz[i]);
  }
  for (size_t i = 0; i < n; ++i) {
    rx[i] -= mn[0];
    ry[i] -= mn[1];
    rz[i] -= mn[2];
  }
  return GLST_OK;
}

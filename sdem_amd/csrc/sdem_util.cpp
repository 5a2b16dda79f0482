// sdem_util.cpp -- seeded initial-condition helpers of libsdem_hip.so.
//
// The reference seeds its particle generator from the system clock
// (GL_MakingAllPrtcl, Prtcl_Variables.f90:225-227), so its initial states
// cannot be reproduced.  These helpers build seeded ones on the host.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../../include/sdem.h"

namespace {

// splitmix64: a small, well-mixed seeded generator
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

int sdem_util_random_pack(int64_t n, const double* radius, const double lo[3], const double hi[3],
                          const int32_t periodic[3], uint64_t seed, int32_t max_attempts,
                          double* pos) {
  if (n < 0 || (n > 0 && (!radius || !pos)) || !lo || !hi || max_attempts < 1) return SDEM_E_ARG;
  if (n == 0) return SDEM_OK;
  double rmax = 0.0;
  for (int64_t i = 0; i < n; i++) {
    if (!(radius[i] > 0.0)) return SDEM_E_ARG;
    rmax = std::max(rmax, radius[i]);
  }
  const int per[3] = {periodic ? periodic[0] : 0, periodic ? periodic[1] : 0, periodic ? periodic[2] : 0};
  double L[3], w[3];
  int nc[3];
  for (int d = 0; d < 3; d++) {
    L[d] = hi[d] - lo[d];
    if (!(L[d] > 2.0 * rmax)) return SDEM_E_ARG;
    nc[d] = std::max(1, (int)std::floor(L[d] / (2.0 * rmax)));
    if (per[d] && nc[d] < 3) return SDEM_E_ARG;
    w[d] = L[d] / nc[d];
  }
  const int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
  std::vector<int64_t> head(ncell, -1), next(n, -1);
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return radius[a] > radius[b]; });
  Rng rng(seed * 0x2545F4914F6CDD1Dull + 0x1234567ull);
  auto cell_of = [&](int d, double x) {
    int c = (int)std::floor((x - lo[d]) / w[d]);
    return c < 0 ? 0 : (c >= nc[d] ? nc[d] - 1 : c);
  };
  for (int64_t q = 0; q < n; q++) {
    const int64_t i = order[q];
    const double r = radius[i];
    bool placed = false;
    for (int a = 0; a < max_attempts && !placed; a++) {
      double p[3];
      for (int d = 0; d < 3; d++) {
        double m = per[d] ? 0.0 : r;
        p[d] = lo[d] + m + rng.uniform() * (L[d] - 2.0 * m);
        if (p[d] >= hi[d]) p[d] = per[d] ? lo[d] : hi[d] - m;
      }
      int c[3] = {cell_of(0, p[0]), cell_of(1, p[1]), cell_of(2, p[2])};
      bool ok = true;
      for (int dz = -1; dz <= 1 && ok; dz++)
        for (int dy = -1; dy <= 1 && ok; dy++)
          for (int dx = -1; dx <= 1 && ok; dx++) {
            int cc[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
            double sh[3] = {0.0, 0.0, 0.0};
            bool skip = false;
            for (int d = 0; d < 3; d++) {
              if (cc[d] < 0) {
                if (!per[d]) { skip = true; break; }
                cc[d] += nc[d]; sh[d] = -L[d];
              } else if (cc[d] >= nc[d]) {
                if (!per[d]) { skip = true; break; }
                cc[d] -= nc[d]; sh[d] = L[d];
              }
            }
            if (skip) continue;
            for (int64_t j = head[((int64_t)cc[2] * nc[1] + cc[1]) * nc[0] + cc[0]]; j >= 0; j = next[j]) {
              double dd = 0.0;
              for (int d = 0; d < 3; d++) {
                double t = p[d] - (pos[3 * j + d] + sh[d]);
                dd += t * t;
              }
              double s = r + radius[j];
              if (dd < s * s) { ok = false; break; }
            }
          }
      if (ok) {
        for (int d = 0; d < 3; d++) pos[3 * i + d] = p[d];
        int64_t key = ((int64_t)c[2] * nc[1] + c[1]) * nc[0] + c[0];
        next[i] = head[key];
        head[key] = i;
        placed = true;
      }
    }
    if (!placed) return SDEM_E_CAPACITY;
  }
  return SDEM_OK;
}

// GuideBin (lambda_amd/csrc/guide.h) against the definition of its columns: phi = atan2(z, x) + pi
// in [c, c + 1) x 2 pi / 16, rows of cos(theta) in steps of 1/4. Directions whose column differs
// must lie within 1e-6 rad of a column boundary (float rounding of the two computations), and the
// directions GuideBinDirection draws in a bin must map back to it (up to the same rounding).
#include "../../lambda_amd/csrc/guide.h"
#include <cmath>
#include <cstdio>
#include <random>
using namespace lgpu;

static uint32_t DefinitionBin(V3 w) {
  const double y = std::fmin(1.0, std::fmax(-1.0, (double)w.y));
  int r = (int)std::floor((y + 1.0) * 4.0);
  r = r < 0 ? 0 : (r > 7 ? 7 : r);
  const double ph = std::atan2((double)w.z, (double)w.x) + M_PI;
  int c = (int)std::floor(ph / (2 * M_PI / 16));
  c = c < 0 ? 0 : (c > 15 ? 15 : c);
  return (uint32_t)(r * 16 + c);
}

int main() {
  std::mt19937 g(7);
  std::normal_distribution<float> n;
  std::uniform_real_distribution<float> u(0.f, 1.f);
  const long N = 4000000;
  long differ = 0, far = 0;
  for (long i = 0; i < N; ++i) {
    const float x = n(g), y = n(g), z = n(g);
    const float l = std::sqrt(x * x + y * y + z * z);
    const V3 w = v3(x / l, y / l, z / l);
    if (GuideBin(w) == DefinitionBin(w)) continue;
    ++differ;
    const double ph = std::atan2((double)w.z, (double)w.x) + M_PI, step = 2 * M_PI / 16;
    const double dphi = std::fabs(ph / step - std::round(ph / step)) * step;
    const double dcos = std::fabs(((double)w.y + 1.0) * 4.0 - std::round(((double)w.y + 1.0) * 4.0)) / 4.0;
    if (dphi > 1e-6 && dcos > 1e-6) ++far;
  }
  long back = 0;
  for (uint32_t b = 0; b < 128; ++b)
    for (int i = 0; i < 20000; ++i) back += GuideBin(GuideBinDirection(b, u(g), u(g))) != b;
  std::printf("differ %ld of %ld (away from a boundary: %ld); round trip misses %ld of %d\n", differ, N, far, back,
              128 * 20000);
  return (far == 0 && differ <= N / 100000 && back <= 128 * 20000 / 100000) ? 0 : 1;
}

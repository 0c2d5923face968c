// Learned path guiding: network layout, input encoding, the directional bin distribution and
// the training record, shared by the host, the guiding kernels (guide.hip) and the guided path
// shade kernel (render.hip).
//
// The network maps a shading point and its outgoing direction to a distribution over incident
// directions: 5 inputs (position normalised to the scene bounds, theta/pi and phi/2pi of wo)
// -> frequency encoding (8 octaves, sin and cos: 80 features) -> 5 hidden layers of 64 ReLU
// units -> 128 bin logits and one mixing logit. BF16 weights and activations, FP32 accumulation
// (MFMA), FP32 master weights trained with Adam.
//
// The distribution: 128 world-space bins, 8 equal steps of cos(theta) (y up) by 16 of phi, each
// 4 pi / 128 sr; within a bin uniform in (cos theta, phi), so uniform in solid angle. The shade
// stage samples it as one-sample MIS with the BSDF: with probability c the guide, otherwise the
// BSDF; the pdf of either choice is c q + (1 - c) p_bsdf. The reference has no guiding: this is
// an extension of PathIntegrator::Li's BSDF sampling step (PathIntegrator.cpp:46-67).
#pragma once
#include "dmath.h"

namespace lgpu {

constexpr int kGuideScalars = 5;
constexpr int kGuideOctaves = 8;
constexpr int kGuideIn = kGuideScalars * kGuideOctaves * 2;  // 80
constexpr int kGuideHidden = 64;
constexpr int kGuideBinsCos = 8, kGuideBinsPhi = 16;
constexpr int kGuideBins = kGuideBinsCos * kGuideBinsPhi;    // 128
constexpr int kGuideOut = kGuideBins + 16;                   // 128 logits, mixing logit, 15 unused
constexpr int kGuideLayers = 6;
constexpr float kGuideBinSr = 4.f * 3.14159265358979323846f / (float)kGuideBins;
// mixing probability c = kGuideCMin + kGuideCSpan * sigmoid(z): the BSDF keeps at least 10 %
constexpr float kGuideCMin = 0.05f, kGuideCSpan = 0.85f;

// Layer l maps C_l inputs to R_l outputs: W_l is R_l x C_l row-major, then b_l (R_l).
LGPU_HD constexpr int GuideRows(int l) { return l == kGuideLayers - 1 ? kGuideOut : kGuideHidden; }
LGPU_HD constexpr int GuideCols(int l) { return l == 0 ? kGuideIn : kGuideHidden; }
LGPU_HD constexpr int GuideWOff(int l) {
  int o = 0;
  for (int k = 0; k < l; ++k) o += GuideRows(k) * GuideCols(k);
  return o;
}
constexpr int kGuideWeights = GuideWOff(kGuideLayers);  // 30720
LGPU_HD constexpr int GuideBOff(int l) {
  int o = 0;
  for (int k = 0; k < l; ++k) o += GuideRows(k);
  return o;
}
constexpr int kGuideBiases = GuideBOff(kGuideLayers);  // 464
constexpr int kGuideParams = kGuideWeights + kGuideBiases;

// FP32 -> BF16, round to nearest even (finite inputs).
LGPU_HD uint16_t ToBf16(float f) {
  uint32_t u = AsU32(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
LGPU_HD float FromBf16(uint16_t h) { return AsF32((uint32_t)h << 16); }

// The encoding: feature w (< 16) of input x is sin (even w) or cos (odd w) of 2 pi t, octave
// w / 2, t = fract(2^(octave - 1) x) (the argument reduced exactly); feature f (< 80) of a query
// is feature f % 16 of input f / 16. The kernels evaluate it in guide.hip (EncodeB).

// Network inputs of a shading point: position in the scene bounds, wo = -ray direction.
LGPU_HD void GuideInputs(V3 p, V3 wo, const float bmin[3], const float binv[3], float s[kGuideScalars]) {
  s[0] = Clampf((p.x - bmin[0]) * binv[0], 0.f, 1.f);
  s[1] = Clampf((p.y - bmin[1]) * binv[1], 0.f, 1.f);
  s[2] = Clampf((p.z - bmin[2]) * binv[2], 0.f, 1.f);
  s[3] = acosf(Clampf(wo.y, -1.f, 1.f)) * (1.f / 3.14159265358979323846f);
  float ph = atan2f(wo.z, wo.x);
  if (ph < 0.f) ph += kPi2;
  s[4] = Clampf(ph * (1.f / kPi2), 0.f, 1.f);
}

// Direction -> bin (row-major: cos rows, phi columns) and back. Column c holds phi = atan2(z, x)
// + pi in [c, c + 1) x 2 pi / 16. It is found without an arctangent: the angle a = atan2(|z|, |x|)
// lies in the k-th 22.5-degree step of its quadrant, k = [|z| >= tan(22.5) |x|] + [|z| >= |x|]
// + [|z| >= tan(67.5) |x|], and the signs of x and z place that step among the 16 columns
// (quadrant x >= 0, z >= 0: 8 + k; x < 0, z >= 0: 15 - k; x < 0, z < 0: k; x >= 0, z < 0: 7 - k).
LGPU_HD uint32_t GuideBin(V3 w) {
  static_assert(kGuideBinsPhi == 16, "the column test assumes 16 steps of phi");
  int r = (int)floorf((Clampf(w.y, -1.f, 1.f) + 1.f) * (.5f * kGuideBinsCos));
  r = r < 0 ? 0 : (r > kGuideBinsCos - 1 ? kGuideBinsCos - 1 : r);
  const float ax = fabsf(w.x), az = fabsf(w.z);
  const int k = (az >= 0.41421356237f * ax) + (az >= ax) + (az >= 2.41421356237f * ax);
  const int c = w.z < 0.f ? (w.x < 0.f ? k : 7 - k) : (w.x < 0.f ? 15 - k : 8 + k);
  return (uint32_t)(r * kGuideBinsPhi + c);
}
LGPU_HD V3 GuideBinDirection(uint32_t bin, float u0, float u1) {
  const uint32_t r = bin / kGuideBinsPhi, c = bin % kGuideBinsPhi;
  const float ct = Clampf(-1.f + ((float)r + u0) * (2.f / kGuideBinsCos), -1.f, 1.f);
  const float ph = -3.14159265358979323846f + ((float)c + u1) * (kPi2 / kGuideBinsPhi);
  const float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
  return v3(st * cosf(ph), ct, st * sinf(ph));
}

// Training record of one guided scattering event (48 B):
//   a = {position normalised (3), theta/pi of wo}
//   b = {phi/2pi of wo, bin of wi (bits), BSDF pdf of wi, sample slot (bits)}
//   c = {radiance luminance of the sample when the next vertex is shaded (NaN until then),
//        throughput luminance before the scattering, training weight, 0}
// weight = (L_final - L_after) / beta: the radiance that arrived along wi, over the pdf it was
// sampled with, times f cos (the integrand of the guiding objective at wi, over its sampling pdf).
struct GuideRecord {
  float4 a, b, c;
};

// Device view of one guide for the render kernels.
struct GuideRefs {
  const uint4 *bins;      // 8 x uint4 per path slot: 128 bytes of bin weights (>= 1)
  const float2 *hdr;      // per path slot: {mixing probability c, sum of the bin weights}
  GuideRecord *rec;       // training records of this wave
  uint32_t *rec_count;    // records appended (may pass rec_cap; only the first rec_cap are written)
  uint32_t rec_cap;
  uint32_t rec_threshold; // record a vertex when its hash is below this (uint32 scale)
  uint32_t max_depth;     // guided vertices have path depth below this
  float bmin[3], binv[3];
};

}  // namespace lgpu

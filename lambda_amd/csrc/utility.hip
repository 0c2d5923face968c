// Utility-pass shade stage: UtilityIntegrator::Li with DepthPass, NormalPass and AOVPass
// (UtilityIntegrators.cpp:7-78). One launch shades the camera hits: the colour of the closest
// hit (every surface counts, pass-through ones included, as Scene::Intersect returns them) is
// turned into a reflectance spectrum and added to the sample; the path ends there.
#include "host.h"
#include "shading.h"
#include "trace.h"

namespace lgpu {

namespace {

__global__ void __launch_bounds__(kBlock) k_shade_utility(DevScene sc, RenderParams rp, PathState ps, SampleState ss,
                                                          uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const bool corrected = rp.corrected != 0;
  const uint32_t sid = ps.aux[i].x;
  const float4 ro4 = ps.ray_o[i], rd4 = ps.ray_d[i];
  const float4 hr = ps.hit[i];
  const uint32_t bi = __float_as_uint(hr.w);
  float c[3] = {0.f, 0.f, 0.f};
  if (bi == kNone) {
    if (rp.integrator == LGPU_INTEGRATOR_DEPTH) { c[0] = rp.pass_bg[0]; c[1] = rp.pass_bg[1]; c[2] = rp.pass_bg[2]; }
  } else if (rp.integrator == LGPU_INTEGRATOR_DEPTH) {  // :26-33
    Hit h;
    ProcessHit(sc, ToV3(ro4), ToV3(rd4), hr.x, hr.y, hr.z, bi, corrected, h);
    const float d = powf(Mag(h.point - ToV3(ro4)) / rp.pass_max_depth, rp.pass_power);
    c[0] = c[1] = c[2] = d;
  } else if (rp.integrator == LGPU_INTEGRATOR_NORMAL) {  // :46-53; y is offset by 1 as written
    Hit h;
    ProcessHit(sc, ToV3(ro4), ToV3(rd4), hr.x, hr.y, hr.z, bi, corrected, h);
    const V3 nn = rp.pass_normal_g ? h.normalG : h.normalS;
    c[0] = nn.x * .5f + .5f;
    c[1] = nn.y * .5f + (corrected ? .5f : 1.f);
    c[2] = nn.z * .5f + .5f;
  }
  // AOVPass: the material AOV lookup is commented out in the reference (:72-75): black.
  float4 L = ss.L[sid];
  if (rp.mode) {  // Spectrum::FromRGB(c, Reflectance) at the hero wavelengths
    const float l0 = ro4.w;
    float lam[4];
    for (int k = 0; k < 4; ++k) {
      float l = l0 + 117.5f * (float)k;
      if (l >= 830.f) l -= 470.f;
      lam[k] = l;
    }
    L.x += Uplift(c[0], c[1], c[2], lam[0]); L.y += Uplift(c[0], c[1], c[2], lam[1]);
    L.z += Uplift(c[0], c[1], c[2], lam[2]); L.w += Uplift(c[0], c[1], c[2], lam[3]);
  } else {
    L.x += c[0]; L.y += c[1]; L.z += c[2];
  }
  ss.L[sid] = L;
}

}  // namespace

void LaunchShadeUtility(uint32_t blocks, hipStream_t stream, const DevScene &sc, const RenderParams &rp, PathState ps,
                        SampleState ss, uint32_t n) {
  hipLaunchKernelGGL(k_shade_utility, dim3(blocks), dim3(kBlock), 0, stream, sc, rp, ps, ss, n);
}

}  // namespace lgpu

/* Renders through the reference's lambda* C API (lambda_compat.h over liblambda_gpu.so): a 2x2
 * floor quad seen from above with the normal pass, progressively; checks the floor's normal
 * colour at the image centre and no hit at the top row. Built with -DTRANSFORM_ONLY it checks
 * lambdaMakeAffineTransform and the instance setters alone (no GPU needed). Exit status 0 = pass. */
#define _DEFAULT_SOURCE
#include <math.h>
#include <stdatomic.h>
#include <stdio.h>
#include <unistd.h>

#include "lambda_compat.h"

#ifndef TRANSFORM_ONLY
static atomic_int passes;
static void on_pass(void) { atomic_fetch_add(&passes, 1); }
#endif

static int near(float a, float b, float tol) { return fabsf(a - b) <= tol; }

static int check_transforms(void) {
  float xfm[12], pos[3] = {1, 2, 3}, scale[3] = {2, 3, 4}, none[3] = {0, 0, 0};
  lambdaMakeAffineTransform(xfm, pos, scale, none);
  const float want[12] = {2, 0, 0, 0, 3, 0, 0, 0, 4, 1, 2, 3};
  for (int k = 0; k < 12; ++k)
    if (!near(xfm[k], want[k], 1e-6f)) { printf("affine[%d] = %g, want %g\n", k, xfm[k], want[k]); return 1; }
  float quarter[3] = {0, 0, 1.5707964f}, one[3] = {1, 1, 1};
  lambdaMakeAffineTransform(xfm, none, one, quarter);  /* rotation about z by pi/2 */
  if (!near(xfm[1], -1.f, 1e-6f) || !near(xfm[3], 1.f, 1e-6f) || !near(xfm[8], 1.f, 1e-6f)) {
    printf("rotation: %g %g %g\n", xfm[1], xfm[3], xfm[8]);
    return 1;
  }
  printf("transforms ok\n");
  return 0;
}

int main(void) {
  if (check_transforms()) return 1;
#ifdef TRANSFORM_ONLY
  return 0;
#else
  LAMBDA_Device *dev = lambdaCreateDevice();
  if (!dev) { printf("no device\n"); return 2; }
  static float v[16] = {-1, 0, -1, 0, 1, 0, -1, 0, 1, 0, 1, 0, -1, 0, 1, 0};
  static float n[16] = {0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0};
  static unsigned t[6] = {0, 3, 2, 0, 2, 1};
  LAMBDA_TriangleMesh *mesh = lambdaCreateTriangleMesh(dev, "floor");
  lambdaTriangleMeshAllocData(mesh, 4, 2);
  lambdaTriangleMeshSetBuffer(mesh, LAMBDA_BUFFER_VERTEX, v, 4);
  lambdaTriangleMeshSetBuffer(mesh, LAMBDA_BUFFER_TRIANGLE, t, 2);
  lambdaTriangleMeshSetBuffer(mesh, LAMBDA_BUFFER_VERTEX_NORMAL, n, 4);
  LAMBDA_Proxy *proxy = lambdaCreateProxy(dev, mesh);
  LAMBDA_Instance *inst = lambdaCreateInstance(dev, proxy);
  LAMBDA_Scene *scene = lambdaCreateScene(dev, "scene");
  lambdaAttachObject(scene, inst);
  lambdaCommitScene(scene);
  const int W = 64, H = 64;
  LAMBDA_Film *film = lambdaCreateFilm(dev, W, H);
  float eye[3] = {0, 2, 4};
  LAMBDA_Camera *cam = lambdaCreateCamera(dev, LAMBDA_CAMERA_THIN_LENS, eye, 3.14159265f, -atanf(0.5f));
  LAMBDA_RenderProperties *props = lambdaCreateRenderProperties();
  props->spp = 2;
  props->integrator = LAMBDA_INTEGRATOR_NORMAL;
  LAMBDA_RenderDirective *dir = lambdaCreateRenderDirective(dev, scene, film, cam, props);
  LAMBDA_ProgressiveRenderer *r = lambdaCreateProgressiveRenderer(dev, dir);
  if (!film || !r) { printf("setup failed\n"); return 3; }
  lambdaSetProgressiveRendererCallback(r, on_pass);
  lambdaStartProgressiveRenderer(r);
  for (int k = 0; k < 3000 && atomic_load(&passes) < 3; ++k) usleep(10000);
  lambdaStopProgressiveRenderer(r);
  int w = 0, h = 0;
  const float *px = (const float *)lambdaGetProgressiveRendererData(r, &w, &h);
  const int np = atomic_load(&passes);
  printf("passes %d, image %dx%d\n", np, w, h);
  if (np < 3 || w != W || h != H || !px) return 4;
  const float *c = px + 4 * ((H / 2) * W + W / 2);
  /* the normal pass as the reference writes it: (n.x/2 + 1/2, n.y/2 + 1, n.z/2 + 1/2) for n = (0, 1, 0) */
  printf("centre %g %g %g\n", c[0], c[1], c[2]);
  if (!near(c[0], 0.5f, 1e-3f) || !near(c[1], 1.5f, 1e-3f) || !near(c[2], 0.5f, 1e-3f)) return 5;
  int sky = 0;
  for (int x = 0; x < W; ++x) {
    const float *q = px + 4 * x;  /* one image row: its rays point above the horizon */
    if (q[0] == 0.f && q[1] == 0.f && q[2] == 0.f) ++sky;
  }
  const float *q0 = px, *q1 = px + 4 * ((H - 1) * W);
  printf("first row %g %g %g, last row %g %g %g, empty pixels in the first row %d\n", q0[0], q0[1], q0[2], q1[0], q1[1],
         q1[2], sky);
  int skyLast = 0;
  for (int x = 0; x < W; ++x) {
    const float *q = px + 4 * ((H - 1) * W + x);
    if (q[0] == 0.f && q[1] == 0.f && q[2] == 0.f) ++skyLast;
  }
  if (sky != W && skyLast != W) return 6;  /* one edge row sees past the floor */
  lambdaReleaseProgressiveRenderer(r);
  lambdaReleaseDevice(dev);
  printf("smoke ok\n");
  return 0;
#endif
}

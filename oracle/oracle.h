/*
 * oracle.h — CPU restatement of the Zoophish/Lambda render path, used ONLY as
 * test infrastructure: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it as the checker. The product library never links it.
 *
 * It consumes the same scene description structs as the product's C ABI
 * (include/lambda_gpu.h) so that both are driven from one scene builder, and
 * re-implements the reference's algorithm independently: per-path recursive
 * PathIntegrator::Li, pull-evaluated shader graphs, pointer-based light tree,
 * binned-SAH BVH.
 */
#ifndef LAMBDA_ORACLE_H
#define LAMBDA_ORACLE_H
#include <stdint.h>
#include "../include/lambda_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene_s *oracle_scene;

oracle_scene oracle_scene_create(void);
void oracle_scene_destroy(oracle_scene s);
int oracle_scene_add_texture(oracle_scene s, const lgpu_texture_desc *d);
int oracle_scene_add_medium(oracle_scene s, const lgpu_medium_desc *d);
int oracle_scene_add_material(oracle_scene s, const lgpu_material_desc *d);
int oracle_scene_add_mesh(oracle_scene s, const lgpu_mesh_desc *d);
int oracle_scene_add_proxy(oracle_scene s, const lgpu_mesh_desc *d);
int oracle_scene_add_instance(oracle_scene s, uint32_t proxy, const float xfm[12], uint32_t material);
int oracle_scene_add_light(oracle_scene s, const lgpu_light_desc *d);
int oracle_scene_commit(oracle_scene s, uint32_t light_strategy);
int oracle_light_tree(oracle_scene s, float *nodes, uint32_t max_nodes);

/* Per-path radiance, 5 floats per path (see lgpu_trace_paths). */
int oracle_trace_paths(oracle_scene s, const lgpu_render_desc *desc, uint32_t width, uint32_t height, uint32_t n,
                       const uint32_t *pixels, const uint32_t *samples, float *out, int threads);
/* Film render over a pixel window [x0,x0+w) x [y0,y0+h) of a width x height
 * image: FP64 RGB sums (3 per window pixel) and counts. Tiles of 16x16 are
 * distributed dynamically over `threads` worker threads. */
int oracle_render(oracle_scene s, const lgpu_render_desc *desc, uint32_t width, uint32_t height,
                  uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, int threads,
                  double *rgb_sum, uint32_t *counts);
/* Closest hit / occlusion on the oracle BVH (same record layout as lgpu_intersect). */
int oracle_intersect(oracle_scene s, uint32_t n, const float *rays, uint32_t *hits);
int oracle_occluded(oracle_scene s, uint32_t n, const float *rays, uint32_t *occ);

/* Building blocks exposed for closed-form tests. */
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float oracle_fr_dielectric(float cos_i, float eta_i, float eta_t);
void oracle_fr_conductor(float cos_i, const float eta_i[3], const float eta_t[3], const float k[3], float out[3]);
float oracle_microfacet_d(int ggx, const float wh[3], float ax, float ay);
float oracle_microfacet_lambda(int ggx, const float w[3], float ax, float ay);
void oracle_blackbody_rgb(float temperature, float out[3]);
float oracle_blackbody_normalised(float lambda_nm, float temperature);
void oracle_cmf(float lambda_nm, float out_xyz[3]);
float oracle_uplift(const float rgb[3], float lambda_nm);
float oracle_hg_p(float cos_theta, float g);
/* Piecewise1D (Piecewise.cpp:9-44): returns offset; pdf, remapped out. */
uint32_t oracle_piecewise1d_sample(const float *f, uint32_t n, float u, float *pdf);
/* Media (HomogeneousMedium.cpp:12-92 plus the grid medium): mode 0 RGB, 1 HERO4 (lambda[4]).
 * sample_* out: t, pdf, interaction (0/1), weight[4]. */
int oracle_medium_tr(oracle_scene s, uint32_t med, const float o[3], const float d[3], float t_far, int mode,
                     const float *lambda, float out[4]);
int oracle_medium_sample_distance(oracle_scene s, uint32_t med, const float o[3], const float d[3], float t_far,
                                  float u0, float u1, int mode, const float *lambda, float out[7]);
float oracle_medium_pdf_distance(oracle_scene s, uint32_t med, const float o[3], const float d[3], float t, int mode,
                                 const float *lambda, int corrected);
int oracle_medium_sample_equiangular(oracle_scene s, uint32_t med, const float o[3], const float d[3],
                                     const float hit_point[3], float t_far, const float light[3], float eps, int mode,
                                     const float *lambda, int corrected, float out[7]);
float oracle_medium_pdf_equiangular(const float o[3], const float d[3], const float light[3], float t_far, float t,
                                    int corrected);
/* ManyLightSampler pick and pdf at a point (tree strategy): pdfs[nTreeLights + 1] (last = env). */
int oracle_light_pdfs(oracle_scene s, const float p[3], const float n[3], int corrected, float *pdfs);
int oracle_light_sample(oracle_scene s, const float p[3], const float n[3], int corrected, uint32_t pixel,
                        uint32_t sample, float *pdf);
/* HenyeyGreenstein::Sample_p (HenyeyGreenstein.cpp:21-36) with explicit u; returns pdf. */
int oracle_eval_node(oracle_scene s, uint32_t material, uint32_t node, const float ev[11], int mode, const float *lambda,
                     float out[4]);
void oracle_post(const float *in, uint32_t w, uint32_t h, const lgpu_post_desc *d, float *out32, uint8_t *out8);
void oracle_denoise(uint32_t w, uint32_t h, const float *colour, const float *albedo, const float *normal,
                    const lgpu_denoise_desc *d, float *out);
/* A material's BxDF in the world frame (normal +y): n samples (Philox pixel, sample i), 11 floats
 * each: wiL[3], pdf, f[4], delta flag, spawn side along the normal, eta; and f[4] + pdf of n pairs. */
int oracle_bxdf_sample_n(oracle_scene s, uint32_t material, const float woL[3], int mode, const float *lambda,
                         int corrected, uint32_t pixel, uint32_t n, float eta, float *out);
int oracle_bxdf_eval_n(oracle_scene s, uint32_t material, uint32_t n, const float *woL, const float *wiL, int mode,
                       const float *lambda, int corrected, float *out);
/* Camera::GenerateRay at n image points uv[2n] (aperture draws from Philox (pixel, i)): o[3], d[3]. */
int oracle_generate_ray_n(const lgpu_camera *cam, uint32_t n, const float *uv, uint32_t pixel, int corrected, float *out);
/* Light::Sample_Li of a scene light at p with normal n: count draws of wi[3], pdf, Li[4]. */
int oracle_light_sample_li_n(oracle_scene s, uint32_t light, const float p[3], const float n[3], uint32_t count,
                             int mode, const float *lambda, int corrected, uint32_t pixel, float *out);
/* EnvironmentLight::PDF_Li for n world directions. */
int oracle_env_pdf_n(oracle_scene s, uint32_t light, uint32_t n, const float *wi, float *out);
/* Learned guiding network (restated from its documented layout): parameter count, quantised
 * forward (129 floats per query), objective gradient (rounded = 1: the device's BF16 steps;
 * 0: exact) returning the loss, and the exact loss with double parameters. */
int oracle_guide_num_params(void);
void oracle_guide_forward(const float *params, uint32_t n, const float *inputs, float *out);
float oracle_guide_gradient(const float *params, uint32_t n, const float *records, int rounded, float *grad);
double oracle_guide_loss64(const double *params, uint32_t n, const float *records);
float oracle_tree_importance(const float node[10], const float p[3], const float n[3], int corrected);
float oracle_hg_sample(float g, const float wo[3], float u0, float u1, int corrected, float wi[3]);

#ifdef __cplusplus
}
#endif
#endif

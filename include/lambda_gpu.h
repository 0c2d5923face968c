/*
 * lambda_gpu.h — C ABI of the MI355X wavefront spectral path tracer.
 *
 * Plain C: opaque handles, plain pointers and sizes, status codes, no
 * exceptions across the boundary. Every scene buffer passed in is copied;
 * the caller keeps ownership of its memory (the reference's
 * lambdaTriangleMeshSetBuffer took ownership, Lambda.cpp:211-244; this ABI
 * does not).
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the Zoophish/Lambda source tree).
 *
 * Threading: one render in flight per (device, stream); calls on one device
 * are serialised by the caller.
 */
#ifndef LAMBDA_GPU_H
#define LAMBDA_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LGPU_NONE 0xFFFFFFFFu
#define LGPU_ABI_VERSION 8u

typedef enum lgpu_status {
  LGPU_OK = 0,
  LGPU_ERR_INVALID_ARGUMENT = 1,
  LGPU_ERR_HIP = 2,
  LGPU_ERR_STATE = 3,
  LGPU_ERR_OUT_OF_MEMORY = 4,
  LGPU_ERR_UNSUPPORTED = 5,
  LGPU_ERR_GRAPH = 6,
  LGPU_ERR_NO_DEVICE = 7
} lgpu_status;

typedef struct lgpu_device_s *lgpu_device;
typedef struct lgpu_scene_s *lgpu_scene;
typedef struct lgpu_film_s *lgpu_film;

/* ------------------------------------------------------------------------
 * Shader-node networks (reference: src/shading/graph/ShaderGraph.h:72-327,
 * include/lambda/Material.h:42-75). A material is a small DAG of nodes; each
 * node has one typed output and up to four inputs that name other nodes of
 * the same material. The graph is compiled to bytecode at
 * lgpu_scene_add_material().
 * --------------------------------------------------------------------- */
typedef enum lgpu_node_type {
  /* inputs (GraphInputs.cpp) */
  LGPU_NODE_SCALAR = 1,         /* out scalar = p[0] */
  LGPU_NODE_VEC2 = 2,           /* out vec2 = p[0..1] */
  LGPU_NODE_VEC3 = 3,           /* out vec3 = p[0..2] */
  LGPU_NODE_RGB = 4,            /* out colour = p[0..2] (RGBInput) */
  LGPU_NODE_BLACKBODY = 5,      /* out spectrum; in[0] temperature (scalar) */
  LGPU_NODE_IMAGE_COLOUR = 6,   /* out colour; p[0] = texture id */
  LGPU_NODE_IMAGE_SCALAR = 7,   /* out scalar; p[0] = texture id */
  LGPU_NODE_UV = 8,             /* out vec2 = surface uv (SurfaceInput) */
  LGPU_NODE_FRESNEL = 9,        /* FresnelInput (GraphInputs.cpp:183-198): out scalar = Schlick
                                   with n1 = the path's eta, n2 = in[0] (p[0] when unset),
                                   cos = woL.y */
  LGPU_NODE_SPECTRAL = 10,      /* SpectralInput (GraphInputs.cpp:55-71): out spectrum; samples
                                   in texture p[0] (1 channel, width x height values in order),
                                   evenly spaced over [p[1], p[2]] nm */
  LGPU_NODE_IMAGE_CHANNEL = 11, /* ImageTextureChannelInput: out scalar; p[0] texture, p[1]
                                   channel 0..3 (alpha 1 unless the texture is RGBA) */
  LGPU_NODE_SPECTRAL_TEXTURE = 12, /* SpectralTextureInput: out spectrum; p[0] texture of RGB spectra */
  LGPU_NODE_NORMAL = 13,        /* SurfaceInfoInput: out vec3 = shading normal */
  /* maths / converters (GraphMaths.cpp, GraphConverters.cpp) */
  LGPU_NODE_SCALAR_MATHS = 16,  /* out scalar = in[0] op in[1]; p[0] = op */
  LGPU_NODE_VECTOR_MATHS = 17,  /* out vec3 = in[0] op in[1]; p[0] = op */
  LGPU_NODE_DOT = 18,           /* out scalar */
  LGPU_NODE_LENGTH = 19,        /* out scalar */
  LGPU_NODE_MERGE_RGB = 20,     /* MergeRGBA: out colour from scalars in[0..2], alpha in[3] (1 when unset) */
  LGPU_NODE_SEPARATE_R = 21,    /* SeparateRGBA: out scalar: channel p[0] (0..3) of a colour */
  LGPU_NODE_CROSS = 22,         /* out vec3 = in[0] x in[1] */
  LGPU_NODE_SEPARATE_XYZ = 23,  /* out scalar: component p[0] of a vec3 */
  LGPU_NODE_MERGE_XYZ = 24,     /* out vec3 from three scalars */
  LGPU_NODE_SCALAR_TO_COLOUR = 25, /* out colour (x, x, x, 1) */
  /* textures (GraphTexture.cpp); at a medium interaction the 3-D variant on the event point */
  LGPU_NODE_CHECKER = 32,       /* out scalar; p[0] = scale */
  LGPU_NODE_PERLIN = 33,        /* out scalar; p[0] = scale (3-D variant: 0, as written) */
  LGPU_NODE_VALUE_NOISE = 34,   /* out scalar; p[0] = scale (3-D variant: 0, as written) */
  LGPU_NODE_OCTAVE_NOISE = 35,  /* out scalar; p[0] scale, p[1] octaves, p[2] base texture node
                                   type (CHECKER/PERLIN/VALUE_NOISE/VORONOI), p[3] base scale */
  LGPU_NODE_VORONOI = 36,       /* out scalar; p[0] = scale */
  /* BxDF nodes (GraphBxDF.cpp) */
  LGPU_NODE_LAMBERTIAN = 64,    /* in[0] albedo */
  LGPU_NODE_OREN_NAYAR_BRDF = 65, /* in[0] albedo, in[1] sigma (radians) */
  LGPU_NODE_OREN_NAYAR_BTDF = 66, /* in[0] albedo, in[1] sigma */
  LGPU_NODE_MICROFACET_BRDF = 67, /* in[0] albedo, in[1] roughness;
                                     p[0] distribution (0 Beckmann, 1 GGX);
                                     p[1] fresnel (0 dielectric, 1 conductor);
                                     p[2] dielectric etaT;
                                     p[3..5] conductor eta RGB, p[6..8] k RGB;
                                     p[9] != 0: roughness socket is alpha */
  LGPU_NODE_FRESNEL_BSDF = 68,  /* in[0] albedo, in[1] ior */
  LGPU_NODE_SPECULAR_BRDF = 69, /* in[0] albedo */
  LGPU_NODE_SPECULAR_BTDF = 70, /* in[0] albedo; p[0] etaT */
  LGPU_NODE_GHOST_BTDF = 71,    /* in[0] alpha */
  LGPU_NODE_MIX_BSDF = 72,      /* in[0] bxdf a, in[1] bxdf b, in[2] ratio */
  LGPU_NODE_ROUGH_DIELECTRIC = 73 /* project extension (no reference BTDF,
                                     Microfacet.h:28): in[0] albedo,
                                     in[1] alpha; p[0] ior (GGX, Walter 2007) */
} lgpu_node_type;

typedef enum lgpu_maths_op {
  LGPU_OP_ADD = 0, LGPU_OP_SUBTRACT = 1, LGPU_OP_MULTIPLY = 2, LGPU_OP_DIVIDE = 3
} lgpu_maths_op;

typedef struct lgpu_node {
  uint32_t type;      /* lgpu_node_type */
  uint32_t in[4];     /* node indices within the material, LGPU_NONE = unset */
  float p[12];        /* constant parameters */
} lgpu_node;

/* reference: Material {mediaBoundary, bxdf, light} (src/shading/Material.h:18-32)
 * plus MeshLight {emission, intensity} (src/lighting/MeshLight.h:13-16). */
typedef struct lgpu_material_desc {
  uint32_t num_nodes;
  const lgpu_node *nodes;
  uint32_t bxdf_root;        /* node index of the BxDF, LGPU_NONE = pass-through */
  uint32_t emission_root;    /* node index of the emission socket, LGPU_NONE = not a light */
  float emission_intensity;  /* MeshLight::intensity */
  uint32_t medium_interior;  /* medium id or LGPU_NONE (MediaBoundary) */
  uint32_t medium_exterior;
} lgpu_material_desc;

/* reference: TriangleMesh (src/core/TriangleMesh.h:14-41) */
typedef struct lgpu_mesh_desc {
  uint32_t num_vertices;
  uint32_t num_triangles;
  const float *positions;    /* 3 * num_vertices */
  const uint32_t *indices;   /* 3 * num_triangles */
  const float *normals;      /* 3 * num_vertices or NULL (flat shading) */
  const float *tangents;     /* 4 * num_vertices (w = bitangent sign) or NULL */
  const float *uvs;          /* 2 * num_vertices or NULL */
  uint32_t material;         /* material id */
} lgpu_mesh_desc;

typedef enum lgpu_light_type {
  LGPU_LIGHT_POINT = 1,       /* PointLight.cpp */
  LGPU_LIGHT_SPOT = 2,        /* Spotlight.cpp */
  LGPU_LIGHT_ENVIRONMENT = 3  /* EnvironmentLight.cpp */
} lgpu_light_type;

typedef struct lgpu_light_desc {
  uint32_t type;
  float position[3];
  float axis[3];             /* spot axis */
  float cone_angle;          /* spot: radians */
  float falloff_start;       /* spot: radians */
  float intensity;
  float rgb[3];              /* emission colour (used when temperature == 0) */
  float temperature;         /* > 0: blackbody emission */
  /* environment: equirectangular FP32 RGB, row-major, width x height */
  uint32_t env_width, env_height;
  const float *env_rgb;
  float env_offset[2];       /* EnvironmentLight::offset (phi, theta) */
} lgpu_light_desc;

typedef enum lgpu_medium_type {
  LGPU_MEDIUM_HOMOGENEOUS = 1, /* HomogeneousMedium.cpp */
  LGPU_MEDIUM_GRID = 2         /* project extension: FP32 density grid */
} lgpu_medium_type;

typedef struct lgpu_medium_desc {
  uint32_t type;
  float sigma_a[3];
  float sigma_s[3];
  float g;                   /* Henyey-Greenstein asymmetry */
  /* grid: density in [0,1] scales sigma_a/sigma_s; z-major nx*ny*nz */
  uint32_t nx, ny, nz;
  const float *density;
  float bounds_min[3], bounds_max[3];
} lgpu_medium_desc;

typedef struct lgpu_texture_desc {
  uint32_t width, height, channels; /* channels: 1, 3 or 4 (RGBA) */
  const float *data;                /* row-major */
} lgpu_texture_desc;

typedef enum lgpu_light_strategy {
  LGPU_LIGHTS_POWER = 0,   /* PowerLightSampler (LightSampler.cpp:14-35) */
  LGPU_LIGHTS_TREE = 1     /* ManyLightSampler (ManyLightSampler.cpp) */
} lgpu_light_strategy;

/* reference: PinholeCamera/ThinLensCamera/SphericalCamera (Camera.cpp:7-71),
 * CircularAperture/BladeAperture (Aperture.cpp:5-27). Axes are given
 * explicitly; lgpu_camera_set_rotation() fills them as Camera::SetRotation. */
typedef enum lgpu_camera_type {
  LGPU_CAMERA_PINHOLE = 0, LGPU_CAMERA_THIN_LENS = 1, LGPU_CAMERA_SPHERICAL = 2
} lgpu_camera_type;
typedef enum lgpu_aperture_type {
  LGPU_APERTURE_NONE = 0, LGPU_APERTURE_CIRCULAR = 1, LGPU_APERTURE_BLADE = 2
} lgpu_aperture_type;

typedef struct lgpu_camera {
  uint32_t type;
  float origin[3];
  float x_hat[3], y_hat[3], z_hat[3];
  float tan_fov;            /* tan(fov / 2) */
  float aspect;             /* height / width */
  float focal_length;       /* thin lens */
  uint32_t aperture;        /* lgpu_aperture_type */
  float aperture_size;
  uint32_t blades;
  float offset_phi, offset_theta; /* spherical */
} lgpu_camera;

typedef enum lgpu_integrator {
  LGPU_INTEGRATOR_PATH = 0,          /* PathIntegrator.cpp:15-88 */
  LGPU_INTEGRATOR_DIRECT = 1,        /* DirectLightingIntegrator.cpp:13-29 */
  LGPU_INTEGRATOR_VOLPATH = 2,       /* VolumetricPathIntegrator.cpp:20-162 */
  LGPU_INTEGRATOR_MIS_VOLPATH = 3,   /* MISVolumetricPathIntegrator.cpp:71-215 */
  LGPU_INTEGRATOR_DEPTH = 4,         /* DepthPass::P (UtilityIntegrators.cpp:26-33) */
  LGPU_INTEGRATOR_NORMAL = 5,        /* NormalPass::P (UtilityIntegrators.cpp:46-53) */
  LGPU_INTEGRATOR_AOV = 6            /* AOVPass::P (UtilityIntegrators.cpp:66-77) */
} lgpu_integrator;

typedef enum lgpu_spectral_mode {
  LGPU_SPECTRAL_RGB = 0,   /* the reference's RGBSpectrum (Spectrum.h:24) */
  LGPU_SPECTRAL_HERO4 = 1  /* 4 hero wavelengths per path over [360, 830) nm */
} lgpu_spectral_mode;

typedef enum lgpu_estimator_mode {
  LGPU_ESTIMATOR_FAITHFUL = 0,  /* reference semantics, defects included */
  LGPU_ESTIMATOR_CORRECTED = 1  /* defects fixed (unbiased estimator) */
} lgpu_estimator_mode;

typedef struct lgpu_guide_s *lgpu_guide;

typedef struct lgpu_render_desc {
  lgpu_scene scene;
  lgpu_film film;
  lgpu_camera camera;
  uint32_t integrator;      /* lgpu_integrator */
  uint32_t spp_begin;       /* sample indices [spp_begin, spp_end) per pixel */
  uint32_t spp_end;
  uint32_t max_bounces;     /* PathIntegrator::maxBounces */
  uint32_t min_bounces;     /* PathIntegrator::minBounces (RR starts after) */
  uint32_t spectral_mode;
  uint32_t estimator_mode;
  uint32_t seed;            /* Philox key word 0 */
  uint32_t path_capacity;   /* paths in flight (the pool, which stays allocated on the device handle, about
                               400-460 B per path): 0 = up to 134217728, no more than this render's samples,
                               held to a quarter of free device memory; a value sizes the pool as given even
                               for a smaller render (so an untimed render can prepare it) */
  uint32_t flags;           /* LGPU_RENDER_COUNT_TRAVERSAL: count BVH nodes / triangles visited */
  /* utility passes (UtilityIntegrators.cpp:26-53); the reference's defaults are power 1,
   * background (1, 0, 1), shading normals */
  float pass_max_depth;     /* DepthPass::maxDepth (> 0) */
  float pass_power;         /* DepthPass::power */
  float pass_background[3]; /* DepthPass::background, returned on a miss */
  uint32_t pass_geometric_normal; /* NormalPass::normalG */
  /* learned path guiding (no reference counterpart: an extension of PathIntegrator::Li's BSDF
   * sampling step, PathIntegrator.cpp:46-67). NULL = plain BSDF sampling. With a guide, guidable
   * vertices (no delta lobe) sample one-sample MIS of the guide's distribution and the BSDF;
   * path integrator and corrected estimator only. */
  lgpu_guide guide;
  uint32_t guide_train_steps; /* Adam steps on each wave's records once the wave is done (0: none) */
} lgpu_render_desc;

#define LGPU_RENDER_COUNT_TRAVERSAL 1u

typedef struct lgpu_render_stats {
  double ms_total;          /* host wall time of the call */
  double ms_camera, ms_extend, ms_shade, ms_shadow; /* summed kernel times */
  uint64_t samples;
  uint64_t extension_rays;
  uint64_t shadow_rays;
  uint64_t nodes_visited;   /* with LGPU_RENDER_COUNT_TRAVERSAL: inner BVH2 nodes fetched (extension rays) */
  uint64_t tris_tested;     /* with LGPU_RENDER_COUNT_TRAVERSAL: triangles tested (extension rays) */
  uint32_t waves;           /* path-pool refills */
  uint32_t launches_extend, launches_shade, launches_shadow;
  uint64_t voxel_steps;     /* with LGPU_RENDER_COUNT_TRAVERSAL: density-grid voxels visited by medium walks */
  uint64_t nodes4_fetched;  /* with LGPU_RENDER_COUNT_TRAVERSAL: 64-B 4-wide node records the render's walk fetches */
  uint64_t tris4_fetched;   /* with LGPU_RENDER_COUNT_TRAVERSAL: 48-B triangle records the render's walk fetches */
  double ms_guide_infer;    /* guided renders: summed inference kernel time */
  double ms_guide_train;    /* guided renders: host wall time of record finalisation and training */
  uint64_t guide_records;   /* guided renders: training records written */
} lgpu_render_stats;

/* ---------------- device (Lambda.h:37-40 lambdaCreateDevice/Release) ---- */
lgpu_status lgpu_device_create(int hip_device, lgpu_device *out);
lgpu_status lgpu_device_destroy(lgpu_device dev);
/* Message of the last failing call on this thread. */
const char *lgpu_last_error(void);
uint32_t lgpu_abi_version(void);

/* ---------------- scene (Scene.cpp:18-26,109-121; Lambda.cpp:168-244) --- */
lgpu_status lgpu_scene_create(lgpu_device dev, lgpu_scene *out);
lgpu_status lgpu_scene_destroy(lgpu_scene scene);
lgpu_status lgpu_scene_add_texture(lgpu_scene scene, const lgpu_texture_desc *desc, uint32_t *id);
lgpu_status lgpu_scene_add_medium(lgpu_scene scene, const lgpu_medium_desc *desc, uint32_t *id);
/* Material.cpp:94-195 lambdaCreateShaderNode + Material: graph -> bytecode */
lgpu_status lgpu_scene_add_material(lgpu_scene scene, const lgpu_material_desc *desc, uint32_t *id);
/* Scene::AddObject with TriangleMesh; an emissive material makes a MeshLight */
lgpu_status lgpu_scene_add_mesh(lgpu_scene scene, const lgpu_mesh_desc *desc, uint32_t *object_id);
/* InstanceProxy(object) + InstanceProxy::Commit (Instance.cpp:7-16): a mesh kept for
 * instancing; it is not part of the scene by itself. desc->material is the default material of
 * its instances (Instance::Commit uses the proxy object's material when it has none). */
lgpu_status lgpu_scene_add_proxy(lgpu_scene scene, const lgpu_mesh_desc *desc, uint32_t *proxy_id);
/* Instance(proxy) + Instance::Commit + Scene::AddObject (Instance.cpp:18-34, Scene.cpp:114-121):
 * the proxy's mesh under the affine transform xfm, 3x4 column-major as affine3 stores it
 * (affine3.h:18-33: columns x, y, z, then the translation). material = LGPU_NONE takes the
 * proxy's. Instances are flattened into world space at this call (normals by the inverse
 * transpose, tangents by the linear part); an emissive material makes the instance a mesh
 * light. Returns the instance's object id (shared with lgpu_scene_add_mesh's ids). */
lgpu_status lgpu_scene_add_instance(lgpu_scene scene, uint32_t proxy_id, const float xfm[12], uint32_t material,
                                    uint32_t *object_id);
/* Scene::AddLight for point / spot / environment lights */
lgpu_status lgpu_scene_add_light(lgpu_scene scene, const lgpu_light_desc *desc, uint32_t *id);
/* Scene::Commit + LightSampler::Commit: builds BVH and light sampler, uploads. */
lgpu_status lgpu_scene_commit(lgpu_scene scene, uint32_t light_strategy);
/* Introspection for tests: counts after commit. */
lgpu_status lgpu_scene_info(lgpu_scene scene, uint32_t *num_triangles, uint32_t *num_bvh_nodes,
                            uint32_t *num_lights, uint32_t *num_tree_nodes);
/* In-place parameter updates of a committed scene (no reference counterpart: the inverse-
 * rendering loop of the adjoint render changes textures and density grids between iterations
 * without a re-commit). texels: the texture's w * h * channels floats; density: the grid's
 * nx * ny * nz floats (host memory). A texture that an emission socket reads cannot be updated in
 * place (LGPU_ERR_UNSUPPORTED: it feeds the light sampler; re-commit instead). */
lgpu_status lgpu_scene_update_texture(lgpu_scene scene, uint32_t texture_id, const float *texels);
lgpu_status lgpu_scene_update_density(lgpu_scene scene, uint32_t medium_id, const float *density);
/* Light-tree export for tests: per node {center xyz, diag, axis xyz, thetaO,
 * thetaE, power, child0, child1 (as float bits), first, count, parent} = 15 floats. */
lgpu_status lgpu_scene_light_tree(lgpu_scene scene, float *nodes, uint32_t max_nodes);

/* ---------------- camera helper (Camera.cpp:14-35) ---------------------- */
void lgpu_camera_init(lgpu_camera *cam, uint32_t type, const float origin[3], float width, float height);
void lgpu_camera_set_rotation(lgpu_camera *cam, float phi, float theta);
void lgpu_camera_set_forwards(lgpu_camera *cam, const float forwards[3]);
void lgpu_camera_set_fov(lgpu_camera *cam, float fov);

/* ---------------- film (Film.h:19-45; Lambda.h:121-126) ----------------- */
lgpu_status lgpu_film_create(lgpu_device dev, uint32_t width, uint32_t height, lgpu_film *out);
lgpu_status lgpu_film_destroy(lgpu_film film);
lgpu_status lgpu_film_clear(lgpu_film film);
/* FP64 accumulation: rgb_sum[3*px] and sample counts[px]. Either may be NULL. */
lgpu_status lgpu_film_read(lgpu_film film, double *rgb_sum, uint32_t *counts);
/* Film::ToRGBTexture: RGBA32F mean per pixel (alpha = 1). */
lgpu_status lgpu_film_read_rgba(lgpu_film film, float *rgba);
/* The inverse of lgpu_film_read (no reference counterpart: Denoise::SetData, Denoise.cpp:13-44,
 * takes any image): restores sums and counts from host arrays, e.g. a saved film to resume or an
 * image to post-process or denoise. counts may be NULL: every count is then 1. */
lgpu_status lgpu_film_write(lgpu_film film, const double *rgb_sum, const uint32_t *counts);

/* ---------------- render (MosaicRenderer::Render / TileRenderer /
 * lambdaStartProgressiveRenderer: Render.cpp:43-86, MosaicRenderer.cpp:22-80) */
lgpu_status lgpu_render(const lgpu_render_desc *desc, lgpu_render_stats *stats);
/* lgpu_render without any allocation (no counterpart in the reference): for callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid it).
 * The path pool lives in a fixed 64 MiB device array per GPU, held until the process ends, so a
 * frame renders in more, smaller waves; desc->path_capacity (0 = as many paths as the array holds
 * for this scene) beyond that is rejected with LGPU_ERR_INVALID_ARGUMENT. Every call on one GPU
 * shares its array: calls that overlap on the same GPU (from different device objects) race. */
lgpu_status lgpu_render_noalloc(const lgpu_render_desc *desc, lgpu_render_stats *stats);

/* ---------------- batched renders (MosaicRenderer::Render's tile dispatch, MosaicRenderer.cpp:22-80
 * and Render.cpp:11-39, across many scenes): every job (scene, film, camera, spp range) renders
 * from one shared path pool whose segments are sized in proportion to the jobs' paths. Each chunk
 * is one hipGraph holding every job's whole wave (camera, bounce iterations with device-side
 * queue counts, film update) on several independent branches; chunks are captured while the
 * previous ones run. Jobs take the path or volumetric path integrators with the corrected
 * estimator and no guide; films must be distinct. Each film equals lgpu_render's for the same job
 * (same per-path radiance; the FP64 sums group samples per chunk, so they agree to rounding). */
#define LGPU_BATCH_AUTOTUNE 1u /* time pool capacities from 2^23 up to one chunk x 4, 8, 16 branches on whole-batch trials; cached */
typedef struct lgpu_batch_options {
  uint32_t path_capacity;  /* paths in flight over all jobs (0: 2^24, or the autotuner's choice) */
  uint32_t flags;          /* LGPU_BATCH_* */
  uint32_t max_iterations; /* bounce iterations captured per chunk (0: 2 max_bounces + 8); paths still
                              alive after them fail the call with LGPU_ERR_STATE */
  uint32_t branches;       /* independent graph branches (capture streams), 1..16 (0: 8) */
} lgpu_batch_options;
typedef struct lgpu_batch_stats {
  uint64_t samples;
  double ms_total;         /* the call, autotuning excluded */
  double ms_render;        /* first capture to completion */
  double ms_capture;       /* host time capturing and instantiating the chunk graphs */
  double ms_autotune;      /* trials of this call (0 when the choice was cached) */
  uint32_t chunks, graph_nodes, path_capacity, branches, iterations, jobs;
} lgpu_batch_stats;
lgpu_status lgpu_render_batch(const lgpu_render_desc *descs, uint32_t count, const lgpu_batch_options *options,
                              lgpu_batch_stats *stats);

/* ---------------- path-replay adjoint (no reference counterpart: the differentiation extension
 * of VolumetricPathIntegrator::Li, VolumetricPathIntegrator.cpp:20-162, and through it of
 * PathIntegrator::Li, PathIntegrator.cpp:15-88). The render is traced twice per wave with the same
 * random numbers: the primal pass accumulates the film as lgpu_render does, the replay adds
 * d/dtheta sum_px,c d_image[px][c] * mean_px,c to the gradients, where mean is the film mean of
 * this call's samples (spp_end - spp_begin per pixel). Sampling (directions, distances, light
 * picks, MIS weights, Russian roulette) is detached, as path replay backpropagation takes it.
 * Requires LGPU_INTEGRATOR_VOLPATH, LGPU_ESTIMATOR_CORRECTED, LGPU_SPECTRAL_RGB and no guide.
 * Parameters:
 *   textures: the albedo of Lambertian and microfacet lobes and the roughness of isotropic
 *     microfacet lobes, when the socket is one image lookup (IMAGE_COLOUR / IMAGE_SCALAR /
 *     IMAGE_CHANNEL) through at most one conversion; other reads of a texture (emission, mix
 *     ratios, procedural inputs) are constants, and a texture read both ways is refused.
 *     Layout: every texture's texels in texture id order (w * h * channels floats each).
 *   density: grid media, in medium id order (nx * ny * nz floats each; homogeneous media none).
 *   emission: per material, the gradient of its emission_intensity (mesh lights).
 * Gradients are added to the arrays (not cleared). */
#define LGPU_ADJOINT_DEVICE_POINTERS 1u /* pointers are device memory (this library's HIP runtime) */
typedef struct lgpu_adjoint_desc {
  const float *d_image;   /* width * height * 3, the loss gradient per film mean value */
  float *grad_textures;   /* or NULL */
  float *grad_density;    /* or NULL */
  float *grad_emission;   /* one per material, or NULL */
  uint32_t flags;         /* LGPU_ADJOINT_* */
  uint32_t reserved;
} lgpu_adjoint_desc;
lgpu_status lgpu_render_adjoint(const lgpu_render_desc *desc, const lgpu_adjoint_desc *adjoint,
                                lgpu_render_stats *stats);

/* ---------------- learned path guiding (no reference counterpart; an extension of the BSDF
 * sampling step, PathIntegrator.cpp:46-67). A 6-layer, width-64 MLP (BF16 MFMA, FP32
 * accumulation) maps a shading point and its outgoing direction (5 inputs, 8-octave frequency
 * encoding) to a distribution over 128 world-space direction bins and the probability c of
 * sampling it instead of the BSDF; trained online with Adam on the renderer's own radiance
 * records (the objective -w log(c q + (1 - c) p_bsdf), w the radiance along the sampled
 * direction over its pdf times f cos). */
typedef struct lgpu_guide_desc {
  uint32_t seed;            /* weight initialisation and batch selection */
  uint32_t record_capacity; /* training records kept per wave (0: 4194304) */
  uint32_t batch;           /* records per Adam step (0: 16384) */
  float learning_rate;      /* Adam step size (0: 0.002) */
  uint32_t max_depth;       /* guide vertices of path depth below this only (0: every depth) */
} lgpu_guide_desc;
lgpu_status lgpu_guide_create(lgpu_device dev, const lgpu_guide_desc *desc, lgpu_guide *out);
lgpu_status lgpu_guide_destroy(lgpu_guide guide);
/* Number of FP32 parameters: W1 (64x80), b1, W2..W5 (64x64), b2..b5, W6 (144x64: 128 bin
 * logits, the mixing logit, 15 unused rows), b6; weights of every layer first (row-major, layer
 * order), then the biases. */
uint32_t lgpu_guide_num_params(void);
lgpu_status lgpu_guide_get_params(lgpu_guide guide, float *params);
/* Replaces the parameters (and resets Adam's moments). */
lgpu_status lgpu_guide_set_params(lgpu_guide guide, const float *params);
/* Network output for n queries of 5 inputs in [0, 1] (position in the scene box, theta/pi, phi/2pi
 * of wo): 129 floats each, the 128 bin probabilities of the quantised table the renderer samples
 * and the mixing probability c. */
lgpu_status lgpu_guide_forward(lgpu_guide guide, uint32_t n, const float *inputs, float *out);
/* Gradient of the guiding objective over n records of 8 floats (5 inputs, bin, BSDF pdf, weight)
 * into grad (lgpu_guide_num_params floats) and the loss; parameters unchanged. */
lgpu_status lgpu_guide_gradient(lgpu_guide guide, uint32_t n, const float *records, float *grad, float *loss);
/* Adam steps on the records of the last guided wave; mean loss per step. */
lgpu_status lgpu_guide_train(lgpu_guide guide, uint32_t steps, float *loss);
lgpu_status lgpu_guide_stats(lgpu_guide guide, uint32_t *records, uint32_t *steps, float *last_loss);

/* ---------------- post-processing (ToneMap.cpp:5-73, PostProcessing.cpp:16-45,
 * Texture::SaveToImageFile Texture.cpp:19-81) ------------------------------------- */
typedef enum lgpu_tonemap { LGPU_TONEMAP_NONE = 0, LGPU_TONEMAP_ACES_FITTED = 1 } lgpu_tonemap;
typedef struct lgpu_post_desc {
  uint32_t tonemap;          /* ToneMap::method */
  uint32_t clamp;            /* ToneMap::clamp: channels to [0, 1] after the tone map */
  uint32_t gamma;            /* SaveToImageFile gamma correction c^(1/2.2), 8-bit output only */
  uint32_t estimator_mode;   /* faithful: ACES matrices applied transposed and the convolution's
                                edge wrap, as written; corrected: rows and clamp-to-edge */
  const float *kernel;       /* ConvolutionFilter kernel, kernel_w x kernel_h weights, or NULL */
  uint32_t kernel_w, kernel_h;
  float kernel_intensity;    /* ConvolutionFilter::intensity */
} lgpu_post_desc;
/* Film mean -> convolution -> tone map -> clamp on the device; rgba receives 4 floats per pixel,
 * rgba8 the gamma-corrected PackChannels bytes (alpha 255). Either may be NULL. */
lgpu_status lgpu_film_post(lgpu_film film, const lgpu_post_desc *desc, float *rgba, uint8_t *rgba8);
/* Denoise::SetData + Denoise::Process (Denoise.cpp:13-67; OIDN's filter with colour, albedo and
 * normal images). OIDN is not available: the filter is the edge-avoiding A-Trous wavelet transform
 * (Dammertz et al. 2010) over the film means, `iterations` passes of the 5x5 B3-spline kernel with
 * taps 2^i apart, weighted by exp(-|dc|^2/sc_i^2 - |dn|^2/sn^2 - |da|^2/sa^2) against the centre
 * pixel, sc_i = sigma_colour 2^-i; a sigma of 0 drops its term. albedo and normal may be NULL.
 * With an albedo film and demodulate != 0 the filter runs on colour / albedo (channels with
 * albedo > 1e-3) and multiplies the albedo back. The result goes to `out` as sums with counts of
 * 1 (out may be the colour film), ready for lgpu_film_post. */
typedef struct lgpu_denoise_desc {
  uint32_t iterations;       /* 1..12, typically 5 */
  float sigma_colour;        /* colour edge-stopping sigma (halved every pass) */
  float sigma_normal;        /* normal-image edge-stopping sigma */
  float sigma_albedo;        /* albedo-image edge-stopping sigma */
  uint32_t demodulate;       /* filter colour / albedo (needs an albedo film) */
} lgpu_denoise_desc;
lgpu_status lgpu_film_denoise(lgpu_film colour, lgpu_film albedo, lgpu_film normal, const lgpu_denoise_desc *desc,
                              lgpu_film out);
/* 8-bit RGBA PNG, rows in film order (stored deflate blocks). */
lgpu_status lgpu_write_png(const char *path, const uint8_t *rgba8, uint32_t width, uint32_t height);

/* ---------------- OBJ ingestion (AssetImporter::Import with aiProcess_Triangulate |
 * aiProcess_CalcTangentSpace, AssetImporter.cpp:5-15; MeshImport::LoadMeshVertexBuffers,
 * MeshImport.cpp:11-92). Host only, no device needed. ------------------------------- */
typedef struct lgpu_obj_s *lgpu_obj;
typedef struct lgpu_obj_material {
  const char *name;          /* usemtl / newmtl name ("" for faces without one) */
  float kd[3], ke[3];        /* MTL diffuse and emissive colours, when given */
  uint32_t has_kd, has_ke;
} lgpu_obj_material;
/* Parses an OBJ file (and the MTL files it names, for Kd / Ke): one mesh per run of faces of one
 * object and material, unshared corners, triangulated, uvs flipped to (1 - u, 1 - v); normals and
 * tangents (w = bitangent sign) only when every corner has a normal and a uv. */
lgpu_status lgpu_obj_load(const char *path, lgpu_obj *out);
lgpu_status lgpu_obj_counts(lgpu_obj obj, uint32_t *num_meshes, uint32_t *num_materials);
/* Mesh i as a descriptor for lgpu_scene_add_mesh (arrays owned by obj until lgpu_obj_destroy);
 * desc->material is the index of the OBJ material, to be mapped to a scene material id. */
lgpu_status lgpu_obj_mesh(lgpu_obj obj, uint32_t i, lgpu_mesh_desc *desc, const char **name);
lgpu_status lgpu_obj_material_info(lgpu_obj obj, uint32_t i, lgpu_obj_material *out);
lgpu_status lgpu_obj_destroy(lgpu_obj obj);

/* ---------------- progressive rendering (ProgressiveRender.cpp:13-62;
 * lambdaCreate/Release/Start/Stop/SetCallback/GetData ProgressiveRenderer,
 * Lambda.h:186-202, Lambda.cpp:431-458) ----------------------------------- */
typedef struct lgpu_progressive_s *lgpu_progressive;
typedef void (*lgpu_progressive_callback)(void *user);
/* desc is copied; its spp range [b, e) sets the pass: pass k renders sample indices
 * [b + k (e - b), b + (k + 1) (e - b)) of every pixel into desc->film. */
lgpu_status lgpu_progressive_create(const lgpu_render_desc *desc, lgpu_progressive *out);
/* Stops a running renderer (the pass in flight completes) and frees it. */
lgpu_status lgpu_progressive_destroy(lgpu_progressive p);
/* Called on the worker thread after each pass, once the new image can be read. */
lgpu_status lgpu_progressive_set_callback(lgpu_progressive p, lgpu_progressive_callback cb, void *user);
/* Starts the worker and returns at once; a second start while running does nothing. */
lgpu_status lgpu_progressive_start(lgpu_progressive p);
/* Returns after the pass in flight has completed; the film keeps what was rendered. */
lgpu_status lgpu_progressive_stop(lgpu_progressive p);
/* Film::Clear and restart at the first pass (between passes when running). */
lgpu_status lgpu_progressive_clear(lgpu_progressive p);
/* Copies the latest completed RGBA32F image (4 floats per pixel, film size) and the number of
 * passes in it; either may be NULL. Returns the worker's error if a pass failed. */
lgpu_status lgpu_progressive_read(lgpu_progressive p, float *rgba, uint32_t *passes);
/* 1 while the worker runs. */
int lgpu_progressive_running(lgpu_progressive p);

/* Per-path radiance for an explicit list of (pixel, sample) pairs. out holds
 * 5 floats per path: radiance lanes 0..3 (RGB in lanes 0..2, or the four hero
 * wavelengths) and lambda0 (HERO4) or 0. desc->film may be NULL. */
lgpu_status lgpu_trace_paths(const lgpu_render_desc *desc, uint32_t num_paths,
                             const uint32_t *pixels, const uint32_t *samples, float *out);

/* Closest-hit queries (Scene::Intersect, Scene.cpp:28-43): rays are 8 floats
 * {ox,oy,oz,tmin, dx,dy,dz,tmax}; hits are 4 words {t, u, v, prim} with
 * prim = LGPU_NONE on a miss (t, u, v as float bits). */
lgpu_status lgpu_intersect(lgpu_scene scene, uint32_t num_rays, const float *rays, uint32_t *hits);
/* Any-hit queries (Scene::MutualVisibility, Scene.cpp:83-99): out[i] = 1 if occluded. */
lgpu_status lgpu_occluded(lgpu_scene scene, uint32_t num_rays, const float *rays, uint32_t *occluded);

/* Philox4x32-10 on the device: counters (4 words each) and one key (2 words) -> 4 words each. */
lgpu_status lgpu_philox(lgpu_device dev, uint32_t n, const uint32_t *counters, const uint32_t key[2], uint32_t *out);

/* Measurement utility, no reference counterpart: `iters` launches of a plain
 * 16-byte-per-lane device copy of `bytes` bytes; *ms receives the mean launch
 * time. Used to calibrate HBM byte counters and the achievable bandwidth. */
lgpu_status lgpu_copy_bench(lgpu_device dev, uint64_t bytes, uint32_t iters, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* LAMBDA_GPU_H */

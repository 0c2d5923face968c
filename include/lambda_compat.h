/* lambda_compat.h - the reference's lambda* C API (include/lambda/Lambda.h:37-202) over lgpu_*.
 *
 * The same names, argument lists and enum values as the reference header, implemented by
 * liblambda_gpu.so on top of the lgpu_* entry points (lambda_gpu.h), so a program written against
 * the reference's C API builds and renders unchanged. Where the reference implementation
 * (src/api/Lambda.cpp) does not do what its header says, this one does what the header says:
 *   - lambdaCreateProxy binds the mesh (the reference leaves the proxy empty);
 *   - lambdaCreateCamera applies _phi and _theta (the reference ignores them);
 *   - lambdaMakeAffineTransform writes _xfm (position, scale, then X, Y and Z rotations; the
 *     reference rotates about X three times into a local copy);
 *   - LAMBDA_BUFFER_TEXTURE_COORDINATE sets the texture coordinates (the reference overwrites
 *     the tangents).
 * Kept as the reference has them: vertex, normal and tangent buffers hold 4 floats per vertex
 * (the 16-byte SSE Vec3, vec3.h:99-101), triangles 3 unsigned, texture coordinates 2 floats;
 * buffers are read at lambdaCommitScene (the caller keeps them alive until then); every object
 * gets the default material (white Lambertian); SetScale and SetPosition write the matrix's
 * diagonal and translation entries, SetEulerAngles replaces the whole matrix (Transformable.h:
 * 18-33); LAMBDA_INTEGRATOR_AOV renders with the path integrator (Lambda.cpp:369-371); cameras are
 * created with aspect 1 and fov 1.1 (Camera.cpp:7-12), thin lenses with focal length 1 and a
 * circular aperture of size 0.
 * The progressive renderer renders spp samples per pass; lambdaGetProgressiveRendererData points
 * at an RGBA32F copy of the latest pass that the next call to it refreshes.
 */
#ifndef LAMBDA_COMPAT_H
#define LAMBDA_COMPAT_H

#include <stddef.h>

#ifdef __cplusplus
#define LAMBDA_INT_ENUM : int
extern "C" {
#else
#define LAMBDA_INT_ENUM
#endif

typedef struct LAMBDA_Device LAMBDA_Device;
typedef struct LAMBDA_Scene LAMBDA_Scene;
typedef struct LAMBDA_TriangleMesh LAMBDA_TriangleMesh;
typedef struct LAMBDA_Proxy LAMBDA_Proxy;
typedef struct LAMBDA_Instance LAMBDA_Instance;
typedef struct LAMBDA_Film LAMBDA_Film;
typedef struct LAMBDA_Camera LAMBDA_Camera;
typedef struct LAMBDA_RenderDirective LAMBDA_RenderDirective;
typedef struct LAMBDA_ProgressiveRenderer LAMBDA_ProgressiveRenderer;

typedef enum LAMBDA_Type LAMBDA_INT_ENUM { LAMBDA_SCENE, LAMBDA_TRIANGLE_MESH, LAMBDA_TEXTURE, LAMBDA_MATERIAL } LAMBDA_Type;
typedef enum LAMBDA_Buffer LAMBDA_INT_ENUM {
  LAMBDA_BUFFER_VERTEX,
  LAMBDA_BUFFER_TRIANGLE,
  LAMBDA_BUFFER_VERTEX_NORMAL,
  LAMBDA_BUFFER_VERTEX_TANGENT,
  LAMBDA_BUFFER_TEXTURE_COORDINATE
} LAMBDA_Buffer;
typedef enum LAMBDA_Integrator LAMBDA_INT_ENUM {
  LAMBDA_INTEGRATOR_PATH,
  LAMBDA_INTEGRATOR_VOLPATH,
  LAMBDA_INTEGRATOR_DIRECT,
  LAMBDA_INTEGRATOR_NORMAL,
  LAMBDA_INTEGRATOR_DEPTH,
  LAMBDA_INTEGRATOR_AOV
} LAMBDA_Integrator;
typedef enum LAMBDA_LightStrategy LAMBDA_INT_ENUM { LAMBDA_LIGHT_STRATEGY_POWER, LAMBDA_LIGHT_STRATEGY_TREE } LAMBDA_LightStrategy;
typedef enum LAMBDA_CameraType LAMBDA_INT_ENUM { LAMBDA_CAMERA_THIN_LENS, LAMBDA_CAMERA_SPHERICAL } LAMBDA_CameraType;

/* Render settings (Lambda.h:160-168): samples per pixel per pass; tile size and thread count are
 * accepted and unused (the GPU renders whole-image waves). */
typedef struct LAMBDA_RenderProperties {
  unsigned spp;
  unsigned tileSizeX;
  unsigned tileSizeY;
  unsigned numThreads;
  LAMBDA_Integrator integrator;
  LAMBDA_LightStrategy lightStrategy;
} LAMBDA_RenderProperties;

/* Device 0 of this process's HIP devices; NULL if it cannot be opened. */
LAMBDA_Device *lambdaCreateDevice(void);
/* Frees the device and every object created on it. */
void lambdaReleaseDevice(LAMBDA_Device *_device);
/* Forgets the resource's name (the object itself lives until the device is released). */
void lambdaReleaseResource(LAMBDA_Device *_device, char *_name, LAMBDA_Type _type);

LAMBDA_Scene *lambdaCreateScene(LAMBDA_Device *_device, char *_name);
/* Builds the scene from its attached instances (BVH, lights) for rendering. */
void lambdaCommitScene(LAMBDA_Scene *_scene);
void lambdaReleaseScene(LAMBDA_Scene *_scene);
void lambdaAttachObject(LAMBDA_Scene *_scene, LAMBDA_Instance *_instance);
void lambdaDetachObject(LAMBDA_Scene *_scene, LAMBDA_Instance *_instance);

LAMBDA_TriangleMesh *lambdaCreateTriangleMesh(LAMBDA_Device *_device, char *_name);
void lambdaReleaseTriangleMesh(LAMBDA_TriangleMesh *_mesh);
void lambdaTriangleMeshAllocData(LAMBDA_TriangleMesh *_mesh, size_t _numVertices, size_t numTriangles);
/* _num must equal the allocated vertex (or triangle) count; the call is ignored otherwise. */
void lambdaTriangleMeshSetBuffer(LAMBDA_TriangleMesh *_mesh, LAMBDA_Buffer _bufferType, void *_ptr, size_t _num);

LAMBDA_Proxy *lambdaCreateProxy(LAMBDA_Device *_device, LAMBDA_TriangleMesh *_mesh);
void lambdaReleaseProxy(LAMBDA_Proxy *_proxy);
LAMBDA_Instance *lambdaCreateInstance(LAMBDA_Device *_device, LAMBDA_Proxy *_proxy);
void lambdaReleaseInstance(LAMBDA_Instance *_instance);
/* 3x4 column-major affine (columns x, y, z, translation), as affine3 stores it (affine3.h:18-33). */
void lambdaMakeAffineTransform(float _xfm[12], float _position[3], float _scale[3], float _eulerAngles[3]);
void lambdaSetTransform(LAMBDA_Instance *_instance, float *_xfm);
void lambdaSetPosition(LAMBDA_Instance *_instance, float _position[3]);
void lambdaSetScale(LAMBDA_Instance *_instance, float _scale[3]);
void lambdaSetEulerAngles(LAMBDA_Instance *_instance, float _eulerAngles[3]);

LAMBDA_Film *lambdaCreateFilm(LAMBDA_Device *_device, int _width, int _height);
void lambdaReleaseFilm(LAMBDA_Film *_film);

LAMBDA_Camera *lambdaCreateCamera(LAMBDA_Device *_device, LAMBDA_CameraType _cameraType, float _pos[3], float _phi,
                                  float _theta);
void lambdaReleaseCamera(LAMBDA_Camera *_camera);
void lambdaSetCameraPosition(LAMBDA_Camera *_camera, float _pos[3]);
void lambdaSetCameraAngles(LAMBDA_Camera *_camera, float _phi, float _theta);
void lambdaSetCameraAxes(LAMBDA_Camera *_camera, float _xHat[3], float _yHat[3], float _zHat[3]);
void lambdaSetCameraFocalLength(LAMBDA_Camera *_camera, float _focalLength);
void lambdaSetCameraApertureSize(LAMBDA_Camera *_camera, float _size);

/* Defaults: 1 spp, 16x16 tiles, automatic threads, path integrator, light tree. Freed with the
 * device. */
LAMBDA_RenderProperties *lambdaCreateRenderProperties(void);
LAMBDA_RenderDirective *lambdaCreateRenderDirective(LAMBDA_Device *_device, LAMBDA_Scene *_scene, LAMBDA_Film *_film,
                                                    LAMBDA_Camera *_camera, LAMBDA_RenderProperties *_properties);
void lambdaReleaseRenderDirective(LAMBDA_RenderDirective *_directive);
void lambdaSetCamera(LAMBDA_RenderDirective *_directive, LAMBDA_Camera *_camera);
void lambdaSetScene(LAMBDA_RenderDirective *_directive, LAMBDA_Scene *_scene);

LAMBDA_ProgressiveRenderer *lambdaCreateProgressiveRenderer(LAMBDA_Device *_device, LAMBDA_RenderDirective *_directive);
void lambdaReleaseProgressiveRenderer(LAMBDA_ProgressiveRenderer *_renderer);
void lambdaSetProgressiveRendererCallback(LAMBDA_ProgressiveRenderer *_renderer, void (*_callback)());
void *lambdaGetProgressiveRendererData(LAMBDA_ProgressiveRenderer *_renderer, int *_width, int *_height);
void lambdaStartProgressiveRenderer(LAMBDA_ProgressiveRenderer *_renderer);
void lambdaStopProgressiveRenderer(LAMBDA_ProgressiveRenderer *_renderer);

#ifdef __cplusplus
}
#endif
#endif

"""ctypes mirror of include/lambda_gpu.h (structs and enums).

Shared by the product binding (lambda_amd.gpu) and by the test-side oracle
binding, so both are driven from the same scene description.
"""
import ctypes as C

ABI_VERSION = 8  # LGPU_ABI_VERSION of include/lambda_gpu.h (struct layouts below)
NONE = 0xFFFFFFFF

# lgpu_status
OK, ERR_INVALID_ARGUMENT, ERR_HIP, ERR_STATE, ERR_OUT_OF_MEMORY, ERR_UNSUPPORTED, ERR_GRAPH, ERR_NO_DEVICE = range(8)

# node types
NODE_SCALAR, NODE_VEC2, NODE_VEC3, NODE_RGB, NODE_BLACKBODY = 1, 2, 3, 4, 5
NODE_IMAGE_COLOUR, NODE_IMAGE_SCALAR, NODE_UV, NODE_FRESNEL = 6, 7, 8, 9
NODE_SPECTRAL, NODE_IMAGE_CHANNEL, NODE_SPECTRAL_TEXTURE, NODE_NORMAL = 10, 11, 12, 13
NODE_SCALAR_MATHS, NODE_VECTOR_MATHS, NODE_DOT, NODE_LENGTH, NODE_MERGE_RGB, NODE_SEPARATE_R = 16, 17, 18, 19, 20, 21
NODE_CROSS, NODE_SEPARATE_XYZ, NODE_MERGE_XYZ, NODE_SCALAR_TO_COLOUR = 22, 23, 24, 25
NODE_CHECKER, NODE_PERLIN, NODE_VALUE_NOISE, NODE_OCTAVE_NOISE, NODE_VORONOI = 32, 33, 34, 35, 36
# value types reported by the oracle's node evaluation
VT_SCALAR, VT_VEC2, VT_VEC3, VT_COLOUR, VT_SPECTRUM = 0, 1, 2, 3, 4
NODE_LAMBERTIAN, NODE_OREN_NAYAR_BRDF, NODE_OREN_NAYAR_BTDF, NODE_MICROFACET_BRDF = 64, 65, 66, 67
NODE_FRESNEL_BSDF, NODE_SPECULAR_BRDF, NODE_SPECULAR_BTDF, NODE_GHOST_BTDF = 68, 69, 70, 71
NODE_MIX_BSDF, NODE_ROUGH_DIELECTRIC = 72, 73

OP_ADD, OP_SUBTRACT, OP_MULTIPLY, OP_DIVIDE = 0, 1, 2, 3

LIGHT_POINT, LIGHT_SPOT, LIGHT_ENVIRONMENT = 1, 2, 3
MEDIUM_HOMOGENEOUS, MEDIUM_GRID = 1, 2
LIGHTS_POWER, LIGHTS_TREE = 0, 1
CAMERA_PINHOLE, CAMERA_THIN_LENS, CAMERA_SPHERICAL = 0, 1, 2
APERTURE_NONE, APERTURE_CIRCULAR, APERTURE_BLADE = 0, 1, 2
INTEGRATOR_PATH, INTEGRATOR_DIRECT, INTEGRATOR_VOLPATH, INTEGRATOR_MIS_VOLPATH = 0, 1, 2, 3
INTEGRATOR_DEPTH, INTEGRATOR_NORMAL, INTEGRATOR_AOV = 4, 5, 6
SPECTRAL_RGB, SPECTRAL_HERO4 = 0, 1
ESTIMATOR_FAITHFUL, ESTIMATOR_CORRECTED = 0, 1

F3 = C.c_float * 3
FP = C.POINTER(C.c_float)
UP = C.POINTER(C.c_uint32)


class Node(C.Structure):
    _fields_ = [("type", C.c_uint32), ("in_", C.c_uint32 * 4), ("p", C.c_float * 12)]


class MaterialDesc(C.Structure):
    _fields_ = [("num_nodes", C.c_uint32), ("nodes", C.POINTER(Node)), ("bxdf_root", C.c_uint32),
                ("emission_root", C.c_uint32), ("emission_intensity", C.c_float),
                ("medium_interior", C.c_uint32), ("medium_exterior", C.c_uint32)]


class MeshDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32), ("positions", FP),
                ("indices", UP), ("normals", FP), ("tangents", FP), ("uvs", FP), ("material", C.c_uint32)]


class LightDesc(C.Structure):
    _fields_ = [("type", C.c_uint32), ("position", F3), ("axis", F3), ("cone_angle", C.c_float),
                ("falloff_start", C.c_float), ("intensity", C.c_float), ("rgb", F3), ("temperature", C.c_float),
                ("env_width", C.c_uint32), ("env_height", C.c_uint32), ("env_rgb", FP),
                ("env_offset", C.c_float * 2)]


class MediumDesc(C.Structure):
    _fields_ = [("type", C.c_uint32), ("sigma_a", F3), ("sigma_s", F3), ("g", C.c_float),
                ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("density", FP),
                ("bounds_min", F3), ("bounds_max", F3)]


class TextureDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("channels", C.c_uint32), ("data", FP)]


class Camera(C.Structure):
    _fields_ = [("type", C.c_uint32), ("origin", F3), ("x_hat", F3), ("y_hat", F3), ("z_hat", F3),
                ("tan_fov", C.c_float), ("aspect", C.c_float), ("focal_length", C.c_float),
                ("aperture", C.c_uint32), ("aperture_size", C.c_float), ("blades", C.c_uint32),
                ("offset_phi", C.c_float), ("offset_theta", C.c_float)]


class RenderDesc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("film", C.c_void_p), ("camera", Camera), ("integrator", C.c_uint32),
                ("spp_begin", C.c_uint32), ("spp_end", C.c_uint32), ("max_bounces", C.c_uint32),
                ("min_bounces", C.c_uint32), ("spectral_mode", C.c_uint32), ("estimator_mode", C.c_uint32),
                ("seed", C.c_uint32), ("path_capacity", C.c_uint32), ("flags", C.c_uint32),
                ("pass_max_depth", C.c_float), ("pass_power", C.c_float), ("pass_background", C.c_float * 3),
                ("pass_geometric_normal", C.c_uint32), ("guide", C.c_void_p), ("guide_train_steps", C.c_uint32)]


class GuideDesc(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("record_capacity", C.c_uint32), ("batch", C.c_uint32),
                ("learning_rate", C.c_float), ("max_depth", C.c_uint32)]

RENDER_COUNT_TRAVERSAL = 1

ADJOINT_DEVICE_POINTERS = 1

BATCH_AUTOTUNE = 1


class BatchOptions(C.Structure):
    _fields_ = [("path_capacity", C.c_uint32), ("flags", C.c_uint32), ("max_iterations", C.c_uint32),
                ("branches", C.c_uint32)]


class BatchStats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("ms_total", C.c_double), ("ms_render", C.c_double),
                ("ms_capture", C.c_double), ("ms_autotune", C.c_double), ("chunks", C.c_uint32),
                ("graph_nodes", C.c_uint32), ("path_capacity", C.c_uint32), ("branches", C.c_uint32),
                ("iterations", C.c_uint32), ("jobs", C.c_uint32)]


class AdjointDesc(C.Structure):
    _fields_ = [("d_image", C.c_void_p), ("grad_textures", C.c_void_p), ("grad_density", C.c_void_p),
                ("grad_emission", C.c_void_p), ("flags", C.c_uint32), ("reserved", C.c_uint32)]


TONEMAP_NONE, TONEMAP_ACES_FITTED = 0, 1


class PostDesc(C.Structure):
    _fields_ = [("tonemap", C.c_uint32), ("clamp", C.c_uint32), ("gamma", C.c_uint32), ("estimator_mode", C.c_uint32),
                ("kernel", FP), ("kernel_w", C.c_uint32), ("kernel_h", C.c_uint32), ("kernel_intensity", C.c_float)]


class DenoiseDesc(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("sigma_colour", C.c_float), ("sigma_normal", C.c_float),
                ("sigma_albedo", C.c_float), ("demodulate", C.c_uint32)]


def denoise_desc(iterations=5, sigma_colour=1.0, sigma_normal=0.1, sigma_albedo=0.1, demodulate=True):
    """lgpu_denoise_desc (edge-avoiding A-Trous filter parameters)."""
    d = DenoiseDesc()
    d.iterations, d.sigma_colour, d.sigma_normal, d.sigma_albedo = iterations, sigma_colour, sigma_normal, sigma_albedo
    d.demodulate = int(demodulate)
    return d


def post_desc(tonemap=TONEMAP_ACES_FITTED, clamp=True, gamma=True, estimator=ESTIMATOR_CORRECTED, kernel=None,
              intensity=1.0):
    """lgpu_post_desc; returns (desc, keep-alive array)."""
    import numpy as np
    d = PostDesc()
    d.tonemap, d.clamp, d.gamma, d.estimator_mode = tonemap, int(clamp), int(gamma), estimator
    k = None
    if kernel is not None:
        k = np.ascontiguousarray(kernel, dtype=np.float32)
        d.kernel = k.ctypes.data_as(FP)
        d.kernel_h, d.kernel_w = k.shape
        d.kernel_intensity = intensity
    return d, k


class ObjMaterial(C.Structure):
    _fields_ = [("name", C.c_char_p), ("kd", C.c_float * 3), ("ke", C.c_float * 3), ("has_kd", C.c_uint32),
                ("has_ke", C.c_uint32)]


class RenderStats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_camera", C.c_double), ("ms_extend", C.c_double),
                ("ms_shade", C.c_double), ("ms_shadow", C.c_double), ("samples", C.c_uint64),
                ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("nodes_visited", C.c_uint64),
                ("tris_tested", C.c_uint64), ("waves", C.c_uint32), ("launches_extend", C.c_uint32),
                ("launches_shade", C.c_uint32), ("launches_shadow", C.c_uint32), ("voxel_steps", C.c_uint64),
                ("nodes4_fetched", C.c_uint64), ("tris4_fetched", C.c_uint64), ("ms_guide_infer", C.c_double),
                ("ms_guide_train", C.c_double), ("guide_records", C.c_uint64)]

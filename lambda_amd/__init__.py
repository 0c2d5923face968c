"""lambda_amd: MI355X-native wavefront spectral path tracer with the render
surface of the Lambda renderer (scene, materials as node networks, camera,
film, path integrator), driven through the C ABI in include/lambda_gpu.h."""
from . import abi, gpu, scene  # noqa: F401

__all__ = ["abi", "gpu", "scene"]

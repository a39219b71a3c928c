"""xdg_amd: MI355X-native geometry-query backend for XDG-style geometry.

Submodules:
  philox    counter-based RNG and sampling protocols (host side)
  geometry  synthetic surface models and tetrahedral meshes
  capi      ctypes binding of the C-ABI library libxdg_hip.so (HIP kernels)
  raytracer batched Python front-end over the C-ABI (HipRayTracer)
  autograd  torch.autograd.Function around the ray-distance adjoint
"""
__all__ = ["philox", "geometry"]

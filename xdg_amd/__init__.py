"""xdg_amd: MI355X-native geometry-query backend for XDG-style geometry.

Submodules:
  philox    counter-based RNG and sampling protocols (host side)
  geometry  synthetic surface models and tetrahedral meshes
  capi      ctypes binding of the C-ABI library libxdg_hip.so (HIP kernels)
  adjoint   torch.autograd.Function around the ray-distance adjoint
  overlap   the overlap check's map built from the batched GPU queries
  h5m       DAGMC .h5m geometry read through libxdg_h5m.so (HDF5 C API)
"""
__all__ = ["philox", "geometry"]

"""MI355X-native real-time Kohn-Sham propagation hot path of SALMON.

Modules:
  system  synthetic periodic systems in SALMON's memory layout
  hip     ctypes binding of the C ABI (include/salmon_hip.h)
  rt      real-time propagation driver on the GPU
  build   in-tree build of libsalmon_hip.so for gfx950
"""
__all__ = ["system", "hip", "rt", "build"]

"""MI355X-native SIMPLE pressure-velocity hot path (gfx950 HIP kernels behind a C ABI).

``chuanflow_amd.native`` binds include/cfx.h and include/cfx_host.h; importing
it requires the built library (``python -m chuanflow_amd.build``).
"""

__all__ = ["native"]

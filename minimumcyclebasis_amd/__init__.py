"""MI355X-native de Pina support-vector phase loop for exact Minimum Cycle Basis.

Layers:
  * ``host``       - C++ preprocessing (``libmcb_host.so``): graph input, degree-two contraction,
                     FVS, spanning tree, per-root shortest-path trees, ranked candidates.
  * ``device``     - gfx950 HIP kernels (``libmcb_hip.so``): tree-path parity, min-rank odd
                     candidate search, cycle vectors and the GF(2) support-vector update, plus
                     the reference-compatible ``gpu_struct`` chunked parity interface.
  * ``generators`` - seeded synthetic graphs of the measured configurations.
"""
from .host import Problem  # noqa: F401

__all__ = ["Problem"]

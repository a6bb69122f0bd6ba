"""tmcore_amd: MI355X-native TRiSK shallow-water hot path.

Host-side Python mirror of the C-ABI in include/: mesh generation, fit weights,
initial states (libtmhost.so) and the HIP device path (libtmgpu.so).
"""
from .mesh import Mesh, Fit, seeded_state, steady_geostrophic_state, RADIUS, G, OMEGA  # noqa: F401

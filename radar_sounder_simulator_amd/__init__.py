"""MI355X-native coherent facet kernel of soundersim (radar sounder clutter).

Drop-in for the reference's hot path: the same Python call surface
(``coherent_cluttergram``, ``refracted_cluttergram``, ``apply_waveform``,
``simulate``) over hand-written gfx950 HIP kernels reached through a C ABI
(``include/soundersim_hip.h``) with torch-ROCm tensors as HBM buffers.
"""

from .kernels import coherent_cluttergram, refracted_cluttergram  # noqa: F401
from .physics import fresnel_normal  # noqa: F401
from .scene import Facets, NavTrack, gaussian_surface, grid_facets  # noqa: F401
from .waveform import apply_waveform, compressed_pulse  # noqa: F401

__all__ = ["coherent_cluttergram", "refracted_cluttergram", "apply_waveform",
           "compressed_pulse", "fresnel_normal", "Facets", "NavTrack",
           "gaussian_surface", "grid_facets"]

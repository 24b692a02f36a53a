"""pyspectral_amd: an MI355X (gfx950) backend for pyspectral's IF-RK4 pseudo-spectral step.

The package mirrors the reference's solver interface (``Config``, ``SpectralSolver``, the
``Variant`` registry, diagnostics and the ``run`` driver) and implements the ``hip/*`` variants
on hand-written HIP kernels reached through the C ABI in ``include/psx.h`` (``libpsx.so``).
"""

__version__ = "0.1.0"

"""MI355X-native VolumetricReSTIR render pass: HIP kernels for gfx950 behind a C ABI
(include/vrestir.h), with a thin Python mirror of the reference pass interface."""
from . import abi  # noqa: F401

__all__ = ["abi"]

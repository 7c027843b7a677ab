"""MI355X-native staggered-grid pseudo-acoustic VTI FD path (HIP, gfx950).

The compute lives in libvti.so (include/vti.h); ``geophysics_amd.vti`` is
its Python binding.  There is no CPU fallback: without the built library or
a GPU the calls raise.
"""
from . import vti  # noqa: F401

__all__ = ["vti"]

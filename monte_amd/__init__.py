"""monte_amd: MI355X (gfx950) engine for cone-beam CT Monte Carlo projection and
FDK reconstruction, a drop-in for the two hot paths of Tomato27/Monte.

The compute lives in libmonte.so (hand-written HIP kernels behind the C ABI in
include/monte.h); this package is its Python face over torch device tensors.
Importing it without the built library raises: there is no CPU path.
"""
# torch first: its wheel ships its own libamdhip64.so.7, and libmonte.so must
# bind to that same HIP runtime (same soname, so the loader reuses whichever
# was loaded first).  Loading libmonte.so first would pull in a second HIP
# runtime beside torch's, and streams and pointers would not carry across.
import torch  # noqa: F401,E402

from . import _lib
from ._lib import MonteError

_lib.lib()  # fail loudly here if libmonte.so is missing

from .context import Context, default_context  # noqa: E402
from .geometry import Geometry  # noqa: E402
from . import geometry, phantoms, tables  # noqa: E402

__all__ = ["Context", "default_context", "Geometry", "MonteError", "geometry", "phantoms", "tables"]
__version__ = "0.1.0"

"""pictor_amd -- MI355X (gfx950) particle step of the PICTOR particle-in-cell code.

Boris push, charge-conserving zigzag deposition, Yee FDTD and periodic halos
as hand-written HIP kernels behind a C ABI (include/pictor_hip.h); this
package holds the Python host that drives them through ctypes.
"""
from ._lib import EXPORTS, LIBPATH, PictorError, lib  # noqa: F401
from .simulation import DIAG_KEYS, Simulation, setup_weibel  # noqa: F401

"""MI355X (gfx950) event-generation inner loop of GRANIITTI's <C> continuum class.

The compute path is the HIP library libgra_hip.so behind the C ABI of
include/gra_hip.h; this package holds its ctypes mirror (abi), the build
script (build) and the host-side process / integrator classes shaped like the
reference's MContinuum / MGraniitti (process).
"""
from . import abi  # noqa: F401

"""hipfdtd — MI355X-native FDTD engine behind the reference solver's Python API.

``HIPFDTD`` is a drop-in for ``OpenCLFDTD`` and ``HIPNear2FarMonitor`` for
``OpenCLNear2FarMonitor``; the hot path (Yee + CPML updates, sources, face
DFT, near-to-far transform) runs in hand-written HIP kernels
(``libhipfdtd.so``, C ABI in ``include/hipfdtd.h``).
"""

__version__ = "1.0.0"

from .constants import C0, EPS0, ETA0, MU0
from .engine import HIPFDTD
from .monitors import HIPNear2FarMonitor
from .plugin import StepCallback

# Reference spellings, so code written against opencl_fdtd_solver can switch
# by changing only the import.
OpenCLFDTD = HIPFDTD
OpenCLNear2FarMonitor = HIPNear2FarMonitor

__all__ = [
    "C0",
    "MU0",
    "EPS0",
    "ETA0",
    "HIPFDTD",
    "HIPNear2FarMonitor",
    "OpenCLFDTD",
    "OpenCLNear2FarMonitor",
    "StepCallback",
    "__version__",
]

"""fd_adi_amd -- MI355X-native ADI finite-difference engine for Heston options.

Drop-in for the GPU path of johntyree/fd_adi: ``FiniteDifferenceEngineGPU``
(scheme API), ``BandedOperatorGPU`` (operator algebra) and ``SizedArrayPtr``
(device arrays), backed by hand-written FP64 HIP kernels for gfx950 in
``libfdadi.so`` (C ABI: include/fdadi.h).  ``batch.HestonBatch`` steps many
options with uneven grids in one launch per ADI stage.
"""

from . import _lib  # noqa: F401  (raises ImportError if libfdadi.so is missing)
from .SizedArrayPtr import SizedArrayPtr  # noqa: F401
from . import BandedOperatorGPU  # noqa: F401
from . import FiniteDifferenceEngineGPU  # noqa: F401
from .batch import HestonBatch  # noqa: F401
from .Option import HestonOption, HestonBarrierOption, Option, BarrierOption  # noqa: F401
from .Grid import Grid  # noqa: F401

__version__ = "1.0.0"

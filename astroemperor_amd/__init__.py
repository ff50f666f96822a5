"""MI355X-native (gfx950) Keplerian radial-velocity likelihood and
parallel-tempering engine with EMPEROR's (ReddTea/astroemperor) interface.

Compute runs in libemp_mi355x.so (hand-written HIP); this package is the host
side: data staging, model layout, the reddemcee-style sampler and the
``Simulation`` facade.
"""
__version__ = "0.1.0"

from ._abi import EmpError, device_info  # noqa: F401
from .blocks import (AccelerationBlock, JitterBlock, KeplerianBlock, MagneticCycleBlock,  # noqa: F401
                     MOAVBlock, OffsetBlock, SAIBlock, SinusoidBlock, SmartSetter)
from .data import DataWrapper, RVData  # noqa: F401
from .model import ReddModel  # noqa: F401
from .backend import SavedRun  # noqa: F401
from .moves import GradientMove, StretchMove  # noqa: F401
from .sampler import PTSampler  # noqa: F401
from .simulation import Simulation  # noqa: F401
from .systems import SystemsPT  # noqa: F401

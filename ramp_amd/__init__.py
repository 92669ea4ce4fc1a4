"""ramp_amd: MI355X-native implementation of RAMP's neutron ray-tracing path.

Drop-in surface of the reference package (mcramp/__init__.py): the
Instrument class, frame_rotate and every component class by its reference
name.  Execution runs on the GPU through libramp.so (hand-written HIP
kernels behind a C-ABI); there is no CPU execution path.
"""
from .components import *  # noqa: F401,F403
from .components import SOURCES, GEOMS, SCATS  # noqa: F401
from .instrument import Context, Instrument, ExecutionBlock, frame_rotate  # noqa: F401
from .batch import BatchInstrument  # noqa: F401

__version__ = "0.1.0"

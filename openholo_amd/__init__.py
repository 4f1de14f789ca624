"""openholo_amd: MI355X-native point-cloud hologram generation.

The product is the HIP library ``libopenholo_amd.so`` (C-ABI declared in
``include/openholo_amd.h``) and the C++ drop-in ``ophPointCloud`` class built
on it.  This Python package is a thin ctypes front end used by the tests and
the benchmark; it never computes a hologram itself.
"""
from . import synth  # noqa: F401

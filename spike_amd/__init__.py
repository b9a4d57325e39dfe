"""Spike on MI355X: a gfx950-native backend for the Spike spiking-network simulator.

The product is the C ABI in include/spike_hip.h (libspike_hip.so, hand-written HIP
kernels) and the API-compatible Spike C++ front-end built on it.  This Python package
holds the build recipe, a ctypes binding of the ABI and synthetic network builders used
by the tests and the benchmark.
"""
from .netdesc import NetDesc, Rule  # noqa: F401

__all__ = ["NetDesc", "Rule"]

"""MI355X-native implementation of SolTrace's coretrace trace loop.

The product is the C-ABI library ``coretrace_api.so`` (built from
``soltrace_amd/csrc``) exporting the reference's stapi.h; this package holds
its ctypes binding, a scene description that replays into any stapi library,
and the synthetic workloads used by the benchmarks and tests.
"""
from . import stapi, scene, workloads  # noqa: F401

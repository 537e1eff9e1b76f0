"""empanada-mi355x: MI355X-native fingerprint -> partition -> suffix/prefix overlap path.

The product is the C ABI library libempanada_hip.so (include/empanada.h) and the `empanada`
command line built on it; this package holds their sources (csrc/), the ctypes bindings used by
tests and benchmarks (native.py) and the synthetic workload generator (gen.py)."""
from . import native  # noqa: F401

__all__ = ["native"]

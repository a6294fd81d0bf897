"""MI355X-native spectral transport for gemstone rendering (the film and metric
hot path of the indicatrix tracer), behind the C ABI in include/ixg.h."""
from . import abi  # noqa: F401

__all__ = ["abi"]

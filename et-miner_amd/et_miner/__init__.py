"""et_miner on MI355X: exact frequent-itemset mining with the counting path on hand-written gfx950 HIP kernels.

The public surface mirrors the reference's ``et_miner`` package for the GPU counting path:
``apriori(..., use_gpu=True, bitvecs=, streaming=, chunk_size=, use_generator_pruning=)`` and
``apriori_streaming``. Kernels live in ``libetm.so`` (C ABI in ``include/etm.h``), loaded with ctypes.
"""

from .core.apriori import apriori
from .core.profiling import ProfilingSession
from .streaming.son import apriori_streaming

__all__ = ["apriori", "apriori_streaming", "ProfilingSession", "gpu_available"]
__version__ = "0.1.0"


def gpu_available() -> bool:
    """True when libetm.so loads and a HIP device is visible."""
    try:
        from ._lib import device_count

        return device_count() > 0
    except Exception:
        return False

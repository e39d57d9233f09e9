"""boitata_amd — an MI355X-native kriging engine for boitatá's estimation hot path.

The product is the HIP library behind include/boitata_gpu.h (built in-tree by
``python -m boitata_amd.build``); this package holds its Python boundary:
``Engine`` (the C ABI) and the drop-in estimator classes that mirror
python/boitata/estimation.py of the reference.
"""
from ._lib import EngineUnavailable  # noqa: F401

__version__ = "0.1.0"

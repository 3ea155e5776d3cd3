"""mlxlearn_amd -- exact brute-force neighbour search on AMD MI355X (gfx950).

Drop-in surface of the reference's neighbour path: ``NearestNeighbors``
(``kneighbors`` / ``radius_neighbors``), ``KNeighborsClassifier``,
``KNeighborsRegressor``, ``RadiusNeighborsClassifier``,
``RadiusNeighborsRegressor`` and ``KMeans`` nearest-centre assignment, backed by
hand-written HIP kernels behind a C ABI (``include/mlxlearn_amd.h``).
"""
from ._config import Config, config_context, get_config, reset_config, set_config
from ._diagnostics import (BackendEvent, clear_backend_diagnostics, diagnostics_summary, get_backend_diagnostics,
                           get_last_backend_event)
from ._index import DeviceNeighborIndex, NeighborIndex, RadiusCSR, build_index, gpu_available
from .cluster import KMeans
from .neighbors import (KNeighborsClassifier, KNeighborsRegressor, NearestNeighbors, RadiusNeighborsClassifier,
                        RadiusNeighborsRegressor)

__all__ = ["DeviceNeighborIndex", "NeighborIndex", "RadiusCSR", "build_index", "gpu_available", "KMeans",
           "KNeighborsClassifier", "KNeighborsRegressor", "NearestNeighbors", "RadiusNeighborsClassifier",
           "RadiusNeighborsRegressor", "Config", "config_context", "get_config", "reset_config", "set_config",
           "BackendEvent", "clear_backend_diagnostics", "diagnostics_summary", "get_backend_diagnostics",
           "get_last_backend_event"]
__version__ = "0.1.0"

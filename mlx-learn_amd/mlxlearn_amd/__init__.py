"""mlxlearn_amd -- exact brute-force neighbour search on AMD MI355X (gfx950).

Drop-in surface of the reference's neighbour path: ``NearestNeighbors``
(``kneighbors`` / ``radius_neighbors``), ``KNeighborsClassifier``,
``KNeighborsRegressor``, ``RadiusNeighborsClassifier``,
``RadiusNeighborsRegressor`` and ``KMeans`` nearest-centre assignment, backed by
hand-written HIP kernels behind a C ABI (``include/mlxlearn_amd.h``).
"""
from ._index import DeviceNeighborIndex, NeighborIndex, RadiusCSR, build_index, gpu_available
from .cluster import KMeans
from .neighbors import (KNeighborsClassifier, KNeighborsRegressor, NearestNeighbors, RadiusNeighborsClassifier,
                        RadiusNeighborsRegressor)

__all__ = ["DeviceNeighborIndex", "NeighborIndex", "RadiusCSR", "build_index", "gpu_available", "KMeans",
           "KNeighborsClassifier", "KNeighborsRegressor", "NearestNeighbors", "RadiusNeighborsClassifier",
           "RadiusNeighborsRegressor"]
__version__ = "0.1.0"

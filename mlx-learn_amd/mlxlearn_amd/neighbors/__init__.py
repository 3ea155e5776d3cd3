"""Exact Euclidean neighbour estimators (reference ``src/mlxlearn/neighbors/__init__.py:9-22``)."""
from ._base import compute_weights
from ._estimators import (KNeighborsClassifier, KNeighborsRegressor, NearestNeighbors, RadiusNeighborsClassifier,
                          RadiusNeighborsRegressor)

__all__ = ["KNeighborsClassifier", "KNeighborsRegressor", "NearestNeighbors", "RadiusNeighborsClassifier",
           "RadiusNeighborsRegressor", "compute_weights"]

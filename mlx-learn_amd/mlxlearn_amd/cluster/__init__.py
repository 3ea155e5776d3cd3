"""k-means with exact GPU nearest-centre assignment."""
from ._k_means import KMeans, assign_device, update_device

__all__ = ["KMeans", "assign_device", "update_device"]

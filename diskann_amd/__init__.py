"""diskann_amd: MI355X-native (gfx950) search-and-quantize path for DiskANN.

The compute path is the HIP library libdiskann_mi355x.so behind the C ABI in
include/diskann_mi355x.h; this package mirrors the reference's search, PQ and distance
interfaces on top of it.
"""
from .errors import (ANNError, CapacityError, DimensionMismatchError, KnnSearchError, PQError,
                     StorageError)
from .index import (DiskANNIndex, FullPrecision, Hybrid, Knn, Metric, Quantized, SearchResult,
                    SearchGraph, Server, Strategy)
from . import minmax, prescore, quantization, scalar, spherical, storage

__all__ = [
    "ANNError", "CapacityError", "DimensionMismatchError", "KnnSearchError", "PQError",
    "StorageError", "storage", "prescore",
    "DiskANNIndex", "FullPrecision", "Hybrid", "Knn", "Metric", "Quantized", "SearchResult",
    "SearchGraph", "Server", "Strategy", "quantization", "scalar", "spherical", "minmax",
]

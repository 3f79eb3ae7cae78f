"""Status codes of the C ABI mapped to exceptions that mirror the reference's
ANNErrorKind / KnnSearchError (diskann/src/error/ann_error.rs:770-830,
diskann/src/graph/search/knn_search.rs)."""

OK = 0
INVALID_ARG = 1
K_ZERO = 2
L_ZERO = 3
L_LESS_THAN_K = 4
BEAM_WIDTH_ZERO = 5
DIM_MISMATCH = 6
OUT_OF_MEMORY = 7
DEVICE = 8
PQ = 9
KMEANS = 10
INDEX = 11
CAPACITY = 12
NOT_READY = 13
IO = 14


class ANNError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("[%d] %s" % (status, message))
        self.status = status


class KnnSearchError(ANNError, ValueError):
    """KZero / LZero / LLessThanK / BeamWidthZero."""


class DimensionMismatchError(ANNError, ValueError):
    pass


class PQError(ANNError):
    pass


class CapacityError(ANNError):
    pass


class StorageError(ANNError, OSError):
    """ANNErrorKind::IOError: a file is missing, short or malformed."""


_KIND = {
    K_ZERO: KnnSearchError, L_ZERO: KnnSearchError, L_LESS_THAN_K: KnnSearchError,
    BEAM_WIDTH_ZERO: KnnSearchError, DIM_MISMATCH: DimensionMismatchError, PQ: PQError,
    KMEANS: PQError, CAPACITY: CapacityError, IO: StorageError,
}


def check(status, lib):
    if status != OK:
        msg = lib.dkm_last_error().decode("utf-8", "replace")
        raise _KIND.get(status, ANNError)(status, msg)

"""Constants of the reference package (minicore/util/constants.py:1-7)."""
from collections import namedtuple

# sensitivity methods of CoresetSampler.make_sampler (coreset.h:31-46)
SENSDICT = {"BFL": 0, "FL": 1, "LFKF": 2, "VX": 3, "LBK": 4}
CSR = namedtuple("CSR", ["data", "indices", "indptr", "shape", "nnz"])

KMCRSV = 4294967295  # -1u, asking for reservoir kmeans++ sampling.

__all__ = ["SENSDICT", "CSR", "KMCRSV"]

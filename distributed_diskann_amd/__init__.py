"""MI355X-native Vamana/DiskANN index build and beam search.

The compute path is libddg.so (HIP kernels for gfx950, C-ABI in
include/ddg.h); this package is its Python host side."""
from ._lib import DdgError, lib
from .index import (BuildParams, Graph, Index, PointRange, QueryParams, beam_search, checkRecall, knn_index,
                    nn_result, recall_with_ties, searchAll, stats)

__all__ = ["DdgError", "lib", "BuildParams", "Graph", "Index", "PointRange", "QueryParams", "beam_search", "checkRecall",
           "knn_index", "nn_result", "recall_with_ties", "searchAll", "stats"]

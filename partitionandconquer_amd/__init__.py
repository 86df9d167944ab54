"""partitionandconquer_amd -- MI355X (gfx950) partitioned neighbour-set intersection:
triangle counting, local clustering coefficient and Jaccard similarity.

The compute path is libpac.so (HIP kernels behind the C ABI in include/pac.h); this
package is its Python binding.  See pac.py.
"""
from .pac import (  # noqa: F401
    DISPATCH_FIXED, DISPATCH_LEARNED, FORCE_BITMAP, FORCE_GALLOP, FORCE_MERGE,
    LAUNCH_GRAPH, LAUNCH_PERSISTENT, LAUNCH_PLAIN, MLP_NPARAMS, MODE_JACCARD, MODE_LCC, MODE_TC, NUM_TILES, STRATEGIES,
    PACKED_ALWAYS, PACKED_AUTO, PACKED_OFF,
    Context, Graph, PacError, PartitionedGraph, contiguous_partition, lib,
    random_partition,
)

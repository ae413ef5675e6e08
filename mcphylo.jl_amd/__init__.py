"""MI355X-native pruning likelihood and branch-length gradient with the
PhyloDist API of MCPhylo.jl.

The package is laid out in directory `mcphylo.jl_amd/`, which is not a valid
Python identifier; load it under a module name of your choice, e.g.

    import importlib.util, sys
    spec = importlib.util.spec_from_file_location(
        "mcphylo_amd", "mcphylo.jl_amd/__init__.py",
        submodule_search_locations=["mcphylo.jl_amd"])
    mcp = importlib.util.module_from_spec(spec)
    sys.modules["mcphylo_amd"] = mcp
    spec.loader.exec_module(mcp)

All likelihood work runs in libmcphylo_hip.so (HIP, gfx950); there is no CPU
fallback.
"""

from ._lib import McpArgumentError, McpDimensionMismatch, McpError, build, load_library
from .engine import Batch, Model, Plan
from .models import GTR, JC, ComplexEigensystemError, Restriction, freeK, setmatrix
from .nexus import ParseCSV, ParseNexus, compress_patterns, create_csvdf, datafortree
from .phylodist import (MultiplePhyloDist, PhyloDist, __logpdf, clear_plan_cache, gradlogpdf,
                        gradlogpdf_params, logpdf, maximum, minimum, size)
from .rates import discrete_gamma_rates, mean_boundaries, median_boundaries
from . import priors, sampler, surrogate, synthetic
from .surrogate import Surrogate
from .tree import (Node, ParseNewick, TreeArrays, get_branchlength_vector, get_leaves, get_mother,
                   number_nodes, post_order, pre_order, set_branchlength_vector, to_newick,
                   tree_arrays, tree_from_parent)

__all__ = [
    "build", "load_library", "McpError", "McpArgumentError", "McpDimensionMismatch",
    "Plan", "Batch", "Model", "Surrogate",
    "JC", "GTR", "Restriction", "freeK", "setmatrix", "ComplexEigensystemError",
    "ParseNexus", "ParseCSV", "create_csvdf", "datafortree", "compress_patterns",
    "PhyloDist", "MultiplePhyloDist", "logpdf", "gradlogpdf", "gradlogpdf_params", "__logpdf",
    "clear_plan_cache",
    "size", "minimum", "maximum",
    "discrete_gamma_rates", "mean_boundaries", "median_boundaries",
    "Node", "ParseNewick", "TreeArrays", "tree_arrays", "tree_from_parent", "post_order",
    "pre_order", "get_leaves", "get_mother", "number_nodes", "get_branchlength_vector",
    "set_branchlength_vector", "to_newick",
]

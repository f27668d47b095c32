"""sfamd: host-side mirror of StructureFunctions.jl's pair-histogram API over the gfx950 backend.

The package exposes the reference's operator types, bin specifications, metrics, result
containers and calculate_structure_function[!] entry points; the pair loop runs in
libsfmi355x.so (HIP, gfx950) through its C ABI.
"""
from .bins import (AbstractBinEdges, BinEdges, InfPaddedBinEdges, LinearBinEdges, LogBinEdges,
                   LogBinEdges_from_log_edges, as_edges, midpoints, n_histogram_bins)
from .dispatch import (MI355XBackend, calculate_structure_function, calculate_structure_function_batch,
                       calculate_structure_function_inplace, calculate_structure_functions_fused)
from .errors import ArgumentError, BackendError, DimensionMismatch
from .metrics import Euclidean, Haversine, SphericalAngle
from .objects import (StructureFunction, StructureFunction2DSumsAndCounts, StructureFunctionSumsAndCounts,
                      marginalize)
from .operators import *  # noqa: F401,F403
from .operators import get_structure_function_type, order
from .single_pass import (SINGLE_PASS_N, SINGLE_PASS_OPERATORS, calculate_structure_functions_single_pass,
                          calculate_structure_functions_single_pass_2d,
                          calculate_structure_functions_single_pass_2d_inplace,
                          calculate_structure_functions_single_pass_inplace)

calculate_structure_function_bang = calculate_structure_function_inplace

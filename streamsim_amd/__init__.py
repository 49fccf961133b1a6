"""MI355X-native facet radiosity and energy-balance engine.

A drop-in for the FacetRT / FacetEB device path of HiStream
(models/histream/src/facetrt, models/histream/src/faceteb): ray-cast form
factors, a multi-band two-sided Jacobi SpMM and a coupled FP64 energy-balance
Newton step, all as hand-written HIP kernels for gfx950 behind the C-ABI of
include/facet_hip.h.
"""
from .engine import FacetEngine, GraphDiagnostics, SolveResult  # noqa: F401

__all__ = ["FacetEngine", "GraphDiagnostics", "SolveResult"]

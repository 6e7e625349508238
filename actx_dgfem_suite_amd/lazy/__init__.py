"""Tracing array context: DG operators written against a grudge-shaped layer
are traced into an einsum DAG, normalised as the reference optimizer does,
matched to an equation and compiled to the HIP kernels."""

from actx_dgfem_suite_amd.lazy.context import HIPDGFEMLazyArrayContext, LazyCompiled
from actx_dgfem_suite_amd.lazy.match import MatchedOperator, Unmatched, match_dg_operator
from actx_dgfem_suite_amd.lazy.normalize import normalize

__all__ = ["HIPDGFEMLazyArrayContext", "LazyCompiled", "MatchedOperator", "Unmatched",
           "match_dg_operator", "normalize"]

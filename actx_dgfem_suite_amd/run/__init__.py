"""Benchmark runner: ``python -m actx_dgfem_suite_amd.run``."""

"""Host evaluation of a lazy DAG with numpy.

It serves as a checker only:
- the runner's value check evaluates the traced operator with it, the
  counterpart of the reference's check against its ``NumpyArrayContext``
  (run/__main__.py:77-96);
- the tests check the traced operators and the normalisation with it.

No compiled callable evaluates through it: ``HIPDGFEMLazyArrayContext``
runs the matched HIP plan and nothing else, and fails without a GPU.
"""

from __future__ import annotations

import numpy as np

from actx_dgfem_suite_amd.lazy import ir


def evaluate(roots, values: dict) -> list:
    """Values of ``roots``; ``values`` maps placeholder names to arrays."""
    memo: dict[int, np.ndarray] = {}
    for node in ir.walk(roots):
        if isinstance(node, ir.Placeholder):
            v = values[node.name]
        elif isinstance(node, ir.DataWrapper):
            v = node.data
        elif isinstance(node, ir.Constant):
            v = node.value
        elif isinstance(node, ir.Einsum):
            v = np.einsum(node.spec, *[memo[id(a)] for a in node.args])
        elif isinstance(node, ir.Take):
            v = np.asarray(memo[id(node.array)]).reshape(-1)[node.index]
        elif isinstance(node, ir.Stack):
            v = np.stack([memo[id(a)] for a in node.arrays])
        else:
            a = [memo[id(x)] for x in node.args]
            v = {
                "+": lambda: a[0] + a[1], "-": lambda: a[0] - a[1],
                "*": lambda: a[0] * a[1], "/": lambda: a[0] / a[1],
                "neg": lambda: -a[0], "sqrt": lambda: np.sqrt(a[0]),
                "abs": lambda: np.abs(a[0]), "maximum": lambda: np.maximum(a[0], a[1]),
                "minimum": lambda: np.minimum(a[0], a[1]),
                "where": lambda: np.where(np.asarray(a[0]) > 0.5, a[1], a[2]),
            }[node.op]()
        memo[id(node)] = v
    return [memo[id(r)] for r in roots]

"""Array-expression DAG that :class:`HIPDGFEMLazyArrayContext` traces.

The node kinds follow the pytato node types the reference optimizer works on
(arraycontext/dgfem_optimizer.py:181-245): ``Placeholder`` (a traced
argument), ``DataWrapper`` (a constant array), ``Einsum``, ``Elementwise`` (an
``IndexLambda`` of one scalar operation), ``Take`` (the folded gather
``a.reshape(-1)[index]`` of constants_folder.py:59-100) and ``Stack``.  A
pytato front-end would map its nodes one to one onto these.

Nodes are immutable and compared by identity; shapes are checked when a node
is made, so a malformed operator fails at trace time, not in a kernel.
"""

from __future__ import annotations

import re

import numpy as np

ELEMENTWISE_OPS = ("+", "-", "*", "/", "neg", "sqrt", "abs", "maximum", "minimum", "where")


class Array:
    """A lazy array of a fixed shape (float64 unless ``dtype`` says so)."""

    shape: tuple
    dtype = np.float64

    @property
    def ndim(self) -> int:
        return len(self.shape)

    # {{{ arithmetic builds Elementwise nodes

    def __add__(self, other):
        return elementwise("+", self, other)

    def __radd__(self, other):
        return elementwise("+", other, self)

    def __sub__(self, other):
        return elementwise("-", self, other)

    def __rsub__(self, other):
        return elementwise("-", other, self)

    def __mul__(self, other):
        return elementwise("*", self, other)

    def __rmul__(self, other):
        return elementwise("*", other, self)

    def __truediv__(self, other):
        return elementwise("/", self, other)

    def __rtruediv__(self, other):
        return elementwise("/", other, self)

    def __neg__(self):
        return elementwise("neg", self)

    def __pow__(self, k):
        if k == 2:
            return elementwise("*", self, self)
        raise NotImplementedError(f"power {k}")

    # }}}

    def __repr__(self):
        return f"{type(self).__name__}{self.shape}"


class Constant(Array):
    """A scalar constant (shape ())."""

    def __init__(self, value: float):
        self.value = float(value)
        self.shape = ()

    def __repr__(self):
        return f"Constant({self.value!r})"


class Placeholder(Array):
    def __init__(self, name: str, shape: tuple):
        self.name = name
        self.shape = tuple(shape)

    def __repr__(self):
        return f"Placeholder({self.name!r}, {self.shape})"


class DataWrapper(Array):
    def __init__(self, data: np.ndarray, name: str = ""):
        self.data = np.asarray(data)
        self.shape = self.data.shape
        self.dtype = self.data.dtype
        self.name = name

    def __repr__(self):
        return f"DataWrapper({self.name!r}, {self.shape})"


class Elementwise(Array):
    """One scalar operation applied entrywise; operands of shape () broadcast."""

    def __init__(self, op: str, args: tuple):
        if op not in ELEMENTWISE_OPS:
            raise NotImplementedError(f"elementwise operation {op!r}")
        shapes = {a.shape for a in args if a.shape != ()}
        if len(shapes) > 1:
            raise ValueError(f"{op}: operand shapes differ: {sorted(shapes)}")
        self.op = op
        self.args = tuple(args)
        self.shape = shapes.pop() if shapes else ()

    def __repr__(self):
        return f"Elementwise({self.op!r}, {self.shape})"


class Einsum(Array):
    def __init__(self, spec: str, args: tuple):
        m = re.fullmatch(r"([a-z,]+)->([a-z]*)", spec)
        if not m:
            raise ValueError(f"einsum spec {spec!r}")
        ins = m.group(1).split(",")
        if len(ins) != len(args):
            raise ValueError(f"einsum {spec!r}: {len(args)} operands")
        sizes: dict[str, int] = {}
        for idx, a in zip(ins, args):
            if len(idx) != a.ndim:
                raise ValueError(f"einsum {spec!r}: operand {a!r} has rank {a.ndim}")
            for ch, n in zip(idx, a.shape):
                if sizes.setdefault(ch, n) != n:
                    raise ValueError(f"einsum {spec!r}: index {ch} has sizes {sizes[ch]}, {n}")
        self.spec = spec
        self.inputs = ins
        self.output = m.group(2)
        self.args = tuple(args)
        self.shape = tuple(sizes[ch] for ch in self.output)

    def __repr__(self):
        return f"Einsum({self.spec!r}, {self.shape})"


class Take(Array):
    """``array.reshape(-1)[index]``: a gather through a constant index map."""

    def __init__(self, array: Array, index: np.ndarray):
        index = np.asarray(index, dtype=np.int64)
        size = int(np.prod(array.shape))
        if index.size and (index.min() < 0 or index.max() >= size):
            raise ValueError("take: index out of range")
        self.array = array
        self.index = index
        self.shape = index.shape

    def __repr__(self):
        return f"Take({self.array!r} -> {self.shape})"


class Stack(Array):
    """The arrays stacked along a new leading axis."""

    def __init__(self, arrays):
        arrays = tuple(arrays)
        shapes = {a.shape for a in arrays}
        if len(shapes) != 1:
            raise ValueError(f"stack: shapes differ: {sorted(shapes)}")
        self.arrays = arrays
        self.shape = (len(arrays), *arrays[0].shape)

    def __repr__(self):
        return f"Stack({len(self.arrays)} x {self.arrays[0].shape})"


def as_array(x) -> Array:
    if isinstance(x, Array):
        return x
    if np.isscalar(x):
        return Constant(float(x))
    raise TypeError(f"not an array expression: {type(x)}")


def elementwise(op: str, *args) -> Array:
    return Elementwise(op, tuple(as_array(a) for a in args))


def children(node: Array) -> tuple:
    if isinstance(node, (Elementwise, Einsum)):
        return node.args
    if isinstance(node, Take):
        return (node.array,)
    if isinstance(node, Stack):
        return node.arrays
    return ()


def walk(roots) -> list:
    """Every node reachable from ``roots``, each once, children first."""
    seen: set[int] = set()
    out: list = []
    stack = [(r, False) for r in reversed(list(roots))]
    while stack:
        node, expanded = stack.pop()
        if id(node) in seen:
            continue
        if expanded:
            seen.add(id(node))
            out.append(node)
            continue
        stack.append((node, True))
        for c in reversed(children(node)):
            if id(c) not in seen:
                stack.append((c, False))
    return out


def placeholders(root: Array) -> list:
    return [n for n in walk([root]) if isinstance(n, Placeholder)]

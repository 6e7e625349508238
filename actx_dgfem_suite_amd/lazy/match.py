"""Matching a normalised DG-operator DAG to the HIP kernels.

After :func:`~actx_dgfem_suite_amd.lazy.normalize.normalize`, each output field
of an operator the kernels implement is a linear combination of three einsum
forms (the reference optimizer's canonical batched einsums,
arraycontext/mass_inverse_fuser.py:20-74, materialization_policy.py:281-293):

* ``xre,rij,xej->ei`` (G, D, Stack of three pointwise volume fluxes);
* ``re,rij,ej->ei``   (G_x, D, one pointwise volume flux along axis x);
* ``ifj,fe,fej->ei``  (LIFT, Fscale, one pointwise face flux).

The matcher reads the constants out of them structurally:
- the derivative matrices and LIFT;
- ``dr/dx`` from the fused metric;
- ``Fscale``;
- the face-node maps from the gathers' index arrays;
- the normals and the boundary mask from the face flux's constant leaves.

It identifies the pointwise parts by evaluating them at sampled states
against each equation's template (:mod:`.templates`). The equation's
parameter comes from one probe at which the template is affine in it, and
is then checked at every sample. The host never evaluates a right-hand side:
a matched operator becomes a :class:`~actx_dgfem_suite_amd.plan.DGPlan`. A
DAG that does not fit raises ``NotImplementedError``, the reference's
convention for an unmatched DAG (arraycontext/dgfem_optimizer.py:220-224).
"""

from __future__ import annotations

from dataclasses import dataclass
from itertools import permutations

import numpy as np

from actx_dgfem_suite_amd.discretization import NFACES
from actx_dgfem_suite_amd.lazy import templates
from actx_dgfem_suite_amd.lazy.ir import (
    Array,
    Constant,
    DataWrapper,
    Einsum,
    Elementwise,
    Placeholder,
    Stack,
    Take,
)
from actx_dgfem_suite_amd.lazy.normalize import Normalized

_NP_TO_ORDER = {4: 1, 10: 2, 20: 3, 35: 4}
NSAMP = 64
RTOL = 1e-11


class Unmatched(NotImplementedError):
    """The DAG is not one of the operators the kernels implement."""


def _fail(msg: str):
    raise Unmatched(f"unmatched DG operator DAG: {msg}")


# {{{ linear structure

def linear_terms(node: Array, coef: float = 1.0) -> list:
    """``[(coef, node)]`` with the sum, difference, negation and constant
    scaling around einsum terms taken apart."""
    if isinstance(node, Elementwise):
        if node.op == "+":
            return linear_terms(node.args[0], coef) + linear_terms(node.args[1], coef)
        if node.op == "-":
            return linear_terms(node.args[0], coef) + linear_terms(node.args[1], -coef)
        if node.op == "neg":
            return linear_terms(node.args[0], -coef)
        if node.op == "*":
            a, b = node.args
            if isinstance(a, Constant):
                return linear_terms(b, coef * a.value)
            if isinstance(b, Constant):
                return linear_terms(a, coef * b.value)
    return [(coef, node)]

# }}}


# {{{ pointwise evaluation at samples

def evaluate(node: Array, env: dict, memo: dict | None = None):
    """Value of a pointwise expression at samples; leaves come from ``env``
    (keyed by node id)."""
    memo = {} if memo is None else memo
    key = id(node)
    if key in memo:
        return memo[key]
    if key in env:
        val = env[key]
    elif isinstance(node, Constant):
        val = node.value
    elif isinstance(node, Elementwise):
        a = [evaluate(x, env, memo) for x in node.args]
        op = node.op
        if op == "+":
            val = a[0] + a[1]
        elif op == "-":
            val = a[0] - a[1]
        elif op == "*":
            val = a[0] * a[1]
        elif op == "/":
            val = a[0] / a[1]
        elif op == "neg":
            val = -a[0]
        elif op == "sqrt":
            val = np.sqrt(a[0])
        elif op == "abs":
            val = np.abs(a[0])
        elif op == "maximum":
            val = np.maximum(a[0], a[1])
        elif op == "minimum":
            val = np.minimum(a[0], a[1])
        else:   # where
            val = np.where(np.asarray(a[0]) > 0.5, a[1], a[2])
    else:
        _fail(f"{node!r} is not pointwise in the traced state")
    memo[key] = val
    return val


def _leaves(node: Array, out: list, seen: set):
    if id(node) in seen:
        return
    seen.add(id(node))
    if isinstance(node, Elementwise):
        for a in node.args:
            _leaves(a, out, seen)
    elif not isinstance(node, Constant):
        out.append(node)


def leaves(nodes) -> list:
    out, seen = [], set()
    for n in nodes:
        _leaves(n, out, seen)
    return out


def _where_conditions(nodes) -> set:
    conds, seen = set(), set()
    stack = list(nodes)
    while stack:
        n = stack.pop()
        if id(n) in seen:
            continue
        seen.add(id(n))
        if isinstance(n, Elementwise):
            if n.op == "where":
                conds.add(id(n.args[0]))
            stack.extend(n.args)
    return conds

# }}}


@dataclass
class MatchedOperator:
    """Everything ``dg_plan_create`` takes, read out of the DAG."""
    equation: str
    param: float
    order: int
    d_weak: np.ndarray          # (3, Np, Np)
    d_strong: np.ndarray        # (3, Np, Np)
    lift: np.ndarray            # (Np, 4 Nfp)
    face_nodes: np.ndarray      # (4, Nfp)
    geometry: np.ndarray        # (K, 25)
    ext_map: np.ndarray         # (K, 4 Nfp) int32, boundary -(own)-1
    mass_inv: np.ndarray        # (Np, Np)
    face_mass: np.ndarray       # (Np, 4 Nfp)

    @property
    def nelements(self) -> int:
        return self.geometry.shape[0]

    def discretization(self):
        return ExtractedDiscretization(self)


class _Ref:
    def __init__(self, m: MatchedOperator):
        self.d_weak, self.d_strong, self.lift = m.d_weak, m.d_strong, m.lift
        self.face_nodes, self.mass_inv, self.face_mass = m.face_nodes, m.mass_inv, m.face_mass
        self.order = m.order
        self.np = m.d_weak.shape[1]
        self.nfp = m.face_nodes.shape[1]


class ExtractedDiscretization:
    """The arrays a :class:`DGPlan` reads, from a matched DAG (no node
    coordinates: the plan's work order is the DAG's element order)."""

    dim = 3
    nodes = None

    def __init__(self, m: MatchedOperator):
        self.matched = m
        self.ref = _Ref(m)
        self.order = m.order
        self.nelements = m.nelements
        self.np = self.ref.np
        self.nfp = self.ref.nfp

    def packed_geometry(self) -> np.ndarray:
        return self.matched.geometry

    def packed_ext_map(self) -> np.ndarray:
        return self.matched.ext_map


def _same(arrays, what):
    a0 = arrays[0]
    for a in arrays[1:]:
        if a.shape != a0.shape or not np.allclose(a, a0, rtol=1e-13, atol=1e-13 * np.abs(a0).max()):
            _fail(f"the {what} differ between terms")
    return a0


def match_dg_operator(norm: Normalized, state: list) -> MatchedOperator:
    """``state``: the traced placeholders in field order."""
    if norm.unfused:
        _fail("an inverse mass multiplies something other than a DG einsum")
    if len(norm.mass_inv) != 1:
        _fail("expected one reference inverse mass")
    minv = norm.mass_inv[0]
    nf = len(norm.outputs)
    cands = [eq for eq, (n, _, _) in templates.SHAPES.items() if n == nf]
    if not cands or len(state) != nf:
        _fail(f"{nf} output fields and {len(state)} traced fields")
    field_of = {id(p): k for k, p in enumerate(state)}
    if len(field_of) != nf or not all(isinstance(p, Placeholder) for p in state):
        _fail("the state must be distinct placeholders")

    # {{{ terms per output field

    vol: list = [[] for _ in range(nf)]     # (coef, G (3, K), D, expr)
    surf: list = [[] for _ in range(nf)]    # (coef, expr)
    lifts, fscales, diffs = [], [], []
    for c, out in enumerate(norm.outputs):
        if out.shape != state[c].shape:
            _fail(f"output {c} has shape {out.shape}")
        for coef, t in linear_terms(out):
            if not isinstance(t, Einsum) or not all(isinstance(a, DataWrapper) for a in t.args[:2]):
                _fail(f"term {t!r} of field {c} is not a DG einsum with constant operators")
            g, d, operand = t.args
            if t.spec == "xre,rij,xej->ei":
                if not isinstance(operand, Stack) or len(operand.arrays) != 3:
                    _fail("divergence operand is not three stacked fluxes")
                diffs.append(d.data)
                for x in range(3):
                    vol[c].append((coef, g.data[x], d.data, operand.arrays[x]))
            elif t.spec == "re,rij,ej->ei":
                diffs.append(d.data)
                vol[c].append((coef, g.data, d.data, operand))
            elif t.spec == "ifj,fe,fej->ei":
                lifts.append(g.data)
                fscales.append(d.data)
                surf[c].append((coef, operand))
            else:
                _fail(f"einsum {t.spec!r}")
    if not lifts or not diffs:
        _fail("no surface or no volume terms")
    dmat = _same(diffs, "derivative matrices")
    lift3 = _same(lifts, "LIFT operators")
    fscale = _same(fscales, "Fscale factors")
    npn, nfaces, nfp = lift3.shape
    if nfaces != NFACES or npn not in _NP_TO_ORDER or dmat.shape != (3, npn, npn):
        _fail(f"operator shapes {dmat.shape}, {lift3.shape}")
    order = _NP_TO_ORDER[npn]
    k = fscale.shape[1]

    # }}}

    # {{{ metric and axes

    axes: list = []       # distinct metric blocks G_x (3, K)
    for c in range(nf):
        for i, (coef, g, d, e) in enumerate(vol[c]):
            for a, ga in enumerate(axes):
                if np.allclose(g, ga, rtol=1e-13, atol=1e-15):
                    break
            else:
                axes.append(g)
                a = len(axes) - 1
            vol[c][i] = (coef, a, e)
    if len(axes) != 3:
        _fail(f"{len(axes)} distinct metric blocks (expected 3)")

    # }}}

    # {{{ face leaves: traces, normals, boundary mask

    flux_exprs = [e for c in range(nf) for _, e in surf[c]]
    fleaves = leaves(flux_exprs)
    conds = _where_conditions(flux_exprs)
    interior, exterior, normals, masks = [], [], [], []
    elem_base = np.arange(k)[None, :, None] * npn
    fmask = None
    ext_index = None
    for lf in fleaves:
        if isinstance(lf, Take):
            if not isinstance(lf.array, Placeholder) or id(lf.array) not in field_of:
                _fail("a face gather of something other than a state field")
            if lf.index.shape != (NFACES, k, nfp):
                _fail("face gather of an unexpected shape")
            local = lf.index - elem_base
            if np.all(local == local[:, :1, :]) and local.min() >= 0 and local.max() < npn:
                if fmask is not None and not np.array_equal(fmask, local[:, 0, :]):
                    _fail("two interior face maps")
                fmask = local[:, 0, :]
                interior.append(lf)
            else:
                if ext_index is not None and not np.array_equal(ext_index, lf.index):
                    _fail("two exterior face maps")
                ext_index = lf.index
                exterior.append(lf)
        elif isinstance(lf, DataWrapper) and lf.shape == (NFACES, k, nfp):
            (masks if id(lf) in conds else normals).append(lf)
        else:
            _fail(f"face flux leaf {lf!r}")
    if fmask is None or len(normals) != 3 or len(masks) > 1:
        _fail("face fluxes need an interior gather, three normal components and at most "
              "one boundary mask")
    bmask = masks[0].data > 0.5 if masks else np.zeros((NFACES, k, nfp), dtype=bool)
    if masks and not np.all((masks[0].data == 0) | (masks[0].data == 1)):
        _fail("boundary mask is not 0/1")
    nvals = [n.data for n in normals]
    for nv in nvals:
        if not np.all(nv == nv[:, :, :1]):
            _fail("normals vary across a face (not affine)")

    # }}}

    # {{{ identify the equation and its parameter by sampling

    rng = np.random.default_rng(1234)
    found = None
    for eq in cands:
        _, weak, _ = templates.SHAPES[eq]
        q = templates.sample_states(eq, NSAMP, rng)
        qm = templates.sample_states(eq, NSAMP, rng)
        qp = templates.sample_states(eq, NSAMP, rng)
        nrm = rng.standard_normal((3, NSAMP))
        nrm /= np.linalg.norm(nrm, axis=0)
        bs = (rng.random(NSAMP) < 0.3).astype(np.float64)

        def vol_target(env_q, perm_axes):
            out = np.zeros((nf, 3) + env_q.shape[1:])
            env = {id(p): env_q[field_of[id(p)]] for p in state}
            memo: dict = {}
            for c in range(nf):
                for coef, a, e in vol[c]:
                    out[c, perm_axes[a]] += coef * evaluate(e, env, memo)
            return out

        def surf_target(qmv, qpv, nv, bv, perm_n):
            env = {}
            for lf in interior:
                env[id(lf)] = qmv[field_of[id(lf.array)]]
            for lf in exterior:
                env[id(lf)] = qpv[field_of[id(lf.array)]]
            for i, lf in enumerate(normals):
                env[id(lf)] = nv[perm_n[i]]
            for lf in masks:
                env[id(lf)] = bv
            memo: dict = {}
            return np.stack([sum(coef * evaluate(e, env, memo) for coef, e in surf[c])
                             for c in range(nf)])

        kind, pf, pa, qprobe = templates.parameter_probe(eq)
        for perm_axes in permutations(range(3)):
            # volume: fit on the probe if it is a volume probe, then check
            if kind == "volume":
                tv = vol_target(qprobe, perm_axes)[pf, pa, 0]
                t0 = templates.volume(eq, qprobe, 0.0)[pf, pa, 0]
                t1 = templates.volume(eq, qprobe, 1.0)[pf, pa, 0]
                p = (tv - t0) / (t1 - t0)
                if not np.allclose(vol_target(q, perm_axes), templates.volume(eq, q, p),
                                   rtol=RTOL, atol=RTOL * max(1.0, abs(p))):
                    continue
            for perm_n in permutations(range(3)):
                if kind == "surface":
                    qpp = qprobe.copy()
                    qpp[pf] = 1.0
                    n1 = np.array([[1.0], [0.0], [0.0]])
                    b0 = np.zeros(1)
                    tv = surf_target(qprobe, qpp, n1, b0, perm_n)[pf, 0]
                    t0 = templates.surface(eq, qprobe, qpp, n1, b0, 0.0)[pf, 0]
                    t1 = templates.surface(eq, qprobe, qpp, n1, b0, 1.0)[pf, 0]
                    if t1 == t0:
                        continue
                    p = (tv - t0) / (t1 - t0)
                    if not np.allclose(vol_target(q, perm_axes), templates.volume(eq, q, p),
                                       rtol=RTOL, atol=RTOL * max(1.0, abs(p))):
                        continue
                with np.errstate(all="ignore"):
                    got = surf_target(qm, qp, nrm, bs, perm_n)
                    want = templates.surface(eq, qm, qp, nrm, bs, p)
                if np.all(np.isfinite(got)) and np.allclose(got, want, rtol=RTOL,
                                                            atol=RTOL * np.abs(want).max()):
                    found = (eq, weak, float(p), perm_axes, perm_n)
                    break
            if found:
                break
        if found:
            break
    if not found:
        _fail(f"no {'/'.join(cands)} operator fits the sampled volume and face fluxes")
    eq, weak, p, perm_axes, perm_n = found

    # }}}

    # {{{ the plan's arrays

    rx = np.empty((k, 3, 3))                        # rx[e, x, r]
    for a, g in enumerate(axes):
        rx[:, perm_axes[a], :] = g.T
    nrm_face = np.empty((k, NFACES, 3))
    for i, nv in enumerate(nvals):
        nrm_face[:, :, perm_n[i]] = nv[:, :, 0].T
    geometry = np.empty((k, 25))
    geometry[:, :9] = rx.reshape(k, 9)
    geometry[:, 9:] = np.concatenate([nrm_face, fscale.T[:, :, None]], axis=2).reshape(k, 16)
    own = (elem_base + fmask[:, None, :])
    if ext_index is None:
        if not np.all(bmask):
            _fail("interior faces without an exterior gather")
        ext_index = own
    ext = np.where(bmask, -own - 1, ext_index)      # (F, K, Nfp)
    ext_map = np.ascontiguousarray(ext.transpose(1, 0, 2).reshape(k, NFACES * nfp),
                                   dtype=np.int32)
    mass = np.linalg.inv(minv)
    if weak:
        d_weak = dmat
        d_strong = np.stack([(mass @ dmat[r] @ minv).T for r in range(3)])
    else:
        d_strong = dmat
        d_weak = np.stack([minv @ dmat[r].T @ mass for r in range(3)])
    # a derivative annihilates constants
    if np.abs(d_strong.sum(axis=2)).max() > 1e-9 * np.abs(d_strong).max():
        _fail("derivative matrices do not annihilate constants")
    lift = lift3.reshape(npn, NFACES * nfp)
    return MatchedOperator(
        equation=eq, param=p, order=order, d_weak=d_weak, d_strong=d_strong, lift=lift,
        face_nodes=np.ascontiguousarray(fmask, dtype=np.int32), geometry=geometry,
        ext_map=ext_map, mass_inv=minv, face_mass=mass @ lift)

    # }}}

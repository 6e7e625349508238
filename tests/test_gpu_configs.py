"""GPU: every benchmark configuration at its full size.

* wave P3 (30^3 x 6 = 162,000 tets), Euler P2 (40^3 x 6 = 384,000 tets) and
  Maxwell P4 (25^3 x 6 = 93,750 tets): the FP64 RHS of the seeded benchmark
  state against the oracle on the whole mesh, per field;
* Euler P2 at 40^3: the mixed forms (FP32 and BF16x3 contractions on an FP32
  state) against the FP64 form, per field, and the FP64 VJP against the
  FP32-storage VJP (two different adjoint kernels);
* wave P4 at 46^3 x 6 = 584,016 tets (20.4M nodes per field): polynomial
  exactness and linearity of the FP64 RHS, and the gradient of the
  final-time misfit after 200 LSERK4 steps (checkpoints every 20) with a
  BF16x3 forward and an FP32 adjoint against the FP64 forward and adjoint.

Each case prints its name before it touches the GPU.
"""

from functools import lru_cache

import numpy as np
import pytest
import torch

from actx_dgfem_suite_amd.plan import DGPlan
from actx_dgfem_suite_amd.timestepping import LSERK4, cfl_timestep, misfit_gradient
from actx_dgfem_suite_amd.workloads import WORKLOADS, initial_state
from oracle import dg_oracle
from tests._helpers import field_errors, oracle_rhs

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
NF = {"wave": 4, "euler": 5, "maxwell": 6}


def _dev(a, device, dtype=None):
    out = [torch.from_numpy(np.ascontiguousarray(f)).to(device) for f in a]
    return out if dtype is None else [t.to(dtype) for t in out]


def _np(fields):
    return np.stack([t.double().cpu().numpy() for t in fields])


@lru_cache(maxsize=2)
def _workload(name):
    w = WORKLOADS[name]
    d = w.discretization()
    return w, d, initial_state(w, d)


@pytest.mark.parametrize("name", ["wave_p3", "euler_p2", "maxwell_p4"])
def test_full_size_fp64_matches_oracle(name, device):
    w, d, q = _workload(name)
    print(f"[case] full-size oracle {name} K={d.nelements} P{w.order}", flush=True)
    plan = DGPlan(d, w.equation, w.param, device=device)
    got = _np(plan.rhs(_dev(q, device)))
    torch.cuda.synchronize()
    del plan
    od = dg_oracle.build_discretization(d.mesh.vertices, d.mesh.elements, d.ref.nodes)
    ref = oracle_rhs(w.equation, od, q, w.param)
    if w.equation == "euler":
        # The uniform x-flow's transverse momenta have an RHS that is zero up
        # to the rounding of their pressure terms (volume and face parts of
        # size p / h that cancel): measure each field against the larger of
        # its RHS and its volume term.
        flux = dg_oracle.euler_volume_flux(q, w.param)
        errs = []
        for c in range(len(q)):
            vol = dg_oracle.inverse_mass(od, sum(dg_oracle.weak_local_d_dx(od, x, flux[c][x])
                                                 for x in range(3)))
            den = max(np.linalg.norm(ref[c]), np.linalg.norm(vol))
            errs.append(float(np.linalg.norm(got[c] - ref[c]) / den))
    else:
        errs = field_errors(got, ref)
    print(f"  rel L2 per field: {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < TOL_FP64, errs


def test_euler_p2_vjp_full_size(device):
    """The FP64 Euler VJP (the MFMA adjoint kernel on 16-element tiles) on
    the whole 40^3 mesh against the FP32-storage VJP, which runs the batch
    adjoint kernel (one thread per node, FP64 arithmetic): independent code
    on the same FP32-representable state and cotangent.  They differ by the
    FP32 rounding of the latter's workspace and result (measured 4.6e-8)."""
    w, d, q = _workload("euler_p2")
    print(f"[case] full-size euler_p2 vjp K={d.nelements}", flush=True)
    plan = DGPlan(d, w.equation, w.param, device=device)
    q32 = _dev(q, device, torch.float32)
    y32 = _dev(np.random.default_rng(7).standard_normal((5, d.nelements, d.np)), device,
               torch.float32)
    v64 = _np(plan.rhs_vjp([t.double() for t in q32], [t.double() for t in y32]))
    v32 = _np(plan.rhs_vjp(q32, y32))
    torch.cuda.synchronize()
    del plan
    errs = field_errors(v32, v64)
    print(f"  rel L2 per field: {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < 1e-6, errs


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-4), ("bf16x3", 1e-4)])
def test_euler_p2_mixed_full_size(precision, tol, device):
    """The Euler benchmark at 40^3 x 6 tets: per-field relative L2 of the
    mixed form against the FP64 form (a field whose FP64 RHS is below 1e-3
    of the whole RHS, the transverse momenta of this x-flow, is measured
    against the whole RHS)."""
    w, d, q = _workload("euler_p2")
    print(f"[case] full-size mixed euler_p2 {precision} K={d.nelements}", flush=True)
    plan = DGPlan(d, w.equation, w.param, device=device)
    ref = plan.rhs(_dev(q, device))
    got = plan.rhs(_dev(q, device, torch.float32), precision=precision)
    torch.cuda.synchronize()
    errs = field_errors(_np(got), _np(ref))
    print(f"  rel L2 per field vs FP64: {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) < tol, errs


def _ridge_fields(x, nf, seed, order=4, nterms=6):
    """Fields that are sums of ridge polynomials c (a.x + b)^order, which
    span the degree-``order`` polynomials, with their exact gradients."""
    rng = np.random.default_rng(seed)
    vals, grads = [], []
    for _ in range(nf):
        v = np.zeros(x.shape[1:])
        g = [np.zeros(x.shape[1:]) for _ in range(3)]
        for _ in range(nterms):
            a = rng.standard_normal(3)
            b, c = rng.standard_normal(), rng.standard_normal()
            s = a[0] * x[0] + a[1] * x[1] + a[2] * x[2] + b
            s3 = s ** (order - 1)
            v += c * s3 * s
            for k in range(3):
                g[k] += (c * order * a[k]) * s3
        vals.append(v)
        grads.append(g)
    return np.stack(vals), grads


def test_wave_p4_large_fp64_exactness_and_linearity(device):
    """wave P4 at 584,016 tets: degree-4 continuous states give the exact
    (div v, grad u) on every interior element, and L is linear."""
    w, d, q0 = _workload("wave_p4_large")
    print(f"[case] full-size wave_p4_large exactness K={d.nelements}", flush=True)
    plan = DGPlan(d, w.equation, w.param, device=device)
    q, g = _ridge_fields(d.nodes, 4, seed=3)
    got = _np(plan.rhs(_dev(q, device)))
    inner = np.nonzero(np.all(d.connectivity.nbr_elem >= 0, axis=1))[0]
    exact = [g[1][0] + g[2][1] + g[3][2], g[0][0], g[0][1], g[0][2]]
    for c in range(4):
        scale = np.abs(exact[c][inner]).max()
        err = np.abs(got[c][inner] - exact[c][inner]).max()
        print(f"  field {c}: max err {err:.2e} (scale {scale:.2e})", flush=True)
        assert err < 1e-10 * max(scale, 1.0), (c, err, scale)
    del got, exact, g
    rng = np.random.default_rng(11)
    y = rng.standard_normal(q0.shape)
    a, b = 0.75, -1.25
    lx = _np(plan.rhs(_dev(q0, device)))
    ly = _np(plan.rhs(_dev(y, device)))
    lxy = _np(plan.rhs(_dev(a * q0 + b * y, device)))
    rel = np.linalg.norm(lxy - (a * lx + b * ly)) / np.linalg.norm(a * lx + b * ly)
    print(f"  linearity rel L2 {rel:.2e}", flush=True)
    assert rel < 1e-13


def test_wave_p4_large_mixed_trajectory_gradient(device):
    """The wave adjoint workload at its full size: 200 LSERK4 steps at the
    CFL step with checkpoints every 20, objective 1/2 ||u(T) - u_tgt||^2
    (u_tgt from seed 1); the gradient with a BF16x3 forward and an FP32
    adjoint against the FP64 forward and FP64 adjoint, per field."""
    w, d, q0 = _workload("wave_p4_large")
    print(f"[case] full-size wave_p4_large trajectory gradient K={d.nelements}", flush=True)
    plan = DGPlan(d, w.equation, w.param, device=device)
    dt = cfl_timestep(d, 1.0)
    q = _dev(q0, device)
    target = _dev(initial_state(w, d, seed=1)[:1], device)
    j64, g64 = misfit_gradient(LSERK4(plan, dt), q, target, 200, 20)
    print(f"  FP64: J = {j64:.10e}", flush=True)
    jm, gm = misfit_gradient(LSERK4(plan, dt, precision="bf16x3"), q, target, 200, 20,
                             adjoint_dtype=torch.float32)
    torch.cuda.synchronize()
    assert all(t.dtype == torch.float32 for t in gm)
    errs = [float(torch.linalg.norm(a.double() - b) / torch.linalg.norm(b))
            for a, b in zip(gm, g64)]
    print(f"  BF16x3/FP32: J = {jm:.10e}; grad rel L2 per field "
          f"{['%.2e' % e for e in errs]}", flush=True)
    assert abs(jm - j64) <= 1e-5 * j64
    assert max(errs) < 3e-5, errs

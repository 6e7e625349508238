"""``instantiate_actx_t`` / ``finish_command_queue`` for the HIP array
contexts (the operator-describing and the tracing one), mirroring
measure.py:16-55 of the reference (which builds a pyopencl queue + memory pool
or a JAX context, and raises ``NotImplementedError`` for any other context
type)."""

from __future__ import annotations

import gc

from actx_dgfem_suite_amd.arraycontext import HIPDGFEMArrayContext


def instantiate_actx_t(actx_t: type):
    from actx_dgfem_suite_amd.lazy import HIPDGFEMLazyArrayContext
    gc.collect()
    if isinstance(actx_t, type) and issubclass(actx_t, (HIPDGFEMArrayContext,
                                                        HIPDGFEMLazyArrayContext)):
        return actx_t()
    raise NotImplementedError(actx_t)


def finish_command_queue(actx) -> None:
    from actx_dgfem_suite_amd.lazy import HIPDGFEMLazyArrayContext
    if isinstance(actx, (HIPDGFEMArrayContext, HIPDGFEMLazyArrayContext)):
        actx.queue.finish()
    else:
        raise NotImplementedError(type(actx))

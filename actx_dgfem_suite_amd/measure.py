"""``instantiate_actx_t`` / ``finish_command_queue`` for the HIP array
context, mirroring measure.py:16-55 of the reference (which builds a pyopencl
queue + memory pool or a JAX context, and raises ``NotImplementedError`` for
any other context type)."""

from __future__ import annotations

import gc

from actx_dgfem_suite_amd.arraycontext import HIPDGFEMArrayContext


def instantiate_actx_t(actx_t: type) -> HIPDGFEMArrayContext:
    gc.collect()
    if isinstance(actx_t, type) and issubclass(actx_t, HIPDGFEMArrayContext):
        return actx_t()
    raise NotImplementedError(actx_t)


def finish_command_queue(actx) -> None:
    if isinstance(actx, HIPDGFEMArrayContext):
        actx.queue.finish()
    else:
        raise NotImplementedError(type(actx))

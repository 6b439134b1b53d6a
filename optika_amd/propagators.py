"""Sequence drivers for ray propagators.

Mirrors ``optika/propagators.py:19-103``.  The reference folds
``propagator.propagate_rays`` over the sequence in a Python loop, allocating
a full set of ray arrays per surface.  Here the whole sequence is compiled
into one system record (``_compile.py``) and traced by one fused kernel
launch: every ray is read once, carried through every surface in registers
and written once (plus once per surface when accumulating).
"""

from __future__ import annotations

from typing import Sequence

__all__ = [
    "propagate_rays",
    "accumulate_rays",
    "AbstractPropagator",
    "AbstractRayPropagator",
    "AbstractLightPropagator",
]


class AbstractPropagator:
    """An object that can propagate information."""


class AbstractRayPropagator(AbstractPropagator):
    """An object that interacts with light rays."""

    def propagate_rays(self, rays):
        raise NotImplementedError


class AbstractLightPropagator(AbstractRayPropagator):
    """An object that interacts with light."""


def _as_list(propagators) -> list:
    if hasattr(propagators, "propagate_rays") and not isinstance(propagators, (list, tuple)):
        return [propagators]
    return list(propagators)


def propagate_rays(propagators: Sequence, rays, precision: str = "f64"):
    """Propagate ``rays`` through every propagator in order
    (``optika/propagators.py:19-41``)."""
    from . import _trace

    return _trace.trace_sequence(_as_list(propagators), rays, accumulate=False, precision=precision)


def accumulate_rays(propagators: Sequence, rays, axis: str = "surface", precision: str = "f64"):
    """Propagate ``rays`` and keep the rays after every propagator
    (``optika/propagators.py:44-74``).  The result's tensors gain a leading
    axis of length ``len(propagators)`` (the reference's ``axis``)."""
    from . import _trace

    return _trace.trace_sequence(_as_list(propagators), rays, accumulate=True, precision=precision)

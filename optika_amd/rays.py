"""Ray containers.

``RayVectorArray`` mirrors ``optika.rays.RayVectorArray``
(``optika/rays/_ray_vectors.py:239-294``): wavelength (nm), position (mm),
direction (direction cosines), intensity, attenuation (1/mm), index of
refraction and the ``unvignetted`` mask.  Here every field is either a
scalar (broadcast to every ray, the reference's defaults are intensity 1,
attenuation 0, index 1, unvignetted True) or a 1-D device tensor of length
``N``.  Vectors hold one tensor per component, so the whole state is
structure-of-arrays: each component is its own contiguous array, which is
exactly what the trace kernel reads and writes with coalesced accesses.

``RayFunctionArray`` mirrors ``optika/rays/_ray_functions.py:24-32``.
"""

from __future__ import annotations

import dataclasses
import math
from typing import Any

import numpy as np
import torch

from .vectors import Cartesian3dVectorArray, as_vector3

__all__ = ["RayVectorArray", "RayFunctionArray"]


def _numel(v) -> int | None:
    if isinstance(v, torch.Tensor) and v.ndim > 0:
        return v.numel()
    if isinstance(v, np.ndarray) and v.ndim > 0:
        return v.size
    return None


@dataclasses.dataclass(eq=False)
class RayVectorArray:
    wavelength: Any = 0.0
    position: Any = dataclasses.field(default_factory=Cartesian3dVectorArray)
    direction: Any = dataclasses.field(default_factory=lambda: Cartesian3dVectorArray(0, 0, 1))
    intensity: Any = 1.0
    attenuation: Any = 0.0
    index_refraction: Any = 1.0
    unvignetted: Any = True

    def __post_init__(self):
        self.position = as_vector3(self.position)
        self.direction = as_vector3(self.direction)

    # -- shape ---------------------------------------------------------------
    def _fields(self):
        return (
            self.wavelength,
            *self.position,
            *self.direction,
            self.intensity,
            self.attenuation,
            self.index_refraction,
            self.unvignetted,
        )

    @property
    def num(self) -> int:
        """Number of rays (1 when every field is scalar)."""
        sizes = {s for s in (_numel(f) for f in self._fields()) if s is not None}
        if len(sizes) > 1:
            raise ValueError(f"ray fields have inconsistent lengths {sorted(sizes)}")
        return sizes.pop() if sizes else 1

    @property
    def device(self) -> torch.device:
        for f in self._fields():
            if isinstance(f, torch.Tensor):
                return f.device
        return torch.device("cuda")

    @property
    def n(self):
        """Complex index ``n + i alpha lambda / (4 pi)``
        (``optika/rays/_ray_vectors.py:73-87``); attenuation in 1/mm and
        wavelength in nm, so the product carries a factor 1e-6."""
        k = self.attenuation * self.wavelength / (4 * math.pi) * 1e-6
        return self.index_refraction + k * 1j

    def copy_shallow(self) -> "RayVectorArray":
        return dataclasses.replace(self)

    def replace(self, **kwargs) -> "RayVectorArray":
        return dataclasses.replace(self, **kwargs)

    def __getitem__(self, index) -> "RayVectorArray":
        def take(v):
            if isinstance(v, (torch.Tensor, np.ndarray)) and v.ndim > 0:
                return v[index]
            return v

        return RayVectorArray(
            wavelength=take(self.wavelength),
            position=Cartesian3dVectorArray(*(take(c) for c in self.position)),
            direction=Cartesian3dVectorArray(*(take(c) for c in self.direction)),
            intensity=take(self.intensity),
            attenuation=take(self.attenuation),
            index_refraction=take(self.index_refraction),
            unvignetted=take(self.unvignetted),
        )

    def to(self, device=None, dtype=None) -> "RayVectorArray":
        """Materialise every field as a 1-D tensor on ``device`` with ``dtype``
        (the mask stays boolean)."""
        n = self.num
        device = torch.device(device) if device is not None else self.device
        dtype = dtype or torch.float64

        def full(v, dt):
            if isinstance(v, torch.Tensor) and v.ndim > 0:
                return v.to(device=device, dtype=dt).contiguous()
            if isinstance(v, np.ndarray) and v.ndim > 0:
                return torch.as_tensor(v).to(device=device, dtype=dt).contiguous()
            if isinstance(v, torch.Tensor):
                v = v.item()
            return torch.full((n,), v, dtype=dt, device=device)

        return RayVectorArray(
            wavelength=full(self.wavelength, dtype),
            position=Cartesian3dVectorArray(*(full(c, dtype) for c in self.position)),
            direction=Cartesian3dVectorArray(*(full(c, dtype) for c in self.direction)),
            intensity=full(self.intensity, dtype),
            attenuation=full(self.attenuation, dtype),
            index_refraction=full(self.index_refraction, dtype),
            unvignetted=full(self.unvignetted, torch.bool),
        )

    def numpy(self) -> dict[str, np.ndarray]:
        """Host copies of every field, broadcast to ``num`` rays."""
        n = self.num

        def host(v, dt):
            if isinstance(v, torch.Tensor):
                v = v.detach().cpu().numpy()
            return np.broadcast_to(np.asarray(v, dtype=dt), (n,)).copy()

        return dict(
            wavelength=host(self.wavelength, np.float64),
            position=np.stack([host(c, np.float64) for c in self.position]),
            direction=np.stack([host(c, np.float64) for c in self.direction]),
            intensity=host(self.intensity, np.float64),
            attenuation=host(self.attenuation, np.float64),
            index_refraction=host(self.index_refraction, np.float64),
            unvignetted=host(self.unvignetted, bool),
        )


@dataclasses.dataclass(eq=False)
class RayFunctionArray:
    """A map from input coordinates to traced rays."""

    inputs: Any = None
    outputs: Any = None

    def copy_shallow(self) -> "RayFunctionArray":
        return dataclasses.replace(self)

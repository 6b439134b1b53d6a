"""Optical constants ``n + i k`` of chemicals from Windt/IMD ``.nk`` tables.

Mirrors ``optika.chemicals.Chemical`` (``optika/chemicals/_chemicals.py:67-144``):
the table file is ``{formula}[_{table}].nk`` (``a-`` prefixed when
amorphous), with wavelengths in Angstroms, and ``n(wavelength)`` is a linear
interpolation of ``n + i k`` that clamps outside the table.

Unlike the reference, which re-reads the text file on every call, a table is
parsed once and cached.  The device kernels receive the table's rows inside
the compiled system (``_compile.py``) and interpolate per ray.

Tables shipped in ``data/nk``: Mo, Si, SiO2, SiC, Cr (copied from the
reference's ``optika/chemicals/nk`` data directory).  More tables can be
found through the ``OPTK_NK_PATH`` environment variable (a ``:``-separated
list of directories).
"""

from __future__ import annotations

import dataclasses
import functools
import os
import pathlib

import numpy as np

__all__ = ["AbstractChemical", "Chemical", "nk_table"]

_DATA = pathlib.Path(__file__).parent / "data" / "nk"


@functools.lru_cache(maxsize=None)
def nk_table(filename: str) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Parse a ``.nk`` file: ``(wavelength_angstrom, n, k)`` float64 arrays."""
    dirs = [_DATA]
    extra = os.environ.get("OPTK_NK_PATH")
    if extra:
        dirs = [pathlib.Path(p) for p in extra.split(":") if p] + dirs
    for d in dirs:
        path = d / filename
        if path.exists():
            break
    else:
        raise FileNotFoundError(
            f"optical-constant table {filename!r} not found in {[str(d) for d in dirs]}"
        )
    rows = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith(";"):
                continue
            rows.append([float(v) for v in s.split()[:3]])
    data = np.asarray(rows, dtype=np.float64)
    w, n, k = data[:, 0].copy(), data[:, 1].copy(), data[:, 2].copy()
    if np.any(np.diff(w) <= 0):
        raise ValueError(f"wavelengths of {filename} are not strictly increasing")
    return w, n, k


class AbstractChemical:
    formula: str
    is_amorphous: bool
    table: None | str

    @property
    def file_nk(self) -> str:
        name = f"{self.formula}"
        if self.table is not None:
            name = f"{name}_{self.table}"
        if self.is_amorphous:
            name = f"a-{name}"
        return f"{name}.nk"

    @property
    def nk(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        return nk_table(self.file_nk)

    def n(self, wavelength):
        """Complex index of refraction at ``wavelength`` (nm), host-side."""
        w, n, k = self.nk
        x = np.asarray(wavelength, dtype=np.float64) * 10.0
        return np.interp(x, w, n) + 1j * np.interp(x, w, k)


@dataclasses.dataclass(eq=False)
class Chemical(AbstractChemical):
    formula: str = ""
    is_amorphous: bool = False
    table: None | str = None

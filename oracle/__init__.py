"""CPU float64 restatement of the reference's sequential ray-trace path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``optika_amd`` imports, calls or
executes anything under ``oracle/``; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
use it, and there only as the checker (and, in ``bench.py``, as the
reference-style CPU timing).  The product's trace path runs only on the
device and raises when its HIP library is missing.

The restatement is numpy code written in the reference's vectorized style
(one array operation per named-array operation), following the reference's
expression order so that float64 results can be compared with the device
kernels at 1e-12 relative.  Every function cites the reference file:line it
follows.

Where the reference's arithmetic lives in its third-party substrate
(``named_arrays ~= 2.1``, not vendored and not importable here), this module
fixes an explicit convention and says so:

* transformations: standard active rotations, lists applied in order
  (pinned by the geometry of ``optika/systems/_sequential.py:1853-1912``),
  a list composed into one affine map ``p -> m p + t`` (matrix products in
  list order) and applied as such, inverse ``m^T (p - t)``; whether
  ``named_arrays`` applies a list as one composed map or element by element
  is not visible here (the two differ by rounding only: ~1e-12 mm after a
  metre-long path through a conic whose intercept cancels catastrophically,
  e.g. a tilted and decentred spectrograph) -- pinned against the element
  by element application to 1e-12 and by inverse-of-apply identity;
* ``na.interp``: ``numpy.interp`` (clamping) -- pinned by the IMD golden
  reflectivity curves of ``optika/materials/_tests/_data``;
* ``na.optimize.root_secant`` (toroidal intercept): replaced by a Newton
  iteration to float64 convergence -- pinned by on-surface residuals and by
  the toroid-with-equal-radii == sphere closed form, NOT by the reference's
  own iteration count (parity with the reference's secant stopping point is
  unpinned);
* ``na.geometry.point_in_polygon``: crossing-number rule -- pinned by brute
  force against a winding-number count away from edges; ties on an edge are
  unpinned;
* ``na.histogram``: ``numpy.histogramdd`` edge rules with ``numpy.linspace``
  edges -- pinned against ``numpy.histogramdd`` itself; parity with the
  reference's histogram on exact bin edges is unpinned;
* matrix ``power``: binary powering (``numpy.linalg.matrix_power``'s
  algorithm) -- pinned against repeated multiplication and the IMD curves;
  the reference's exact algorithm is unpinned.

A reference quirk kept on purpose: the triangular groove efficiency uses
``gamma / ((pi gamma / 2)^2 + i^2)`` as the reference's code does
(``optika/rulings/_rulings.py:903``), not the ``- i^2`` of its docstring
(Magnusson Table 1, ``:876-877``).  With the code's sign the orders do not
sum to one (about 0.90 at gamma = 0.3); parity follows the code, and
"fixing" the sign here or in the kernels would break it.  The pins
(``tests/test_oracle_pins.py``) fix the code's form by its gamma -> 0
limit and by values derived by hand at gamma = 2 / pi.
"""

from .trace import *  # noqa: F401,F403

"""The two systems the reference's stop-solve tests use
(optika/systems/_sequential_test.py:536-672), from optika_amd.workloads."""

from optika_amd.workloads import RADIUS_FIELD_GRAZING, RADIUS_FIELD_NEWTONIAN  # noqa: F401
from optika_amd.workloads import stop_grazing as grazing  # noqa: F401
from optika_amd.workloads import stop_newtonian as newtonian  # noqa: F401

"""The C ABI: the library builds, loads (no GPU needed) and exports every
entry point declared in include/optk.h with matching record layouts."""

import ctypes
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "optk.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int32_t|int64_t|size_t|const char \*)\s*(optk_\w+)\(", text,
                                 re.MULTILINE)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("optk_abi_version", "optk_last_error", "optk_sizeof", "optk_system_create",
              "optk_system_destroy", "optk_trace", "optk_trace_grid", "optk_eval", "optk_collect",
              "optk_batch_create", "optk_batch_destroy", "optk_batch_trace", "optk_batch_autotune"):
        assert n in names


def test_library_loads_and_exports_every_symbol():
    from optika_amd import _lib

    if not _lib.library_path().exists():
        pytest.skip("liboptk.so not built (run python -m optika_amd._build)")
    L = _lib.lib()  # checks ABI version and every record size
    for name in declared_functions():
        assert hasattr(L, name), name
    assert L.optk_abi_version() == 1


def test_invalid_arguments_are_rejected_without_a_device():
    from optika_amd import _lib

    if not _lib.library_path().exists():
        pytest.skip("liboptk.so not built")
    L = _lib.lib()
    assert L.optk_system_create(None, None) == -1
    assert b"null" in L.optk_last_error()
    d = _lib.SystemDesc()
    d.num_surfaces = 1
    d.num_paths = 1
    handle = ctypes.c_void_p()
    # surfaces pointer missing -> rejected before any device call
    assert L.optk_system_create(ctypes.byref(d), ctypes.byref(handle)) == -1
    d.num_paths = 99
    d.num_surfaces = 0
    assert L.optk_system_create(ctypes.byref(d), ctypes.byref(handle)) == -1
    assert b"num_paths" in L.optk_last_error()
    assert L.optk_sizeof(b"no_such_record") == -1


def test_batch_arguments_are_rejected_without_a_device():
    from optika_amd import _lib

    if not _lib.library_path().exists():
        pytest.skip("liboptk.so not built")
    L = _lib.lib()
    handle = ctypes.c_void_p()
    assert L.optk_batch_create(None, ctypes.byref(handle)) == -1
    desc = _lib.BatchDesc()
    assert L.optk_batch_create(ctypes.byref(desc), ctypes.byref(handle)) == -1  # no base
    base = _lib.SystemDesc()
    desc.base = ctypes.pointer(base)
    desc.num_instances = 0
    assert L.optk_batch_create(ctypes.byref(desc), ctypes.byref(handle)) == -1
    assert b"num_instances" in L.optk_last_error()
    assert L.optk_batch_trace(None, 0, None, None, None, None, None, 0, None) == -1
    assert L.optk_batch_destroy(None) == 0


def test_polynomial_arguments_are_rejected_without_a_device():
    import numpy as np

    from optika_amd import _lib, surrogate

    if not _lib.library_path().exists():
        pytest.skip("liboptk.so not built")
    L = _lib.lib()
    codes = surrogate._EXP_CODES.copy()
    buf = ctypes.c_double()
    p = ctypes.addressof(buf)
    assert L.optk_poly_eval(p, p, 1, 0, codes.ctypes.data, p, None) == -1
    assert b"num_targets" in L.optk_last_error()
    bad = codes.copy()
    bad[3] = 3 | (3 << 2)  # degree 6
    assert L.optk_poly_eval(p, p, 1, 3, bad.ctypes.data, p, None) == -1
    assert b"degree" in L.optk_last_error()
    assert L.optk_poly_normal(p, None, p, 1, 3, codes.ctypes.data, p, p, 1 << 30, None) == -1
    assert L.optk_poly_normal(p, p, p, 1, 3, codes.ctypes.data, p, p, 8, None) == -1
    assert b"workspace" in L.optk_last_error()
    assert L.optk_poly_workspace(10) >= 8 * 56 * 67
    assert np.uint32 == codes.dtype.type


def test_stop_newton_arguments_are_rejected_without_a_device():
    from optika_amd import _lib

    if not _lib.library_path().exists():
        pytest.skip("liboptk.so not built")
    L = _lib.lib()
    a = _lib.StopArgs()
    assert L.optk_stop_newton(None, None, ctypes.byref(a), 1, None) == -1
    assert b"null" in L.optk_last_error()


def test_pupil_and_expose_arguments_are_rejected_without_a_device():
    from optika_amd import _lib

    if not _lib.library_path().exists():
        pytest.skip("liboptk.so not built")
    L = _lib.lib()
    assert L.optk_pupil_reduce(None, 1, None) == -1
    a = _lib.PupilArgs()
    a.per_segment = 4
    assert L.optk_pupil_reduce(ctypes.byref(a), 1, None) == -1
    assert b"unvignetted" in L.optk_last_error()
    assert L.optk_pupil_reduce(ctypes.byref(a), -1, None) == -1
    e = _lib.ExposeArgs()
    assert L.optk_expose(ctypes.byref(e), 4, None) == -1
    assert b"image" in L.optk_last_error()
    buf = (ctypes.c_double * 4)()
    e.photons = e.electrons = ctypes.addressof(buf)
    e.scale = -1.0
    assert L.optk_expose(ctypes.byref(e), 4, None) == -1
    assert b"non-negative" in L.optk_last_error()
    # zero-size calls succeed without touching a device
    e.scale = 1.0
    assert L.optk_expose(ctypes.byref(e), 0, None) == 0


def test_collect_arguments_are_rejected_without_a_device():
    from optika_amd import _lib

    if not _lib.library_path().exists():
        pytest.skip("liboptk.so not built")
    L = _lib.lib()
    assert L.optk_collect(None, 1, None) == -1
    a = _lib.CollectArgs()
    buf = (ctypes.c_double * 4)()
    p = ctypes.addressof(buf)
    a.position[0] = a.position[1] = p
    a.direction_z = a.wavelength = a.intensity = a.wavelength_edges = p
    a.image = a.moment = p
    a.n_wavelength, a.nx, a.ny, a.precision = 1, 2048, 2048, _lib.F64
    a.x_lo, a.x_hi, a.y_lo, a.y_hi = -15.36, 15.36, -15.36, 15.36
    # an empty range (a range merely narrow for its offset from the origin is
    # accepted: its bins come from a search over the edges)
    a.x_hi = a.x_lo
    assert L.optk_collect(ctypes.byref(a), 4, None) == -1
    assert b"empty pixel range" in L.optk_last_error()
    # zero rays: accepted without touching a device
    a.x_lo, a.x_hi = -15.36, 15.36
    assert L.optk_collect(ctypes.byref(a), 0, None) == 0


def test_denormalize_arguments_are_rejected_without_a_device():
    from optika_amd import _lib

    if not _lib.library_path().exists():
        pytest.skip("liboptk.so not built")
    L = _lib.lib()
    assert L.optk_denormalize(None, 1, None) == -1
    a = _lib.DenormalizeArgs()
    assert L.optk_denormalize(ctypes.byref(a), 1, None) == -1
    assert b"table" in L.optk_last_error()
    a.n_table = 1
    assert L.optk_denormalize(ctypes.byref(a), -1, None) == -1
    buf = (ctypes.c_double * 4)()
    a.in_[0] = ctypes.addressof(buf)
    assert L.optk_denormalize(ctypes.byref(a), 4, None) == -1
    assert b"missing" in L.optk_last_error()
    # no components: nothing to do, no device call
    a.in_[0] = None
    assert L.optk_denormalize(ctypes.byref(a), 4, None) == 0


def test_noalloc_collect_rejects_oversized_calls_without_a_device():
    """optk_collect_noalloc refuses, before any device work, a call whose
    tiled workspace exceeds its scratch (the pointers are never read)."""
    from optika_amd import _lib

    if not _lib.library_path().exists():
        pytest.skip("liboptk.so not built")
    L = _lib.lib()
    a = _lib.CollectArgs()
    fake = 0x1000
    a.position[0] = a.position[1] = a.direction_z = a.wavelength = a.intensity = fake
    a.wavelength_edges = a.image = a.moment = fake
    a.n_wavelength, a.nx, a.ny = 1, 2048, 2048
    a.x_lo, a.x_hi, a.y_lo, a.y_hi = -15.36, 15.36, -15.36, 15.36
    for prec, n_ok in ((_lib.F32, 3_272_688), (_lib.F64, 2_981_872)):
        a.precision = prec
        # the largest call the scratch holds, and one ray more
        assert L.optk_collect_workspace(ctypes.byref(a), n_ok) <= _lib.COLLECT_SCRATCH_BYTES
        assert L.optk_collect_workspace(ctypes.byref(a), n_ok + 1) > _lib.COLLECT_SCRATCH_BYTES
        assert L.optk_collect_noalloc(ctypes.byref(a), n_ok + 1, None) == -1
        assert b"scratch" in L.optk_last_error()
    assert L.optk_collect_noalloc(ctypes.byref(a), 1 << 31, None) == -1
    assert L.optk_collect_noalloc(None, 10, None) == -1
    a.image = 0
    assert L.optk_collect_noalloc(ctypes.byref(a), 10, None) == -1
    assert b"missing" in L.optk_last_error()

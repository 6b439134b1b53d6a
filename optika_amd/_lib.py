"""ctypes binding of liboptk.so (the C ABI declared in include/optk.h).

The structures below mirror the C records field for field; their sizes are
checked against ``optk_sizeof`` when the library is loaded.  torch is
imported first so that the library binds to the HIP runtime torch already
loaded (one HIP runtime per process).

There is no fallback: if the library is missing or fails to load, every
trace entry point raises.
"""

from __future__ import annotations

import ctypes
import os
import pathlib

import torch  # noqa: F401  (must precede loading liboptk.so)

__all__ = ["lib", "check", "OptkError", "Affine", "Surface", "Layer", "Block", "Stack",
           "Chemical", "SystemDesc", "Rays", "Grid", "EvalArgs", "CollectArgs", "BatchDesc",
           "Tune"]

_LIB_PATH = pathlib.Path(os.environ.get(
    "OPTK_LIBRARY", pathlib.Path(__file__).resolve().parent / "liboptk.so"))

MAX_ORDERS = 8
MAX_POLY = 8
MAX_BLOCKS = 8

F64 = 0
F32 = 1
TRACE_APPLY_POST = 1
TRACE_GENERIC = 4  # the generic kernels even where a specialised one exists

OP_SAG_VALUE = 0
OP_SAG_NORMAL = 1
OP_SAG_INTERCEPT = 2
OP_APERTURE = 3
OP_SNELL = 4
OP_MULTILAYER = 5
OP_SPACING = 6
OP_INCIDENT = 7
OP_RULINGS_EFF = 8
OP_MATERIAL_EFF = 9
OP_INDEX = 10
OP_SNELL_SCALAR = 11

c_double = ctypes.c_double
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_void_p = ctypes.c_void_p


class Affine(ctypes.Structure):
    _fields_ = [("m", c_double * 9), ("t", c_double * 3), ("kind", c_int32), ("pad", c_int32)]


class Surface(ctypes.Structure):
    _fields_ = [
        ("transform", Affine),
        ("sag_kind", c_int32), ("pad0", c_int32),
        ("sag_p", c_double * 4),
        ("sag_c", c_double),
        ("sag_transform", Affine),
        ("material_kind", c_int32), ("is_mirror", c_int32),
        ("material_p", c_double * 6),
        ("material_table", c_int32), ("material_table_len", c_int32),
        ("stack", c_int32), ("pad1", c_int32),
        ("rulings_kind", c_int32), ("spacing_kind", c_int32),
        ("num_orders", c_int32), ("num_terms", c_int32),
        ("orders", c_int32 * MAX_ORDERS),
        ("rulings_p", c_double * 2),
        ("spacing_constant", c_double),
        ("wavelength_scale", c_double),
        ("spacing_normal", c_double * 3),
        ("poly_coeff", c_double * MAX_POLY),
        ("poly_power", c_int32 * MAX_POLY),
        ("holo_x1", c_double * 3),
        ("holo_x2", c_double * 3),
        ("holo_wavelength", c_double),
        ("holo_diverging1", c_int32), ("holo_diverging2", c_int32),
        ("spacing_transform", Affine),
        ("rulings_table", c_int32), ("rulings_table_len", c_int32),
        ("aperture_kind", c_int32), ("aperture_flags", c_int32),
        ("aperture_p", c_double * 4),
        ("polygon_offset", c_int32), ("polygon_count", c_int32),
        ("aperture_transform", Affine),
    ]


class Layer(ctypes.Structure):
    _fields_ = [("chemical", c_int32), ("profile_kind", c_int32),
                ("thickness", c_double), ("profile_width", c_double)]


class Block(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("first_layer", c_int32),
                ("num_layers", c_int32), ("num_periods", c_int32)]


class Stack(ctypes.Structure):
    _fields_ = [("num_blocks", c_int32), ("substrate_layer", c_int32),
                ("blocks", Block * MAX_BLOCKS)]


class Chemical(ctypes.Structure):
    _fields_ = [("offset", c_int32), ("count", c_int32)]


class SystemDesc(ctypes.Structure):
    _fields_ = [
        ("surfaces", ctypes.POINTER(Surface)),
        ("num_surfaces", c_int32), ("num_paths", c_int32),
        ("layers", ctypes.POINTER(Layer)),
        ("num_layers", c_int32), ("num_stacks", c_int32),
        ("stacks", ctypes.POINTER(Stack)),
        ("chemicals", ctypes.POINTER(Chemical)),
        ("num_chemicals", c_int32), ("pad", c_int32),
        ("tables", ctypes.POINTER(c_double)),
        ("num_tables", c_int64),
        ("pre", Affine * 2),
        ("pre_is_inverse", c_int32 * 2),
        ("post", Affine),
    ]


class Rays(ctypes.Structure):
    _fields_ = [
        ("position", c_void_p * 3),
        ("direction", c_void_p * 3),
        ("wavelength", c_void_p),
        ("intensity", c_void_p),
        ("attenuation", c_void_p),
        ("index_refraction", c_void_p),
        ("unvignetted", c_void_p),
        ("wavelength_default", c_double),
        ("intensity_default", c_double),
        ("attenuation_default", c_double),
        ("index_default", c_double),
        ("unvignetted_default", c_int32), ("pad", c_int32),
    ]


class Grid(ctypes.Structure):
    _fields_ = [
        ("field", c_void_p * 2),
        ("pupil", c_void_p * 2),
        ("wavelength", c_void_p),
        ("intensity", c_void_p),
        ("field_default", c_double * 2),
        ("pupil_default", c_double * 2),
        ("wavelength_default", c_double),
        ("intensity_default", c_double),
        ("object_at_infinity", c_int32), ("pad", c_int32),
    ]


class EvalArgs(ctypes.Structure):
    _fields_ = [
        ("in_", c_void_p * 12),
        ("out", c_void_p * 12),
        ("out_mask", c_void_p),
        ("surface", c_int32), ("iparam", c_int32),
        ("stack", c_int32), ("pad", c_int32),
    ]


class CollectArgs(ctypes.Structure):
    _fields_ = [
        ("position", c_void_p * 2),
        ("direction_z", c_void_p),
        ("wavelength", c_void_p),
        ("intensity", c_void_p),
        ("unvignetted", c_void_p),
        ("wavelength_edges", c_void_p),
        ("n_wavelength", c_int32), ("nx", c_int32), ("ny", c_int32), ("precision", c_int32),
        ("x_lo", c_double), ("x_hi", c_double), ("y_lo", c_double), ("y_hi", c_double),
        ("image", c_void_p),
        ("moment", c_void_p),
        ("workspace", c_void_p),
        ("workspace_bytes", ctypes.c_uint64),
    ]


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("base", ctypes.POINTER(SystemDesc)),
        ("surfaces", ctypes.POINTER(Surface)),
        ("pre", ctypes.POINTER(Affine)),
        ("post", ctypes.POINTER(Affine)),
        ("num_instances", c_int32), ("pad", c_int32),
    ]


class Tune(ctypes.Structure):
    _fields_ = [("block", c_int32), ("rays_per_thread", c_int32), ("blocks_per_cu", c_int32),
                ("pad", c_int32), ("ms", c_double)]


class MlpArgs(ctypes.Structure):
    _fields_ = [("f32", c_void_p * 8), ("u16", c_void_p * 4), ("rows", c_int64), ("count", c_int64),
                ("cols", c_int32), ("inner", c_int32), ("flag", c_int32), ("pad", c_int32),
                ("scalars", ctypes.c_float * 8)]


SURROGATE_POLY = 56
SURROGATE_MAX_APERTURES = 4
SURROGATE_MAX_TARGETS = 3 + 2 * SURROGATE_MAX_APERTURES
SURROGATE_MAX_VERTICES = 16


class SurrogateAperture(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("inverted", c_int32), ("num_vertices", c_int32),
                ("pad", c_int32), ("p", ctypes.c_float * 4),
                ("vx", ctypes.c_float * SURROGATE_MAX_VERTICES),
                ("vy", ctypes.c_float * SURROGATE_MAX_VERTICES)]


class ExposeArgs(ctypes.Structure):
    _fields_ = [("photons", c_void_p), ("electrons", c_void_p), ("scale", ctypes.c_double),
                ("read_noise", ctypes.c_double), ("seed", ctypes.c_uint64), ("noise", c_int32),
                ("pad", c_int32)]


class PupilArgs(ctypes.Structure):
    _fields_ = [("x", c_void_p), ("y", c_void_p), ("intensity", c_void_p), ("unvignetted", c_void_p),
                ("weight", c_void_p), ("per_segment", c_int64), ("count", c_void_p),
                ("sum_x", c_void_p), ("sum_y", c_void_p), ("all_x", c_void_p), ("all_y", c_void_p),
                ("throughput", c_void_p)]


class DenormalizeArgs(ctypes.Structure):
    _fields_ = [("in_", c_void_p * 4), ("out", c_void_p * 4), ("lo", c_void_p * 4), ("ptp", c_void_p * 4),
                ("table_index", c_void_p), ("n_table", c_int64)]


class StopArgs(ctypes.Structure):
    _fields_ = [("position", c_void_p * 3), ("direction", c_void_p * 3), ("wavelength", c_void_p),
                ("target", c_void_p * 2), ("x", c_void_p * 2), ("iterations", c_void_p),
                ("z", c_void_p), ("tol", ctypes.c_double), ("dx", ctypes.c_double),
                ("free_component", c_int32), ("target_component", c_int32),
                ("max_iter", c_int32), ("pad", c_int32)]


STOP_FREE_DIRECTION, STOP_FREE_POSITION = 0, 1
STOP_TARGET_POSITION, STOP_TARGET_DIRECTION = 0, 1


class SurrogateDesc(ctypes.Structure):
    _fields_ = [("w1", c_void_p), ("b1", c_void_p), ("w1_16", c_void_p), ("w16", c_void_p * 3),
                ("b", c_void_p * 3),
                ("w5", c_void_p), ("b5", c_void_p), ("w5_16", c_void_p), ("poly", c_void_p),
                ("error_map", c_void_p),
                ("in_mid", ctypes.c_float * 5), ("in_inv_half", ctypes.c_float * 5),
                ("scale", ctypes.c_float * 3), ("edge_margin", ctypes.c_float),
                ("poly_exp", ctypes.c_uint32 * SURROGATE_POLY),
                ("num_targets", c_int32), ("num_apertures", c_int32),
                ("apertures", SurrogateAperture * SURROGATE_MAX_APERTURES)]


class SurrogateIO(ctypes.Structure):
    _fields_ = [("field_x", c_void_p), ("field_y", c_void_p), ("pupil_x", c_void_p),
                ("pupil_y", c_void_p), ("wavelength", c_void_p), ("x", c_void_p), ("y", c_void_p),
                ("intensity", c_void_p), ("unvignetted", c_void_p), ("fallback", c_void_p),
                ("fallback_list", c_void_p), ("fallback_count", c_void_p),
                ("fallback_capacity", c_int64), ("n", c_int64)]


MLP_BIAS_SILU = 0
MLP_DENSE_IN = 1
MLP_DENSE_OUT = 2
MLP_MSE_GRAD = 3
MLP_DENSE_OUT_BACK = 4
MLP_SILU_BACK = 5
MLP_OUTER_GRAD = 6
MLP_COLUMN_SUM = 7
MLP_ADAM = 8
MLP_TO_BF16 = 9
MLP_TRANSPOSE16 = 10

GEMM_BIAS_SILU = 1
GEMM_SILU_BACK = 2
BATCH_STATS = 8
BATCH_GRAPH = 2
# device scratch of optk_collect_noalloc (bytes, one array per GPU)
COLLECT_SCRATCH_BYTES = 64 << 20


class OptkError(RuntimeError):
    """An error reported by liboptk.so."""


_lib = None


def lib() -> ctypes.CDLL:
    """Load liboptk.so once, declare its prototypes and check record sizes."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise OptkError(
            f"{_LIB_PATH} is missing: build it with `python -m optika_amd._build` "
            "(the trace runs only on the GPU; there is no CPU fallback)"
        )
    L = ctypes.CDLL(str(_LIB_PATH))
    L.optk_abi_version.restype = c_int32
    L.optk_last_error.restype = ctypes.c_char_p
    L.optk_sizeof.restype = c_int64
    L.optk_sizeof.argtypes = [ctypes.c_char_p]
    L.optk_system_create.restype = c_int32
    L.optk_system_create.argtypes = [ctypes.POINTER(SystemDesc), ctypes.POINTER(c_void_p)]
    L.optk_system_destroy.restype = c_int32
    L.optk_system_destroy.argtypes = [c_void_p]
    L.optk_system_specialised.restype = c_int32
    L.optk_system_specialised.argtypes = [c_void_p]
    L.optk_trace.restype = c_int32
    L.optk_trace.argtypes = [c_void_p, c_int32, ctypes.POINTER(Rays), ctypes.POINTER(Rays),
                             ctypes.POINTER(Rays), c_int64, c_int32, c_void_p]
    L.optk_trace_grid.restype = c_int32
    L.optk_trace_grid.argtypes = [c_void_p, c_int32, ctypes.POINTER(Grid), ctypes.POINTER(Rays),
                                  ctypes.POINTER(Rays), c_int64, c_int32, c_void_p]
    L.optk_eval.restype = c_int32
    L.optk_eval.argtypes = [c_void_p, c_int32, ctypes.POINTER(EvalArgs), c_int64, c_void_p]
    L.optk_collect.restype = c_int32
    L.optk_collect.argtypes = [ctypes.POINTER(CollectArgs), c_int64, c_void_p]
    L.optk_collect_workspace.restype = ctypes.c_uint64
    L.optk_collect_workspace.argtypes = [ctypes.POINTER(CollectArgs), c_int64]
    L.optk_collect_noalloc.restype = c_int32
    L.optk_collect_noalloc.argtypes = [ctypes.POINTER(CollectArgs), c_int64, c_void_p]
    L.optk_batch_create.restype = c_int32
    L.optk_batch_create.argtypes = [ctypes.POINTER(BatchDesc), ctypes.POINTER(c_void_p)]
    L.optk_batch_destroy.restype = c_int32
    L.optk_batch_destroy.argtypes = [c_void_p]
    L.optk_batch_trace.restype = c_int32
    L.optk_batch_trace.argtypes = [c_void_p, c_int32, ctypes.POINTER(Grid), ctypes.POINTER(c_int64),
                                   ctypes.POINTER(Rays), c_void_p, ctypes.POINTER(Tune), c_int32,
                                   c_void_p]
    L.optk_batch_autotune.restype = c_int32
    L.optk_batch_autotune.argtypes = [c_void_p, c_int32, ctypes.POINTER(Grid), ctypes.POINTER(c_int64),
                                      c_void_p, c_int32, c_void_p, ctypes.POINTER(Tune)]
    L.optk_gemm_bf16.restype = c_int32
    L.optk_gemm_bf16.argtypes = [c_int32, c_int32, c_int64, c_int64, c_int64, c_void_p, c_int64,
                                 c_void_p, c_int64, c_void_p, c_int64, c_int32, c_void_p, c_void_p]
    L.optk_gemm_bf16_fused.restype = c_int32
    L.optk_gemm_bf16_fused.argtypes = [c_int32, c_int32, c_int64, c_int64, c_int64, c_void_p, c_int64,
                                       c_void_p, c_int64, c_void_p, c_int64, c_int32, c_void_p,
                                       c_void_p, c_void_p, c_void_p]
    L.optk_gemm_bf16_silu_back_sums.restype = c_int32
    L.optk_gemm_bf16_silu_back_sums.argtypes = [c_int64, c_void_p, c_int64, c_void_p, c_int64, c_void_p,
                                                c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    L.optk_mlp_op.restype = c_int32
    L.optk_mlp_op.argtypes = [c_int32, ctypes.POINTER(MlpArgs), c_void_p]
    L.optk_surrogate_infer.restype = c_int32
    L.optk_surrogate_infer.argtypes = [ctypes.POINTER(SurrogateDesc), ctypes.POINTER(SurrogateIO),
                                       c_void_p]
    L.optk_expose.restype = c_int32
    L.optk_expose.argtypes = [ctypes.POINTER(ExposeArgs), c_int64, c_void_p]
    L.optk_pupil_reduce.restype = c_int32
    L.optk_pupil_reduce.argtypes = [ctypes.POINTER(PupilArgs), c_int64, c_void_p]
    L.optk_denormalize.restype = c_int32
    L.optk_denormalize.argtypes = [ctypes.POINTER(DenormalizeArgs), c_int64, c_void_p]
    L.optk_stop_newton.restype = c_int32
    L.optk_stop_newton.argtypes = [c_void_p, c_void_p, ctypes.POINTER(StopArgs), c_int64, c_void_p]
    L.optk_poly_workspace.restype = ctypes.c_size_t
    L.optk_poly_workspace.argtypes = [c_int64]
    L.optk_poly_normal.restype = c_int32
    L.optk_poly_normal.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_void_p,
                                   c_void_p, ctypes.c_size_t, c_void_p]
    L.optk_poly_eval.restype = c_int32
    L.optk_poly_eval.argtypes = [c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_void_p, c_void_p]
    if L.optk_abi_version() != 1:
        raise OptkError(f"liboptk ABI {L.optk_abi_version()} != 1")
    for name, cls in [("optk_affine", Affine), ("optk_surface", Surface), ("optk_layer", Layer),
                      ("optk_block", Block), ("optk_stack", Stack), ("optk_chemical", Chemical),
                      ("optk_system_desc", SystemDesc), ("optk_rays", Rays), ("optk_grid", Grid),
                      ("optk_eval_args", EvalArgs), ("optk_collect_args", CollectArgs),
                      ("optk_batch_desc", BatchDesc), ("optk_tune", Tune),
                      ("optk_mlp_args", MlpArgs), ("optk_surrogate_desc", SurrogateDesc),
                      ("optk_surrogate_io", SurrogateIO), ("optk_stop_args", StopArgs),
                      ("optk_pupil_args", PupilArgs), ("optk_denormalize_args", DenormalizeArgs),
                      ("optk_expose_args", ExposeArgs)]:
        size = L.optk_sizeof(name.encode())
        if size != ctypes.sizeof(cls):
            raise OptkError(f"record {name}: C size {size} != binding size {ctypes.sizeof(cls)}")
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().optk_last_error().decode(errors="replace")
        raise OptkError(f"liboptk error {rc}: {msg}")


def library_path() -> pathlib.Path:
    return _LIB_PATH

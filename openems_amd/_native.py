"""ctypes binding of ``lib/libopenems_hip.so`` (C ABI in include/ems_operator.h
and include/ems_hip.h).  Importing this module never falls back to anything:
a missing library raises ImportError naming the build command."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libopenems_hip.so")

vp = C.c_void_p
fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)
up = C.POINTER(C.c_uint)
ip = C.POINTER(C.c_int)
usp = C.POINTER(C.c_ushort)


class ProbeDesc(C.Structure):
    _fields_ = [("type", C.c_int), ("nelem", C.c_uint), ("comp", up), ("pos", up), ("sign", ip),
                ("divisor", dp), ("interval", C.c_uint), ("fd_interval", C.c_uint),
                ("start_ts", C.c_uint), ("stop_ts", C.c_uint), ("capacity", C.c_uint)]


# (name, restype, argtypes) for every exported symbol
OPERATOR_API = [
    ("emsop_last_error", C.c_char_p, []),
    ("emsop_create", C.c_int, [C.POINTER(vp)]),
    ("emsop_destroy", None, [vp]),
    ("emsop_set_grid", C.c_int, [vp, dp, C.c_uint, dp, C.c_uint, dp, C.c_uint, C.c_double]),
    ("emsop_set_background", C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double]),
    ("emsop_add_material", C.c_int, [vp, dp, dp, dp, dp, C.c_int, dp, dp]),
    ("emsop_add_material_grid", C.c_int, [vp, up, up, dp, dp, C.c_double, C.c_double, C.c_int]),
    ("emsop_eps_weights", C.c_int, [vp, C.c_int, up, up, dp]),
    ("emsop_add_metal", C.c_int, [vp, C.c_int, dp, dp]),
    ("emsop_add_metal_curve", C.c_int, [vp, C.c_uint, dp]),
    ("emsop_add_excitation_curve", C.c_int, [vp, C.c_int, dp, C.c_double, C.c_uint, dp]),
    ("emsop_add_metal_cylinder", C.c_int, [vp, C.c_int, dp, dp, C.c_double, C.c_double]),
    ("emsop_add_excitation_cylinder", C.c_int, [vp, C.c_int, dp, C.c_double, C.c_int, dp, dp, C.c_double, C.c_double]),
    ("emsop_set_excitation_weight", C.c_int, [vp, C.POINTER(C.c_char_p)]),
    ("emsop_eval_function", C.c_int, [C.c_char_p, C.c_char_p, C.c_uint, dp, C.c_uint, dp]),
    ("emsop_add_lumped", C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, dp, dp]),
    ("emsop_add_excitation", C.c_int, [vp, C.c_int, dp, C.c_double, C.c_int, dp, dp]),
    ("emsop_set_boundary", C.c_int, [vp, ip, up, dp]),
    ("emsop_set_signal", C.c_int, [vp, C.c_int, C.c_double, C.c_double]),
    ("emsop_set_custom_signal", C.c_int, [vp, C.c_char_p, C.c_double]),
    ("emsop_set_timestep", C.c_int, [vp, C.c_int, C.c_double, C.c_double]),
    ("emsop_set_cell_constant_material", C.c_int, [vp, C.c_int]),
    ("emsop_set_threads", C.c_int, [vp, C.c_int]),
    ("emsop_build", C.c_int, [vp, C.c_uint]),
    ("emsop_num_lines", C.c_int, [vp, up]),
    ("emsop_timestep", C.c_double, [vp]),
    ("emsop_nyquist", C.c_uint, [vp]),
    ("emsop_signal_period", C.c_double, [vp]),
    ("emsop_signal_length", C.c_uint, [vp]),
    ("emsop_signal_period_steps", C.c_int, [vp]),
    ("emsop_signal_volt", fp, [vp]),
    ("emsop_signal_curr", fp, [vp]),
    ("emsop_coeff", fp, [vp, C.c_int]),
    ("emsop_ec", fp, [vp, C.c_int]),
    ("emsop_effmat_at", C.c_int, [vp, C.c_int, up, dp]),
    ("emsop_exc_count", C.c_uint, [vp, C.c_int]),
    ("emsop_exc_get", C.c_int, [vp, C.c_int, up, usp, fp, up]),
    ("emsop_upml_count", C.c_uint, [vp]),
    ("emsop_upml_get", C.c_int, [vp, C.c_uint, up, up, C.POINTER(fp)]),
    ("emsop_set_dispersion", C.c_int, [vp, C.c_uint, dp]),
    ("emsop_set_debye", C.c_int, [vp, C.c_uint, dp]),
    ("emsop_lorentz_orders", C.c_uint, [vp]),
    ("emsop_lorentz_get", C.c_int, [vp, C.c_uint, up, ip, C.POINTER(up), C.POINTER(fp)]),
    ("emsop_rlc_count", C.c_uint, [vp]),
    ("emsop_rlc_get", C.c_int, [vp, C.POINTER(up), C.POINTER(ip), C.POINTER(fp)]),
    ("emsop_mur_count", C.c_uint, [vp]),
    ("emsop_mur_get", C.c_int, [vp, C.c_uint, ip, up, up, up, C.POINTER(fp), C.POINTER(fp)]),
    ("emsop_disc_line", C.c_double, [vp, C.c_int, C.c_uint, C.c_int]),
    ("emsop_edge_length", C.c_double, [vp, C.c_int, C.c_uint, C.c_int]),
    ("emsop_snap_to_mesh", C.c_int, [vp, dp, C.c_int, up, ip]),
    ("emsop_snap_box", C.c_int, [vp, dp, dp, C.c_int, up, up, ip, ip]),
    ("emsop_find_path", C.c_int, [vp, dp, dp, C.c_uint, up, ip, up]),
    ("emsop_release_build_arrays", None, [vp]),
]

ENGINE_API = [
    ("ems_last_error", C.c_char_p, []),
    ("ems_device_count", C.c_int, []),
    ("ems_create", C.c_int, [C.POINTER(vp), up, fp, fp, fp, fp, C.c_int]),
    ("ems_create_layout", C.c_int, [C.POINTER(vp), up, fp, fp, fp, fp, C.c_int, C.c_int]),
    ("ems_create_sse_compressed", C.c_int, [C.POINTER(vp), up, up, C.c_uint, C.POINTER(fp), C.POINTER(fp),
                                            C.POINTER(fp), C.POINTER(fp), C.c_int, C.c_int]),
    ("ems_compress_operator", C.c_int, [up, fp, fp, fp, fp, usp, fp, fp, C.c_uint]),
    ("ems_expand_sse_compressed", C.c_int, [up, up, C.c_uint, C.POINTER(fp), C.POINTER(fp), C.POINTER(fp),
                                            C.POINTER(fp), fp, fp, fp, fp]),
    ("ems_get_layout", C.c_int, [vp, ip, up, ip, up]),
    ("ems_set_fused", C.c_int, [vp, C.c_int]),
    ("ems_set_steady_state", C.c_int, [vp, C.c_uint, C.c_uint, ip, up, C.c_uint]),
    ("ems_steady_state", C.c_int, [vp, dp, up, up, dp, C.c_uint]),
    ("ems_get_fused", C.c_int, [vp, ip, up]),
    ("ems_read_field", C.c_int, [vp, C.c_int, C.c_int, up, up, fp]),
    ("ems_write_field", C.c_int, [vp, C.c_int, C.c_int, up, up, fp]),
    ("ems_get_field_at", C.c_int, [vp, C.c_int, C.c_int, up, C.POINTER(dp), dp]),
    ("ems_set_quad", C.c_int, [vp, C.c_int]),
    ("ems_get_quad", C.c_int, [vp, ip, up]),
    ("ems_set_excitation", C.c_int, [vp, fp, fp, C.c_uint, C.c_int, C.c_uint, up, usp, fp, up,
                                     C.c_uint, up, usp, fp, up]),
    ("ems_add_upml", C.c_int, [vp, up, up, fp, fp, fp, fp, fp, fp]),
    ("ems_add_mur", C.c_int, [vp, C.c_int, C.c_uint, C.c_uint, fp, fp, C.c_uint]),
    ("ems_add_lumped_rlc", C.c_int, [vp, C.c_uint, up, ip, C.POINTER(fp)]),
    ("ems_add_lorentz", C.c_int, [vp, C.c_uint, up, fp, fp, fp, fp, fp, fp, C.c_int, C.c_int]),
    ("ems_init", C.c_int, [vp]),
    ("ems_iterate", C.c_int, [vp, C.c_uint]),
    ("ems_get_numts", C.c_uint, [vp]),
    ("ems_set_async", C.c_int, [vp, C.c_int]),
    ("ems_set_graph", C.c_int, [vp, C.c_uint]),
    ("ems_get_graph", C.c_int, [vp, up, up, up]),
    ("ems_sync", C.c_int, [vp]),
    ("ems_energy_noalloc", C.c_int, [vp, C.c_void_p]),
    ("ems_get_volt", C.c_int, [vp, fp]),
    ("ems_get_curr", C.c_int, [vp, fp]),
    ("ems_set_volt", C.c_int, [vp, fp]),
    ("ems_set_curr", C.c_int, [vp, fp]),
    ("ems_add_probe", C.c_int, [vp, C.POINTER(ProbeDesc), ip]),
    ("ems_probe_results", C.c_int, [vp, C.c_int, up, up, dp]),
    ("ems_add_fd_dump", C.c_int, [vp, C.c_int, C.c_int, up, up, C.c_uint, C.c_uint, C.c_uint, C.c_uint, fp,
                                  C.POINTER(dp), ip]),
    ("ems_add_td_dump", C.c_int, [vp, C.c_int, C.c_int, up, up, C.c_uint, C.c_uint, C.c_uint, C.c_uint,
                                  C.POINTER(dp), ip]),
    ("ems_td_dump_results", C.c_int, [vp, C.c_int, up, up, fp]),
    ("ems_fd_dump_results", C.c_int, [vp, C.c_int, fp]),
    ("ems_energy", C.c_int, [vp, dp]),
    ("ems_last_timing", C.c_int, [vp, dp, dp, dp]),
    ("ems_set_profiling", C.c_int, [vp, C.c_int]),
    ("ems_destroy", None, [vp]),
    ("ems_batch_create", C.c_int, [C.POINTER(vp), C.POINTER(vp), C.c_uint]),
    ("ems_batch_set_config", C.c_int, [vp, C.c_uint, up, C.c_uint]),
    ("ems_batch_get_config", C.c_int, [vp, up, up, up, up]),
    ("ems_batch_autotune", C.c_int, [vp, C.c_uint, dp]),
    ("ems_batch_set_grouping", C.c_int, [vp, C.c_ulonglong]),
    ("ems_batch_get_grouping", C.c_int, [vp, C.POINTER(C.c_ulonglong), up]),
    ("ems_batch_iterate", C.c_int, [vp, C.c_uint]),
    ("ems_batch_set_graph", C.c_int, [vp, C.c_uint]),
    ("ems_batch_get_graph", C.c_int, [vp, up]),
    ("ems_batch_set_streams", C.c_int, [vp, C.c_uint]),
    ("ems_batch_get_streams", C.c_int, [vp, up]),
    ("ems_batch_last_ms", C.c_int, [vp, dp]),
    ("ems_batch_destroy", None, [vp]),
    ("ems_adj_create", C.c_int, [C.POINTER(vp), vp, C.c_uint, up, up, dp, C.c_double, C.c_double, up, up, C.c_uint]),
    ("ems_adj_forward", C.c_int, [vp, C.c_uint, dp]),
    ("ems_adj_backward", C.c_int, [vp, dp, dp]),
    ("ems_adj_destroy", None, [vp]),
]

_LIB = None


def lib():
    """Load the native library (raises ImportError when it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing; build it with `python -m openems_amd.build` "
                          "or __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    for name, res, args in OPERATOR_API + ENGINE_API:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


class NativeError(RuntimeError):
    pass


def check_op(rc):
    if rc < 0:
        raise NativeError(f"operator error {rc}: {lib().emsop_last_error().decode()}")
    return rc


def check(rc):
    if rc < 0:
        raise NativeError(f"engine error {rc}: {lib().ems_last_error().decode()}")
    return rc


def ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))

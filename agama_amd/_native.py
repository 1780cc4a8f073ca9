"""ctypes binding of libagama_amd.so (the C ABI declared in include/agama_amd.h).

The library is built in-tree by agama_amd.build; importing this module loads it and fails
loudly when it is missing - there is no CPU fallback for the product path."""
import ctypes
import os

# PyTorch-ROCm bundles its own HIP runtime under the same soname as the system one.  If this
# library were loaded first, torch would later load its copy as a second runtime in the same
# process, which then finds no GPU; loading torch first makes both use one runtime.
try:
    import torch  # noqa: F401
except ImportError:
    pass

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libagama_amd.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p


class AgamaAmdError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"agama_amd native library not found at {LIB_PATH}; build it with "
            "`python agama_amd/build.py` (hipcc for gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "agx_last_error": (ctypes.c_char_p, []),
        "agx_potential_create": (_vp, [ctypes.c_char_p]),
        "agx_potential_from_coefs": (_vp, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]),
        "agx_potential_destroy": (None, [_vp]),
        "agx_potential_write": (ctypes.c_int, [_vp, ctypes.c_char_p]),
        "agx_build_coef_file": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p]),
        "agx_debug_tables": (ctypes.c_long, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _vp]),
        "agx_potential_coefs": (ctypes.c_int, [_vp, _ip, _ip, _vp, _vp, _vp]),
        "agx_eval_potential": (ctypes.c_double, [_vp, _vp, ctypes.c_double, _vp, _vp]),
        "agx_eval_many": (ctypes.c_int, [_vp, ctypes.c_long, _vp, ctypes.c_double, _vp, _vp, ctypes.c_int]),
        "agx_eval_many_device": (ctypes.c_int, [_vp, ctypes.c_long, _vp, ctypes.c_double, _vp, _vp,
                                                ctypes.c_int, _vp]),
        "agx_orbit_var_batch": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp,
                                               _vp, _vp, _vp, _vp, ctypes.c_int]),
        "agx_orbit_var_batch_device": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp,
                                                      ctypes.c_double, ctypes.c_double, ctypes.c_long, _vp, _vp,
                                                      _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
        "agx_lyapunov_fit": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp]),
        "agx_eval_hess_many": (ctypes.c_int, [_vp, ctypes.c_long, _vp, ctypes.c_double, _vp, _vp, _vp,
                                              ctypes.c_int]),
        "agx_eval_hess_many_device": (ctypes.c_int, [_vp, ctypes.c_long, _vp, ctypes.c_double, _vp, _vp, _vp,
                                                     ctypes.c_int, _vp]),
        "agx_orbit_batch": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp,
                                           _vp, _vp, ctypes.c_int]),
        "agx_orbit_batch_device": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp,
                                                  ctypes.c_double, ctypes.c_double, ctypes.c_long,
                                                  _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
        "agx_orbit_batch_device_noalloc": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp,
                                                          ctypes.c_double, ctypes.c_double, ctypes.c_long,
                                                          _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int,
                                                          _vp]),
        "agx_orbit_record_batch": (_vp, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_long, ctypes.c_int, _vp, _vp, _vp, _vp,
                                         _vp, _vp, _vp, ctypes.c_int]),
        "agx_orbit_record_fetch": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
        "agx_orbit_record_num_series": (ctypes.c_long, [_vp]),
        "agx_orbit_record_series_info": (ctypes.c_int, [_vp, _vp, _vp]),
        "agx_orbit_record_node_times": (ctypes.c_int, [_vp, ctypes.c_long, _vp]),
        "agx_orbit_record_nodes": (ctypes.c_int, [_vp, ctypes.c_long, _vp]),
        "agx_orbit_record_interp": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, ctypes.c_int, _vp]),
        "agx_orbit_record_free": (None, [_vp]),
        "agx_set_record_rows": (ctypes.c_int, [ctypes.c_long]),
        "agx_set_construction": (ctypes.c_int, [ctypes.c_int]),
        "agx_set_surrogate_pilot": (ctypes.c_int, [ctypes.c_int]),
        "agx_surrogate_last_pilot": (ctypes.c_int, [_vp, _vp]),
        "agx_target_density_sphharm": (_vp, [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int]),
        "agx_target_num_values": (ctypes.c_int, [_vp]),
        "agx_target_density_cylindrical": (_vp, [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int]),
        "agx_target_kinem_shell": (_vp, [ctypes.c_int, _vp, ctypes.c_int]),
        "agx_target_density_classic": (_vp, [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double]),
        "agx_target_destroy": (None, [_vp]),
        "agx_orbit_target_batch": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_long, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]),
        "agx_orbit_target_traj_batch": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_long, ctypes.c_int, _vp, _vp, _vp, _vp,
                                                   _vp, _vp, _vp, ctypes.c_int]),
        "agx_set_orbit_launch": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
        "agx_set_fetch": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
        "agx_surrogate_train": (_vp, [_vp, ctypes.c_char_p]),
        "agx_surrogate_create": (_vp, [_vp, _vp, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_double, ctypes.c_double]),
        "agx_surrogate_destroy": (None, [_vp]),
        "agx_surrogate_info": (ctypes.c_int, [_vp, _vp, _vp]),
        "agx_surrogate_loss_history": (ctypes.c_long, [_vp, _vp, ctypes.c_long]),
        "agx_surrogate_eval_many": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp]),
        "agx_surrogate_eval_many_device": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp]),
        "agx_surrogate_network": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp]),
        "agx_surrogate_train_grad": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp]),
        "agx_surrogate_orbit_batch": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, ctypes.c_double,
                                                     ctypes.c_double, ctypes.c_long, ctypes.c_double, _vp, _vp,
                                                     _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "agx_surrogate_orbit_batch_device": (ctypes.c_int, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp,
                                                            ctypes.c_double, ctypes.c_double, ctypes.c_long,
                                                            ctypes.c_double, _vp, _vp, _vp, _vp, _vp, _vp,
                                                            _vp, _vp, _vp, _vp]),
        "agx_debug_gemm": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]),
        "agx_set_orbit_waves": (ctypes.c_int, [ctypes.c_int]),
        "agx_set_orbit_groups": (ctypes.c_int, [ctypes.c_int, ctypes.c_long, ctypes.c_long]),
        "agx_set_orbit_probe": (ctypes.c_int, [ctypes.c_int, ctypes.c_long]),
        "agx_get_orbit_schedule": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "agx_get_orbit_config": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
        "agx_orbit_graph_create": (_vp, [_vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                         _vp, ctypes.c_int]),
        "agx_orbit_graph_launch": (ctypes.c_int, [_vp, _vp]),
        "agx_orbit_graph_destroy": (None, [_vp]),
        "agama_getError": (ctypes.c_char_p, []),
        "agama_createPotential": (_vp, [ctypes.c_char_p]),
        "agama_deletePotential": (None, [_vp]),
        "agama_evalPotential": (ctypes.c_double, [_vp, _vp, ctypes.c_double, _vp, _vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()

EXPORTED = ["agx_last_error", "agx_potential_create", "agx_potential_from_coefs",
            "agx_potential_destroy", "agx_potential_write", "agx_build_coef_file",
            "agx_debug_tables", "agx_potential_coefs",
            "agx_eval_potential", "agx_eval_many", "agx_eval_many_device", "agx_eval_hess_many",
            "agx_eval_hess_many_device", "agx_orbit_var_batch", "agx_orbit_var_batch_device", "agx_lyapunov_fit", "agx_orbit_batch",
            "agx_orbit_batch_device", "agx_orbit_batch_device_noalloc", "agx_orbit_record_batch", "agx_orbit_record_fetch",
            "agx_orbit_record_num_series", "agx_orbit_record_series_info", "agx_orbit_record_node_times",
            "agx_orbit_record_interp", "agx_orbit_record_free", "agx_set_record_rows", "agx_set_construction", "agx_set_surrogate_pilot", "agx_surrogate_last_pilot", "agx_target_density_sphharm", "agx_target_density_cylindrical", "agx_target_kinem_shell", "agx_target_density_classic", "agx_target_num_values",
            "agx_target_destroy", "agx_orbit_target_batch", "agx_orbit_target_traj_batch", "agx_set_orbit_launch", "agx_set_fetch", "agx_set_orbit_groups", "agx_set_orbit_probe", "agx_get_orbit_schedule", "agx_orbit_record_nodes", "iniacceleration",
            "inipotential", "potential_double", "potential_float",
            "agx_surrogate_train", "agx_surrogate_create", "agx_surrogate_destroy",
            "agx_surrogate_info", "agx_surrogate_loss_history", "agx_surrogate_eval_many",
            "agx_surrogate_eval_many_device", "agx_surrogate_network", "agx_surrogate_train_grad",
            "agx_surrogate_orbit_batch", "agx_surrogate_orbit_batch_device", "agx_debug_gemm",
            "agama_getError", "agama_createPotential", "agama_deletePotential",
            "agama_evalPotential"]


def last_error():
    msg = lib.agx_last_error()
    return msg.decode() if msg else ""


def check(rc, what):
    if rc != 0:
        raise AgamaAmdError(f"{what} failed (code {rc}): {last_error()}")


def ptr(a):
    """data pointer of a numpy array (or None)"""
    return None if a is None else a.ctypes.data

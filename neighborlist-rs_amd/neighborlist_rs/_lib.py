"""ctypes binding of libnlmi355 (include/nl_api.h, include/nl_engine_ext.h).

The shared library is built in-tree by neighborlist-rs_amd/build.py.  There
is no CPU fallback: if the library is missing, importing this module raises.
"""
import ctypes as C
import importlib.util
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnlmi355.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libnlmi355.so not found at {LIB_PATH}: build it with "
        "`python neighborlist-rs_amd/build.py` (the engine has no CPU fallback)")


def _torch_runtime():
    """Binds the process to torch's bundled HIP runtime when torch is
    installed, without importing torch.

    torch ships its own libamdhip64 / libhsa-runtime64 (sonames
    libamdhip64.so.7 / libhsa-runtime64.so.1, the ones this library needs)
    and loads them by file name from its own directory.  Loaded here first
    and globally, they satisfy this library's dependencies by soname, and a
    later `import torch` finds the same files already mapped: the process
    holds one HIP runtime whichever package is imported first.  Without this,
    loading the library first maps the system runtime and torch then maps a
    second one, which finds no device."""
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return []
    if spec is None or not spec.submodule_search_locations:
        return []
    tlib = os.path.join(list(spec.submodule_search_locations)[0], "lib")
    held = []
    for name in ("libhsa-runtime64.so", "libamdhip64.so"):
        path = os.path.join(tlib, name)
        if os.path.exists(path):
            held.append(C.CDLL(path, mode=C.RTLD_GLOBAL))
    return held


_runtime = _torch_runtime()
lib = C.CDLL(LIB_PATH)

dp = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
fp = C.POINTER(C.c_float)
res_p = C.c_void_p
res_pp = C.POINTER(C.c_void_p)

_SIGS = {
    # nl_api.h
    "nl_search_single": (C.c_int, [dp, C.c_size_t, dp, u8p, C.c_double, C.c_int, res_pp]),
    "nl_search_single_multi": (C.c_int, [dp, C.c_size_t, dp, u8p, dp, C.c_size_t, C.c_int,
                                         res_pp]),
    "nl_search_batch": (C.c_int, [dp, C.c_size_t, i32p, C.c_size_t, dp, u8p, u8p, C.c_double,
                                  C.c_int, res_pp]),
    "nl_search_batch_multi": (C.c_int, [dp, C.c_size_t, i32p, C.c_size_t, dp, u8p, u8p, dp,
                                        C.c_size_t, C.c_int, res_pp]),
    "nl_result_num_lists": (C.c_size_t, [res_p]),
    "nl_result_num_edges": (C.c_int64, [res_p, C.c_size_t]),
    "nl_result_copy": (C.c_int, [res_p, C.c_size_t, i64p, i64p, i32p]),
    "nl_result_device": (C.c_int, [res_p, C.c_size_t, C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_void_p)]),
    "nl_result_free": (None, [res_p]),
    "nl_get_brute_force_threshold": (C.c_size_t, []),
    "nl_set_brute_force_threshold": (None, [C.c_size_t]),
    "nl_get_parallel_threshold": (C.c_size_t, []),
    "nl_set_parallel_threshold": (None, [C.c_size_t]),
    "nl_get_stack_threshold": (C.c_size_t, []),
    "nl_set_stack_threshold": (None, [C.c_size_t]),
    "nl_last_error": (C.c_char_p, []),
    "nl_version": (C.c_char_p, []),
    # nl_engine_ext.h
    "nl_search_batch_device": (C.c_int, [C.c_void_p, C.c_size_t, i64p, C.c_size_t, dp, u8p, u8p,
                                         dp, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_int, res_pp]),
    "nl_search_systems": (C.c_int, [dp, C.c_size_t, i64p, C.c_size_t, dp, u8p, u8p, dp,
                                    C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    res_pp]),
    "nl_search_ex": (C.c_int, [C.c_void_p, C.c_size_t, i64p, C.c_size_t, dp, u8p, u8p, dp,
                               C.c_size_t, C.c_int, C.c_int, C.c_uint, C.c_void_p, res_pp]),
    "nl_result_dlpack": (C.c_int, [res_p, C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]),
    "nl_result_system_ranges": (C.c_int, [res_p, i64p, i32p]),
    "nl_result_copy_dist": (C.c_int, [res_p, C.c_size_t, dp]),
    "nl_result_copy_disp": (C.c_int, [res_p, C.c_size_t, dp]),
    "nl_set_timing": (None, [C.c_int]),
    "nl_last_timing": (C.c_int, [fp, fp, fp, fp, i64p, i64p]),
    "nl_set_cell_div": (None, [C.c_int]),
    "nl_get_cell_div": (C.c_int, []),
    "nl_set_small_path": (None, [C.c_int]),
    "nl_get_small_path": (C.c_int, []),
    "nl_set_bin_order": (None, [C.c_int]),
    "nl_set_fused_build": (None, [C.c_int]),
    "nl_get_fused_build": (C.c_int, []),
    "nl_get_bin_order": (C.c_int, []),
    "nl_set_fp32_filter": (None, [C.c_int]),
    "nl_get_fp32_filter": (C.c_int, []),
    "nl_set_work_split": (None, [C.c_int, C.c_int]),
    "nl_get_work_split": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "nl_set_autotune": (None, [C.c_int]),
    "nl_get_autotune": (C.c_int, []),
    "nl_clear_tuning": (None, []),
    "nl_tune_report": (C.c_size_t, [C.c_char_p, C.c_size_t]),
    "nl_plan_create": (C.c_int, [C.c_void_p, C.c_size_t, i64p, C.c_size_t, dp, u8p, u8p, dp,
                                 C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int64,
                                 C.POINTER(C.c_void_p)]),
    "nl_plan_run": (C.c_int, [C.c_void_p]),
    "nl_plan_num_edges": (C.c_int64, [C.c_void_p, C.c_size_t]),
    "nl_plan_device": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_void_p)]),
    "nl_plan_copy": (C.c_int, [C.c_void_p, C.c_size_t, i64p, i64p, i32p]),
    "nl_plan_stats": (C.c_int, [C.c_void_p, i64p, i64p, i64p]),
    "nl_plan_free": (None, [C.c_void_p]),
    "nl_plan_create_ex": (C.c_int, [C.c_void_p, C.c_size_t, i64p, C.c_size_t, dp, u8p, u8p, dp,
                                    C.c_size_t, C.c_int, C.c_int, C.c_uint, C.c_int64,
                                    C.POINTER(C.c_void_p)]),
    "nl_plan_array": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]),
    "nl_plan_system_ranges": (C.c_int, [C.c_void_p, i64p, i32p]),
    "nl_server_create": (C.c_void_p, [C.c_int, C.c_int64, C.c_int64, C.c_int64]),
    "nl_server_submit": (C.c_int64, [C.c_void_p, dp, C.c_size_t, dp, u8p, C.c_double]),
    "nl_server_submit_ex": (C.c_int64, [C.c_void_p, dp, C.c_size_t, dp, u8p, C.c_double,
                                        C.c_uint]),
    "nl_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_double]),
    "nl_server_result_info": (C.c_int, [C.c_void_p, C.c_int64, i64p, C.POINTER(C.c_double),
                                        C.POINTER(C.c_int)]),
    "nl_server_take": (C.c_int, [C.c_void_p, C.c_int64, i64p, i64p, i32p]),
    "nl_server_result_device": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p),
                                          C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "nl_server_release": (C.c_int, [C.c_void_p, C.c_int64]),
    "nl_server_stats": (C.c_int, [C.c_void_p, i64p, i64p, i64p, i64p, i64p]),
    "nl_server_allocations": (C.c_int, [C.c_void_p, i64p]),
    "nl_server_free": (None, [C.c_void_p]),
    "nl_server_create_ex": (C.c_void_p, [C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_uint]),
    "nl_server_take_ex": (C.c_int, [C.c_void_p, C.c_int64, i64p, i64p, i32p, dp, dp]),
    "nl_server_result_arrays": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p),
                                          C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                          C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "nl_skin_create": (C.c_int, [C.c_void_p, C.c_size_t, i64p, C.c_size_t, dp, u8p, u8p, dp, dp,
                                 C.c_size_t, C.c_int, C.c_int, C.c_uint, C.c_void_p,
                                 C.POINTER(C.c_void_p)]),
    "nl_skin_update": (C.c_int, [C.c_void_p, C.c_void_p, dp, C.c_void_p, C.POINTER(C.c_int)]),
    "nl_skin_check": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, dp, C.POINTER(C.c_int)]),
    "nl_skin_num_edges": (C.c_int64, [C.c_void_p, C.c_size_t]),
    "nl_skin_device": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_void_p)]),
    "nl_skin_stats": (C.c_int, [C.c_void_p, i64p, i64p, dp]),
    "nl_skin_free": (None, [C.c_void_p]),
    "nl_search_noalloc": (C.c_int, [C.c_void_p, C.c_size_t, i64p, C.c_size_t, dp, u8p, u8p, dp,
                                    C.c_size_t, C.c_int, C.c_int, C.c_uint, C.c_void_p,
                                    C.c_void_p, C.c_int64, i64p, i64p, C.c_void_p]),
    "nl_cell_inverse": (C.c_int, [dp, dp]),
    "nl_cell_wrap": (C.c_int, [dp, u8p, dp, dp]),
    "nl_device_malloc": (C.c_void_p, [C.c_size_t]),
    "nl_host_alloc": (C.c_void_p, [C.c_size_t]),
    "nl_host_free": (None, [C.c_void_p]),
    "nl_device_free": (C.c_int, [C.c_void_p]),
    "nl_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "nl_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "nl_memcpy_d2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "nl_device_synchronize": (C.c_int, []),
    "nl_set_device": (C.c_int, [C.c_int]),
    "nl_device_count": (C.c_int, []),
}

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def exported_symbols():
    return sorted(_SIGS)


def last_error() -> str:
    return lib.nl_last_error().decode()

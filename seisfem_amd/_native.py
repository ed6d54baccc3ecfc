"""ctypes bindings of libseisfem_amd.so (C ABI declared in include/seisfem_host.h
and include/seisfem_gpu.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
seisfem_amd/csrc``).  There is no fallback: importing a GPU entry point without
the library raises immediately.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# SEISFEM_AMD_LIB selects another build of the same library (the sanitized host
# build of tests/sanitize.sh)
LIB_PATH = os.environ.get("SEISFEM_AMD_LIB") or os.path.join(_HERE, "libseisfem_amd.so")

_i = C.c_int
_d = C.c_double
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)
_pf = C.POINTER(C.c_float)
_vp = C.c_void_p


class ShotDesc(C.Structure):
    """Mirror of sfg_shot_desc (include/seisfem_gpu.h)."""
    _fields_ = [
        ("precision", _i), ("nshots", _i), ("src_node", _pi), ("nsteps", _pi),
        ("max_steps", _i), ("rec_num", _i), ("rec_node", _pi),
        ("dt", _d), ("f0", _d), ("t0", _d), ("amplitude", _d), ("angle_deg", _d),
        ("block", _i), ("shots_per_lane_group", _i), ("guard_every", _i), ("use_graph", _i),
        ("xcd_remap", _i), ("src_count", _pi), ("src_nodes", _pi), ("src_weight", _pd),
        ("kernel_variant", _i), ("row_order", _pi), ("sell_sigma", _i),
    ]


_lib = None


def lib():
    """Loads libseisfem_amd.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first "
                           "(there is no CPU fallback for the GPU path)")
    L = C.CDLL(LIB_PATH)
    sig = {
        # host setup
        "sfh_element_order": (_i, [_i]),
        "sfh_mesh_node_num": (_i, [_i, _i, _i]),
        "sfh_mesh_element_num": (_i, [_i, _i, _i]),
        "sfh_mesh_element": (_i, [_i, _i, _i, _pi]),
        "sfh_mesh_xy": (_i, [_i, _i, _i, _d, _d, _d, _d, _pd, _pd]),
        "sfh_assemble_coo": (_i, [_i, _i, _i, _pi, _pd, _pd, _pd, _pi, _pi, C.POINTER(_pd)]),
        "sfh_lumped_mass": (_i, [_i, _i, _i, _pi, _pd, _pd, _pd, _pd]),
        "sfh_coo2csr": (_i, [_i, _i, _pi, _pi, _pd, _pi, _pi, _pd, _pi]),
        "__coo2csr_lib_MOD_coo2csr_canonical": (None, [_pi, _pi, _pi, _pi, _pi, _pd, _pi, _pi, _pd]),
        "sfh_node_location": (_i, [_i, _d, _pd, _pd, _d, _d]),
        "sfh_station_line": (None, [_i, _d, _d, _d, _d, _pd, _pd]),
        "sfh_dirichlet_flags": (None, [_i, _pd, _pd, _d, _d, _d, _d, _d, _i, _pi]),
        "sfh_locality_order": (_i, [_i, _pd, _pd, _pi]),
        "sfh_csr_union": (_i, [_i, _i, C.POINTER(_pi), C.POINTER(_pi), C.POINTER(_pd), _i, _pi, _pi,
                               C.POINTER(_pd)]),
        # GPU
        "sfg_last_error": (C.c_char_p, []),
        "sfg_set_device": (_i, [_i]),
        "csr_matvec": (None, [_i, _pi, _pi, _pd, _pd, _pd]),
        "abc_mpml": (None, [_i, _i, _i, _pi, C.POINTER(_pd), _i, _i, _d, _d, _d, _d, _d, _d,
                            _i, _i, _i, _i, _pd, _pd, _pd, _pd, _pd, _pd]),
        "seismic_source": (_d, [_d, _d, _d, _d]),
        "sfg_operator_create": (_i, [C.POINTER(_vp), _i, C.POINTER(_pi), C.POINTER(_pi),
                                     C.POINTER(_pd)]),
        "sfg_operator_nnz": (_i, [_vp]),
        "sfg_operator_destroy": (None, [_vp]),
        "sfg_medium_create": (_i, [C.POINTER(_vp), _vp, C.POINTER(_pd), _pd, _pi, C.POINTER(_pd)]),
        "sfg_medium_destroy": (None, [_vp]),
        "sfg_medium_set_row_scale": (_i, [_vp, _pd]),
        "sfg_run_create": (_i, [C.POINTER(_vp), _vp, _vp, C.POINTER(ShotDesc)]),
        "sfg_run_create_media": (_i, [C.POINTER(_vp), _vp, C.POINTER(_vp), C.POINTER(ShotDesc)]),
        "sfg_run_advance": (_i, [_vp, _i]),
        "sfg_run_sync": (_i, [_vp]),
        "sfg_run_reset": (_i, [_vp]),
        "sfg_run_seismograms": (_i, [_vp, _pd, _pd]),
        "sfg_run_snapshot": (_i, [_vp, _i, _pd, _pd]),
        "sfg_run_snapshot_noalloc": (_i, [_vp, _i, _pd, _pd]),
        "sfg_run_energy": (_i, [_vp, _pd, _i]),
        "sfg_run_last_ms_per_step": (_d, [_vp]),
        "sfg_run_step_bytes": (_d, [_vp, _i]),
        "sfg_run_destroy": (None, [_vp]),
        "sfg_estimate_dt": (_i, [_vp, _vp, _i, _d, _pd, _pd]),
        "sfg_stream_copy": (_i, [C.c_size_t, _i, _pd, _pd]),
        "sfg_multi_create": (_i, [C.POINTER(_vp), C.POINTER(_vp), _i]),
        "sfg_multi_advance": (_i, [_vp, _i]),
        "sfg_multi_last_ms_per_step": (_d, [_vp]),
        "sfg_multi_set_persistent": (_i, [_vp, _i]),
        "sfg_multi_step_bytes": (_d, [_vp, _i]),
        "sfg_multi_destroy": (None, [_vp]),
        "elastic_wave": (None, [C.c_char_p, _i, _i, _i, _pi, C.POINTER(_pd), _i, _i, _i, _d, _d, _d, _d,
                                _d, _d, _d, _d, _i, _i, _i, _pi, _i, _pi, C.c_char_p, _i, _i]),
        "seisfem_run": (_i, [C.c_char_p]),
        "sfh_driver_options": (None, [_i, _i]),
        "sfh_driver_status": (_i, []),
        # surrogate MLP (include/seisfem_surrogate.h)
        "sfg_mlp_last_error": (C.c_char_p, []),
        "sfg_mlp_create": (_i, [C.POINTER(_vp), _i, _i, _i, _i, _i, C.c_ulonglong]),
        "sfg_mlp_destroy": (None, [_vp]),
        "sfg_mlp_param_count": (C.c_longlong, [_vp]),
        "sfg_mlp_set_data": (_i, [_vp, _pf, _pf, _i]),
        "sfg_mlp_train_step": (_i, [_vp, _pi, C.c_float, _pd]),
        "sfg_mlp_train": (_i, [_vp, _i, C.c_float, C.c_longlong, C.c_ulonglong, _pd, _pd]),
        "sfg_mlp_predict": (_i, [_vp, _pf, _i, _pf, _pd]),
        "sfg_mlp_gradients": (_i, [_vp, _pf, _pf, _pd, _i, _pf, _pf]),
        "sfg_mlp_get_layer": (_i, [_vp, _i, _pf, _pf]),
        "sfg_mlp_set_layer": (_i, [_vp, _i, _pf, _pf]),
        "sfg_gemm_bf16_nt": (_i, [_i, _i, _i, _pf, _pf, _pf]),
        "sfg_gemm_bf16_time": (_i, [_i, _i, _i, _i, C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error():
    return lib().sfg_last_error().decode(errors="replace")


class NativeError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"{where} failed with code {code}: {last_error()}")
        self.code = code


def check(code, where):
    if code < 0:
        raise NativeError(code, where)
    return code

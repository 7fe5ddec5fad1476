"""ctypes binding of ``liblsbflow_amd.so`` (the C-ABI declared in include/lsbflow.h).

There is no CPU fallback: importing this module without the built library, or
calling a GPU entry point without a HIP device, raises.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsbflow_amd.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p


class LsbConfig(C.Structure):
    """Mirror of ``lsb_config`` (field order must match include/lsbflow.h)."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("mapping", C.c_int), ("bctype", C.c_int), ("forcing", C.c_int),
        ("sgsflag", C.c_int), ("sgsmodel", C.c_int), ("dealiasflag", C.c_int),
        ("dealias_ref_compat", C.c_int), ("dct_exact", C.c_int), ("top_bc", C.c_int), ("dyn_avg", C.c_int), ("nut_ceil0", C.c_int),
        ("ifrz", C.c_int), ("nstart", C.c_int),
        ("precision", C.c_int), ("device", C.c_int),
        ("xlen", C.c_double), ("ylen", C.c_double), ("zlen", C.c_double), ("x0", C.c_double),
        ("u0", C.c_double), ("xnu", C.c_double), ("blfac", C.c_double), ("prtl", C.c_double),
        ("a", C.c_double), ("b", C.c_double), ("cs", C.c_double), ("kc", C.c_double),
        ("filt_order", C.c_double), ("filt_amp", C.c_double),
        ("samp", C.c_double), ("istartbound", C.c_double),
        ("dt", C.c_double),
    ]


# name -> (restype, argtypes)
_SIGS = {
    "lsb_create": (C.c_int, [C.POINTER(LsbConfig), C.POINTER(_vp)]),
    "lsb_destroy": (C.c_int, [_vp]),
    "lsb_last_error": (C.c_char_p, []),
    "lsb_set_grid": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double]),
    "lsb_set_mats": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "lsb_set_paral": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "lsb_get_grid": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "lsb_get_mats": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "lsb_build_cheb_zder_matrix": (C.c_int, [C.c_int, C.c_double, _dp, _dp, _dp, _dp]),
    "lsb_dct_diff": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, C.c_int]),
    "lsb_diagonalize_matrix": (C.c_int, [_dp, C.c_int, _dp, _dp, _dp]),
    "lsb_horfft": (C.c_int, [_vp, _dp, _dp, C.c_int, C.c_int]),
    "lsb_norm": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "lsb_dataflip_xtoy": (C.c_int, [_vp, _dp, _dp]),
    "lsb_dataflip_ytox": (C.c_int, [_vp, _dp, _dp]),
    "lsb_dfdx": (C.c_int, [_vp, _dp, _dp]),
    "lsb_dfdy": (C.c_int, [_vp, _dp, _dp]),
    "lsb_dfdz": (C.c_int, [_vp, _dp, _dp]),
    "lsb_d2fdz2": (C.c_int, [_vp, _dp, _dp]),
    "lsb_dpdz": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, C.c_int]),
    "lsb_divuv": (C.c_int, [_vp, _dp, _dp, _dp]),
    "lsb_div": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "lsb_lapl": (C.c_int, [_vp, _dp, _dp]),
    "lsb_divcheck": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "lsb_solve_helmholtz": (C.c_int, [_vp, _dp, C.c_double, C.c_int, _dp, _dp, _dp]),
    "lsb_solve_nhom_helmholtz": (C.c_int, [_vp, _dp, _dp, C.c_double, _dp, _dp, _dp]),
    "lsb_tau_sgs": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "lsb_static_smagorinsky": (C.c_int, [_vp, C.c_double, _dp, _dp]),
    "lsb_dynamic_smagorinsky": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "lsb_sigma_model": (C.c_int, [_vp, _dp, _dp]),
    "lsb_interscale_model": (C.c_int, [_vp, _dp, _dp]),
    "lsb_filter_q_3pt": (C.c_int, [_vp, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]),
    "lsb_filterqspec": (C.c_int, [_vp, _dp, C.c_int]),
    "lsb_specfilter_exp_xy": (C.c_int, [_vp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int]),
    "lsb_setup_dealiasing": (C.c_int, [_vp, _ip, _ip]),
    "lsb_dealiasing_2d_fftw": (C.c_int, [_vp, _dp, _dp, _dp]),
    "lsb_nlterm_skewsym": (C.c_int, [_vp, _dp, _dp]),
    "lsb_state_set": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_double]),
    "lsb_state_get": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp]),
    "lsb_timestep": (C.c_int, [_vp, C.c_int]),
    "lsb_navierstokes_update": (C.c_int, [_vp]),
    "lsb_state_device_ptrs": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "lsb_stream": (_vp, [_vp]),
    "lsb_horfft_dev": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "lsb_horfft_dev_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "lsb_autotune": (C.c_int, [_vp]),
    "lsb_ensemble_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.POINTER(_vp)]),
    "lsb_ensemble_timestep": (C.c_int, [_vp, C.c_int]),
    "lsb_ensemble_synchronize": (C.c_int, [_vp]),
    "lsb_ensemble_info": (C.c_int, [_vp, _ip, _ip, _ip]),
    "lsb_ensemble_destroy": (C.c_int, [_vp]),
    "lsb_solve_helmholtz_mixed_dev": (C.c_int, [_vp, _vp, C.c_double, C.c_int, C.c_int, _vp]),
    "lsb_autotune_get": (C.c_int, [_vp, _ip, _fp, C.c_int]),
    "lsb_autotune_set": (C.c_int, [_vp, _ip, C.c_int]),
    "lsb_dfdz_dev": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "lsb_solve_helmholtz_dev": (C.c_int, [_vp, _vp, C.c_double, C.c_int, _vp]),
    "lsb_profile_begin": (C.c_int, [_vp]),
    "lsb_profile_begin_classes": (C.c_int, [_vp, C.c_uint]),
    "lsb_profile_end": (C.c_int, [_vp, _dp, _ip, _dp, _dp]),
    "lsb_synchronize": (C.c_int, [_vp]),
    "lsb_snapshot_record": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_double,
                                      _fp]),
    "lsb_state_snapshot": (C.c_int, [_vp, C.c_int, C.c_char_p, _fp]),
    "lsb_dealiasf_2d_fftw": (C.c_int, [_vp, _dp, _dp, _dp]),
    "lsb_test_fft": (C.c_int, [_vp, _dp, _dp]),
    "lsb_zinteg": (C.c_int, [_vp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "lsb_set_time_average": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "lsb_state_snapshot_average": (C.c_int, [_vp, C.c_char_p, _fp]),
    "lsb_state_get_time_average": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "lsb_read_snapshot": (C.c_int, [_vp, C.c_char_p, _dp, _ip, _dp, _dp]),
    "lsb_write_restart": (C.c_int, [_vp, C.c_char_p, _fp]),
    "lsb_read_restart": (C.c_int, [_vp, C.c_char_p, _fp]),
    "lsb_maxcalc": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                              _ip]),
    "lsb_state_maxcalc": (C.c_int, [_vp, C.c_double, C.c_double, _dp, _ip]),
    "lsb_nancheck": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _ip]),
    "lsb_state_nancheck": (C.c_int, [_vp, _ip]),
    "lsb_wall_shear": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "lsb_state_wall_shear": (C.c_int, [_vp, _dp, _dp, _dp]),
    "lsb_set_adaptive_dt": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double]),
    "lsb_sponge_lambda": (C.c_int, [C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
    "lsb_set_sponge": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double]),
    "lsb_filterall3pt": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "lsb_tns_automatic_filtering": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip]),
    "lsb_state_filterall3pt": (C.c_int, [_vp]),
    "lsb_state_tns_filter": (C.c_int, [_vp, _dp, _dp, _ip]),
    "lsb_set_filtering": (C.c_int, [_vp, C.c_int, C.c_int]),
    "lsb_ener": (C.c_int, [_vp, _dp, _dp, _dp, C.c_int, _dp]),
    "lsb_lintest": (C.c_int, [_vp, _dp, _dp, _dp, C.c_int, C.c_double, _dp, _dp, _dp, _dp]),
    "lsb_state_lintest": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "lsb_set_temperature": (C.c_int, [_vp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double]),
    "lsb_solve_orrsomm": (C.c_int, [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                    C.c_double, _dp, C.c_int, _dp, _dp, _dp]),
    "lsb_state_set_temperature": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "lsb_state_get_temperature": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
}


def exported_symbols():
    """Names of every C-ABI entry point this binding declares."""
    return sorted(_SIGS)


def load(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise RuntimeError(
            f"lsbflow_amd: native library {path} is missing; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)  # AttributeError if the library lacks the symbol
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


class LsbError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().lsb_last_error()
        raise LsbError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

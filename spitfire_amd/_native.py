"""Loader for libgriffon_amd.so (the C-ABI over the gfx950 HIP kernels).

The library is built in-tree by spitfire_amd.build. There is no fallback: if it is missing or fails to load,
every compute entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, '_lib', 'libgriffon_amd.so')

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_L = ctypes.POINTER(ctypes.c_long)
_V = ctypes.c_void_p
_int, _dbl, _long = ctypes.c_int, ctypes.c_double, ctypes.c_long

# (name, argtypes) for every symbol declared in include/griffon_amd.h
SIGNATURES = {
    'gf_last_error': [],
    'gf_device_count': [],
    'gf_mech_create': [ctypes.POINTER(_V)],
    'gf_mech_destroy': [_V],
    'gf_mechanism_set_gas_constant': [_V, _dbl],
    'gf_mechanism_set_ref_pressure': [_V, _dbl],
    'gf_mechanism_set_ref_temperature': [_V, _dbl],
    'gf_mechanism_add_species': [_V, _dbl],
    'gf_mechanism_resize_heat_capacity_data': [_V],
    'gf_mechanism_add_const_cp': [_V, _int, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl],
    'gf_mechanism_add_nasa7_cp': [_V, _int, _dbl, _dbl, _dbl, _D, _D],
    'gf_mechanism_add_reaction': [_V, _int, _int, _int, _I, _I, _int, _I, _I, _D, _int, _I, _D, _dbl, _D, _D, _int,
                                  _I, _D],
    'gf_mech_n_species': [_V],
    'gf_mech_n_reactions': [_V],
    'gf_production_rates': [_V, _dbl, _dbl, _D, _D],
    'gf_prod_rates_primitive_sensitivities': [_V, _dbl, _dbl, _D, _int, _D],
    'gf_reactor_rhs_isobaric': [_V, _D, _dbl, _dbl, _D, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _int, _int, _D],
    'gf_reactor_jac_isobaric': [_V, _D, _dbl, _dbl, _D, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _int, _int, _int, _int,
                                _D, _D],
    'gf_reactor_rhs_isochoric': [_V, _D, _dbl, _dbl, _D, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _int, _int, _D],
    'gf_reactor_jac_isochoric': [_V, _D, _dbl, _dbl, _D, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _int, _int, _int, _D,
                                 _D],
    'gf_mixture_molecular_weight': [_V, _D, _D],
    'gf_mole_fractions': [_V, _D, _D],
    'gf_ideal_gas_density': [_V, _dbl, _dbl, _D, _D],
    'gf_ideal_gas_pressure': [_V, _dbl, _dbl, _D, _D],
    'gf_cp_mix': [_V, _dbl, _D, _D],
    'gf_cv_mix': [_V, _dbl, _D, _D],
    'gf_species_cp': [_V, _dbl, _D],
    'gf_species_cv': [_V, _dbl, _D],
    'gf_enthalpy_mix': [_V, _dbl, _D, _D],
    'gf_energy_mix': [_V, _dbl, _D, _D],
    'gf_species_enthalpies': [_V, _dbl, _D],
    'gf_species_energies': [_V, _dbl, _D],
    'gf_dcpdT_species': [_V, _dbl, _D, _D],
    'gf_flamelet_stencils': [_V, _D, _int, _D, _D, _D, _D, _D, _D, _D],
    'gf_flamelet_jac_indices': [_V, _int, _I, _I],
    'gf_flamelet_rhs': [_V, _D, _dbl, _D, _D, _int, _D, _D, _D, _D, _int, _D, _D, _D, _D, _D, _D, _int, _int, _int,
                        _D],
    'gf_flamelet_jacobian': [_V, _D, _dbl, _D, _D, _int, _D, _D, _D, _D, _int, _D, _D, _D, _D, _D, _D, _int, _dbl,
                             _int, _dbl, _int, _int, _int, _int, _int, _D, _D],
    'gf_btddod_full_factorize': [_D, _int, _int, _D, _I],
    'gf_btddod_full_solve': [_D, _D, _I, _D, _int, _int, _D],
    'gf_btddod_scale_and_add_diagonal': [_D, _dbl, _D, _dbl, _int, _int],
    'gf_set_chem_launch_shape': [_int],
    'gf_batch_masked_update': [_int, _int, _V, _V, _V, _V],
    'gf_batch_adjoint_stage_prep': [_int, _int, _V, _V, _V, _V, _V, _V, _V, _V, _dbl, _dbl, _V, _V, _V],
    'gf_batch_adjoint_stage_finish': [_int, _int, _V, _V, _V, _V, _V, _V, _V, _dbl, _dbl, _V, _V, _V],
    'gf_batch_stage_residual': [_int, _int, _V, _dbl, _V, _V, _V, _V, _V, _dbl, _V, _V, _int, _V],
    'gf_batch_reactor_rhs_isobaric': [_V, _int, _V, _long, _long, _V, _long, _V, _long, _long, _V],
    'gf_batch_reactor_jac_isobaric': [_V, _int, _V, _long, _long, _V, _long, _V, _long, _long, _V, _long, _V],
    'gf_batch_reactor_general': [_V, _int, _int, _V, _long, _V, _long, _dbl, _dbl, _V, _dbl, _dbl, _dbl, _dbl,
                                 _dbl, _dbl, _int, _int, _int, _V, _V, _V],
    'gf_batch_reactor_general_noalloc': [_V, _int, _int, _V, _long, _V, _long, _dbl, _dbl, _V, _dbl, _dbl, _dbl,
                                         _dbl, _dbl, _dbl, _int, _int, _int, _V, _V, _V],
    'gf_batch_btddod_full_factorize': [_int, _V, _int, _V, _V, _V, _V, _V, _V, _V],
    'gf_batch_btddod_full_solve': [_int, _V, _int, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V],
    'gf_batch_reactor_jac_isobaric_terms': [_V, _int, _V, _long, _long, _V, _long, _V, _long, _long, _V, _long, _V, _V, _V, _V, _V],
    'gf_batch_btddod_inverse_factorize': [_int, _V, _int, _V, _V, _V, _V, _int, _V],
    'gf_batch_mixture_properties': [_V, _int, _V, _V, _V, _V, _V],
    'gf_batch_eigenvalue_bound': [_V, _int, _int, _V, _dbl, _V, _V, _V],
    'gf_batch_btddod_inverse_solve': [_int, _V, _int, _V, _V, _V, _V, _int, _V, _V, _V, _V],
    'gf_psitc_options_default': [_V],
    'gf_flamelet_batch_create': [_V, _int, _I, _dbl, _D, _D, _D, _D, _int, _int, ctypes.POINTER(_V)],
    'gf_flamelet_batch_destroy': [_V],
    'gf_flamelet_batch_set_heat_loss': [_V, _D, _D, _D, _D, _int],
    'gf_flamelet_batch_set_lewis_numbers': [_V, _D],
    'gf_flamelet_batch_set_state': [_V, _D],
    'gf_flamelet_batch_get_state': [_V, _D],
    'gf_flamelet_batch_rhs': [_V, _D],
    'gf_flamelet_batch_steady_psitc': [_V, _V, _I, _I, _I, _D, _D, ctypes.POINTER(ctypes.c_longlong)],
    'gf_flamelet_batch_steady_newton': [_V, _V, _I, _I, _I, _D, ctypes.POINTER(ctypes.c_longlong)],
    'gf_flamelet_batch_transient_rhs': [_V, _I, _V, _V],
    'gf_flamelet_batch_transient_setup': [_V, _I, _V, _D, _int],
    'gf_flamelet_batch_transient_solve': [_V, _I, _V, _V],
}


class PsitcOptions(ctypes.Structure):
    """gf_psitc_options (include/griffon_amd.h)"""
    _fields_ = [('tolerance', _dbl), ('max_iterations', _int), ('ds_init', _dbl), ('ds_init_decrease', _dbl),
                ('adaptive_restart', _int), ('max_recursion_depth', _int), ('diffusion_factor', _dbl),
                ('global_ds', _int), ('ds_safety', _dbl), ('ds_ramp', _dbl), ('ds_max', _dbl),
                ('max_factor_line_search', _dbl), ('max_allowed_residual', _dbl), ('min_allowable_state_var', _dbl),
                ('lu_precision', _int), ('log_rate', _int)]


_lib = None


class NativeError(RuntimeError):
    pass


def lib():
    """Load (once) and return the native library; raises if it is not built.

    PyTorch-ROCm bundles its own HIP/HSA runtime (torch/lib/libamdhip64.so, SONAME libamdhip64.so.7). torch is
    imported first so that the library's libamdhip64.so.7 dependency resolves to that already-loaded runtime:
    a process must hold exactly one HSA runtime, or the second one finds no GPU.
    """
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(LIB_PATH + ' is not built; run `python -m spitfire_amd.build`')
        try:
            import torch  # noqa: F401  (loads torch's HIP runtime before ours)
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(l, name)
            f.argtypes = args
            f.restype = {'gf_last_error': ctypes.c_char_p, 'gf_psitc_options_default': None}.get(name, ctypes.c_int)
        _lib = l
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().gf_last_error().decode()
        raise NativeError('griffon_amd error %d: %s' % (rc, msg))


def dptr(a):
    return a.ctypes.data_as(_D)


def iptr(a):
    return a.ctypes.data_as(_I)

"""MI355X implementation of Spitfire's Griffon chemistry path and the drivers around it.

The top-level names follow the reference package's (src/spitfire/__init__.py) for the parts built here; they are
imported on first use, so `import spitfire_amd` loads neither torch nor the HIP library.
"""
import importlib

_EXPORTS = {
    'Dimension': 'library', 'Library': 'library',
    'explosive_mode_analysis': 'analysis',
    'ChemicalMechanismSpec': 'mechanism',
    'Flamelet': 'flamelet', 'FlameletSpec': 'flamelet', 'FlameletBatch': 'flamelet',
    'compute_dissipation_rate': 'flamelet',
    'HomogeneousReactor': 'reactors',
    'ReactorBatch': 'reactor_batch',
    'build_adiabatic_eq_library': 'tabulation', 'build_adiabatic_bs_library': 'tabulation',
    'build_adiabatic_slfm_library': 'tabulation', 'build_unreacted_library': 'tabulation',
    'build_nonadiabatic_defect_bs_library': 'tabulation', 'build_nonadiabatic_defect_eq_library': 'tabulation',
    'build_nonadiabatic_defect_transient_slfm_library': 'tabulation',
    'build_nonadiabatic_defect_steady_slfm_library': 'tabulation',
    'PyCombustionKernels': 'griffon',
    'odesolve': 'time_integration', 'SimpleNewtonSolver': 'time_integration', 'PIController': 'time_integration',
    'KennedyCarpenterS6P4Q3': 'time_integration', 'BackwardEulerS1P1Q1': 'time_integration',
}

__all__ = sorted(_EXPORTS)


def __getattr__(name):
    if name in _EXPORTS:
        value = getattr(importlib.import_module('.' + _EXPORTS[name], __name__), name)
        globals()[name] = value
        return value
    raise AttributeError("module 'spitfire_amd' has no attribute %r" % name)

"""The C-ABI library loads and exports every entry point include/griffon_amd.h declares (no compute on the CPU)."""
import os
import re

import pytest

from tests.conftest import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, 'include', 'griffon_amd.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(gf_[A-Za-z0-9_]+)\s*\(', txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ('gf_production_rates', 'gf_reactor_jac_isobaric', 'gf_flamelet_jacobian', 'gf_btddod_full_factorize',
              'gf_batch_reactor_jac_isobaric', 'gf_batch_btddod_full_solve'):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from spitfire_amd import _native
    lib = _native.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes signature table covers the header exactly
    assert sorted(_native.SIGNATURES) == declared_symbols()


def test_mechanism_builds_without_device(mech_c1_data):
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    m = ChemicalMechanismSpec(mech_c1_data)
    assert m.n_species == 10 and m.n_reactions == 30
    assert m.griffon.n_species == 10


def test_compute_without_device_fails_loudly(mech_c1_data):
    if gpu_available():
        pytest.skip('a GPU is present')
    import numpy as np
    from spitfire_amd._native import NativeError
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    m = ChemicalMechanismSpec(mech_c1_data)
    with pytest.raises(NativeError):
        m.griffon.production_rates(1000., 1., np.ones(10) / 10, np.zeros(10))


def test_setter_errors_raise(mech_c1_data):
    from spitfire_amd._native import NativeError
    from spitfire_amd.griffon import PyCombustionKernels
    g = PyCombustionKernels()
    g.mechanism_set_element_mw_map({'H': 1.008})
    g.mechanism_add_element('H')
    g.mechanism_add_species('H', {'H': 1.})
    with pytest.raises(ValueError):
        g.mechanism_add_species('H', {'H': 1.})
    g.mechanism_add_species('H2', {'H': 2.})
    with pytest.raises(NativeError):  # stoichiometric coefficient outside 1..3
        g.mechanism_add_reaction_simple({'H': 4}, {'H2': 2}, True, 1., 0., 0.)

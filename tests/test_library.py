"""Library / Dimension (src/spitfire/chemistry/library.py) on the CPU: grid attributes, property assignment,
slicing, squeeze / swapaxes / copies, and persistence, including a library file written by the reference itself
(tests/tabulation/adiabatic_slfm/gold.pkl, kept as tests/golden/ref_adiabatic_slfm_gold.pkl and read with the
data-only pickle reader)."""
import os
import pickle

import numpy as np
import pytest

from spitfire_amd.library import Dimension, Library, LibraryIndexError

GOLD = os.path.join(os.path.dirname(__file__), 'golden')


def _lib():
    z = Dimension('mixture_fraction', np.linspace(0., 1., 5))
    x = Dimension('dissipation_rate_stoich', np.logspace(0, 1, 3), True)
    lib = Library(z, x)
    lib['temperature'] = 300. + 1000. * lib.mixture_fraction_grid * np.log(lib.dissipation_rate_stoich_grid + 1.)
    lib['pressure'] = 101325.
    return lib


def test_grid_attributes_and_properties():
    lib = _lib()
    assert lib.shape == (5, 3) and lib.size == 15
    assert lib.dim_names == ['mixture_fraction', 'dissipation_rate_stoich']
    assert lib.mixture_fraction_npts == 5 and lib.dissipation_rate_stoich_max == 10.
    assert lib.dim('dissipation_rate_stoich').log_scaled
    np.testing.assert_array_equal(lib.mixture_fraction_grid[:, 1], lib.mixture_fraction_values)
    assert np.all(lib['pressure'] == 101325.)
    assert 'temperature' in lib and 'density' not in lib
    with pytest.raises(ValueError):
        lib['bad'] = np.zeros(4)
    with pytest.raises(TypeError):
        lib['bad'] = 'text'
    lib.remove('pressure')
    assert lib.props == ['temperature']


def test_slices_squeeze_swapaxes_copies():
    lib = _lib()
    col = lib[:, 1]
    assert col.shape == (5, 1)
    np.testing.assert_array_equal(col['temperature'][:, 0], lib['temperature'][:, 1])
    sq = Library.squeeze(col)
    assert sq.dim_names == ['mixture_fraction'] and sq['temperature'].shape == (5,)
    pt = Library.squeeze(lib[2, 0])
    assert isinstance(pt, dict) and pt['properties']['temperature'] == lib['temperature'][2, 0]
    sw = Library.swapaxes(lib, 0, 1)
    assert sw.dim_names == ['dissipation_rate_stoich', 'mixture_fraction']
    np.testing.assert_array_equal(sw['temperature'], lib['temperature'].T)
    with pytest.raises(LibraryIndexError):
        lib[1]
    shallow, deep = Library.copy(lib), Library.deepcopy(lib)
    lib['temperature'][0, 0] = -1.
    assert shallow['temperature'][0, 0] == -1. and deep['temperature'][0, 0] != -1.
    lib.scale_dimension('mixture_fraction', 2.)
    assert lib.mixture_fraction_max == 2.


def test_save_load_round_trip_with_mechanism(tmp_path):
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    m = ChemicalMechanismSpec(cantera_input=os.path.join(GOLD, 'h2-burke.yaml'), group_name='h2-burke')
    lib = _lib()
    lib.extra_attributes['mech_spec'] = m
    lib.extra_attributes['note'] = 'seeded'
    path = str(tmp_path / 'lib.pkl')
    lib.save_to_file(path)
    with open(path, 'rb') as f:  # a plain dictionary of arrays and strings: our own file
        assert isinstance(pickle.load(f), dict)
    back = Library.load_from_file(path)
    assert back.dim_names == lib.dim_names and back.props == lib.props
    for p in lib.props:
        np.testing.assert_array_equal(back[p], lib[p])
    assert back.dim('dissipation_rate_stoich').log_scaled
    assert back.extra_attributes['note'] == 'seeded'
    m2 = back.extra_attributes['mech_spec']
    assert m2.species_names == m.species_names and m2.n_reactions == m.n_reactions
    assert m2.reference_mech_data() == m.reference_mech_data()


def test_load_reference_library_file():
    """Library.load_from_file on the file the reference's SLFM test wrote (its own Library pickle): dimensions,
    properties and the mechanism it carries."""
    lib = Library.load_from_file(os.path.join(GOLD, 'ref_adiabatic_slfm_gold.pkl'))
    gold = np.load(os.path.join(GOLD, 'ref_slfm_gold.npz'))
    assert lib.dim_names == ['mixture_fraction', 'dissipation_rate_stoich'] and lib.shape == (34, 8)
    np.testing.assert_array_equal(lib.dissipation_rate_stoich_values, gold['dim|dissipation_rate_stoich'])
    # the file predates the log_scaled flag; a dimension without one loads as linear (library.py:172)
    assert not lib.dim('dissipation_rate_stoich').log_scaled
    for k in gold.files:
        if k.startswith('prop|'):
            np.testing.assert_array_equal(lib[k[5:]], gold[k])
    m = lib.extra_attributes['mech_spec']
    assert m.species_names == ['H', 'H2', 'O', 'OH', 'H2O', 'O2', 'HO2', 'H2O2', 'AR', 'HE', 'N2']
    assert m.n_reactions == 27


def test_text_directory(tmp_path):
    lib = _lib()
    out = str(tmp_path / 'txt')
    lib.save_to_text_directory(out)
    with open(os.path.join(out, 'metadata_independent_variables.txt')) as f:
        assert f.read().split() == lib.dim_names
    T = np.loadtxt(os.path.join(out, 'bulkdata_dvar_temperature.txt'))
    np.testing.assert_allclose(T, lib['temperature'].ravel(order='F'), rtol=1e-13)
    with pytest.raises(FileExistsError):
        lib.save_to_text_directory(out)
    lib.save_to_text_directory(out, overwrite=True)

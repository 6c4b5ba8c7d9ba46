"""Flamelet grids and dissipation-rate profiles (restated from flamelet.py:240-341) against closed forms, and
the vectorized mixture thermo used for initial conditions against the scalar routines."""
import numpy as np
import pytest


def test_peters_dissipation_rate_closed_form():
    from spitfire_amd.flamelet import compute_dissipation_rate
    z = np.linspace(0., 1., 101)
    chi = compute_dissipation_rate(z, 7., 'Peters')
    assert abs(chi[50] - 7.) <= 1e-14 * 7.              # erfinv(0) = 0: the maximum sits at Z = 1/2
    assert chi[0] == 0. and chi[-1] == 0.               # erfinv(+-1) = +-inf
    assert np.allclose(chi, chi[::-1], rtol=1e-12, atol=1e-300)
    # the erfc form exp(-2 [erfc^-1(2Z)]^2) is the same curve (erfinv(2Z-1) = -erfcinv(2Z))
    zz = np.array([0.055, 0.2, 0.8])
    from scipy.special import erfcinv
    assert np.allclose(compute_dissipation_rate(zz, 3., 'Peters'), 3. * np.exp(-2. * erfcinv(2. * zz) ** 2),
                       rtol=1e-12)
    assert np.all(compute_dissipation_rate(z, 2., 'constant') == 2.)


@pytest.mark.parametrize('n, zc, b', [(34, 0.0557, 4.), (128, 0.055, 4.), (65, 0.5, 6.)])
def test_clustered_grid(n, zc, b):
    from spitfire_amd.flamelet import clustered_grid
    z, dz = clustered_grid(n, zc, b)
    assert z[0] == 0. and z[-1] == 1. and z.size == n
    assert np.all(dz > 0.) and np.allclose(np.cumsum(dz), z[1:], rtol=0, atol=1e-15)
    # densest spacing at the cluster point
    k = np.argmin(dz)
    assert z[k] <= zc + 2. / n and z[k + 1] >= zc - 2. / n


def test_vectorized_thermo_matches_scalar(mech_53_data):
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    m = ChemicalMechanismSpec(mech_53_data)
    T = np.array([250., 300., 999.9, 1000., 1000.1, 2200., 3500., 4000.])
    H, C = m.species_enthalpies_array(T), m.species_cp_array(T)
    for i, t in enumerate(T):
        assert np.array_equal(H[i], m.species_enthalpies(t))
        assert np.array_equal(C[i], m.species_cp(t))
    rng = np.random.default_rng(3)
    Y = rng.dirichlet(np.full(m.n_species, 0.5), 5)
    h = np.array([np.dot(m.species_enthalpies(t), y) for t, y in zip([400., 900., 1500., 2000., 2900.], Y)])
    Tv = m.temperatures_from_enthalpies(h, Y)
    for k in range(5):
        assert abs(Tv[k] - m.temperature_from_enthalpy(h[k], Y[k])) <= 1e-9

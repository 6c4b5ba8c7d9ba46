"""Flamelet libraries and library analysis on the GPU.

- build_adiabatic_slfm_library (tabulation.py:229-336) on the batched flamelet solver, in the reference's
  sequential continuation and in the batched form, then the property evaluators (analysis.py:77-218): every field
  of the reference's own adiabatic SLFM library (tests/tabulation/adiabatic_slfm) at the reference test's rtol 1e-6,
  atol 1e-6, viscosity excepted (no transport here).
- gf_batch_mixture_properties against the oracle's thermodynamic getters on seeded states of the 53-species
  mechanism, in and out of the NASA7 ranges, given the pressure and given the density.
- explosive_mode_analysis against eigenvalues of the oracle's reactor Jacobians, to the first-order perturbation
  bound of each eigenvalue.
"""
import numpy as np
import pytest

from tests.test_reference_gold import check_slfm_gold, slfm_setup

pytestmark = pytest.mark.gpu


def _props(m, lib):
    from spitfire_amd import analysis as sca
    lib = sca.compute_specific_enthalpy(m, lib)
    lib = sca.compute_isochoric_specific_heat(m, lib)
    lib = sca.compute_isobaric_specific_heat(m, lib)
    lib = sca.compute_density(m, lib)
    return sca.compute_pressure(m, lib)


@pytest.mark.parametrize('continuation, prec', [('sequential', 64), ('batched', 64), ('batched', 32)])
def test_adiabatic_slfm_library_matches_reference_gold(continuation, prec):
    from spitfire_amd.tabulation import build_adiabatic_slfm_library
    print('adiabatic SLFM library, h2-burke, 34 points x 8 dissipation rates, continuation %s, fp%d'
          % (continuation, prec), flush=True)
    m, spec, chis = slfm_setup()
    lib = build_adiabatic_slfm_library(spec, diss_rate_values=chis, verbose=False, continuation=continuation,
                                       lu_precision=prec)
    assert lib.extra_attributes['mech_spec'] is m and lib.dim('dissipation_rate_stoich').log_scaled
    lib = _props(m, lib)
    worst = check_slfm_gold(lib, m)
    print('  largest |difference| / (1e-6 + 1e-6 |gold|): %.3e' % worst, flush=True)


def test_slfm_library_stops_at_extinction():
    """Dissipation rates past extinction: the table ends before the first extinguished flamelet, and
    include_extinguished keeps it."""
    from spitfire_amd.tabulation import build_adiabatic_slfm_library
    m, spec, _ = slfm_setup()
    chis = np.logspace(1, 5, 5)
    print('SLFM past extinction, chi_st', chis, flush=True)
    lib = build_adiabatic_slfm_library(spec, diss_rate_values=chis, verbose=False)
    n = lib.dissipation_rate_stoich_npts
    assert 1 <= n < chis.size
    np.testing.assert_allclose(lib.dissipation_rate_stoich_values, chis[:n], rtol=1e-12)
    T = lib['temperature']
    Tlin = T[0, 0] + (T[-1, 0] - T[0, 0]) * lib.mixture_fraction_values
    assert np.all(np.max(T - Tlin[:, None], axis=0) > 10.)
    lib2 = build_adiabatic_slfm_library(spec, diss_rate_values=chis, verbose=False, include_extinguished=True)
    assert lib2.dissipation_rate_stoich_npts == chis.size
    print('  burning up to chi_st %.3g' % chis[n - 1], flush=True)


def test_unreacted_and_burke_schumann_libraries():
    from spitfire_amd.flamelet import Flamelet, FlameletSpec
    from spitfire_amd.tabulation import build_adiabatic_bs_library, build_unreacted_library
    m, spec, _ = slfm_setup()
    for build, ic in ((build_unreacted_library, 'unreacted'), (build_adiabatic_bs_library, 'Burke-Schumann')):
        lib = build(spec, verbose=False)
        f = Flamelet(FlameletSpec(initial_condition=ic, **spec))
        assert lib.dim_names == ['mixture_fraction'] and lib.mixture_fraction_npts == 34
        np.testing.assert_array_equal(lib['temperature'], f.initial_temperature)
        np.testing.assert_array_equal(lib['mass fraction H2O'], f.initial_mass_fraction('H2O'))


def _random_states(m, n, seed):
    rng = np.random.default_rng(seed)
    T = rng.uniform(250., 3600., n)  # includes temperatures outside the NASA7 ranges
    Y = rng.dirichlet(np.full(m.n_species, 0.5), n)
    p = 10. ** rng.uniform(4., 6.5, n)
    return T, Y, p


def test_batch_mixture_properties_vs_oracle(mech_53_data):
    import torch
    from oracle.oracle import CpuGriffon
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.mechdata import CompiledMechanism
    m = ChemicalMechanismSpec(mech_53_data)
    cpu = CpuGriffon(CompiledMechanism(m.mech_data, order='name'))
    n = 2000
    T, Y, p = _random_states(m, n, 7)
    rho = np.random.default_rng(8).uniform(0.01, 10., n)
    print('batched mixture properties, 53 species, %d states' % n, flush=True)
    dev = torch.device('cuda')
    Ty = torch.from_numpy(np.hstack((T[:, None], Y))).to(dev)
    out = torch.empty((n, 8), dtype=torch.float64, device=dev)
    m.griffon.batch_mixture_properties(Ty, out, p=torch.from_numpy(p).to(dev), rho=torch.from_numpy(rho).to(dev))
    g = out.cpu().numpy()
    ref = np.array([cpu.thermo_getters(T[i], Y[i], p[i], rho[i])[0] for i in range(n)])
    err = np.abs(g - ref) / np.maximum(np.abs(ref), 1e-300)
    print('  max relative difference per field', err.max(axis=0), flush=True)
    assert err[:, :7].max() <= 1e-13
    # dcp/dT is a sum of terms of both signs; compare against the sum of magnitudes
    assert np.max(np.abs(g[:, 7] - ref[:, 7]) / np.maximum(np.abs(ref[:, 7]), 1.)) <= 1e-12
    m.griffon.batch_mixture_properties(Ty, out)  # no pressure or density given: those entries come out zero
    g2 = out.cpu().numpy()
    assert np.all(g2[:, 1:3] == 0.) and np.array_equal(g2[:, [0, 3, 4, 5, 6, 7]], g[:, [0, 3, 4, 5, 6, 7]])


@pytest.mark.parametrize('configuration', ['isobaric', 'isochoric'])
def test_explosive_mode_analysis_vs_oracle(mech_53_data, configuration):
    from scipy.linalg import eig
    from oracle.oracle import CpuGriffon
    from spitfire_amd import analysis as sca
    from spitfire_amd.library import Dimension, Library
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.mechdata import CompiledMechanism
    m = ChemicalMechanismSpec(mech_53_data)
    cpu = CpuGriffon(CompiledMechanism(m.mech_data, order='name'))
    n = 64
    T, Y, p = _random_states(m, n, 11)
    T = np.clip(T, 800., 2600.)
    lib = Library(Dimension('state', np.arange(n, dtype=np.float64)))
    lib['temperature'] = T
    lib['pressure'] = p
    mmw = 1. / (Y / m.molecular_weights[None, :]).sum(axis=1)
    lib['density'] = p * mmw / (8314.462618 * T)
    for i, s in enumerate(m.species_names):
        lib['mass fraction ' + s] = Y[:, i]
    print('CEMA %s, 53 species, %d states' % (configuration, n), flush=True)
    sca.explosive_mode_analysis(m, lib, configuration=configuration, compute_explosion_indices=True,
                                include_secondary_mode=True)
    ns = m.n_species
    worst = 0.
    for i in range(n):
        if configuration == 'isobaric':
            _, J = cpu.reactor_jac_isobaric(np.r_[T[i], Y[i, :-1]], p[i])
            ne = ns
        else:
            _, J = cpu.reactor_jac_isochoric(np.r_[lib['density'][i], T[i], Y[i, :-1]])
            ne = ns + 1
        Jm = J.reshape((ne, ne), order='F')
        lam_all, vl, vr = eig(Jm, left=True, right=True)
        w = np.real(lam_all)
        w[(w > -1.e-4) & (w < 1.e-4)] = -np.inf
        k = np.argmax(w)
        # first-order componentwise perturbation bound: Jacobians that agree entrywise to a relative eps move the
        # eigenvalue by at most eps |y|^T |J| |x| / |y^H x| (y, x its left and right eigenvectors). The GPU and oracle
        # Jacobians agree to 1e-12; the bound is taken at 1e-11. (The norm-wise bound is useless here: ||J|| reaches
        # 1e17 on this stiff mechanism while the explosive eigenvalues are O(1) to O(1e11).)
        yv, xv = vl[:, k], vr[:, k]
        tol = 1e-11 * (np.abs(yv) @ np.abs(Jm) @ np.abs(xv)) / abs(np.vdot(yv, xv)) + 1e-13 * abs(w[k])
        worst = max(worst, abs(lib['cema-lexp1'][i] - w[k]) / tol)
        assert abs(lib['cema-lexp1'][i] - w[k]) <= tol, (i, lib['cema-lexp1'][i], w[k], tol)
    print('  largest |difference| / bound: %.3e' % worst, flush=True)
    ei = np.array([lib['cema-ei1 T']] + [lib['cema-ei1 ' + s] for s in m.species_names[:-1]])
    assert np.all(ei >= 0.) and np.all(ei.sum(axis=0) <= 1. + 1e-12)
    assert 'cema-lexp2' in lib and 'cema-ei2 T' in lib


def _check_library_gold(name, lib, m, rtol, atol):
    """every property of the reference's library file except viscosity (no transport here), and its dimensions"""
    import os
    from spitfire_amd.library import Library
    gold = Library.load_from_file(os.path.join(os.path.dirname(__file__), 'golden', 'ref_%s_gold.pkl' % name))
    assert lib.dim_names == gold.dim_names and lib.shape == gold.shape, (lib.dim_names, lib.shape, gold.shape)
    for d, gd in zip(lib.dims, gold.dims):
        np.testing.assert_allclose(d.values, gd.values, rtol=rtol, atol=atol, err_msg=d.name)
    worst = 0.
    for p in gold.props:
        if p == 'viscosity':
            continue
        np.testing.assert_allclose(lib[p], gold[p], rtol=rtol, atol=atol, err_msg=p)
        worst = max(worst, float(np.max(np.abs(lib[p] - gold[p]) / (atol + rtol * np.abs(gold[p])))))
    print('  %s: largest |difference| / (atol + rtol |gold|) = %.3e' % (name, worst), flush=True)


@pytest.mark.parametrize('name, rtol, atol', [('adiabatic_unreacted', 1e-7, 1e-14),
                                              ('adiabatic_burke_schumann', 3e-3, 1e-7),
                                              ('nonadiabatic_defect_burke_schumann', 3e-3, 1e-7)])
def test_unstrained_libraries_match_reference_gold(name, rtol, atol):
    """The reference's unreacted, Burke-Schumann and Burke-Schumann enthalpy-defect libraries
    (tests/tabulation/<name>: h2-burke, air at 1200 K, H2 at 300 K, 34 points; the reference tests' tolerances, with
    numpy's default rtol of 1e-7 where they give only atol)."""
    from spitfire_amd import tabulation
    m, spec, _ = slfm_setup()
    print('library', name, flush=True)
    build = {'adiabatic_unreacted': tabulation.build_unreacted_library,
             'adiabatic_burke_schumann': tabulation.build_adiabatic_bs_library,
             'nonadiabatic_defect_burke_schumann': tabulation.build_nonadiabatic_defect_bs_library}[name]
    lib = _props(m, build(spec, verbose=False))
    _check_library_gold(name, lib, m, rtol, atol)


def test_nonadiabatic_defect_steady_slfm_library_matches_reference_gold():
    """tests/tabulation/nonadiabatic_defect_steady_slfm: chi_st = logspace(0, 1, 4), rtol 2e-4, atol 1e-4; every
    heat-loss step of the four dissipation rates solved in one batch."""
    from spitfire_amd.tabulation import build_nonadiabatic_defect_steady_slfm_library
    m, spec, _ = slfm_setup()
    print('nonadiabatic defect steady SLFM library', flush=True)
    lib = build_nonadiabatic_defect_steady_slfm_library(spec, verbose=False, diss_rate_values=np.logspace(0, 1, 4),
                                                        diss_rate_log_scaled=True,
                                                        integration_args={'transient_tolerance': 1e-10})
    _check_library_gold('nonadiabatic_defect_steady_slfm', _props(m, lib), m, 2e-4, 1e-4)


def test_nonadiabatic_defect_transient_slfm_library_matches_reference_gold():
    """tests/tabulation/nonadiabatic_defect_transient_slfm: burke-hydrogen mechanism, chi_st = logspace(0, 1, 4),
    transient tolerance 1e-10, rtol 2e-4, atol 1e-4."""
    import os
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.tabulation import build_nonadiabatic_defect_transient_slfm_library
    m = ChemicalMechanismSpec(cantera_input=os.path.join(os.path.dirname(__file__), 'golden', 'burke-hydrogen.yaml'),
                              group_name='gas')
    air = m.stream(stp_air=True)
    air.TP = 1200., 101325.
    fuel = m.stream('TPY', (300., 101325., 'H2:1'))
    spec = {'mech_spec': m, 'oxy_stream': air, 'fuel_stream': fuel, 'grid_points': 34}
    print('nonadiabatic defect transient SLFM library', flush=True)
    lib = build_nonadiabatic_defect_transient_slfm_library(spec, verbose=True, diss_rate_values=np.logspace(0, 1, 4),
                                                           integration_args={'transient_tolerance': 1e-10})
    _check_library_gold('nonadiabatic_defect_transient_slfm', _props(m, lib), m, 2e-4, 1e-4)

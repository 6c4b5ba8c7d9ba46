"""The product's drivers against the reference's own gold vectors (converted by tests/golden/convert_reference_gold.py
from the reference's numpy pickles; no reference code runs here).

- ESDIRK64 / backward Euler + Newton + PI control + odesolve on the reference's A -> B, A + B -> 2C problem with
  LAPACK, diagonal and GMRES linear solvers (tests/time_integration/chemistry_abc_implicit/test.py: atol 1e-8 on
  every time value and state).
- The clustered mixture-fraction grid of the 34-point adiabatic SLFM library (flamelet.py:300-341).
"""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla
from scipy.linalg.lapack import dgetrf, dgetrs

from spitfire_amd.time_integration import BackwardEulerS1P1Q1, KennedyCarpenterS6P4Q3, SimpleNewtonSolver, odesolve

GOLD = os.path.join(os.path.dirname(__file__), 'golden')


class ABCProblem:
    """The reference test's A -> B (k_ab), A + B -> 2C (k_bc) system with three linear-solver choices."""

    def __init__(self, k_ab=1., k_bc=0.2):
        self.k_ab, self.k_bc = k_ab, k_bc
        self.gmres_iter = 0

    def rhs(self, t, c):
        q1 = self.k_ab * c[0]
        q2 = self.k_bc * c[0] * c[1]
        return np.array([-q1 - q2, q1 - q2, 2. * q2])

    def jac(self, c):
        a, b = self.k_bc * c[1], self.k_bc * c[0]
        k = self.k_ab
        return np.array([[-k - a, -b, 0.], [k - a, -b, 0.], [2. * a, 2. * b, 0.]])

    def setup_lapack(self, t, c, prefactor):
        self.lu = dgetrf(prefactor * self.jac(c) - np.eye(3))[:2]

    def solve_lapack(self, r):
        return dgetrs(self.lu[0], self.lu[1], r)[0], 1, True

    def setup_diagonal(self, t, c, prefactor):
        self.dinv = 1. / (prefactor * self.jac(c) - np.eye(3)).diagonal()

    def solve_diagonal(self, r):
        return self.dinv * r, 1, True

    def setup_gmres(self, t, c, prefactor):
        A = prefactor * self.jac(c) - np.eye(3)
        self.op = spla.LinearOperator((3, 3), lambda x: A.dot(x))
        self.pre = spla.LinearOperator((3, 3), lambda r: 1. / A.diagonal() * r)

    def solve_gmres(self, r):
        def count(*args, **kwargs):
            self.gmres_iter += 1
        x, i = spla.gmres(self.op, r, M=self.pre, atol=1.e-8, callback=count, callback_type='legacy')
        return x, self.gmres_iter, not i


@pytest.mark.parametrize('method', [KennedyCarpenterS6P4Q3, BackwardEulerS1P1Q1])
@pytest.mark.parametrize('solver', ['lapack', 'diagonal', 'gmres'])
def test_abc_implicit_gold(method, solver):
    gold = np.load(os.path.join(GOLD, 'ref_abc_gold.npz'))
    p = ABCProblem()
    m = method(SimpleNewtonSolver())
    t, sol = odesolve(p.rhs, np.array([1., 0., 0.]), stop_at_time=10., method=m,
                      linear_setup=getattr(p, 'setup_' + solver), linear_solve=getattr(p, 'solve_' + solver),
                      save_each_step=True)
    key = solver + '-' + m.name
    gt, gs = gold['t|' + key], gold['sol|' + key]
    assert t.shape == gt.shape, (key, t.shape, gt.shape)
    # GMRES stops at SciPy's default relative tolerance (1e-5); with the first-order method, whose error estimate
    # compares two inexact solves, SciPy versions differ by up to 7e-8 in the states (the ESDIRK64 runs agree to 1e-8)
    atol = 1.e-7 if (solver == 'gmres' and method is BackwardEulerS1P1Q1) else 1.e-8
    np.testing.assert_allclose(t, gt, atol=1.e-8, rtol=0.)
    np.testing.assert_allclose(sol, gs, atol=atol, rtol=0.)


def test_pi_controller_step_sequence_matches_gold():
    """The accepted step sizes of the LAPACK ESDIRK64 run are the reference's to roundoff: the PI controller's
    recorded-error convention, ramp limit and first step all enter this sequence."""
    gold = np.load(os.path.join(GOLD, 'ref_abc_gold.npz'))
    p = ABCProblem()
    t, _ = odesolve(p.rhs, np.array([1., 0., 0.]), stop_at_time=10., method=KennedyCarpenterS6P4Q3(SimpleNewtonSolver()),
                    linear_setup=p.setup_lapack, linear_solve=p.solve_lapack, save_each_step=True)
    dts, gdts = np.diff(t), np.diff(gold['t|lapack-Kennedy/Carpenter ESDIRK64'])
    np.testing.assert_allclose(dts, gdts, rtol=1.e-12, atol=0.)
    # ramp-limited growth from the first step of 1e-3 while the error is far below target
    assert abs(dts[0] - 1.e-3) < 1e-18 and abs(dts[1] / dts[0] - 1.1) < 1e-12


def test_clustered_grid_matches_slfm_gold():
    """The 34-point grid of the reference's adiabatic SLFM library (h2-burke, H2 fuel at 300 K vs air at 1200 K):
    sinh clustering about the stoichiometric mixture fraction with the default intensity."""
    import json
    from spitfire_amd.flamelet import clustered_grid
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    gold = np.load(os.path.join(GOLD, 'ref_slfm_gold.npz'))
    with open(os.path.join(GOLD, 'ref_h2_burke_mech_data.json')) as f:
        m = ChemicalMechanismSpec.from_mech_data(json.load(f))
    air = m.stream(stp_air=True)
    air.TP = 1200., 101325.
    fuel = m.stream('TPY', (300., 101325., 'H2:1'))
    zst = m.stoich_mixture_fraction(fuel, air)
    z, _ = clustered_grid(34, zst, 4.)  # FlameletSpec default intensity (flamelet.py:105)
    np.testing.assert_allclose(z, gold['dim|mixture_fraction'], rtol=1.e-13, atol=1.e-15)


def _h2_burke_gold_mech_data():
    import json
    with open(os.path.join(GOLD, 'ref_h2_burke_mech_data.json')) as f:
        return json.load(f)


def test_h2_burke_yaml_matches_reference_extraction():
    """The YAML loader on h2-burke.yaml gives the mechanism data the reference extracted through Cantera (stored in
    the SLFM gold library): species order, atoms, NASA7 data, type-sorted reactions, unit conversions (cal/mol to
    J/kmol), efficiencies, default efficiencies and Troe parameters."""
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    gold = _h2_burke_gold_mech_data()
    m = ChemicalMechanismSpec(cantera_input=os.path.join(GOLD, 'h2-burke.yaml'), group_name='h2-burke')
    md = m.reference_mech_data()
    assert list(md['species']) == list(gold['species'])
    assert md['elements'] == gold['elements']
    for e, w in gold['element_mw_map'].items():
        assert md['element_mw_map'][e] == w
    for n, s in gold['species'].items():
        assert md['species'][n]['atom_map'] == s['atom_map']
        assert md['species'][n]['cp'][0] == s['cp'][0] == 'NASA7'
        np.testing.assert_array_equal(md['species'][n]['cp'][1:4], s['cp'][1:4])
        np.testing.assert_array_equal(md['species'][n]['cp'][4], s['cp'][4])
        np.testing.assert_array_equal(md['species'][n]['cp'][5], s['cp'][5])
    assert len(md['reactions']) == len(gold['reactions']) == 27
    for a, g in zip(md['reactions'], gold['reactions']):
        assert a[0] == g[0] and a[3] == g[3]
        assert a[1] == g[1] and a[2] == g[2]
        np.testing.assert_allclose(a[4:7], g[4:7], rtol=1.e-15, atol=0.)
        if g[0] != 'simple':
            assert a[7] == g[7] and a[8] == g[8]
        if g[0] in ('Lindemann', 'Troe'):
            np.testing.assert_allclose(a[9:12], g[9:12], rtol=1.e-15, atol=0.)
        if g[0] == 'Troe':
            np.testing.assert_array_equal(a[12], g[12])
    assert m.n_species == 11 and m.n_reactions == 27
    np.testing.assert_allclose(m.molecular_weights,
                               ChemicalMechanismSpec.from_mech_data(gold).molecular_weights, rtol=0., atol=0.)


def test_equivalence_ratio_mixture():
    """phi = 1 H2/air: X_H2 : X_O2 : X_N2 = 2 : 1 : 3.74; phi = 0.5 halves the fuel."""
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    m = ChemicalMechanismSpec.from_mech_data(_h2_burke_gold_mech_data())
    air = m.stream(stp_air=True)
    fuel = m.stream('X', 'H2:1')
    i = {n: m.species_index(n) for n in ('H2', 'O2', 'N2')}
    for phi in (1., 0.5, 2.):
        X = m.mix_for_equivalence_ratio(phi, fuel, air).X
        x = np.array([2. * phi, 1., 3.74]) / (2. * phi + 4.74)
        np.testing.assert_allclose([X[i['H2']], X[i['O2']], X[i['N2']]], x, rtol=1.e-14)
        assert abs(np.sum(X) - 1.) < 1e-15


def test_reaction_test_mechanism_yaml():
    """The reference's single-reaction test set (tests/test_mechanisms/reaction_test_mechanism.yaml): irreversible and
    reversible elementary reactions with T^b and kcal/mol activation energies, special orders, three-body, Lindemann
    and Troe. Each parsed reaction's forward rate on the oracle (the FP64 griffon restatement) equals the closed
    form k(T) * prod C^order with k = A T^b exp(-Ea / (R T))."""
    from oracle import oracle
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.mechanism_io import load_cantera_yaml
    from spitfire_amd.mechdata import CompiledMechanism
    data = load_cantera_yaml(os.path.join(GOLD, 'reaction_test_mechanism.yaml'))
    rx = data['reactions']
    assert len(rx) == 16
    assert [r['type'] for r in rx[:13]] == ['simple'] * 13
    assert rx[13]['type'] == 'three-body' and rx[14]['type'] == 'Lindemann' and rx[15]['type'] == 'Troe'
    assert [r['reversible'] for r in rx[:8]] == [False] * 8 and all(r['reversible'] for r in rx[8:13])
    assert rx[1]['orders'] == {'A': 0.5, 'D': 0.5} and rx[3]['orders'] == {'A': 1.5, 'D': 1.5}
    assert rx[7]['Ea'] == 10. * 4184. * 1000.  # 10 kcal/mol in J/kmol
    np.testing.assert_array_equal(rx[15]['Troe-params'], [1., 100., 1000., 2000.])
    R = data['gas_constant']
    T, p = 1300., 101325.
    names = data['species_names']
    Y = np.full(len(names), 1. / len(names))
    for k in range(8):  # irreversible elementary reactions, alone in a mechanism
        one = dict(data, reactions=[rx[k]])
        cpu = oracle.CpuGriffon(CompiledMechanism(one, order='name'))
        spec = ChemicalMechanismSpec(mech_data=one)
        mw = spec.molecular_weights
        rho = p / (R * T * np.sum(Y / mw))
        w = cpu.production_rates(T, rho, Y)
        C = rho * Y / mw
        r = rx[k]
        kf = r['A'] * T ** r['b'] * np.exp(-r['Ea'] / (R * T))
        orders = r.get('orders') or r['reactants']
        q = kf * np.prod([C[names.index(s)] ** o for s, o in orders.items()])
        iA = names.index('A')
        np.testing.assert_allclose(w[iA], -q * mw[iA], rtol=1.e-13)


def _closed_reactor_setup():
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    m = ChemicalMechanismSpec(cantera_input=os.path.join(GOLD, 'h2-burke.yaml'), group_name='h2-burke')
    air = m.stream(stp_air=True)
    fuel = m.stream('X', 'H2:1')
    mix = m.mix_for_equivalence_ratio(1.0, fuel, air)
    mix.TP = 1200., 101325.
    return m, mix


CLOSED_REACTORS = {'cp, adiabatic': ('isobaric', 'adiabatic'), 'cp, isothermal': ('isobaric', 'isothermal'),
                   'cv, adiabatic': ('isochoric', 'adiabatic'), 'cv, isothermal': ('isochoric', 'isothermal')}


def check_closed_reactor_gold(key, lib, m):
    """tests/reactor/closed_reactors/test.py: t, T and Y at rtol 1e-4, atol 1e-8"""
    gold = np.load(os.path.join(GOLD, 'ref_closed_reactors_gold.npz'))
    Y = np.array([lib['mass fraction ' + s] for s in m.species_names])
    gt, gT, gY = gold['t|' + key], gold['T|' + key], gold['Y|' + key]
    assert lib.time_values.shape == gt.shape, (key, lib.time_values.shape, gt.shape)
    np.testing.assert_allclose(lib.time_values, gt, atol=1.e-8, rtol=1.e-4)
    np.testing.assert_allclose(lib['temperature'], gT, atol=1.e-8, rtol=1.e-4)
    np.testing.assert_allclose(Y, gY, atol=1.e-8, rtol=1.e-4)


@pytest.mark.parametrize('key', list(CLOSED_REACTORS))
def test_closed_reactor_gold_on_oracle(key):
    """The reference's h2-burke closed reactors (phi = 1 H2/air at 1200 K, 1 atm) integrated to steady state by the
    product's reactor driver on the oracle's chemistry: every accepted time, temperature and mass fraction of the
    reference's trajectory."""
    from oracle.reactor_oracle import OracleReactor
    m, mix = _closed_reactor_setup()
    conf, heat = CLOSED_REACTORS[key]
    r = OracleReactor(m, mix, configuration=conf, heat_transfer=heat, mass_transfer='closed')
    lib = r.integrate_to_steady()
    check_closed_reactor_gold(key, lib, m)


def slfm_setup():
    """The reference's adiabatic SLFM library test (tests/tabulation/adiabatic_slfm/rebless.py): h2-burke, air at
    1200 K and H2 at 300 K, 1 atm, 34 clustered points, chi_st = logspace(0, 1, 8)."""
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    m = ChemicalMechanismSpec(cantera_input=os.path.join(GOLD, 'h2-burke.yaml'), group_name='h2-burke')
    air = m.stream(stp_air=True)
    air.TP = 1200., 101325.
    fuel = m.stream('TPY', (300., 101325., 'H2:1'))
    return m, {'mech_spec': m, 'oxy_stream': air, 'fuel_stream': fuel, 'grid_points': 34}, np.logspace(0, 1, 8)


SLFM_GOLD_PROPS = ('temperature', 'pressure', 'enthalpy', 'heat capacity cv', 'heat capacity cp', 'density')


def check_slfm_gold(lib, m, props=None):
    """tests/tabulation/adiabatic_slfm/test.py: every property at rtol 1e-6, atol 1e-6 (viscosity excluded: no
    transport here). Returns the largest |difference| / (atol + rtol |gold|) over the checked fields."""
    gold = np.load(os.path.join(GOLD, 'ref_slfm_gold.npz'))
    np.testing.assert_allclose(lib.mixture_fraction_values, gold['dim|mixture_fraction'], rtol=1e-13, atol=0.)
    np.testing.assert_allclose(lib.dissipation_rate_stoich_values, gold['dim|dissipation_rate_stoich'], rtol=1e-13,
                               atol=0.)
    names = list(props if props is not None else SLFM_GOLD_PROPS) + ['mass fraction ' + s for s in m.species_names]
    worst = 0.
    for name in names:
        g = gold['prop|' + name]
        np.testing.assert_allclose(lib[name], g, rtol=1.e-6, atol=1.e-6, err_msg=name)
        worst = max(worst, float(np.max(np.abs(lib[name] - g) / (1e-6 + 1e-6 * np.abs(g)))))
    return worst


def test_slfm_gold_on_oracle_continuation():
    """The reference's SLFM continuation (tabulation.py:271-312: compute_steady_state at tolerance 1e-6 per
    dissipation rate, each from its predecessor's state) run by the product's Flamelet on the oracle's flamelet
    RHS/Jacobian with the reference's default model (enthalpy flux and variable cp on): temperature and mass
    fractions of all eight flamelets against the reference's library. The first flamelet starts from Burke-Schumann
    (the reference: Cantera equilibrium), so it lands within the solver tolerance of the gold; the rest follow the
    reference's continuation step for step."""
    from oracle.flamelet_oracle import OracleFlamelet
    from spitfire_amd.flamelet import FlameletSpec
    from spitfire_amd.library import Dimension, Library
    m, spec, chis = slfm_setup()
    ic = 'Burke-Schumann'
    cols = []
    for c in chis:
        f = OracleFlamelet(FlameletSpec(stoich_dissipation_rate=c, initial_condition=ic, **spec))
        cols.append(f.compute_steady_state(tolerance=1e-6))
        ic = f.current_interior_state
        assert np.max(f.current_temperature - f.linear_temperature) > 10.
    lib = Library(Dimension('mixture_fraction', f.mixfrac_grid), Dimension('dissipation_rate_stoich', chis, True))
    for q in cols[0].props:
        lib[q] = np.stack([c[q] for c in cols], axis=1)
    assert check_slfm_gold(lib, m, props=('temperature', 'pressure')) < 0.05


def test_slfm_gold_thermo_on_oracle():
    """Enthalpy, cp, cv and density of the reference's library (Cantera's SolutionArray at T, p, Y) against the
    oracle's NASA7 getters at the gold's own T, p and Y."""
    from oracle.oracle import CpuGriffon
    from spitfire_amd.mechdata import CompiledMechanism
    m, _, _ = slfm_setup()
    gold = np.load(os.path.join(GOLD, 'ref_slfm_gold.npz'))
    cpu = CpuGriffon(CompiledMechanism(m.mech_data, order='name'))
    T, p = gold['prop|temperature'], gold['prop|pressure']
    Y = np.stack([gold['prop|mass fraction ' + s] for s in m.species_names], axis=-1)
    out = {k: np.empty_like(T) for k in ('enthalpy', 'heat capacity cp', 'heat capacity cv', 'density')}
    for idx in np.ndindex(T.shape):
        y = np.maximum(Y[idx], 0.)
        y /= y.sum()
        sc, _ = cpu.thermo_getters(T[idx], y, p[idx], 0.)
        out['density'][idx], out['heat capacity cp'][idx], out['heat capacity cv'][idx], out['enthalpy'][idx] = \
            sc[1], sc[3], sc[4], sc[5]
    for k, v in out.items():
        np.testing.assert_allclose(v, gold['prop|' + k], rtol=1e-10, atol=1e-6, err_msg=k)

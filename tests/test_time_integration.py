"""ESDIRK64 / Newton / PI step loop (spitfire_amd/time_integration.py) and the reactor integration on the oracle.

The tableau is pinned by the Runge-Kutta order conditions (fourth order for b, third for the embedded b-hat)
and by the observed order on the reference's own nonlinear test problem
(tests/time_integration/test_time_order_nonlinear_problem.py:16-42: q'' = -exp(1+q), whose exact solution has
q(1) = 0). The reactor integration is pinned by the invariants of a closed adiabatic isobaric reactor:
mixture enthalpy and element mass fractions are conserved."""
import numpy as np
import pytest

from spitfire_amd.time_integration import KennedyCarpenterS6P4Q3, PIController, SimpleNewtonSolver, odesolve


def test_esdirk_order_conditions():
    m = KennedyCarpenterS6P4Q3(SimpleNewtonSolver())
    A, b, bh, c = m.A, m.b, m.bh, m.c
    assert np.allclose(c, A.sum(axis=1), rtol=0, atol=1e-15)
    assert np.array_equal(b, A[-1]) and np.all(np.diag(A)[1:] == 0.25) and A[0, 0] == 0.
    conds = [(lambda w: w.sum(), 1.), (lambda w: w @ c, 1 / 2), (lambda w: w @ c ** 2, 1 / 3),
             (lambda w: w @ (A @ c), 1 / 6)]
    conds4 = [(lambda w: w @ c ** 3, 1 / 4), (lambda w: w @ (c * (A @ c)), 1 / 8), (lambda w: w @ (A @ c ** 2), 1 / 12),
              (lambda w: w @ (A @ (A @ c)), 1 / 24)]
    for f, v in conds + conds4:
        assert abs(f(b) - v) <= 1e-14
    for f, v in conds:
        assert abs(f(bh) - v) <= 1e-14
    assert max(abs(f(bh) - v) for f, v in conds4) > 1e-6  # the embedded solution is third order only


class _Const:
    def __init__(self, dt):
        self.dt = dt

    def first_step_size(self):
        return self.dt

    def __call__(self, n, dt, out):
        return self.dt

    def target_error(self):
        return 1.


def _solve_bratu(dt):
    theta = 3.03623184819656
    q0 = np.array([0., theta * np.tanh(0.25 * theta)])
    rhs = lambda t, q: np.array([q[1], -np.exp(1. + q[0])])
    op = {}

    def setup(t, q, pref):
        J = np.array([[0., 1.], [-np.exp(1. + q[0]), 0.]])
        op['M'] = pref * J - np.eye(2)

    def solve(r):
        return np.linalg.solve(op['M'], r)

    method = KennedyCarpenterS6P4Q3(SimpleNewtonSolver(evaluate_jacobian_every_iter=True))
    q, t, _ = odesolve(rhs, q0, stop_at_time=1.0, method=method, step_size=_Const(dt), linear_setup=setup,
                       linear_solve=solve)
    assert abs(t - 1.) < 1e-12
    return abs(q[0])


def test_esdirk_fourth_order_on_reference_problem():
    dts = [0.06, 0.02, 0.01, 0.009, 0.008, 0.006]
    err = [_solve_bratu(dt) for dt in dts]
    orders = [np.log(err[i] / err[i + 1]) / np.log(dts[i] / dts[i + 1]) for i in range(len(dts) - 1)]
    assert 3.9 < np.mean(orders) < 4.1, orders


def test_pi_controller():
    c = PIController(target_error=1e-10, first_step=1e-6, max_step=1., max_ramp=1.1)
    assert c.first_step_size() == 1e-6

    class O:
        temporal_error = 0.

    assert c(0, 1e-6, O()) == pytest.approx(1.1e-6, rel=1e-15)  # a zero error estimate ramps
    O.temporal_error = 1e-8  # error 100x the target: shrink by 100^-ki
    assert c(0, 1e-6, O()) == pytest.approx(1e-6 * (1e-2) ** 0.1333333333, rel=1e-14)
    O.temporal_error = 2e-11
    got = c(1, 1e-6, O())
    assert got == pytest.approx(1e-6 * min(1.1, 5. ** 0.1333333333 * 1. ** 0.06666666667), rel=1e-14)


def _c1_mixture(m):
    from spitfire_amd.mechanism import Stream
    air = m.stream(stp_air=True)
    fuel = m.stream('TPY', (300., 101325., 'H2:1'))
    zst = m.stoich_mixture_fraction(fuel, air)
    y = m.mix_streams([(air, 1 - zst), (fuel, zst)], 'mass', 'HP').Y
    return Stream(m, 1200., 101325., y)


def test_oracle_reactor_conserves_enthalpy_and_elements(mech_c1_data, oracle_lib):
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from oracle.reactor_oracle import OracleReactor
    m = ChemicalMechanismSpec(mech_c1_data)
    mix = _c1_mixture(m)
    r = OracleReactor(m, mix)
    out = r.integrate_to_time(1.e-3)
    Y0, Y1 = mix.Y, r.current_mass_fractions
    h0 = np.dot(m.species_enthalpies(mix.T), Y0)
    h1 = np.dot(m.species_enthalpies(r.current_temperature), Y1)
    assert abs(h1 - h0) <= 1e-9 * np.max(np.abs(m.species_enthalpies(r.current_temperature) * Y1))
    atoms = np.array([[m.mech_data['species'][n]['atoms'].get(e, 0.) for n in m.species_names]
                      for e in m.mech_data['elements']]) / m.molecular_weights[None, :]
    assert np.max(np.abs(atoms @ (Y1 - Y0))) <= 1e-12
    assert r.current_temperature > 2500. and out['temperature'][0] == 1200.
    assert np.all(np.diff(out.time_values) > 0) and abs(out.time_values[-1] - 1e-3) < 1e-15

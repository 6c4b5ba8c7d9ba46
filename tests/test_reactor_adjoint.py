"""Discrete adjoint of batched reactor integrations (spitfire_amd.reactor_batch): the gradient of the integrated
heat release L = int qdot dt with respect to ln A of every reaction.

Pinned by central differences of the frozen-step forward map (the map the discrete adjoint differentiates:
the same accepted step sizes replayed, ln A_r perturbed by +-1e-6), with the product's step loop and adjoint
recursion running on the oracle kernels (oracle/reactor_batch_oracle.py). Bound: normwise over the reactions,
max |adjoint - differences| <= 1e-6 max |differences| (difference noise: eps(L) / 1e-6 ~ 1e-10 relative to L)."""
import numpy as np
import pytest


def c1_reactors(m, T0, phis):
    names = m.species_names
    Y = []
    for phi in phis:
        X = np.zeros(m.n_species)
        X[names.index('H2')], X[names.index('O2')], X[names.index('N2')] = 2. * phi, 1., 3.76
        y = X * m.molecular_weights
        Y.append(y / y.sum())
    return np.asarray(T0, dtype=np.float64), np.stack(Y)


def test_adjoint_matches_differences_of_frozen_step_map(mech_c1_data, oracle_lib):
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from oracle.reactor_batch_oracle import OracleReactorBatch, scaled_mechanism
    m = ChemicalMechanismSpec(mech_c1_data)
    T0, Y0 = c1_reactors(m, [1200., 1500.], [1.0, 0.6])
    t_end = 1.e-4
    rb = OracleReactorBatch(m, T0, Y0)
    rb.integrate(t_end)
    L, G = rb.heat_release_gradient()
    assert np.allclose(L.numpy(), rb.heat_release().numpy(), rtol=1e-12, atol=0)
    G = G.numpy()
    sched = rb.step_sizes()
    # the reactions with the largest sensitivities of each type, plus the smallest nonzero one
    order = np.argsort(-np.abs(G).max(axis=0))
    types = [m.mech_data['reactions'][r]['type'] for r in range(m.n_reactions)]
    pick = []
    for t in ('simple', 'three-body', 'Troe'):
        pick += [r for r in order if types[r] == t][:2]
    eps = 1.e-6
    FD = np.zeros((len(T0), len(pick)))
    for c, r in enumerate(pick):
        Ls = []
        for sgn in (1., -1.):
            b = OracleReactorBatch(ChemicalMechanismSpec(scaled_mechanism(m.mech_data, r, np.exp(sgn * eps))), T0, Y0)
            b.integrate(t_end, dt_schedule=sched)
            Ls.append(b.heat_release().numpy())
        FD[:, c] = (Ls[0] - Ls[1]) / (2. * eps)
    err = np.max(np.abs(G[:, pick] - FD), axis=1) / np.max(np.abs(FD), axis=1)
    print('reactions', pick, 'normwise error per reactor', err)
    assert np.all(err <= 1e-6)


def test_adjoint_matches_forward_sensitivities(mech_c1_data, oracle_lib):
    """The discrete adjoint (reverse recursion over the recorded stages) against forward sensitivities of the
    same frozen-step map (the tangent-linear model, oracle/reactor_batch_oracle.py), for all 30 reactions of the
    10-species mechanism at once: two exact derivatives of one discrete map, so they agree to rounding (bound
    1e-9 normwise per reactor). The two share only the per-stage building blocks (Jacobian, dq/dlnA, the
    heat-release gradient), which the central-difference test above pins."""
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from oracle.reactor_batch_oracle import OracleReactorBatch, tangent_linear_heat_release_gradient
    m = ChemicalMechanismSpec(mech_c1_data)
    T0, Y0 = c1_reactors(m, [1200., 1500.], [1.0, 0.6])
    rb = OracleReactorBatch(m, T0, Y0)
    rb.integrate(1.e-4)
    _, G = rb.heat_release_gradient()
    T = tangent_linear_heat_release_gradient(rb)
    G, T = G.numpy(), T.numpy()
    err = np.max(np.abs(G - T), axis=1) / np.max(np.abs(T), axis=1)
    print('normwise adjoint vs tangent-linear per reactor', err, 'largest |dL/dlnA|', np.abs(T).max(axis=1))
    assert np.all(err <= 1e-9)


def test_log_pre_exponential_round_trip(mech_c1_data):
    """with_log_pre_exponentials(m, log_pre_exponentials(m)) reproduces the mechanism's rates; a shift of one
    reaction's ln A scales exactly that reaction's A (forward A, or the high-pressure A of a falloff reaction)."""
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.reactor_batch import log_pre_exponentials, with_log_pre_exponentials
    m = ChemicalMechanismSpec(mech_c1_data)
    la = log_pre_exponentials(m)
    assert la.shape == (m.n_reactions,)
    same = with_log_pre_exponentials(m, la)
    for a, b in zip(m.mech_data['reactions'], same.mech_data['reactions']):
        k = 'A' if 'A' in a else 'fwd-A'
        assert np.isclose(a[k], b[k], rtol=1e-14, atol=0)  # exp(log(A)): a few ulp
    types = [rx['type'] for rx in m.mech_data['reactions']]
    r = types.index('Troe')
    la2 = la.copy()
    la2[r] += 0.5
    m2 = with_log_pre_exponentials(m, la2)
    rx, rx2 = m.mech_data['reactions'][r], m2.mech_data['reactions'][r]
    assert np.isclose(rx2['fwd-A'] / rx['fwd-A'], np.exp(0.5), rtol=1e-14) and rx2['flf-A'] == rx['flf-A']


@pytest.mark.parametrize('mech', ['mech_c1_data', 'mech_53_data'])
def test_oracle_rate_of_progress_log_a_derivative_matches_differences(mech, request, oracle_lib):
    """The oracle's exact d q_r / d ln A_r (forward or high-pressure A; falloff reactions include the shift of Pr)
    against central differences of the oracle's production rates over mechanisms with A_r scaled by exp(+-1e-4),
    read from the participating species with the least rounding; bound 1e-6 relative plus the rounding of the two
    production rates carried through the difference."""
    from spitfire_amd import synthetic
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.mechdata import CompiledMechanism
    from oracle.reactor_batch_oracle import scaled_mechanism
    m = ChemicalMechanismSpec(request.getfixturevalue(mech))
    ns, nr = m.n_species, m.n_reactions
    x, p, _ = synthetic.random_states(ns, 4, seed=9)
    base = oracle_lib.CpuGriffon(CompiledMechanism(m.mech_data, order='name'))
    gross = base.batch_rhs_scale(np.ascontiguousarray(x), p)
    mw = m.molecular_weights
    names = list(m.species_names)
    N = np.zeros((ns, nr))
    for g, rx in enumerate(m._reactions):
        for s_, c in rx['reactants'].items():
            N[names.index(s_), g] -= c
        for s_, c in rx['products'].items():
            N[names.index(s_), g] += c
    eps = 1.e-4
    fo = m._reaction_file_order
    worst = 0.
    for g in range(nr):
        plus = oracle_lib.CpuGriffon(CompiledMechanism(scaled_mechanism(m.mech_data, int(fo[g]), np.exp(eps)),
                                                       order='name'))
        minus = oracle_lib.CpuGriffon(CompiledMechanism(scaled_mechanism(m.mech_data, int(fo[g]), np.exp(-eps)),
                                                        order='name'))
        for i in range(x.shape[1]):
            T = x[0, i]
            y = np.hstack([x[1:, i], 1. - x[1:, i].sum()])
            rho = p[i] / (m.gas_constant * T * np.sum(y / mw))
            qa = base.rate_of_progress_log_a_derivative(rho, T, y)
            ks = [k for k in range(ns - 1) if N[k, g] != 0.]
            k = min(ks, key=lambda k: gross[i, 1 + k] / (mw[k] * abs(N[k, g])))
            d = (plus.production_rates(T, rho, y)[k] - minus.production_rates(T, rho, y)[k]) / (2. * eps)
            fd = d / (mw[k] * N[k, g])
            noise = 64. * 2.2e-16 * rho * gross[i, 1 + k] / (2. * eps) / (mw[k] * abs(N[k, g]))
            worst = max(worst, abs(qa[g] - fd) / (1e-6 * abs(fd) + noise + 1e-300))
    print('worst error / bound', worst)
    assert worst <= 1.

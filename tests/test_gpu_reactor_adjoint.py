"""GPU parity of the reactor-adjoint path: the chemistry kernel's d(rate of progress)/d(ln A) against central
differences of the oracle's production rates over mechanisms with A scaled by exp(+-1e-5) (53 species, every
reaction type, falloff reactions with the high-pressure A); the batched GPU integration and discrete adjoint
against the same step loop and recursion on the oracle kernels; and the GPU adjoint against central differences
of the GPU's own frozen-step forward map on the 53-species mechanism.

Through an ignition the forward map amplifies perturbations strongly, so differences with a 1e-6 step are
dominated by the Newton tolerance (1e-12) carried through that amplification; the 53-species check uses a 1e-3
step with stage equations solved to 1e-14, where the differences settle (their truncation error is O(1e-6)
relative)."""
import numpy as np
import pytest

from tests.test_reactor_adjoint import c1_reactors

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_rate_constant_sensitivities_match_oracle_differences(mech_53_data, oracle_lib):
    import torch
    from spitfire_amd import synthetic
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.mechdata import CompiledMechanism
    from oracle.reactor_batch_oracle import scaled_mechanism
    m = ChemicalMechanismSpec(mech_53_data)
    ns, nr = m.n_species, m.n_reactions
    x, p, _ = synthetic.random_states(ns, 12, seed=7)  # SoA [ns, n]
    n = x.shape[1]
    dev = torch.device('cuda:0')
    X = torch.tensor(np.ascontiguousarray(x.T), device=dev)
    P = torch.tensor(p, device=dev)
    out = [torch.empty(s, dtype=torch.float64, device=dev) for s in ((n, ns), (n, ns * ns), (n, 4), (n, ns), (n, ns),
                                                                      (n, nr))]
    say('qa kernel, 53 species,', n, 'states')
    m.griffon.batch_reactor_jac_isobaric_terms(X, P, *out)
    torch.cuda.synchronize()
    aux, qa = out[2].cpu().numpy(), out[5].cpu().numpy()
    names = list(m.species_names)
    N = np.zeros((ns, nr))
    for g, rx in enumerate(m._reactions):
        for s_, c in rx['reactants'].items():
            N[names.index(s_), g] -= c
        for s_, c in rx['products'].items():
            N[names.index(s_), g] += c
    eps = 1.e-4
    fo = m._reaction_file_order
    plus = [oracle_lib.CpuGriffon(CompiledMechanism(scaled_mechanism(m.mech_data, int(fo[g]), np.exp(eps)),
                                                    order='name')) for g in range(nr)]
    minus = [oracle_lib.CpuGriffon(CompiledMechanism(scaled_mechanism(m.mech_data, int(fo[g]), np.exp(-eps)),
                                                     order='name')) for g in range(nr)]
    base = oracle_lib.CpuGriffon(CompiledMechanism(m.mech_data, order='name'))
    gross = base.batch_rhs_scale(np.ascontiguousarray(x), p)  # [n, ns]: magnitudes of the terms of each RHS entry
    mw = m.molecular_weights
    types = [m._reactions[g]['type'] for g in range(nr)]
    worst, where = 0., None
    for i in range(n):
        T = x[0, i]
        y = np.hstack([x[1:, i], 1. - x[1:, i].sum()])
        rho = aux[i, 0]
        for g in range(nr):
            ks = [k for k in range(ns - 1) if N[k, g] != 0.]
            # the participating species whose production rate has the least rounding relative to this reaction
            k = min(ks, key=lambda k: gross[i, 1 + k] / (mw[k] * abs(N[k, g])))
            d = (plus[g].production_rates(T, rho, y)[k] - minus[g].production_rates(T, rho, y)[k]) / (2. * eps)
            fd = d / (mw[k] * N[k, g])
            # rounding of the two production rates, carried through the difference, in units of qa
            noise = 64. * 2.2e-16 * rho * gross[i, 1 + k] / (2. * eps) / (mw[k] * abs(N[k, g]))
            e = abs(qa[i, g] - fd) / (1e-6 * abs(fd) + noise + 1e-300)
            if e > worst:
                worst, where = e, (i, g, types[g], qa[i, g], fd, noise, T)
    say('worst (error / bound) %.3e at state, reaction, type, qa, differences, noise, T = %s' % (worst, where))
    assert worst <= 1.


def test_gpu_adjoint_matches_oracle_adjoint_c1(mech_c1_data, oracle_lib):
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.reactor_batch import ReactorBatch
    from oracle.reactor_batch_oracle import OracleReactorBatch
    m = ChemicalMechanismSpec(mech_c1_data)
    T0, Y0 = c1_reactors(m, [1200., 1500.], [1.0, 0.6])
    say('C1 reactors to 0.1 ms: GPU forward + adjoint')
    g = ReactorBatch(m, T0, Y0)
    g.integrate(1.e-4)
    Lg, Gg = g.heat_release_gradient()
    o = OracleReactorBatch(m, T0, Y0)
    o.integrate(1.e-4)
    Lo, Go = o.heat_release_gradient()
    xg, xo = g.final_state.cpu().numpy(), o.final_state.numpy()
    scale = np.r_[1e3, np.ones(m.n_species - 1)]
    e_x = np.max(np.abs(xg - xo) / scale)
    e_L = np.max(np.abs(Lg.cpu().numpy() - Lo.numpy()) / np.abs(Lo.numpy()))
    Gg, Go = Gg.cpu().numpy(), Go.numpy()
    e_G = np.max(np.max(np.abs(Gg - Go), axis=1) / np.max(np.abs(Go), axis=1))
    say('steps gpu %d oracle %d; state %.2e, L %.2e, gradient %.2e' % (g.stats['steps'], o.stats['steps'], e_x, e_L,
                                                                        e_G))
    assert abs(g.stats['steps'] - o.stats['steps']) <= 2
    assert e_x <= 1e-8 and e_L <= 1e-9 and e_G <= 1e-6


def test_gpu_adjoint_matches_gpu_differences_53(mech_53_data):
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.reactor_batch import ReactorBatch
    from oracle.reactor_batch_oracle import scaled_mechanism
    m = ChemicalMechanismSpec(mech_53_data)
    T0, Y0 = c1_reactors(m, [1300.], [1.0])
    t_end = 2.e-4
    say('53 species, one reactor to 0.2 ms: adjoint and frozen-step differences')
    g = ReactorBatch(m, T0, Y0)
    g.integrate(t_end)
    L, G = g.heat_release_gradient()
    G = G.cpu().numpy()[0]
    sched = g.step_sizes()
    types = [rx['type'] for rx in m.mech_data['reactions']]
    order = np.argsort(-np.abs(G))
    pick = []
    for t in ('simple', 'three-body', 'Troe'):
        pick += [r for r in order if types[r] == t][:2]
    eps = 1.e-3
    fd = []
    for r in pick:
        Ls = []
        for sgn in (1., -1.):
            b = ReactorBatch(ChemicalMechanismSpec(scaled_mechanism(m.mech_data, int(r), np.exp(sgn * eps))), T0, Y0)
            b.integrate(t_end, dt_schedule=sched, nonlinear_solve_tolerance=1.e-14)
            Ls.append(float(b.heat_release().cpu().numpy()[0]))
        fd.append((Ls[0] - Ls[1]) / (2. * eps))
    fd = np.array(fd)
    err = np.max(np.abs(G[pick] - fd)) / np.max(np.abs(fd))
    say('steps %d, reactions %s, normwise error %.3e' % (g.stats['steps'], pick, err))
    assert err <= 1e-4


def test_heat_release_autograd_function(mech_c1_data):
    """reactor_batch.heat_release as a torch.autograd.Function: its forward value equals ReactorBatch's
    heat_release() of the same integration, and backward of a weighted sum over the reactors equals the same
    weights applied to the discrete adjoint (heat_release_gradient) of that integration, to 1e-12 normwise
    (the integrations are the same deterministic kernels). A shifted ln A rebuilds the mechanism: the loss
    changes, and its backward is the adjoint of the shifted mechanism."""
    import torch
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.reactor_batch import ReactorBatch, heat_release, log_pre_exponentials, \
        with_log_pre_exponentials
    m = ChemicalMechanismSpec(mech_c1_data)
    T0, Y0 = c1_reactors(m, [1200., 1500., 1350.], [1.0, 0.6, 1.6])
    dev = torch.device('cuda:0')
    wts = torch.tensor([0.3, -1.7, 1.1], dtype=torch.float64, device=dev)
    la0 = log_pre_exponentials(m)
    shift = np.zeros_like(la0)
    shift[[1, 5]] = [0.2, -0.3]
    for la_np in (la0, la0 + shift):
        say('autograd heat release, ln A shifted by', shift if la_np is not la0 else 0)
        la = torch.tensor(la_np, device=dev, requires_grad=True)
        L = heat_release(la, m, T0, Y0, t_end=1.e-4)
        (L * wts).sum().backward()
        rb = ReactorBatch(m if la_np is la0 else with_log_pre_exponentials(m, la_np), T0, Y0)
        rb.integrate(1.e-4)
        L2, G = rb.heat_release_gradient()
        ref = (wts[:, None] * G).sum(dim=0)
        eL = float((L - L2).abs().max() / L2.abs().max())
        eg = float((la.grad - ref).abs().max() / ref.abs().max())
        say('  L', L.tolist(), 'rel diff %.3e, gradient normwise diff %.3e' % (eL, eg))
        assert eL <= 1e-12 and eg <= 1e-12
        if la_np is la0:
            L0 = L.detach().clone()
        else:
            assert float((L.detach() - L0).abs().max()) > 1e-6 * float(L0.abs().max())


def test_graph_captured_newton_iterations_match_eager(mech_c1_data):
    """The stage Newton iterations replayed from a captured HIP graph (two iterations per replay, eager work and
    host synchronizations between replays) give bitwise the eager path's states, step records and counts, for
    controlled steps and for a replayed step schedule."""
    import torch
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.reactor_batch import ReactorBatch
    m = ChemicalMechanismSpec(mech_c1_data)
    rng = np.random.default_rng(5)
    T0, Y0 = c1_reactors(m, rng.uniform(1000., 1400., 64), rng.uniform(0.5, 2., 64))
    out = {}
    for graphs in (False, True):
        say('64 C1 reactors to 1e-4 s, graphs', graphs)
        g = ReactorBatch(m, T0, Y0)
        x = g.integrate(1.e-4, use_graphs=graphs)
        sched = g.step_sizes()
        g2 = ReactorBatch(m, T0, Y0)
        x2 = g2.integrate(1.e-4, dt_schedule=sched, use_graphs=graphs)
        out[graphs] = (x.cpu(), [tuple(a.cpu() for a in rec) for rec in g.records], dict(g.stats), x2.cpu())
        say('  steps', g.stats)
    (a, ra, sa, a2), (b, rb, sb, b2) = out[False], out[True]
    assert torch.equal(a, b) and torch.equal(a2, b2) and sa == sb and len(ra) == len(rb)
    assert all(torch.equal(p, q) for u, v in zip(ra, rb) for p, q in zip(u, v))

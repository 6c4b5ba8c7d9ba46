"""The device-resident batched steady flamelet solver (gf_flamelet_batch_*, spitfire_amd.flamelet.FlameletBatch)
against the oracle: the batched flamelet RHS against the oracle's flamelet_rhs for each flamelet, and converged
steady states against the oracle's own pseudo-transient continuation (the reference algorithm with LAPACK
eigenvalues and LU, oracle/flamelet_oracle.py), both solved to a scaled residual of 1e-8 (the FP64 rounding
floor of the scaled flamelet residual is near 5e-9 on 128-point clustered grids, so 1e-10 is not reachable in
general).

Bound on converged states: max |x_gpu - x_oracle| / scale <= 1e-8, scale 1e3 for temperature and 1 for mass
fractions (the oracle's own solutions at residual 1e-8 and 1e-9 differ by under 1e-10 on these cases)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _streams(m):
    return m.stream(stp_air=True), m.stream('TPY', (300., 101325., 'H2:1 N2:1'))


def _flamelets(m, chis, npts, cls=None, ics=None):
    from spitfire_amd.flamelet import Flamelet, FlameletSpec
    cls = cls or Flamelet
    air, fuel = _streams(m)
    ics = ics or ['Burke-Schumann'] * len(chis)  # FlameletSpec's default initial condition
    return [cls(FlameletSpec(mech_spec=m, oxy_stream=air, fuel_stream=fuel, grid_points=n,
                             stoich_dissipation_rate=c, initial_condition=ic)) for c, n, ic in zip(chis, npts, ics)]


def _scaled(a, b, ns):
    return float(np.max(np.abs(a - b) / np.tile(np.r_[1e3, np.ones(ns - 1)], a.size // ns)))


def _near_steady_guesses(m, chis, npts):
    """steady states (batched psitc) with a smooth 1% temperature bump: guesses Newton converges from"""
    from spitfire_amd.flamelet import FlameletBatch
    ns = m.n_species
    states, conv, _ = FlameletBatch(_flamelets(m, chis, npts)).steady_solve_psitc(tolerance=1e-9)
    assert conv.all()
    out = []
    for s, n in zip(states, npts):
        g = s.copy()
        zi = np.arange(1, n - 1) / (n - 1.)
        g[::ns] *= 1. + 0.01 * np.sin(np.pi * zi)
        out.append(g)
    return states, out


@pytest.mark.parametrize('mech', ['mech_c1_data', 'mech_53_data'])
def test_batch_rhs_matches_oracle(mech, request, oracle_lib):
    from spitfire_amd.flamelet import FlameletBatch
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from oracle.flamelet_oracle import OracleFlamelet
    m = ChemicalMechanismSpec(request.getfixturevalue(mech))
    chis, npts = [0.5, 20., 300.], [9, 34, 17]
    say('batch rhs', mech, npts)
    fb = FlameletBatch(_flamelets(m, chis, npts))
    got = fb.split(fb.rhs())
    for f, g in zip(_flamelets(m, chis, npts, OracleFlamelet), got):
        ref = f.rhs(0., f.initial_interior_state)
        ns = m.n_species
        e = np.max(np.abs(g - ref).reshape(-1, ns).max(axis=1) /
                   np.maximum(np.abs(ref).reshape(-1, ns).max(axis=1), 1e-300))
        say('  nzi %d row err %.3e' % (f.n_interior_points, e))
        assert e <= 1e-12


@pytest.mark.parametrize('mech, chis', [('mech_c1_data', [1., 10., 100., 1000.]), ('mech_53_data', [1., 10.])])
def test_batch_steady_states_match_oracle(mech, chis, request, oracle_lib):
    from spitfire_amd.flamelet import FlameletBatch
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from oracle.flamelet_oracle import OracleFlamelet
    m = ChemicalMechanismSpec(request.getfixturevalue(mech))
    ns = m.n_species
    npts = [34] * len(chis)
    say('batch steady', mech, chis)
    fb = FlameletBatch(_flamelets(m, chis, npts))
    states, conv, its = fb.steady_solve_psitc(tolerance=1e-8)
    say('  gpu iterations', its.tolist(), 'counters', fb.counters, 'residuals', fb.residuals.tolist())
    assert conv.all()
    for f, g in zip(_flamelets(m, chis, npts, OracleFlamelet), states):
        _, it, ok, _ = f.steady_solve_psitc(tolerance=1e-8)
        assert ok
        ref = f.current_interior_state
        scale = np.tile(np.r_[1e3, np.ones(ns - 1)], f.n_interior_points)
        e = np.max(np.abs(g - ref) / scale)
        say('  chi_st %.3g Tmax gpu %.2f oracle %.2f max scaled diff %.3e (oracle %d iterations)'
            % (f.dissipation_rate.max(), g[::ns].max(), ref[::ns].max(), e, it))
        assert e <= 1e-8


@pytest.mark.parametrize('mech, chis', [('mech_c1_data', [1., 10., 100., 1000.]), ('mech_53_data', [1., 10.])])
def test_mixed_precision_lu_converges_to_fp64_solution(mech, chis, request):
    """FP32 block-LU factors with FP64 residuals reach the FP64-LU steady state (to 1e-8 scaled, both solved to
    a scaled residual of 1e-8): the factors only precondition the FP64 residual."""
    from spitfire_amd.flamelet import FlameletBatch
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    m = ChemicalMechanismSpec(request.getfixturevalue(mech))
    ns = m.n_species
    npts = [34] * len(chis)
    say('mixed precision', mech, chis)
    s64, c64, i64 = FlameletBatch(_flamelets(m, chis, npts)).steady_solve_psitc(tolerance=1e-8, lu_precision=64)
    fb = FlameletBatch(_flamelets(m, chis, npts))
    s32, c32, i32 = fb.steady_solve_psitc(tolerance=1e-8, lu_precision=32)
    say('  iterations fp64', i64.tolist(), 'fp32', i32.tolist(), 'residuals', fb.residuals.tolist())
    assert c64.all() and c32.all()
    for a, b_ in zip(s32, s64):
        scale = np.tile(np.r_[1e3, np.ones(ns - 1)], a.size // ns)
        e = np.max(np.abs(a - b_) / scale)
        say('  max scaled diff %.3e' % e)
        assert e <= 1e-8


def test_batch_newton_matches_reference_newton(mech_c1_data, oracle_lib):
    """The batched Newton method (steady_solve_newton: frozen -J refreshed only after a line search, the
    reference's compounding line search, no restarts) against the same method on the oracle flamelets from the
    same guesses. Near-steady guesses: both converge, in the same number of iterations, to states equal to
    1e-10 scaled. Burke-Schumann guesses: the same outcome (converged or not)."""
    from spitfire_amd.flamelet import FlameletBatch
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from oracle.flamelet_oracle import OracleFlamelet
    m = ChemicalMechanismSpec(mech_c1_data)
    ns = m.n_species
    chis, npts = [1., 10., 100.], [24, 31, 24]
    _, guesses = _near_steady_guesses(m, chis, npts)
    ics = guesses + ['Burke-Schumann'] * 3
    chis, npts = chis * 2, npts * 2
    say('batch newton', chis, npts)
    fb = FlameletBatch(_flamelets(m, chis, npts, ics=ics))
    states, conv, its = fb.steady_solve_newton(tolerance=1e-9, max_iterations=38, raise_on_failure=False)
    say('  gpu converged', conv.tolist(), 'iterations', its.tolist(), 'residuals', fb.residuals.tolist())
    for k, f in enumerate(_flamelets(m, chis, npts, OracleFlamelet, ics=ics)):
        _, it, ok = f.steady_solve_newton(tolerance=1e-9, max_iterations=38, verbose=False)
        ref = f.current_interior_state
        say('  flamelet %d: oracle converged %s in %s iterations' % (k, ok, it))
        assert bool(conv[k]) == bool(ok)
        if k < 3:
            assert ok and it == its[k]
            e = _scaled(states[k], ref, ns)
            say('    max scaled diff %.3e' % e)
            assert e <= 1e-10


def test_batch_compute_steady_state_stages(mech_c1_data, oracle_lib):
    """compute_steady_state's cascade. With the default stage limits, near-steady guesses converge in the Newton
    stage and every flamelet converges. With Newton and psitc capped at 2 iterations (no restarts), flamelets
    from Burke-Schumann guesses go on to the ESDIRK64 stage (per-flamelet steps, batched) and converge there;
    the oracle runs the same cascade flamelet by flamelet, and the batch matches it (states to 1e-8 scaled,
    step counts within 5 %; on these cases identical counts and 1.5e-12). All end at the
    oracle's pseudo-transient steady state from the same guess: 1e-8 scaled from the Newton and psitc stages
    (all solved to a residual of 1e-9), 1e-6 from time stepping stopped at a steadiness rate of 1e-9.
    Time stepping stops on |dx/scale|/dt, not on the residual: its final scaled residual can be near 1e-5 in the
    temperature rows (the reference algorithm on the oracle ends the same way on these cases, 5e-6..1.4e-5)
    while the state is within about 1e-12 of the steady state."""
    from spitfire_amd.flamelet import FlameletBatch
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from oracle.flamelet_oracle import OracleFlamelet
    m = ChemicalMechanismSpec(mech_c1_data)
    ns = m.n_species
    _, guesses = _near_steady_guesses(m, [1., 100.], [24, 24])
    cases = [
        ([1., 100., 3., 30.], [24, 24, 24, 28], guesses + ['Burke-Schumann'] * 2, {}, 1e-8),
        ([3., 30., 300.], [24, 28, 20], ['Burke-Schumann'] * 3,
         {'newton_args': {'max_iterations': 2}, 'psitc_args': {'max_iterations': 2, 'adaptive_restart': False}},
         1e-6),
    ]
    for chis, npts, ics, limits, bound in cases:
        say('cascade', chis, npts, limits)
        fb = FlameletBatch(_flamelets(m, chis, npts, ics=ics))
        states, conv, its = fb.compute_steady_state(tolerance=1e-9, **limits)
        say('  methods', fb.method.tolist(), 'steps/iterations', its.tolist(), 'residuals', fb.residuals.tolist(),
            'transient', fb.stage_counters.get('transient'))
        assert conv.all()
        if limits:
            assert list(fb.method) == ['transient'] * len(chis)
        else:
            assert list(fb.method[:2]) == ['newton', 'newton']
        for k, f in enumerate(_flamelets(m, chis, npts, OracleFlamelet, ics=ics)):
            _, it, ok, _ = f.steady_solve_psitc(tolerance=1e-9)
            assert ok
            ref = f.current_interior_state
            e = _scaled(states[k], ref, ns)
            say('  flamelet %d (%s): Tmax %.2f oracle %.2f, max scaled diff %.3e'
                % (k, fb.method[k], states[k][::ns].max(), ref[::ns].max(), e))
            assert e <= bound
            if limits:
                o = _flamelets(m, chis[k:k + 1], npts[k:k + 1], OracleFlamelet, ics=ics[k:k + 1])[0]
                o.compute_steady_state(tolerance=1e-9, **limits)
                eo = _scaled(states[k], o.current_interior_state, ns)
                say('    oracle cascade: %d time steps (batch %d), max scaled diff to the batch %.3e'
                    % (o.info['steps'], its[k], eo))
                # the batch runs the reference step loop per flamelet: the same steps up to rounding
                assert eo <= 1e-8 and abs(o.info['steps'] - its[k]) <= max(2, 0.05 * its[k])


def test_time_stepping_stage_fp32_and_state_errors(mech_c1_data, oracle_lib):
    """The time-stepping stage with FP32 block inverses (FP64 residuals and stage equations) reaches the
    FP64 stage's steady state (both stop at a steadiness rate of 1e-9; 1e-6 scaled, the distance that rate
    leaves); a solve for a flamelet with no factors of its own, and a setup in the other precision while factors
    are held, are refused with GF_ERR_STATE rather than computed."""
    import torch
    from spitfire_amd import _native
    from spitfire_amd.flamelet import FlameletBatch
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    m = ChemicalMechanismSpec(mech_c1_data)
    ns = m.n_species
    chis, npts = [3., 30.], [24, 28]
    out = {}
    for prec in (64, 32):
        say('time stepping, lu precision', prec)
        fb = FlameletBatch(_flamelets(m, chis, npts))
        states, conv, steps = fb.integrate_to_steady(steady_tolerance=1e-9, lu_precision=prec,
                                                     first_time_step=1e-8, transient_tolerance=1e-8,
                                                     max_time_step=1e4)
        say('  steps', steps.tolist(), 'residuals', fb.residuals.tolist())
        assert conv.all()
        out[prec] = states
    for a, b in zip(out[32], out[64]):
        e = _scaled(a, b, ns)
        say('  fp32 vs fp64 max scaled diff %.3e' % e)
        assert e <= 1e-6
    # state errors, from the host side: no kernel runs
    fb = FlameletBatch(_flamelets(m, chis, npts))
    lib = _native.lib()
    dev = torch.device('cuda:0')
    x = torch.tensor(fb._x0, device=dev)
    r = torch.zeros_like(x)
    o = torch.zeros_like(x)
    flags = np.array([1, 0], dtype=np.int32)
    assert lib.gf_flamelet_batch_transient_solve(fb._h, _native.iptr(flags), r.data_ptr(), o.data_ptr()) == 2
    pre = np.array([1e-6, 1e-6])
    _native.check(lib.gf_flamelet_batch_transient_setup(fb._h, _native.iptr(flags), x.data_ptr(), _native.dptr(pre), 64))
    assert lib.gf_flamelet_batch_transient_setup(fb._h, _native.iptr(flags), x.data_ptr(), _native.dptr(pre), 32) == 2
    _native.check(lib.gf_flamelet_batch_transient_solve(fb._h, _native.iptr(flags), r.data_ptr(), o.data_ptr()))
    torch.cuda.synchronize()
    assert bool(torch.isfinite(o).all())


def test_time_stepping_stage_bounds(mech_c1_data):
    """The time-stepping stage's bounds, forced from the host with a steady tolerance no state can meet (1e-30):
    a stall window stops each flamelet once its steadiness measure stops halving, a step cap stops it at that
    many steps, and both report the flamelets as not converged with the reason in transient_stop."""
    from spitfire_amd.flamelet import FlameletBatch
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    m = ChemicalMechanismSpec(mech_c1_data)
    chis, npts = [3., 30.], [24, 28]
    say('time stepping bounds: stall window 25, then a 60-step cap')
    fb = FlameletBatch(_flamelets(m, chis, npts))
    _, conv, steps = fb.integrate_to_steady(steady_tolerance=1e-30, first_time_step=1e-8, transient_tolerance=1e-8,
                                            max_time_step=1e4, maximum_time_step_count=100000, stall_window=25,
                                            raise_on_failure=False)
    say('  stall: steps', steps.tolist(), 'stop', list(fb.transient_stop))
    assert not conv.any() and list(fb.transient_stop) == ['stalled', 'stalled'] and steps.max() < 100000
    fb = FlameletBatch(_flamelets(m, chis, npts))
    _, conv, steps = fb.integrate_to_steady(steady_tolerance=1e-30, first_time_step=1e-8, transient_tolerance=1e-8,
                                            max_time_step=1e4, maximum_time_step_count=60, raise_on_failure=False)
    say('  cap: steps', steps.tolist(), 'stop', list(fb.transient_stop))
    assert not conv.any() and list(fb.transient_stop) == ['step limit', 'step limit'] and steps.tolist() == [61, 61]


@pytest.mark.parametrize('case', ['convection', 'scaled'])
def test_nonadiabatic_batch_steady_states_match_oracle(mech_c1_data, oracle_lib, case):
    """Nonadiabatic flamelets (convection + radiation heat loss, flamelet_kernels.cpp:1121-1138, 1290-1320) solved
    by the batched pseudo-transient continuation, FP64 and FP32 block LU, against the oracle's single-flamelet
    solves: plain heat loss with scalar parameters, and heat loss scaled by the temperature range with per-point
    parameter arrays."""
    from spitfire_amd.flamelet import Flamelet, FlameletBatch, FlameletSpec
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from oracle.flamelet_oracle import OracleFlamelet
    m = ChemicalMechanismSpec(mech_c1_data)
    ns = m.n_species
    air, fuel = m.stream(stp_air=True), m.stream('TPY', (300., 101325., 'H2:1 N2:1'))
    chis, n = [3., 30.], 24

    def spec(chi):
        if case == 'convection':
            hl = dict(convection_temperature=300., radiation_temperature=300., convection_coefficient=1e5,
                      radiative_emissivity=0.5)
        else:
            hl = dict(convection_temperature=300., radiation_temperature=np.full(n - 2, 350.),
                      convection_coefficient=1e8 * (1. + np.linspace(0., 1., n - 2)),
                      radiative_emissivity=np.full(n - 2, 0.3), scale_heat_loss_by_temp_range=True)
        return FlameletSpec(mech_spec=m, oxy_stream=air, fuel_stream=fuel, grid_points=n, stoich_dissipation_rate=chi,
                            heat_transfer='nonadiabatic', **hl)

    refs = []
    for chi in chis:
        o = OracleFlamelet(spec(chi))
        _, it, ok, _ = o.steady_solve_psitc(tolerance=1e-9)
        a = OracleFlamelet(FlameletSpec(mech_spec=m, oxy_stream=air, fuel_stream=fuel, grid_points=n,
                                        stoich_dissipation_rate=chi))
        a.steady_solve_psitc(tolerance=1e-9)
        say('nonadiabatic %s chi %g: oracle %d iterations, Tmax %.1f (adiabatic %.1f)'
            % (case, chi, it, o.current_temperature.max(), a.current_temperature.max()))
        assert ok and a.current_temperature.max() - o.current_temperature.max() > 50.
        refs.append(o.current_interior_state)
    for prec in (64, 32):
        fb = FlameletBatch([Flamelet(spec(c)) for c in chis])
        states, conv, its = fb.steady_solve_psitc(tolerance=1e-9, lu_precision=prec)
        for k, (g, ref) in enumerate(zip(states, refs)):
            e = _scaled(g, ref, ns)
            say('  fp%d flamelet %d: %d iterations, max scaled diff to the oracle %.3e' % (prec, k, its[k], e))
            assert conv[k] and e <= 1e-8


def test_lewis_numbers_batch_matches_oracle(mech_c1_data, oracle_lib):
    """Non-unity Lewis numbers, different for each flamelet (each Flamelet's _lewis_numbers, the array the
    reference passes to flamelet_stencils, flamelet_kernels.cpp:31-48): the batched right-hand side and the batched
    pseudo-transient steady states against the oracle flamelets with the same numbers."""
    from spitfire_amd.flamelet import FlameletBatch
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from oracle.flamelet_oracle import OracleFlamelet
    m = ChemicalMechanismSpec(mech_c1_data)
    ns = m.n_species
    chis, npts = [3., 30.], [20, 26]
    rng = np.random.default_rng(4)
    les = [np.r_[1., rng.uniform(0.3, 2.0, ns - 1)] for _ in chis]

    def with_lewis(fl):
        for f, le in zip(fl, les):
            f._lewis_numbers = le.copy()
            f._setup_stencils()
        return fl

    say('Lewis numbers, batch rhs and psitc vs oracle')
    fb = FlameletBatch(with_lewis(_flamelets(m, chis, npts)))
    got = fb.split(fb.rhs())
    orc = with_lewis(_flamelets(m, chis, npts, OracleFlamelet))
    for f, g in zip(orc, got):
        ref = f.rhs(0., f.initial_interior_state)
        e = np.max(np.abs(g - ref).reshape(-1, ns).max(axis=1) / np.maximum(np.abs(ref).reshape(-1, ns).max(axis=1),
                                                                          1e-300))
        say('  rhs nzi %d row err %.3e' % (f.n_interior_points, e))
        assert e <= 1e-12
    # the Lewis numbers move the right-hand side far beyond that tolerance (the comparison is not vacuous)
    unity = _flamelets(m, chis, npts, OracleFlamelet)[0].rhs(0., orc[0].initial_interior_state)
    d = np.abs(unity - orc[0].rhs(0., orc[0].initial_interior_state)).reshape(-1, ns).max(axis=1)
    assert np.max(d / np.abs(unity).reshape(-1, ns).max(axis=1)) > 1e-6
    states, conv, its = fb.steady_solve_psitc(tolerance=1e-9)
    for k, f in enumerate(orc):
        _, it, ok, _ = f.steady_solve_psitc(tolerance=1e-9)
        e = _scaled(states[k], f.current_interior_state, ns)
        say('  flamelet %d: batch %d iterations, oracle %s, max scaled diff %.3e' % (k, its[k], it, e))
        assert conv[k] and ok and e <= 1e-8

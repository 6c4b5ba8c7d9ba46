"""CPU tests that pin the oracle (oracle/griffon_oracle.cpp) before it is trusted as the checker.

Pins: (1) the golden fixtures written by the reference's own griffon build (tests/golden/make_golden.py), which
the restatement must reproduce to roundoff; (2) mass conservation of the production rates; (3) central finite
differences of the RHS for every Jacobian column (the reference's own check,
tests/griffon/test_rhsjac_isobaric_closed_adiabatic.py:39-61); (4) dense LAPACK (numpy) solves of the
assembled block-tridiagonal matrix and of each diagonal block's eigenvalues; (5) the live reference build on
fresh random states, where it exists (bit-identical expected).
"""
import os

import numpy as np
import pytest

from tests.helpers import rel_err

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _cm(data):
    from spitfire_amd.mechdata import CompiledMechanism
    return CompiledMechanism(data, order='index')  # the reference build names species S000.. (index order)


def _fixture(name):
    return np.load(os.path.join(GOLD, name))


@pytest.mark.parametrize('fx, mech', [('chem_c1.npz', 'mech_c1_data'), ('chem_53.npz', 'mech_53_data')])
def test_mechanism_regenerates(fx, mech, request):
    from tests.golden.make_golden import mech_hash
    f = _fixture(fx)
    assert mech_hash(_cm(request.getfixturevalue(mech))) == str(f['mech_hash'])


@pytest.mark.parametrize('fx, mech', [('chem_c1.npz', 'mech_c1_data'), ('chem_53.npz', 'mech_53_data')])
def test_oracle_chemistry_matches_reference_fixtures(fx, mech, request, oracle_lib):
    f = _fixture(fx)
    o = oracle_lib.CpuGriffon(_cm(request.getfixturevalue(mech)))
    x, p = f['x'], f['p']
    for i in range(x.shape[1]):
        rhs = o.reactor_rhs_isobaric(x[:, i], p[i])
        jr, jac = o.reactor_jac_isobaric(x[:, i], p[i])
        y = np.append(x[1:, i], 1. - np.sum(x[1:, i]))
        w = o.production_rates(x[0, i], f['rho'][i], y)
        sens = o.prod_rates_sens(f['rho'][i], x[0, i], y)
        assert rel_err(rhs, f['rhs'][i], 1e-300) <= 1e-13
        assert rel_err(jr, f['jac_rhs'][i], 1e-300) <= 1e-13
        assert rel_err(jac, f['jac'][i], 1e-300) <= 1e-13
        assert rel_err(w, f['w'][i], 1e-300) <= 1e-13
        assert rel_err(sens, f['sens'][i], 1e-300) <= 1e-13


def test_mass_conservation(mech_53_data, oracle_lib):
    from spitfire_amd import synthetic
    cm = _cm(mech_53_data)
    o = oracle_lib.CpuGriffon(cm)
    st, p, Y = synthetic.random_states(cm.ns, 64, seed=5)
    for i in range(64):
        T = st[0, i]
        rho = p[i] / (np.sum(Y[i] / cm.mw) * T * cm.gas_constant)
        w = o.production_rates(T, rho, Y[i])
        assert abs(np.sum(w)) <= 1e-13 * np.sum(np.abs(w))


def test_jacobian_matches_central_differences(mech_c1_data, oracle_lib):
    """Same check as the reference (FD step 1e-6 in T and Y, |J - FD| / (|J| + 1) < 1e-2)."""
    cm = _cm(mech_c1_data)
    o = oracle_lib.CpuGriffon(cm)
    ns = cm.ns
    rng = np.random.default_rng(2)
    for trial in range(6):
        T = rng.uniform(800., 2200.)
        y = rng.dirichlet(np.ones(ns))
        p = 101325. * rng.uniform(0.5, 5.)
        state = np.hstack([T, y[:-1]])
        _, jac = o.reactor_jac_isobaric(state, p)
        jac = jac.reshape((ns, ns), order='F')
        fd = np.zeros((ns, ns))
        d = 1.e-6
        fd[:, 0] = (o.reactor_rhs_isobaric(state + np.eye(ns)[0] * d, p) -
                    o.reactor_rhs_isobaric(state - np.eye(ns)[0] * d, p)) / (2 * d)
        for k in range(ns - 1):
            e = np.zeros(ns)
            e[1 + k] = d
            fd[:, 1 + k] = (o.reactor_rhs_isobaric(state + e, p) - o.reactor_rhs_isobaric(state - e, p)) / (2 * d)
        assert np.max(np.abs(jac - fd) / (np.abs(jac) + 1.)) < 1.e-2


def _dense_btd(values, nb, bs):
    A = np.zeros((nb * bs, nb * bs))
    for i in range(nb):
        A[i * bs:(i + 1) * bs, i * bs:(i + 1) * bs] = values[i * bs * bs:(i + 1) * bs * bs].reshape((bs, bs),
                                                                                                  order='F')
    sub = values[nb * bs * bs:nb * bs * bs + (nb - 1) * bs]
    sup = values[nb * bs * bs + (nb - 1) * bs:]
    for i in range(1, nb):
        for q in range(bs):
            A[i * bs + q, (i - 1) * bs + q] = sub[(i - 1) * bs + q]
            A[(i - 1) * bs + q, i * bs + q] = sup[(i - 1) * bs + q]
    return A


@pytest.mark.parametrize('fx, mech', [('flamelet_c1.npz', 'mech_c1_data'), ('flamelet_53.npz', 'mech_53_data')])
def test_oracle_flamelet_matches_reference_fixtures(fx, mech, request, oracle_lib):
    f = _fixture(fx)
    o = oracle_lib.CpuGriffon(_cm(request.getfixturevalue(mech)))
    nzi, ns = int(f['nzi']), int(f['ns'])
    st = o.flamelet_stencils(f['dz'], nzi, f['chi'], np.ones(ns))
    for a, k in zip(st, ('cmaj', 'csub', 'csup', 'mco', 'nco')):
        assert rel_err(a, f[k], 1e-300) <= 1e-14
    rows, cols = o.flamelet_jac_indices(nzi)
    assert np.array_equal(rows, f['rows']) and np.array_equal(cols, f['cols'])
    args = (f['state'], float(f['p']), f['oxy'], f['fuel'], nzi, st, f['chi'])
    for key in f.files:
        if key.startswith('rhs_') and 'hl' not in key:
            ef, vcp = int(key[4]), int(key[5])
            assert rel_err(o.flamelet_rhs(*args, incl_ef=ef, incl_vcp=vcp), f[key], 1e-300) <= 1e-13
        if key.startswith('jac_') and key != 'jac_so' and 'hl' not in key:
            ef, vcp = int(key[4]), int(key[5])
            assert rel_err(o.flamelet_jacobian(*args, incl_ef=ef, incl_vcp=vcp)[0], f[key], 1e-300) <= 1e-13
    # nonadiabatic (convection + radiation, plain and scaled by the temperature range; C++ argument order)
    hl = (f['hl_Tc'], f['hl_hc'], f['hl_Tr'], f['hl_hr'])
    for sc in (0, 1):
        kw = dict(heat_loss=hl, incl_ef=1, incl_vcp=1, scaled_hl=sc)
        assert rel_err(o.flamelet_rhs(*args, **kw), f['rhs_hl%d' % sc], 1e-300) <= 1e-13
        assert rel_err(o.flamelet_jacobian(*args, **kw)[0], f['jac_hl%d' % sc], 1e-300) <= 1e-13
    so = o.flamelet_jacobian(*args, heat_loss=hl, incl_ef=1, incl_vcp=1, scaled_hl=1, scale_and_offset=True,
                             prefactor=-2.5e-4)[0]
    assert rel_err(so, f['jac_hl1_so'], 1e-300) <= 1e-13
    # the heat loss moves the temperature rows by far more than the tolerance (the check is not vacuous)
    assert np.max(np.abs(f['rhs_hl0'] - f['rhs_11'])[::ns] / np.abs(f['rhs_11'][::ns])) > 1e-6
    so = o.flamelet_jacobian(*args, scale_and_offset=True, prefactor=-2.5e-4)[0]
    assert rel_err(so, f['jac_so'], 1e-300) <= 1e-13
    # eigenvalue-based pseudo-time scale: numpy (LAPACK dgeev) on the same blocks as the reference
    expeig = o.flamelet_jacobian(*args, compute_eig=True, diffterm=20.)[1]
    assert rel_err(expeig, f['expeig'], 1e-300) <= 1e-8
    # block Thomas: same solution as the reference and as a dense LAPACK solve of the assembled matrix
    for k in ('btd', 'btdr'):  # the psitc matrix, then a random system whose factorization pivots
        d, l, piv = o.btddod_full_factorize(f[k + '_A'], nzi, ns)
        if k == 'btdr':
            assert np.array_equal(piv, f['btdr_piv'])
        x = o.btddod_full_solve(d, l, piv, f['btd_b'], nzi, ns)
        A = _dense_btd(f[k + '_A'], nzi, ns)
        x_np = np.linalg.solve(A, f['btd_b'])
        scale = np.max(np.abs(x_np))
        assert np.max(np.abs(x - f[k + '_x'])) <= 1e-12 * scale
        assert np.max(np.abs(x - x_np)) <= 1e-12 * scale
    assert rel_err(o.btddod_scale_and_add_diagonal(f['jac_00'], -1., f['sad_diag'], 1., nzi, ns), f['sad'],
                   1e-300) <= 1e-15


def test_oracle_bit_identical_to_live_reference(mech_53_data, oracle_lib):
    if not oracle_lib.ref_available():
        pytest.skip('oracle/_ref (the reference build) is not present')
    from spitfire_amd import synthetic
    cm = _cm(mech_53_data)
    o = oracle_lib.CpuGriffon(cm)
    r = oracle_lib.CpuGriffon(cm, ref=True)
    st, p, _ = synthetic.random_states(cm.ns, 48, seed=99)
    ro, jo = o.batch_rhs_jac(st, p, nthreads=1)
    rr, jr = r.batch_rhs_jac(st, p, nthreads=1)
    assert np.array_equal(ro, rr) and np.array_equal(jo, jr)

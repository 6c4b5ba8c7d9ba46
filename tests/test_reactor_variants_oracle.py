"""The oracle's restatement of the reactor variants against the reference build's outputs (tests/golden/reactors_*.npz,
written by make_golden.py from oracle/_ref): isobaric reactors with open / diathermal / isothermal terms and option-1
sensitivities, isochoric reactors (isochoric_reactor_kernels.cpp:18-333), option-1 raw sensitivities
(rates_sensitivities_no_tbaf_sens.cpp) and the thermodynamic getters (griffon.pyx:684-760). Tolerance 1e-13 per
row relative to the row's largest entry (the restatement follows the reference's formulas and summation orders)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
TOL = 1e-13


def _rows(a, b, n):
    A, B = np.asarray(a).reshape(-1, n, n), np.asarray(b).reshape(-1, n, n)
    return float(np.max(np.max(np.abs(A - B), axis=1) / np.maximum(np.max(np.abs(B), axis=1), 1e-300)))


def _vec(a, b):
    a, b = np.atleast_2d(a), np.atleast_2d(b)
    return float(np.max(np.max(np.abs(a - b), axis=1) / np.maximum(np.max(np.abs(b), axis=1), 1e-300)))


@pytest.mark.parametrize('fx, mech', [('reactors_c1.npz', 'mech_c1_data'), ('reactors_53.npz', 'mech_53_data')])
def test_oracle_reactor_variants_match_reference(fx, mech, request, oracle_lib):
    from spitfire_amd.mechdata import CompiledMechanism
    f = np.load(os.path.join(GOLD, fx))
    cpu = oracle_lib.CpuGriffon(CompiledMechanism(request.getfixturevalue(mech), order='name'))
    ns = int(f['ns'])
    Tin, tau, Tf, Ts, hc, eps, sov = f['params']
    err = {}
    for i in range(f['x'].shape[1]):
        st, p, rho = f['x'][:, i], f['p'][i], f['rho'][i]
        for k, (heat, op, ro) in enumerate(f['options']):
            kw = dict(Tin=Tin, yin=f['yin'], tau=tau, Tf=Tf, Ts=Ts, hc=hc, eps=eps, sov=sov, heat=int(heat),
                      open_=bool(op))
            r, j = cpu.reactor_jac_isobaric_opt(st, p, rates_opt=int(ro), **kw)
            err['isobaric rhs'] = max(err.get('isobaric rhs', 0.), _vec(r, f['ib_rhs'][i, k]))
            err['isobaric jac'] = max(err.get('isobaric jac', 0.), _rows(j, f['ib_jac'][i, k], ns))
            sti = np.r_[rho, st]
            r, j = cpu.reactor_jac_isochoric(sti, rho_in=1.1 * rho, rates_opt=int(ro), **kw)
            err['isochoric rhs'] = max(err.get('isochoric rhs', 0.), _vec(r, f['ic_rhs'][i, k]))
            err['isochoric jac'] = max(err.get('isochoric jac', 0.), _rows(j, f['ic_jac'][i, k], ns + 1))
            r = cpu.reactor_rhs_isochoric(sti, rho_in=1.1 * rho, **kw)
            err['isochoric rhs only'] = max(err.get('isochoric rhs only', 0.), _vec(r, f['ic_rhs_only'][i, k]))
        y = np.append(st[1:], 1. - np.sum(st[1:]))
        s1 = cpu.prod_rates_sens_opt(rho, st[0], y, 1)
        err['sens option 1'] = max(err.get('sens option 1', 0.), _rows(s1, f['sens1'][i], ns + 1))
        sc, arrs = cpu.thermo_getters(st[0], y, p, rho)
        err['getters'] = max(err.get('getters', 0.), float(np.max(np.abs(sc - f['thermo_sc'][i]) /
                                                                   np.maximum(np.abs(f['thermo_sc'][i]), 1e-300))),
                             _vec(np.array(arrs), f['thermo_arr'][i]))
    print(fx, err)
    assert max(err.values()) <= TOL, err

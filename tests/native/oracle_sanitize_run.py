"""Sanitizer run of the oracle (test infrastructure): every exported oracle function on the committed reference
fixtures, with liboracle built under AddressSanitizer + UndefinedBehaviorSanitizer (GF_ORACLE_SO). Run by
tests/test_sanitizers.py in a child process with libasan preloaded; any sanitizer report aborts it."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle.oracle import CpuGriffon  # noqa: E402
from spitfire_amd import synthetic  # noqa: E402
from spitfire_amd.mechdata import CompiledMechanism  # noqa: E402

GOLD = os.path.join(ROOT, 'tests', 'golden')


def close(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b).max(), 1e-300))) <= tol


def main():
    for mech, cfx, ffx, rfx in ((synthetic.mechanism_c1(), 'chem_c1.npz', 'flamelet_c1.npz', 'reactors_c1.npz'),
                               (synthetic.mechanism_53(), 'chem_53.npz', 'flamelet_53.npz', 'reactors_53.npz')):
        o = CpuGriffon(CompiledMechanism(mech, order='index'))
        ns = o.ns
        f = np.load(os.path.join(GOLD, cfx))
        x, p = f['x'], f['p']
        n = min(x.shape[1], 6)
        rhs, jac = o.batch_rhs_jac(np.ascontiguousarray(x[:, :n]), p[:n], nthreads=1)
        assert close(rhs, f['rhs'][:n], 1e-10) and close(jac, f['jac'][:n], 1e-10)
        o.batch_rhs_scale(np.ascontiguousarray(x[:, :n]), p[:n])
        for i in range(n):
            st = np.ascontiguousarray(x[:, i])
            y = np.append(x[1:, i], 1. - np.sum(x[1:, i]))
            o.production_rates(x[0, i], f['rho'][i], y)
            o.prod_rates_sens(f['rho'][i], x[0, i], y)
            o.rate_of_progress_log_a_derivative(f['rho'][i], x[0, i], y)
            o.thermo_getters(x[0, i], y, p[i], f['rho'][i])
            o.reactor_rhs_isobaric(st, p[i])
            o.reactor_jac_isobaric(st, p[i])
        r = np.load(os.path.join(GOLD, rfx))
        m = min(r['x'].shape[1], 3) if 'x' in r.files else 0
        for i in range(m):
            st = np.ascontiguousarray(r['x'][:, i])
            yin = np.ascontiguousarray(r['yin']) if 'yin' in r.files else np.full(ns, 1. / ns)
            for heat, op in ((0, 0), (1, 0), (2, 1)):
                kw = dict(Tin=900., yin=yin, tau=1e-3, Tf=400., Ts=500., hc=10., eps=0.5, sov=1e2, heat=heat,
                          open_=bool(op))
                o.reactor_jac_isobaric_opt(st, float(r['p'][i]), rates_opt=0, **kw)
                sti = np.r_[r['rho'][i], st]
                o.reactor_rhs_isochoric(sti, rho_in=1.1 * r['rho'][i], **kw)
                o.reactor_jac_isochoric(sti, rho_in=1.1 * r['rho'][i], **kw)
        g = np.load(os.path.join(GOLD, ffx))
        nzi = int(g['nzi'])
        stn = o.flamelet_stencils(g['dz'], nzi, g['chi'], np.ones(ns))
        o.flamelet_jac_indices(nzi)
        args = (g['state'], float(g['p']), g['oxy'], g['fuel'], nzi, stn, g['chi'])
        assert close(o.flamelet_rhs(*args, incl_ef=1, incl_vcp=1), g['rhs_11'], 1e-10)
        hl = (g['hl_Tc'], g['hl_hc'], g['hl_Tr'], g['hl_hr'])
        assert close(o.flamelet_rhs(*args, heat_loss=hl, incl_ef=1, incl_vcp=1, scaled_hl=1), g['rhs_hl1'], 1e-10)
        jv = o.flamelet_jacobian(*args, heat_loss=hl, incl_ef=1, incl_vcp=1, scaled_hl=1)[0]
        o.flamelet_jacobian(*args, compute_eig=True, diffterm=20.)
        o.flamelet_jacobian(*args, scale_and_offset=True, prefactor=-2.5e-4)
        d, l, piv = o.btddod_full_factorize(g['btdr_A'], nzi, ns)
        o.btddod_full_solve(d, l, piv, g['btd_b'], nzi, ns)
        o.btddod_scale_and_add_diagonal(jv, -1., g['sad_diag'], 1., nzi, ns)
        print(cfx, 'ok', flush=True)


if __name__ == '__main__':
    main()

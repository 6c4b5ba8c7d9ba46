"""The fused per-system kernels of the batched reactor path (gf_reactor.hip) on random inputs.

The Newton-step kernels must give bitwise the elementwise torch form they replace: x -= d where not converged,
r = h (gamma f + prior) - (x - xn), and the convergence test max |r w| < tol with a NaN anywhere in a system
keeping it unconverged (a NaN maximum compares false). The adjoint-stage kernels are checked against the
oracle batch's elementwise restatement (oracle/reactor_batch_oracle.py): the matrix I - h a_ii J^T and the
finish outputs bitwise, the J^T v products to 1e-14 relative (their summation order differs from torch's)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_newton_step_kernels_match_elementwise_form():
    import torch
    from spitfire_amd import _native
    lib = _native.lib()
    dev = torch.device('cuda:0')
    R, n = 300, 53
    rng = np.random.default_rng(5)
    T = lambda *s: torch.tensor(rng.standard_normal(s), device=dev)
    x, d, f, prior, xn = T(R, n), T(R, n), T(R, n), T(R, n), T(R, n)
    h = torch.tensor(10. ** rng.uniform(-9, -3, R), device=dev)
    w = torch.ones(n, dtype=torch.float64, device=dev)
    w[0] = 1.e-3
    conv = torch.tensor(rng.random(R) < 0.3, device=dev)
    say('masked update and stage residual,', R, 'systems of', n)
    ref_x = torch.where(conv[:, None], x, x - d)
    _native.check(lib.gf_batch_masked_update(R, n, x.data_ptr(), d.data_ptr(), conv.data_ptr(), None))
    torch.cuda.synchronize()
    assert torch.equal(x, ref_x)
    # residual: make some systems converge (x close to the stage value), plant a NaN in one of them
    g = 0.25
    target = h[:, None] * (g * f + prior) + xn
    close = torch.tensor(rng.random(R) < 0.5, device=dev)
    x = torch.where(close[:, None], target - 1.e-16 * T(R, n), x)
    x[7, 3] = float('nan')
    close[7] = True
    conv = torch.zeros(R, dtype=torch.bool, device=dev)
    conv[11] = True  # already converged stays converged
    r = torch.empty_like(x)
    tol = 1.e-12
    ref_r = h[:, None] * (g * f + prior) - (x - xn)
    ref_c = conv | (torch.max(torch.abs(ref_r * w), dim=1).values < tol)
    _native.check(lib.gf_batch_stage_residual(R, n, h.data_ptr(), g, f.data_ptr(), prior.data_ptr(), x.data_ptr(),
                                              xn.data_ptr(), w.data_ptr(), tol, r.data_ptr(), conv.data_ptr(), 1, None))
    torch.cuda.synchronize()
    same = (r == ref_r) | (torch.isnan(r) & torch.isnan(ref_r))
    assert bool(same.all())
    assert torch.equal(conv, ref_c)
    assert not bool(conv[7]) and bool(conv[11]) and int(conv.sum()) > 10
    say('  converged systems', int(conv.sum()))


def test_adjoint_stage_kernels_match_oracle_restatement(mech_c1_data):
    import torch
    from spitfire_amd import _native
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from oracle.reactor_batch_oracle import OracleReactorBatch
    from tests.test_reactor_adjoint import c1_reactors
    lib = _native.lib()
    dev = torch.device('cuda:0')
    m = ChemicalMechanismSpec(mech_c1_data)
    T0, Y0 = c1_reactors(m, [1100., 1300., 1500., 1700.], [0.7, 1.0, 1.3, 2.0])
    ob = OracleReactorBatch(m, T0, Y0)
    R, ns, nr = ob.R, ob.ns, ob.nr
    rng = np.random.default_rng(3)
    x = ob.x0.clone()
    f, J = torch.empty((R, ns), dtype=torch.float64), torch.empty((R, ns * ns), dtype=torch.float64)
    aux, cpi, hs = (torch.empty((R, 4), dtype=torch.float64), torch.empty((R, ns), dtype=torch.float64),
                    torch.empty((R, ns), dtype=torch.float64))
    ob._terms(x, f, J, aux, cpi, hs, None)
    dinvm = 1. / ob.mw[:-1] - 1. / ob.mw[-1]
    v = torch.tensor(rng.standard_normal((R, ns)))
    theta = torch.tensor(rng.standard_normal((R, ns)))
    h = torch.tensor(10. ** rng.uniform(-8, -5, R))
    bi, aii = 0.3, 0.25
    # oracle restatement on the CPU
    f64 = dict(dtype=torch.float64)
    rhs_o, M_o, cm_o, L_o = (torch.empty((R, ns), **f64), torch.empty((R, ns * ns), **f64), torch.empty((R, ns), **f64),
                             torch.zeros(R, **f64))
    ob._adjoint_prep(x, f, J, aux, cpi, dinvm, v, h, bi, aii, rhs_o, M_o)
    ob._adjoint_finish(v, theta, aux, hs, f, h, bi, aii, cm_o, L_o)
    # kernels
    G = lambda t: t.to(dev).contiguous()
    xg, fg, Jg, auxg, cpig, hsg, dg, vg, thg, hg, mwg = map(G, (x, f, J, aux, cpi, hs, dinvm, v, theta, h, ob.mw))
    rhs, M, cm = (torch.empty((R, ns), dtype=torch.float64, device=dev), torch.empty((R, ns * ns), dtype=torch.float64,
                  device=dev), torch.empty((R, ns), dtype=torch.float64, device=dev))
    L = torch.zeros(R, dtype=torch.float64, device=dev)
    say('adjoint stage kernels,', R, 'reactors')
    _native.check(lib.gf_batch_adjoint_stage_prep(R, ns, xg.data_ptr(), fg.data_ptr(), Jg.data_ptr(), auxg.data_ptr(),
                                                  cpig.data_ptr(), dg.data_ptr(), vg.data_ptr(), hg.data_ptr(), bi, aii,
                                                  rhs.data_ptr(), M.data_ptr(), None))
    _native.check(lib.gf_batch_adjoint_stage_finish(R, ns, vg.data_ptr(), thg.data_ptr(), auxg.data_ptr(),
                                                    hsg.data_ptr(), mwg.data_ptr(), fg.data_ptr(), hg.data_ptr(), bi,
                                                    aii, cm.data_ptr(), L.data_ptr(), None))
    torch.cuda.synchronize()
    e_rhs = float((rhs.cpu() - rhs_o).abs().max() / rhs_o.abs().max())
    say('  rhs rel diff %.3e' % e_rhs)
    assert e_rhs <= 1e-14
    assert torch.equal(M.cpu(), M_o)
    assert torch.equal(cm.cpu(), cm_o) and torch.equal(L.cpu(), L_o)

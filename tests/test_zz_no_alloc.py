"""The allocation-free batched reactor entry point (gf_batch_reactor_general_noalloc) against the allocating one
(gf_batch_reactor_general) through the C-ABI: bitwise equal outputs on two calls in a row, the second with fewer
states, for isobaric and isochoric reactors with open and diathermal terms; sizes beyond its fixed scratch are
rejected before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _states(spec, n, isochoric, seed):
    from spitfire_amd import synthetic
    ns = spec.n_species
    x, p, _ = synthetic.random_states(ns, n, seed=seed)
    X = np.ascontiguousarray(x.T)
    Y = np.hstack([X[:, 1:], 1. - X[:, 1:].sum(axis=1, keepdims=True)])
    rho = p / (spec.gas_constant * X[:, 0]) / np.sum(Y / spec.molecular_weights, axis=1)
    if isochoric:
        X = np.ascontiguousarray(np.hstack([rho[:, None], X]))
    return X, p


@pytest.mark.parametrize('isochoric', [False, True])
def test_noalloc_matches_allocating_entry_point(isochoric, mech_53_data):
    import torch
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    spec = ChemicalMechanismSpec(mech_53_data)
    g, ns = spec.griffon, spec.n_species
    dev = torch.device('cuda:0')
    yin = torch.tensor(np.random.default_rng(4).dirichlet(np.full(ns, 0.5)), device=dev)
    kw = dict(isochoric=isochoric, inflow_density=0.9, inflow_T=700., inflow_y=yin, tau=2e-3, fluid_T=600.,
              surf_T=650., h_conv=20., eps_rad=0.3, surface_area_over_volume=12., heat_transfer_option=2, open=True,
              rates_sensitivity_option=0)
    for call, n in enumerate((1000, 137)):  # the second call is smaller and reuses the scratch the first filled
        print('no-alloc entry point: isochoric', isochoric, 'call', call, 'states', n, flush=True)
        X, p = _states(spec, n, isochoric, seed=20 + call)
        neq = X.shape[1]
        xt, pt = torch.tensor(X, device=dev), torch.tensor(p, device=dev)
        r0 = torch.zeros((n, neq), dtype=torch.float64, device=dev)
        j0 = torch.zeros((n, neq * neq), dtype=torch.float64, device=dev)
        r1, j1 = torch.full_like(r0, np.nan), torch.full_like(j0, np.nan)
        g.batch_reactor_general(xt, r0, j0, p=pt, **kw)
        torch.cuda.synchronize()
        g.batch_reactor_general_noalloc(xt, r1, j1, p=pt, **kw)
        torch.cuda.synchronize()
        assert torch.equal(r0, r1) and torch.equal(j0, j1)
        assert bool(torch.isfinite(j1).all())
        # RHS only (no Jacobian; the rates-only chemistry kernel, rounded differently from the Jacobian kernel)
        # through the same scratch, against the allocating entry point's RHS-only call
        r2, r3 = torch.full_like(r0, np.nan), torch.full_like(r0, np.nan)
        g.batch_reactor_general(xt, r2, None, p=pt, **kw)
        torch.cuda.synchronize()
        g.batch_reactor_general_noalloc(xt, r3, None, p=pt, **kw)
        torch.cuda.synchronize()
        assert torch.equal(r2, r3)
        assert float(((r2 - r0).abs().max(dim=1).values / r0.abs().max(dim=1).values).max()) < 1e-12


def test_noalloc_rejects_sizes_beyond_scratch(mech_53_data):
    import torch
    from spitfire_amd._native import NativeError
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    spec = ChemicalMechanismSpec(mech_53_data)
    g, ns = spec.griffon, spec.n_species
    dev = torch.device('cuda:0')
    X, p = _states(spec, 8, False, seed=3)
    xt, pt = torch.tensor(X, device=dev), torch.tensor(p, device=dev)
    r = torch.zeros((8, ns), dtype=torch.float64, device=dev)
    g.batch_reactor_general(xt, r, None, p=pt)  # puts the mechanism on the device
    max_n = (32 << 20) // 8 // (3 * ns + 4 + (ns + 1) ** 2)
    assert max_n == 1362
    n = max_n + 1
    big = torch.zeros((n, ns), dtype=torch.float64, device=dev)
    rb = torch.full((n, ns), 7., dtype=torch.float64, device=dev)
    jb = torch.full((n, ns * ns), 7., dtype=torch.float64, device=dev)
    with pytest.raises(NativeError, match='exceeds'):
        g.batch_reactor_general_noalloc(big, rb, jb, p=torch.full((n,), 101325., dtype=torch.float64, device=dev))
    torch.cuda.synchronize()
    assert bool((rb == 7.).all()) and bool((jb == 7.).all())


def test_noalloc_needs_a_resident_mechanism(mech_c1_data):
    import torch
    from spitfire_amd._native import NativeError
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    spec = ChemicalMechanismSpec(mech_c1_data)  # fresh: its tables are not on the device yet
    ns = spec.n_species
    dev = torch.device('cuda:0')
    X, p = _states(spec, 4, False, seed=1)
    r = torch.zeros((4, ns), dtype=torch.float64, device=dev)
    with pytest.raises(NativeError, match='not on the device'):
        spec.griffon.batch_reactor_general_noalloc(torch.tensor(X, device=dev), r, None, p=torch.tensor(p, device=dev))

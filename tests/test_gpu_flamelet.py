"""GPU parity of the flamelet and block-tridiagonal kernels against the reference fixtures and the oracle."""
import os

import numpy as np
import pytest

from tests.helpers import row_norm_err

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
RTOL = 1e-12


def say(*a):
    print(*a, flush=True)


def _rows_err(a, b, n):
    A = np.asarray(a).reshape(-1, n)
    B = np.asarray(b).reshape(-1, n)
    return float(np.max(np.max(np.abs(A - B), axis=1) / np.maximum(np.max(np.abs(B), axis=1), 1e-300)))


def _jac_err(a, b, nzi, ns):
    """row-normwise error over the diagonal blocks and elementwise over the off-diagonals"""
    nbl = nzi * ns * ns
    e_blk = row_norm_err(a[:nbl], b[:nbl], ns)
    off = np.abs(a[nbl:] - b[nbl:]) / np.maximum(np.abs(b[nbl:]), 1e-300)
    return max(e_blk, float(np.max(off)) if off.size else 0.)


@pytest.mark.parametrize('fx, mech', [('flamelet_c1.npz', 'mech_c1_data'), ('flamelet_53.npz', 'mech_53_data')])
def test_flamelet_dropins_match_reference_fixtures(fx, mech, request):
    from spitfire_amd.griffon import (py_btddod_full_factorize, py_btddod_full_solve,
                                      py_btddod_scale_and_add_diagonal)
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    f = np.load(os.path.join(GOLD, fx))
    g = ChemicalMechanismSpec(request.getfixturevalue(mech)).griffon
    nzi, ns = int(f['nzi']), int(f['ns'])
    say('flamelet fixture', fx, nzi, 'points')
    nv = nzi * ns
    cm, cs, cu, mc, nc = [np.zeros(nv) for _ in range(3)] + [np.zeros(nzi), np.zeros(nzi)]
    g.flamelet_stencils(f['dz'], nzi, f['chi'], np.ones(ns), cm, cs, cu, mc, nc)
    for a, k in zip((cm, cs, cu, mc, nc), ('cmaj', 'csub', 'csup', 'mco', 'nco')):
        assert np.max(np.abs(a - f[k]) / np.maximum(np.abs(f[k]), 1e-300)) <= 1e-14
    nel = ns * (nzi * ns + 2 * (nzi - 1))
    rows, cols = np.zeros(nel, dtype=np.int32), np.zeros(nel, dtype=np.int32)
    g.flamelet_jac_indices(nzi, rows, cols)
    assert np.array_equal(rows, f['rows']) and np.array_equal(cols, f['cols'])
    null = np.zeros(1)
    st = f['state']
    for key in f.files:
        if 'hl' in key:
            continue
        if key.startswith('rhs_'):
            ef, vcp = int(key[4]), int(key[5])
            out = np.zeros(nv)
            g.flamelet_rhs(st, float(f['p']), f['oxy'], f['fuel'], True, null, null, null, null, nzi, cm, cs, cu, mc,
                           nc, f['chi'], bool(ef), bool(vcp), False, out)
            e = _rows_err(out, f[key], ns)
            say(key, 'row err %.3e' % e)
            assert e <= RTOL
        if key.startswith('jac_'):
            ef, vcp = (0, 0) if key == 'jac_so' else (int(key[4]), int(key[5]))
            so = key == 'jac_so'
            out = np.zeros(nel)
            g.flamelet_jacobian(st, float(f['p']), f['oxy'], f['fuel'], True, null, null, null, null, nzi, cm, cs, cu,
                                mc, nc, f['chi'], False, 0., so, -2.5e-4 if so else 0., 0, 0, bool(ef), bool(vcp),
                                False, null, out)
            e = _jac_err(out, f[key], nzi, ns)
            say(key, 'err %.3e' % e)
            assert e <= RTOL
    # nonadiabatic, through the reference's argument order (T_convection, T_radiation, h_convection, h_radiation)
    Tc, hc, Tr, hr = f['hl_Tc'], f['hl_hc'], f['hl_Tr'], f['hl_hr']
    for sc in (0, 1):
        out = np.zeros(nv)
        g.flamelet_rhs(st, float(f['p']), f['oxy'], f['fuel'], False, Tc, Tr, hc, hr, nzi, cm, cs, cu, mc, nc,
                       f['chi'], True, True, bool(sc), out)
        e = _rows_err(out, f['rhs_hl%d' % sc], ns)
        say('rhs_hl%d row err %.3e' % (sc, e))
        assert e <= RTOL
        for so in ((False, True) if sc else (False,)):
            out = np.zeros(nel)
            g.flamelet_jacobian(st, float(f['p']), f['oxy'], f['fuel'], False, Tc, Tr, hc, hr, nzi, cm, cs, cu, mc,
                                nc, f['chi'], False, 0., so, -2.5e-4 if so else 0., 0, 0, True, True, bool(sc), null,
                                out)
            key = 'jac_hl%d' % sc + ('_so' if so else '')
            e = _jac_err(out, f[key], nzi, ns)
            say(key, 'err %.3e' % e)
            assert e <= RTOL
    # scale_and_add_diagonal, block Thomas factor + solve
    v = f['jac_00'].copy()
    py_btddod_scale_and_add_diagonal(v, -1., f['sad_diag'], 1., nzi, ns)
    assert np.max(np.abs(v - f['sad']) / np.maximum(np.abs(f['sad']), 1e-300)) <= 1e-15
    for k in ('btd', 'btdr'):  # the psitc matrix, then a random system whose factorization pivots
        d = f[k + '_A'].copy()
        l = np.zeros(nzi * ns * ns)
        piv = np.zeros(nel, dtype=np.int32)
        py_btddod_full_factorize(d, nzi, ns, l, piv)
        if k == 'btdr':
            assert np.array_equal(piv[:nv], f['btdr_piv'][:nv])
        x = np.zeros(nv)
        py_btddod_full_solve(d, l, piv, f['btd_b'], nzi, ns, x)
        e = np.max(np.abs(x - f[k + '_x'])) / np.max(np.abs(f[k + '_x']))
        say(k, 'solve err %.3e' % e)
        assert e <= 1e-11


def test_batched_btd_ragged_matches_oracle(mech_53_data, oracle_lib):
    """Several independent block-tridiagonal systems of different lengths in one launch."""
    import torch
    rng = np.random.default_rng(3)
    bs = 53
    sizes = [1, 2, 7, 31, 64]
    vals, rhs = [], []
    for nb in sizes:
        nel = bs * (nb * bs + 2 * (nb - 1))
        v = rng.standard_normal(nel)
        for i in range(nb):  # diagonally weighted blocks
            v[i * bs * bs:(i + 1) * bs * bs].reshape(bs, bs)[np.diag_indices(bs)] += 20.
        vals.append(v)
        rhs.append(rng.standard_normal(nb * bs))
    voff = np.cumsum([0] + [v.size for v in vals])[:-1]
    loff = np.cumsum([0] + [nb * bs * bs for nb in sizes])[:-1]
    poff = np.cumsum([0] + [nb * bs for nb in sizes])[:-1]
    dev = torch.device('cuda:0')
    V = torch.tensor(np.concatenate(vals), device=dev)
    L = torch.zeros(int(sum(nb * bs * bs for nb in sizes)), dtype=torch.float64, device=dev)
    P = torch.zeros(int(sum(nb * bs for nb in sizes)), dtype=torch.int32, device=dev)
    B = torch.tensor(np.concatenate(rhs), device=dev)
    X = torch.zeros_like(B)
    nbt = torch.tensor(sizes, dtype=torch.int32, device=dev)
    vo, lo, po = [torch.tensor(a.astype(np.int64), device=dev) for a in (voff, loff, poff)]
    from spitfire_amd import _native
    lib = _native.lib()
    s = torch.cuda.current_stream().cuda_stream
    say('batched btd', sizes)
    _native.check(lib.gf_batch_btddod_full_factorize(len(sizes), nbt.data_ptr(), bs, V.data_ptr(), vo.data_ptr(),
                                                     L.data_ptr(), lo.data_ptr(), P.data_ptr(), po.data_ptr(), s))
    _native.check(lib.gf_batch_btddod_full_solve(len(sizes), nbt.data_ptr(), bs, V.data_ptr(), vo.data_ptr(),
                                                 L.data_ptr(), lo.data_ptr(), P.data_ptr(), po.data_ptr(),
                                                 B.data_ptr(), X.data_ptr(), po.data_ptr(), s))
    torch.cuda.synchronize()
    Xh = X.cpu().numpy()
    o = oracle_lib
    from spitfire_amd.mechdata import CompiledMechanism
    cpu = o.CpuGriffon(CompiledMechanism(mech_53_data))
    for k, nb in enumerate(sizes):
        d, l, piv = cpu.btddod_full_factorize(vals[k], nb, bs)
        xr = cpu.btddod_full_solve(d, l, piv, rhs[k], nb, bs)
        e = np.max(np.abs(Xh[poff[k]:poff[k] + nb * bs] - xr)) / np.max(np.abs(xr))
        say('system', nb, 'blocks: err %.3e' % e)
        assert e <= 1e-11


@pytest.mark.parametrize('bs', [1, 5, 13, 33, 53, 64])
@pytest.mark.parametrize('precision', [64, 32])
def test_batched_btd_inverse_path_matches_oracle(bs, precision, mech_53_data, oracle_lib):
    """The throughput form of the block Thomas recurrence (inverse diagonal blocks, chain eliminated from both
    ends): ragged systems of 1..255 blocks, odd and even, against the oracle's block Thomas with LAPACK-style
    LU. FP64 factors agree to 1e-11 of max|x|; FP32 factors (FP64 vectors) to 1e-4, the rounding of FP32
    inverse blocks times the condition of these diagonally weighted systems."""
    import torch
    rng = np.random.default_rng(11 + bs)
    sizes = [1, 2, 3, 4, 7, 64, 255] if bs <= 53 else [1, 2, 5, 64]
    vals, rhs = [], []
    for nb in sizes:
        v = rng.standard_normal(bs * (nb * bs + 2 * (nb - 1)))
        for i in range(nb):
            v[i * bs * bs:(i + 1) * bs * bs].reshape(bs, bs)[np.diag_indices(bs)] += 10. + 4. * np.sqrt(bs)
        vals.append(v)
        rhs.append(rng.standard_normal(nb * bs))
    voff = np.cumsum([0] + [v.size for v in vals])[:-1]
    doff = np.cumsum([0] + [nb * bs * bs for nb in sizes])[:-1]
    xoff = np.cumsum([0] + [nb * bs for nb in sizes])[:-1]
    dev = torch.device('cuda:0')
    V = torch.tensor(np.concatenate(vals), device=dev)
    D = torch.zeros(int(sum(nb * bs * bs for nb in sizes)), dtype=torch.float64 if precision == 64 else torch.float32,
                    device=dev)
    B = torch.tensor(np.concatenate(rhs), device=dev)
    X = torch.zeros_like(B)
    nbt = torch.tensor(sizes, dtype=torch.int32, device=dev)
    vo, do, xo = [torch.tensor(a.astype(np.int64), device=dev) for a in (voff, doff, xoff)]
    from spitfire_amd import _native
    lib = _native.lib()
    s = torch.cuda.current_stream().cuda_stream
    say('inverse btd bs', bs, 'precision', precision, 'sizes', sizes)
    _native.check(lib.gf_batch_btddod_inverse_factorize(len(sizes), nbt.data_ptr(), bs, V.data_ptr(), vo.data_ptr(),
                                                        D.data_ptr(), do.data_ptr(), precision, s))
    _native.check(lib.gf_batch_btddod_inverse_solve(len(sizes), nbt.data_ptr(), bs, V.data_ptr(), vo.data_ptr(),
                                                    D.data_ptr(), do.data_ptr(), precision, B.data_ptr(),
                                                    X.data_ptr(), xo.data_ptr(), s))
    torch.cuda.synchronize()
    Xh = X.cpu().numpy()
    assert np.array_equal(V.cpu().numpy(), np.concatenate(vals))  # the values are not modified
    from spitfire_amd.mechdata import CompiledMechanism
    cpu = oracle_lib.CpuGriffon(CompiledMechanism(mech_53_data))
    tol = 1e-11 if precision == 64 else 1e-4
    for k, nb in enumerate(sizes):
        d, l, piv = cpu.btddod_full_factorize(vals[k], nb, bs)
        xr = cpu.btddod_full_solve(d, l, piv, rhs[k], nb, bs)
        e = np.max(np.abs(Xh[xoff[k]:xoff[k] + nb * bs] - xr)) / np.max(np.abs(xr))
        say('system', nb, 'blocks: err %.3e' % e)
        assert e <= tol


@pytest.mark.parametrize('bs', [12, 13, 53, 64])
@pytest.mark.parametrize('precision', [64, 32])
def test_inverse_factor_kernels_agree_bitwise(bs, precision):
    """The inverse factor has two kernels for block sizes >= 12: columns split over two waves per chain for
    batches of fewer than four systems per compute unit, one wave per chain for larger batches. Both apply
    the same pivots and the same operations to every entry, so the same systems factored in a small batch and
    inside a batch large enough for the other kernel give bitwise the same inverse blocks."""
    import torch
    from spitfire_amd import _native
    rng = np.random.default_rng(7 + bs)
    sizes = [1, 2, 3, 6, 33]
    vals = []
    for nb in sizes:
        v = rng.standard_normal(bs * (nb * bs + 2 * (nb - 1)))
        for i in range(nb):
            v[i * bs * bs:(i + 1) * bs * bs].reshape(bs, bs)[np.diag_indices(bs)] += 10. + 4. * np.sqrt(bs)
        vals.append(v)
    dev = torch.device('cuda:0')
    ncu = torch.cuda.get_device_properties(dev).multi_processor_count
    lib = _native.lib()
    s = torch.cuda.current_stream().cuda_stream
    fdt = torch.float64 if precision == 64 else torch.float32

    def factor(reps):
        sz = sizes * reps
        vv = vals * reps
        voff = np.cumsum([0] + [v.size for v in vv])[:-1]
        doff = np.cumsum([0] + [nb * bs * bs for nb in sz])[:-1]
        V = torch.tensor(np.concatenate(vv), device=dev)
        D = torch.zeros(int(sum(nb * bs * bs for nb in sz)), dtype=fdt, device=dev)
        nbt = torch.tensor(sz, dtype=torch.int32, device=dev)
        vo, do = [torch.tensor(a.astype(np.int64), device=dev) for a in (voff, doff)]
        _native.check(lib.gf_batch_btddod_inverse_factorize(len(sz), nbt.data_ptr(), bs, V.data_ptr(), vo.data_ptr(),
                                                            D.data_ptr(), do.data_ptr(), precision, s))
        torch.cuda.synchronize()
        return D.cpu().numpy()

    reps = -(-4 * ncu // len(sizes))  # at least four systems per compute unit
    say('inverse factor kernels bs', bs, 'precision', precision, 'systems', len(sizes), 'and', len(sizes) * reps)
    small, large = factor(1), factor(reps)
    n1 = small.size
    assert np.all(np.isfinite(small))
    for r in range(reps):
        assert np.array_equal(large[r * n1:(r + 1) * n1], small)


@pytest.mark.parametrize('fx, mech', [('flamelet_c1.npz', 'mech_c1_data'), ('flamelet_53.npz', 'mech_53_data')])
def test_flamelet_eigenvalue_limiter_matches_reference(fx, mech, request, oracle_lib):
    """expeig = max(0, max Re(lambda(chemistry block)) - diffterm) per point (flamelet_kernels.cpp:1329-1341):
    GPU balancing + Hessenberg + Francis QR against the reference's LAPACK dgeev values, to within
    1e-12 * max|lambda| of each block (eigenvalues agree to rounding of the block norm)."""
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.mechdata import CompiledMechanism
    data = request.getfixturevalue(mech)
    f = np.load(os.path.join(GOLD, fx))
    g = ChemicalMechanismSpec(data).griffon
    nzi, ns = int(f['nzi']), int(f['ns'])
    nv = nzi * ns
    cm, cs, cu, mc, nc = [np.zeros(nv) for _ in range(3)] + [np.zeros(nzi), np.zeros(nzi)]
    g.flamelet_stencils(f['dz'], nzi, f['chi'], np.ones(ns), cm, cs, cu, mc, nc)
    null = np.zeros(1)
    nel = ns * (nzi * ns + 2 * (nzi - 1))
    out, expeig, plain = np.zeros(nel), np.zeros(nv), np.zeros(nel)
    say('eigenvalue limiter', fx)
    g.flamelet_jacobian(f['state'], float(f['p']), f['oxy'], f['fuel'], True, null, null, null, null, nzi, cm, cs,
                        cu, mc, nc, f['chi'], True, 20., False, 0., 0, 0, False, False, False, expeig, out)
    g.flamelet_jacobian(f['state'], float(f['p']), f['oxy'], f['fuel'], True, null, null, null, null, nzi, cm, cs,
                        cu, mc, nc, f['chi'], False, 0., False, 0., 0, 0, False, False, False, null, plain)
    assert np.array_equal(out, plain)  # computing eigenvalues leaves the Jacobian unchanged
    o = oracle_lib.CpuGriffon(CompiledMechanism(data, order='name'))
    st = o.flamelet_stencils(f['dz'], nzi, f['chi'], np.ones(ns))
    o.flamelet_jacobian(f['state'], float(f['p']), f['oxy'], f['fuel'], nzi, st, f['chi'], compute_eig=True,
                        diffterm=20.)
    pre = o.last_chemistry_blocks
    for iz in range(nzi):
        rad = np.abs(np.linalg.eigvals(pre[iz * ns * ns:(iz + 1) * ns * ns].reshape((ns, ns), order='F'))).max()
        e = np.abs(expeig[iz * ns:(iz + 1) * ns] - f['expeig'][iz * ns:(iz + 1) * ns]).max()
        say('  point %d expeig %.6e ref %.6e |diff|/rho %.2e' % (iz, expeig[iz * ns], f['expeig'][iz * ns], e / rad))
        assert np.all(expeig[iz * ns:(iz + 1) * ns] == expeig[iz * ns])
        assert e <= 1e-12 * rad


def test_batched_eigenvalue_bound_matches_numpy(mech_53_data):
    """gf_batch_eigenvalue_bound on random, badly scaled and defective-like matrices of sizes 1..64 and on
    53-species chemistry Jacobians: max Re(lambda) against numpy.linalg.eigvals (LAPACK dgeev, what the reference
    calls), to 1e-12 of max |lambda| per block; a shift diffterm = -C (C above every |Re lambda|) keeps the bound
    positive so that the maximum itself is compared."""
    import torch
    from spitfire_amd import _native, synthetic
    from spitfire_amd.mechanism import ChemicalMechanismSpec
    from spitfire_amd.mechdata import CompiledMechanism
    from oracle import oracle
    lib = _native.lib()
    dev = torch.device('cuda:0')
    rng = np.random.default_rng(21)
    cases = []
    for n in (1, 2, 3, 8, 31, 53, 64):
        B = rng.standard_normal((40, n, n))
        B[1::2] *= 10. ** rng.uniform(-4, 4, (20, n, 1)) * 10. ** rng.uniform(-4, 4, (20, 1, n))
        cases.append(B)
    m = ChemicalMechanismSpec(mech_53_data)
    x, p, _ = synthetic.random_states(m.n_species, 200, seed=9)
    _, jac = oracle.CpuGriffon(CompiledMechanism(mech_53_data, order='name')).batch_rhs_jac(x, p)
    cases.append(np.stack([jac[b].reshape(53, 53, order='F') for b in range(200)]))
    for B in cases:
        nb, n = B.shape[0], B.shape[1]
        lam = [np.linalg.eigvals(B[b]) for b in range(nb)]
        C = 2. * max(np.abs(l).max() for l in lam) + 1.
        blocks = torch.tensor(np.ascontiguousarray(np.transpose(B, (0, 2, 1))).ravel(), device=dev)  # column-major
        out = torch.zeros(nb, dtype=torch.float64, device=dev)
        nfail = torch.zeros(1, dtype=torch.int32, device=dev)
        say('eigenvalue bound', nb, 'blocks of', n)
        _native.check(lib.gf_batch_eigenvalue_bound(None, nb, n, blocks.data_ptr(), -C, out.data_ptr(),
                                                    nfail.data_ptr(), torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        got = out.cpu().numpy() - C
        assert int(nfail.item()) == 0
        for b in range(nb):
            ref = lam[b].real.max()
            assert abs(got[b] - ref) <= 1e-12 * C, (n, b, got[b], ref)

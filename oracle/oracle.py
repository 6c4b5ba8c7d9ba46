"""TEST INFRASTRUCTURE ONLY — ctypes bindings to the CPU oracle (liboracle.so, the FP64 restatement
of griffon in griffon_oracle.cpp) and, where it was built, to oracle/_ref/libgriffon_ref.so (the
reference's own griffon sources). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may import this module.

Both libraries expose the same C signatures (prefix gfo_ / gfr_), so `CpuGriffon(mech, ref=True)`
drives the reference build and `CpuGriffon(mech)` the restatement on identical inputs.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GF_ORACLE_SO selects another build of the same source (the sanitizer build, oracle/Makefile target asan)
ORACLE_SO = os.environ.get('GF_ORACLE_SO') or os.path.join(HERE, 'liboracle.so')
REF_SO = os.path.join(HERE, '_ref', 'libgriffon_ref.so')

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_libs = {}


def build(ref=None):
    """Compile liboracle.so (always) and oracle/_ref (when the reference checkout exists)."""
    subprocess.check_call(['make', '-s', '-C', HERE, 'oracle'])
    if ref or (ref is None and os.path.isdir('/root/reference/src/spitfire/griffon')):
        subprocess.check_call(['make', '-s', '-C', HERE, 'ref'])


def ref_available():
    return os.path.exists(REF_SO)


def _lib(ref):
    key = 'ref' if ref else 'oracle'
    if key not in _libs:
        path = REF_SO if ref else ORACLE_SO
        if not os.path.exists(path):
            raise FileNotFoundError(path + ' is not built (run oracle.build())')
        lib = ctypes.CDLL(path)
        p = 'gfr_' if ref else 'gfo_'
        getattr(lib, p + 'mech_create').restype = ctypes.c_void_p
        getattr(lib, p + 'mech_create').argtypes = [ctypes.c_int, _D, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double]
        getattr(lib, p + 'mech_destroy').argtypes = [ctypes.c_void_p]
        _libs[key] = (lib, p)
    return _libs[key]


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(_D)


def _ip(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(_I)


class CpuGriffon:
    """A griffon mechanism on the CPU: the oracle restatement (ref=False) or the reference (ref=True)."""

    def __init__(self, cm, ref=False):
        self.cm = cm
        self.ns = cm.ns
        self.nr = cm.nr
        self.ref = ref
        self.lib, self.p = _lib(ref)
        mw = np.ascontiguousarray(cm.mw, dtype=np.float64)
        self.h = ctypes.c_void_p(self._f('mech_create')(cm.ns, _dp(mw), cm.gas_constant, cm.ref_pressure,
                                                        cm.ref_temperature))
        for i, th in enumerate(cm.thermo):
            if th[0] == 'NASA7':
                lo = np.ascontiguousarray(th[4], dtype=np.float64)
                hi = np.ascontiguousarray(th[5], dtype=np.float64)
                rc = self._f('mech_set_nasa7')(self.h, i, ctypes.c_double(th[1]), ctypes.c_double(th[2]),
                                               ctypes.c_double(th[3]), _dp(lo), _dp(hi))
            else:
                rc = self._f('mech_set_const_cp')(self.h, i, *[ctypes.c_double(v) for v in th[1:]])
            assert rc == 0
        for r in cm.reactions:
            ri, rs = r['reactants']
            pi, ps = r['products']
            ti, te = r['tb']
            oi, ov = r['orders']
            kf = np.ascontiguousarray(r['kfwd'], dtype=np.float64)
            kp = np.ascontiguousarray(r['kpres'], dtype=np.float64)
            tr = np.ascontiguousarray(r['troe'], dtype=np.float64)
            rc = self._f('mech_add_reaction')(self.h, r['type'], int(r['reversible']), len(ri), _ip(ri), _ip(rs),
                                              len(pi), _ip(pi), _ip(ps), _dp(kf), len(ti), _ip(ti),
                                              _dp(np.ascontiguousarray(te, dtype=np.float64)),
                                              ctypes.c_double(r['default_eff']), _dp(kp), _dp(tr), len(oi),
                                              _ip(oi), _dp(np.ascontiguousarray(ov, dtype=np.float64)))
            if rc != 0:
                raise ValueError('reaction rejected by %s (code %d)' % ('reference' if ref else 'oracle', rc))

    def _f(self, name):
        return getattr(self.lib, self.p + name)

    def __del__(self):
        try:
            self._f('mech_destroy')(self.h)
        except Exception:
            pass

    # --- pointwise ---
    def production_rates(self, T, rho, y):
        w = np.zeros(self.ns)
        assert self._f('production_rates')(self.h, ctypes.c_double(T), ctypes.c_double(rho),
                                           _dp(np.ascontiguousarray(y, dtype=np.float64)), _dp(w)) == 0
        return w

    def prod_rates_sens(self, rho, T, y):
        w = np.zeros(self.ns)
        ws = np.zeros((self.ns + 1) ** 2)
        assert self._f('prod_rates_sens')(self.h, ctypes.c_double(rho), ctypes.c_double(T),
                                          _dp(np.ascontiguousarray(y, dtype=np.float64)), _dp(w), _dp(ws)) == 0
        return ws

    def thermo(self, T, y):
        sc = np.zeros(4)
        cpi, hi, cpis = np.zeros(self.ns), np.zeros(self.ns), np.zeros(self.ns)
        self._f('thermo')(self.h, ctypes.c_double(T), _dp(np.ascontiguousarray(y, dtype=np.float64)), _dp(sc),
                          _dp(cpi), _dp(hi), _dp(cpis))
        return sc, cpi, hi, cpis

    def _reactor_args(self, p, Tin, yin, tau, Tf, Ts, hc, eps, sov, heat, open_):
        yin = np.zeros(self.ns) if yin is None else np.ascontiguousarray(yin, dtype=np.float64)
        self._yin_keep = yin
        return [ctypes.c_double(p), ctypes.c_double(Tin), _dp(yin), ctypes.c_double(tau), ctypes.c_double(Tf),
                ctypes.c_double(Ts), ctypes.c_double(hc), ctypes.c_double(eps), ctypes.c_double(sov), int(heat),
                int(open_)]

    def reactor_rhs_isobaric(self, state, p, Tin=0., yin=None, tau=0., Tf=0., Ts=0., hc=0., eps=0., sov=0.,
                             heat=0, open_=False):
        rhs = np.zeros(self.ns)
        st = np.ascontiguousarray(state, dtype=np.float64)
        assert self._f('reactor_rhs_isobaric')(self.h, _dp(st), *self._reactor_args(p, Tin, yin, tau, Tf, Ts, hc,
                                                                                      eps, sov, heat, open_),
                                               _dp(rhs)) == 0
        return rhs

    def reactor_jac_isobaric(self, state, p, Tin=0., yin=None, tau=0., Tf=0., Ts=0., hc=0., eps=0., sov=0.,
                             heat=0, open_=False):
        rhs = np.zeros(self.ns)
        jac = np.zeros(self.ns * self.ns)
        st = np.ascontiguousarray(state, dtype=np.float64)
        assert self._f('reactor_jac_isobaric')(self.h, _dp(st), *self._reactor_args(p, Tin, yin, tau, Tf, Ts, hc,
                                                                                      eps, sov, heat, open_),
                                               _dp(rhs), _dp(jac)) == 0
        return rhs, jac

    def rate_of_progress_log_a_derivative(self, rho, T, y):
        """d q_r / d ln A_r for every reaction, griffon order (the oracle restatement only)"""
        assert not self.ref
        qa = np.zeros(self.nr)
        assert self._f('rate_of_progress_log_a_derivative')(self.h, ctypes.c_double(rho), ctypes.c_double(T),
                                                             _dp(np.ascontiguousarray(y, dtype=np.float64)),
                                                             _dp(qa)) == 0
        return qa

    def prod_rates_sens_opt(self, rho, T, y, option):
        ws = np.zeros((self.ns + 1) ** 2)
        rc = self._f('prod_rates_sens_opt')(self.h, ctypes.c_double(rho), ctypes.c_double(T),
                                            _dp(np.ascontiguousarray(y, dtype=np.float64)), int(option), _dp(ws))
        if rc != 0:
            raise ValueError('rates sensitivity option %d rejected (code %d)' % (option, rc))
        return ws

    def reactor_jac_isobaric_opt(self, state, p, Tin=0., yin=None, tau=0., Tf=0., Ts=0., hc=0., eps=0., sov=0.,
                                 heat=0, open_=False, rates_opt=0):
        rhs = np.zeros(self.ns)
        jac = np.zeros(self.ns * self.ns)
        st = np.ascontiguousarray(state, dtype=np.float64)
        assert self._f('reactor_jac_isobaric_opt')(self.h, _dp(st), *self._reactor_args(p, Tin, yin, tau, Tf, Ts, hc,
                                                                                          eps, sov, heat, open_),
                                                   int(rates_opt), _dp(rhs), _dp(jac)) == 0
        return rhs, jac

    def reactor_rhs_isochoric(self, state, rho_in=0., Tin=0., yin=None, tau=0., Tf=0., Ts=0., hc=0., eps=0., sov=0.,
                              heat=0, open_=False):
        rhs = np.zeros(self.ns + 1)
        st = np.ascontiguousarray(state, dtype=np.float64)
        assert self._f('reactor_rhs_isochoric')(self.h, _dp(st), *self._reactor_args(rho_in, Tin, yin, tau, Tf, Ts, hc,
                                                                                       eps, sov, heat, open_),
                                                _dp(rhs)) == 0
        return rhs

    def reactor_jac_isochoric(self, state, rho_in=0., Tin=0., yin=None, tau=0., Tf=0., Ts=0., hc=0., eps=0., sov=0.,
                              heat=0, open_=False, rates_opt=0):
        rhs = np.zeros(self.ns + 1)
        jac = np.zeros((self.ns + 1) ** 2)
        st = np.ascontiguousarray(state, dtype=np.float64)
        assert self._f('reactor_jac_isochoric')(self.h, _dp(st), *self._reactor_args(rho_in, Tin, yin, tau, Tf, Ts, hc,
                                                                                       eps, sov, heat, open_),
                                                int(rates_opt), _dp(rhs), _dp(jac)) == 0
        return rhs, jac

    def thermo_getters(self, T, y, p, rho):
        """{mmw, density, pressure, cp, cv, h, e, dcpdT} and mole fractions, species cp, cv, h, e, dcp/dT"""
        sc = np.zeros(8)
        arrs = [np.zeros(self.ns) for _ in range(6)]
        assert self._f('thermo_getters')(self.h, ctypes.c_double(T), _dp(np.ascontiguousarray(y, dtype=np.float64)),
                                         ctypes.c_double(p), ctypes.c_double(rho), _dp(sc), *[_dp(a) for a in arrs]) == 0
        return sc, arrs

    # --- batched (closed adiabatic isobaric), state SoA [ns, N] ---
    def batch_rhs_jac(self, stateSoA, p, nthreads=0):
        N = stateSoA.shape[1]
        rhs = np.empty((N, self.ns))
        jac = np.empty((N, self.ns * self.ns))
        st = np.ascontiguousarray(stateSoA, dtype=np.float64)
        pp = np.ascontiguousarray(p, dtype=np.float64)
        assert self._f('batch_rhs_jac')(self.h, N, _dp(st), _dp(pp), _dp(rhs), _dp(jac), nthreads) == 0
        return rhs, jac

    def batch_rhs_scale(self, stateSoA, p):
        """Per-entry magnitude of the terms summed into the RHS (oracle only): |dT/dt| and |dY/dt| scales."""
        assert not self.ref
        N = stateSoA.shape[1]
        sc = np.empty((N, self.ns))
        assert self._f('batch_rhs_scale')(self.h, N, _dp(np.ascontiguousarray(stateSoA, dtype=np.float64)),
                                          _dp(np.ascontiguousarray(p, dtype=np.float64)), _dp(sc)) == 0
        return sc

    def batch_rhs(self, stateSoA, p, nthreads=0):
        N = stateSoA.shape[1]
        rhs = np.empty((N, self.ns))
        st = np.ascontiguousarray(stateSoA, dtype=np.float64)
        pp = np.ascontiguousarray(p, dtype=np.float64)
        assert self._f('batch_rhs')(self.h, N, _dp(st), _dp(pp), _dp(rhs), nthreads) == 0
        return rhs

    # --- flamelet ---
    def flamelet_stencils(self, dz, nzi, chi, invLe):
        n = self.ns
        out = [np.zeros(nzi * n) for _ in range(3)] + [np.zeros(nzi), np.zeros(nzi)]
        self._f('flamelet_stencils')(self.h, _dp(np.ascontiguousarray(dz)), nzi, _dp(np.ascontiguousarray(chi)),
                                     _dp(np.ascontiguousarray(invLe)), *[_dp(o) for o in out])
        return out

    def flamelet_jac_indices(self, nzi):
        nel = self.ns * (nzi * self.ns + 2 * (nzi - 1))
        rows = np.zeros(nel, dtype=np.int32)
        cols = np.zeros(nel, dtype=np.int32)
        self._f('flamelet_jac_indices')(self.h, nzi, _ip(rows), _ip(cols))
        return rows, cols

    def _hl(self, heat_loss):
        if heat_loss is None:
            z = np.zeros(1)
            return 1, [z, z, z, z]
        Tc, hc, Tr, hr = [np.ascontiguousarray(a, dtype=np.float64) for a in heat_loss]
        return 0, [Tc, hc, Tr, hr]

    def flamelet_rhs(self, state, p, oxy, fuel, nzi, stencils, chi, heat_loss=None, incl_ef=False, incl_vcp=False,
                     scaled_hl=False):
        """heat_loss = (T_conv, h_conv, T_rad, h_rad) in the C++ order, or None for adiabatic."""
        out = np.zeros(nzi * self.ns)
        adi, hl = self._hl(heat_loss)
        cm, cs, cu, mc, nc = [np.ascontiguousarray(a) for a in stencils]
        self._f('flamelet_rhs')(self.h, _dp(np.ascontiguousarray(state)), ctypes.c_double(p),
                                _dp(np.ascontiguousarray(oxy)), _dp(np.ascontiguousarray(fuel)), adi,
                                *[_dp(a) for a in hl], nzi, _dp(cm), _dp(cs), _dp(cu), _dp(mc), _dp(nc),
                                _dp(np.ascontiguousarray(chi)), int(incl_ef), int(incl_vcp), int(scaled_hl), _dp(out))
        return out

    def flamelet_jacobian(self, state, p, oxy, fuel, nzi, stencils, chi, heat_loss=None, compute_eig=False,
                          diffterm=0., scale_and_offset=False, prefactor=0., incl_ef=False, incl_vcp=False,
                          scaled_hl=False):
        n = self.ns
        nel = n * (nzi * n + 2 * (nzi - 1))
        out = np.zeros(nel)
        expeig = np.zeros(nzi * n)
        adi, hl = self._hl(heat_loss)
        cm, cs, cu, mc, nc = [np.ascontiguousarray(a) for a in stencils]
        st = np.ascontiguousarray(state)
        args = [_dp(st), ctypes.c_double(p), _dp(np.ascontiguousarray(oxy)), _dp(np.ascontiguousarray(fuel)), adi,
                *[_dp(a) for a in hl], nzi, _dp(cm), _dp(cs), _dp(cu), _dp(mc), _dp(nc),
                _dp(np.ascontiguousarray(chi))]
        if self.ref:
            self._f('flamelet_jacobian')(self.h, *args, int(compute_eig), ctypes.c_double(diffterm),
                                         int(scale_and_offset), ctypes.c_double(prefactor), int(incl_ef),
                                         int(incl_vcp), int(scaled_hl), _dp(expeig), _dp(out))
        else:
            pre = np.zeros(nzi * n * n) if compute_eig else None
            self._f('flamelet_jacobian')(self.h, *args, int(scale_and_offset), ctypes.c_double(prefactor),
                                         int(incl_ef), int(incl_vcp), int(scaled_hl), _dp(pre), _dp(out))
            self.last_chemistry_blocks = pre
            if compute_eig:
                # eigenvalues of each pre-diffusion chemistry block, LAPACK dgeev via numpy
                # (flamelet_kernels.cpp:1329-1341 calls dgeev on the same block)
                for iz in range(nzi):
                    blk = pre[iz * n * n:(iz + 1) * n * n].reshape((n, n), order='F')
                    lam = np.linalg.eigvals(blk)
                    expeig[iz * n:(iz + 1) * n] = max(0., np.max(np.maximum(lam.real - diffterm, 0.)))
        return out, expeig

    def btddod_full_factorize(self, values, nb, bs):
        d = np.ascontiguousarray(values, dtype=np.float64).copy()
        l = np.zeros(nb * bs * bs)
        piv = np.zeros(nb * bs, dtype=np.int32)
        self._f('btddod_full_factorize')(_dp(d), nb, bs, _dp(l), _ip(piv))
        return d, l, piv

    def btddod_full_solve(self, d, l, piv, rhs, nb, bs):
        x = np.zeros(nb * bs)
        self._f('btddod_full_solve')(_dp(d), _dp(l), _ip(piv), _dp(np.ascontiguousarray(rhs, dtype=np.float64)), nb,
                                     bs, _dp(x))
        return x

    def btddod_scale_and_add_diagonal(self, values, a, diag, b, nb, bs):
        v = np.ascontiguousarray(values, dtype=np.float64).copy()
        self._f('btddod_scale_and_add_diagonal')(_dp(v), ctypes.c_double(a),
                                                 _dp(np.ascontiguousarray(diag, dtype=np.float64)),
                                                 ctypes.c_double(b), nb, bs)
        return v

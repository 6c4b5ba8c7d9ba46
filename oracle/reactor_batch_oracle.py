"""TEST INFRASTRUCTURE ONLY — spitfire_amd.reactor_batch.ReactorBatch with its chemistry and linear algebra on
the CPU oracle (oracle/griffon_oracle.cpp, SciPy LU) instead of the GPU: the checker for the batched reactor
integration and its discrete adjoint. The step loop and the adjoint recursion are the product's; the two fused
Newton-step kernels are restated here elementwise (x -= d where unconverged; the stage residual and its
convergence test).

d(q_r)/d(ln A_r) comes from the oracle's exact per-reaction derivative (gfo_rate_of_progress_log_a_derivative),
itself pinned against central differences of the oracle's production rates over mechanisms with A_r scaled by
exp(+-eps) (tests/test_reactor_adjoint.py); qa='differences' uses those differences directly:
d w_k / d ln A_r = M_k nu_kr dq_r/dln A_r for any species k in reaction r."""
import copy

import numpy as np
from scipy.linalg import lu_factor, lu_solve

from oracle.oracle import CpuGriffon
from spitfire_amd.mechdata import CompiledMechanism
from spitfire_amd.reactor_batch import ReactorBatch


def scaled_mechanism(mech_data, file_index, factor):
    """mech_data with the forward (high-pressure, for falloff) pre-exponential of one reaction scaled"""
    d = copy.deepcopy(mech_data)
    rx = d['reactions'][file_index]
    key = 'A' if 'A' in rx else 'fwd-A'
    rx[key] = rx[key] * factor
    return d


class _CpuLU:
    def __init__(self, R, ns):
        self.R, self.ns = R, ns

    def factor(self, values, stream):
        V = values.numpy().reshape(self.R, self.ns, self.ns)  # [reactor, column, row]
        self.f = [lu_factor(V[r].T) for r in range(self.R)]

    def solve(self, rhs, out, stream):
        import torch
        r = rhs.numpy()
        out.copy_(torch.from_numpy(np.stack([lu_solve(self.f[i], r[i]) for i in range(self.R)])))
        return out


class OracleReactorBatch(ReactorBatch):
    _device = 'cpu'

    def __init__(self, mech_spec, T0, Y0, pressure=101325., qa_eps=1.e-5, qa='exact'):
        self._cpu = CpuGriffon(CompiledMechanism(mech_spec.mech_data, order='name'))
        self._qa_eps = qa_eps
        self._qa_mode = qa
        self._perturbed = None
        self._pressure = float(pressure)
        super().__init__(mech_spec, T0, Y0, pressure)

    def _make_lu(self):
        return _CpuLU(self.R, self.ns)

    def _adjoint_prep(self, x, f, J, aux, cpi, dinvm, v, h, bi, aii, rhs, M):
        """rhs = h (J^T v + b_i grad qdot), grad qdot = rho cp dfT/dx + f_T (cp drho/dx + rho dcp/dx); M = I - h a_ii
        J^T (column-major), elementwise"""
        import torch
        R, ns = self.R, self.ns
        rho, cp, dcp, mmw = aux[:, 0], aux[:, 1], aux[:, 2], aux[:, 3]
        Jc = J.view(R, ns, ns)  # Jc[r, column, row]
        fT = f[:, 0]
        drho = torch.empty((R, ns), dtype=torch.float64)
        drho[:, 0] = -rho / x[:, 0]
        drho[:, 1:] = -(rho * mmw)[:, None] * dinvm[None, :]
        dcpv = torch.empty((R, ns), dtype=torch.float64)
        dcpv[:, 0] = dcp
        dcpv[:, 1:] = cpi[:, :-1] - cpi[:, -1:]
        gq = (rho * cp)[:, None] * Jc[:, :, 0] + fT[:, None] * (cp[:, None] * drho + rho[:, None] * dcpv)
        rhs.copy_(h[:, None] * (torch.bmm(Jc, v[:, :, None])[:, :, 0] + bi * gq))
        if M is not None:
            eye = torch.eye(ns, dtype=torch.float64)
            M.copy_((eye - (h * aii)[:, None, None] * Jc.transpose(1, 2)).reshape(R, ns * ns))

    def _adjoint_finish(self, v, theta, aux, hs, f, h, bi, aii, cm, L):
        """wv = v + a_ii theta; C from wv and the species enthalpies; cm = C M; L += h b_i rho cp f_T"""
        import torch
        rho, cp = aux[:, 0], aux[:, 1]
        wv = v + aii * theta
        wT = wv[:, 0] + bi * rho * cp
        C = torch.zeros((self.R, self.ns), dtype=torch.float64)
        C[:, :-1] = wv[:, 1:] / rho[:, None]
        C = C - (wT / (rho * cp))[:, None] * hs
        cm.copy_(C * self.mw[None, :])
        L += h * bi * rho * cp * f[:, 0]

    def _masked_update(self, x, d, conv):
        import torch
        x.copy_(torch.where(conv[:, None], x, x - d))

    def _stage_residual(self, h, g, f, prior, x, xn, r, conv, test, tol):
        import torch
        r.copy_(h[:, None] * (g * f + prior) - (x - xn))
        if test:
            conv |= torch.max(torch.abs(r * self.w), dim=1).values < tol

    def _soa(self, x):
        return np.ascontiguousarray(x.numpy().T)

    def _rhs(self, x, out):
        import torch
        out.copy_(torch.from_numpy(self._cpu.batch_rhs(self._soa(x), np.full(self.R, self._pressure))))
        return out

    def _jac(self, x, rhs, jac):
        import torch
        r, j = self._cpu.batch_rhs_jac(self._soa(x), np.full(self.R, self._pressure))
        rhs.copy_(torch.from_numpy(r))
        jac.copy_(torch.from_numpy(j))

    def _full_y(self, xr):
        return np.hstack([xr[1:], 1. - np.sum(xr[1:])])

    def _terms(self, x, rhs, jac, aux, cpi, h, qa):
        import torch
        self._jac(x, rhs, jac)
        X = x.numpy()
        Ru = self._mech.mech_data['gas_constant']
        A, C, H = np.zeros((self.R, 4)), np.zeros((self.R, self.ns)), np.zeros((self.R, self.ns))
        for r in range(self.R):
            y = self._full_y(X[r])
            sc, cp_i, h_i, _ = self._cpu.thermo(X[r, 0], y)
            A[r] = [self._pressure * sc[0] / (Ru * X[r, 0]), sc[1], sc[2], sc[0]]
            C[r], H[r] = cp_i, h_i
        aux.copy_(torch.from_numpy(A))
        cpi.copy_(torch.from_numpy(C))
        h.copy_(torch.from_numpy(H))
        if qa is not None:
            qa.copy_(torch.from_numpy(self._qa(X, A[:, 0])))

    def _qa(self, X, rho):
        if self._qa_mode == 'exact':
            return np.stack([self._cpu.rate_of_progress_log_a_derivative(rho[r], X[r, 0], self._full_y(X[r]))
                             for r in range(self.R)])
        if self._perturbed is None:
            md = self._mech.mech_data
            e = self._qa_eps
            self._perturbed = [
                (CpuGriffon(CompiledMechanism(scaled_mechanism(md, int(fi), np.exp(e)), order='name')),
                 CpuGriffon(CompiledMechanism(scaled_mechanism(md, int(fi), np.exp(-e)), order='name')))
                for fi in self.file_order]
        N = self.N.numpy()
        mw = self.mw.numpy()
        out = np.zeros((self.R, self.nr))
        for r in range(self.R):
            y = self._full_y(X[r])
            for g, (plus, minus) in enumerate(self._perturbed):
                d = (plus.production_rates(X[r, 0], rho[r], y) - minus.production_rates(X[r, 0], rho[r], y)) / (
                    2. * self._qa_eps)
                k = int(np.argmax(np.abs(N[:, g])))
                out[r, g] = d[k] / (mw[k] * N[k, g])
        return out

    def heat_release(self):
        import torch
        f = torch.empty((self.R, self.ns), dtype=torch.float64)
        J = torch.empty((self.R, self.ns * self.ns), dtype=torch.float64)
        aux = torch.empty((self.R, 4), dtype=torch.float64)
        cpi = torch.empty((self.R, self.ns), dtype=torch.float64)
        hs = torch.empty((self.R, self.ns), dtype=torch.float64)
        L = torch.zeros(self.R, dtype=torch.float64)
        for h, Xs, clip in self.records:
            for i in range(6):
                self._terms(Xs[i].contiguous(), f, J, aux, cpi, hs, None)
                L += h * self.b[i] * aux[:, 0] * aux[:, 1] * f[:, 0]
        return L


def tangent_linear_heat_release_gradient(rb):
    """dL/dln A [R, nr] (declared reaction order) of a recorded integration by forward sensitivities: the
    tangent-linear model of the same frozen-step map the discrete adjoint differentiates, carried forward
    through every step for all reactions at once (the forward-sensitivity reference for the adjoint).

    Per step (h, stage states X_i, clip mask) and reaction r, with J_i the Jacobian and f_p,i = dq_r/dlnA_r *
    df/dq_r at X_i:
        dK_0 = J_0 dx_n + f_p,0,
        (I - h a_ii J_i) dK_i = J_i (dx_n + h sum_{j<i} a_ij dK_j) + f_p,i,     dX_i = dx_n + h sum_{j<=i} a_ij dK_j,
        dx_{n+1} = dx_n + h sum_i b_i dK_i, zero in clipped components,
        dL += h sum_i b_i (grad(qdot)_i . dX_i + rho cp (f_p,i)_T),
    where f_T = -sum_k h_k w_k / (rho cp), f_Yk = w_k / rho and w_k = M_k sum_r N_kr q_r."""
    import torch
    R, ns, nr = rb.R, rb.ns, rb.nr
    A, b = rb.A, rb.b
    f = torch.empty((R, ns), dtype=torch.float64)
    J = torch.empty((R, ns * ns), dtype=torch.float64)
    aux = torch.empty((R, 4), dtype=torch.float64)
    cpi = torch.empty((R, ns), dtype=torch.float64)
    hs = torch.empty((R, ns), dtype=torch.float64)
    qa = torch.empty((R, nr), dtype=torch.float64)
    N, mw = rb.N, rb.mw
    dinvm = 1. / mw[:-1] - 1. / mw[-1]
    eye = torch.eye(ns, dtype=torch.float64)
    dx = torch.zeros((R, ns, nr), dtype=torch.float64)
    dL = torch.zeros((R, nr), dtype=torch.float64)
    for h, Xs, clip in rb.records:
        dK = []
        for i in range(6):
            x = Xs[i].contiguous()
            rb._terms(x, f, J, aux, cpi, hs, qa)
            rho, cp, dcp, mmw = aux[:, 0], aux[:, 1], aux[:, 2], aux[:, 3]
            Jm = J.view(R, ns, ns).transpose(1, 2)  # [reactor, row, column]
            dfdq = torch.zeros((R, ns, nr), dtype=torch.float64)  # df/dq_r, griffon reaction order
            MN = mw[:, None] * N  # [species, reaction]
            dfdq[:, 1:, :] = MN[None, :-1, :] / rho[:, None, None]
            dfdq[:, 0, :] = -(hs @ MN) / (rho * cp)[:, None]
            fp = dfdq * qa[:, None, :]
            base = dx + h[:, None, None] * sum(A[i, j] * dK[j] for j in range(i)) if i else dx
            rhs = torch.bmm(Jm, base) + fp
            if i == 0:
                dKi = rhs
            else:
                M = eye[None] - (h * A[i, i])[:, None, None] * Jm
                dKi = torch.linalg.solve(M, rhs)
            dXi = base + (h * A[i, i])[:, None, None] * dKi
            drho = torch.empty((R, ns), dtype=torch.float64)
            drho[:, 0] = -rho / x[:, 0]
            drho[:, 1:] = -(rho * mmw)[:, None] * dinvm[None, :]
            dcpv = torch.empty((R, ns), dtype=torch.float64)
            dcpv[:, 0] = dcp
            dcpv[:, 1:] = cpi[:, :-1] - cpi[:, -1:]
            gq = (rho * cp)[:, None] * Jm[:, 0, :] + f[:, 0:1] * (cp[:, None] * drho + rho[:, None] * dcpv)
            dL += (h * b[i])[:, None] * (torch.einsum('rk,rkp->rp', gq, dXi) + (rho * cp)[:, None] * fp[:, 0, :])
            dK.append(dKi)
        dx = dx + h[:, None, None] * sum(b[j] * dK[j] for j in range(6))
        dx = torch.where(clip[:, :, None], torch.zeros_like(dx), dx)
    out = torch.empty_like(dL)
    out[:, torch.as_tensor(rb.file_order)] = dL
    return out

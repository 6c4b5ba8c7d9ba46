"""Many closed adiabatic isobaric reactors integrated together on the GPU, and the discrete adjoint of the
integrated heat release with respect to the pre-exponential factor of every reaction.

Forward: the step loop of HomogeneousReactor.integrate with its defaults (reactors.py:429-678: ESDIRK64,
methods.py:441-612; frozen-Jacobian Newton, nonlinear.py:185-268, tolerance 1e-12 on the scaled residual;
PI control, stepcontrol.py:50-117; a Jacobian setup every step; negative entries clipped after every step),
run for all reactors in lock step. Each reactor keeps its own time, step size, Newton convergence and clipping.
The chemistry (RHS, Jacobian) runs in the batched HIP kernel; every dt*gamma*J - I is inverted by the batched
register-resident Gauss-Jordan kernel (partial pivoting, one block per system) and each Newton solve is a
matrix-vector product with that inverse.

Loss: L = integral over [0, t_end] of qdot = -sum_k h_k w_k (= rho cp dT/dt for a closed adiabatic isobaric
reactor, isobaric_reactor_kernels.cpp:23), discretised by the Runge-Kutta weights of each step, i.e. as an
extra state Q' = qdot integrated by the same method.

Adjoint: with the accepted step sizes frozen and the stage equations taken as solved exactly, the discrete
adjoint of one step k_i = f(X_i), X_i = x_n + dt sum_j a_ij k_j, x_{n+1} = x_n + dt sum_i b_i k_i is
(Hager, Numer. Math. 87, 2000; Sandu, 2006)
    theta_i = dt J_i^T w_i + dt b_i grad(qdot)_i,   w_i = b_i lambda_{n+1} + sum_j a_ji theta_j,
    lambda_n = lambda_{n+1} + sum_i theta_i,
    dL/dp += dt sum_i [ (df/dp)_i^T w_i + b_i (dqdot/dp)_i ],
solved stage by stage from the last (the tableau is lower triangular, so each implicit stage needs one solve
with I - dt a_ii J_i^T). A clipped component has zero derivative through the clip. For p = ln A_r,
df/dp = (df/dq_r) dq_r/dln A_r, with dq_r/dln A_r from the chemistry kernel (the rate of progress itself for
elementary and three-body reactions; for falloff reactions A is the high-pressure factor and the falloff
function's dependence on it is included).

heat_release(log_A, ...) wraps forward integration and adjoint as a torch.autograd.Function of ln A; the stage
states of every step are the checkpoints the backward pass runs over.
"""
import numpy as np

from .time_integration import KennedyCarpenterS6P4Q3, SimpleNewtonSolver


class _BatchedLU:
    """Solves with R independent ns x ns column-major blocks: each block inverted once per factor() by the
    register-resident Gauss-Jordan kernel with partial pivoting (the one-block case of the batched
    block-tridiagonal inverse form), each solve one matrix-vector product with the stored FP64 inverse."""

    def __init__(self, R, ns, device):
        import torch
        from . import _native
        self._lib = _native.lib()
        self._check = _native.check
        self.R, self.ns = R, ns
        i64 = dict(dtype=torch.int64, device=device)
        self.nb = torch.ones(R, dtype=torch.int32, device=device)
        self.voff = torch.arange(R, **i64) * ns * ns
        self.xoff = torch.arange(R, **i64) * ns
        self.inv = torch.zeros(R * ns * ns, dtype=torch.float64, device=device)
        self.values = None

    def factor(self, values, stream):
        """values: contiguous [R, ns*ns] column-major blocks (left unmodified)"""
        self.values = values
        self._check(self._lib.gf_batch_btddod_inverse_factorize(self.R, self.nb.data_ptr(), self.ns,
                                                                values.data_ptr(), self.voff.data_ptr(),
                                                                self.inv.data_ptr(), self.voff.data_ptr(), 64,
                                                                stream))

    def solve(self, rhs, out, stream):
        self._check(self._lib.gf_batch_btddod_inverse_solve(self.R, self.nb.data_ptr(), self.ns,
                                                            self.values.data_ptr(), self.voff.data_ptr(),
                                                            self.inv.data_ptr(), self.voff.data_ptr(), 64,
                                                            rhs.data_ptr(), out.data_ptr(), self.xoff.data_ptr(),
                                                            stream))
        return out


class ReactorBatch:
    """R closed adiabatic isobaric reactors of one mechanism at one pressure, on cuda:0.

    T0: [R] initial temperatures; Y0: [R, ns] initial mass fractions; pressure in Pa."""
    _device = 'cuda:0'

    def __init__(self, mech_spec, T0, Y0, pressure=101325.):
        import torch
        self._mech = mech_spec
        self._g = mech_spec.griffon
        self.dev = torch.device(self._device)
        ns = mech_spec.n_species
        self.ns, self.nr = ns, mech_spec.n_reactions
        T0 = np.asarray(T0, dtype=np.float64)
        Y0 = np.asarray(Y0, dtype=np.float64)
        self.R = T0.size
        x0 = np.hstack([T0[:, None], Y0[:, :-1]])
        self.x0 = torch.tensor(x0, device=self.dev)
        self.p = torch.tensor([float(pressure)], dtype=torch.float64, device=self.dev)
        self.w = torch.ones(ns, dtype=torch.float64, device=self.dev)
        self.w[0] = 1.e-3  # norm weighting 1/variable scale (reactors.py:311-312)
        m = KennedyCarpenterS6P4Q3(SimpleNewtonSolver())
        self.A, self.b, self.bh, self.gamma = m.A, m.b, m.bh, m.gamma
        # net stoichiometry (products - reactants) in griffon reaction order, molecular weights
        names = list(mech_spec.species_names)
        idx = {n: i for i, n in enumerate(names)}
        N = np.zeros((ns, self.nr))
        for g, rx in enumerate(mech_spec._reactions):
            for s, c in rx['reactants'].items():
                N[idx[s], g] -= c
            for s, c in rx['products'].items():
                N[idx[s], g] += c
        self.N = torch.tensor(N, device=self.dev)
        self.mw = torch.tensor(mech_spec.molecular_weights, device=self.dev)
        self.file_order = np.asarray(mech_spec._reaction_file_order)
        self.lu = self._make_lu()
        self.records = None

    def _make_lu(self):
        return _BatchedLU(self.R, self.ns, self.dev)

    # ---------------- kernels ----------------
    def _stream(self):
        import torch
        return torch.cuda.current_stream().cuda_stream if self.dev.type == 'cuda' else None

    def _rhs(self, x, out):
        self._g.batch_reactor_rhs_isobaric(x, self.p, out, layout='aos')
        return out

    def _jac(self, x, rhs, jac):
        self._g.batch_reactor_jac_isobaric(x, self.p, rhs, jac, layout='aos')

    def _adjoint_prep(self, x, f, J, aux, cpi, dinvm, v, h, bi, aii, rhs, M):
        from . import _native
        _native.check(_native.lib().gf_batch_adjoint_stage_prep(
            self.R, self.ns, x.data_ptr(), f.data_ptr(), J.data_ptr(), aux.data_ptr(), cpi.data_ptr(),
            dinvm.data_ptr(), v.data_ptr(), h.data_ptr(), float(bi), float(aii), rhs.data_ptr(),
            None if M is None else M.data_ptr(), self._stream()))

    def _adjoint_finish(self, v, theta, aux, hs, f, h, bi, aii, cm, L):
        from . import _native
        _native.check(_native.lib().gf_batch_adjoint_stage_finish(
            self.R, self.ns, v.data_ptr(), theta.data_ptr(), aux.data_ptr(), hs.data_ptr(), self.mw.data_ptr(),
            f.data_ptr(), h.data_ptr(), float(bi), float(aii), cm.data_ptr(), L.data_ptr(), self._stream()))

    def _masked_update(self, x, d, conv):
        from . import _native
        _native.check(_native.lib().gf_batch_masked_update(self.R, self.ns, x.data_ptr(), d.data_ptr(),
                                                           conv.data_ptr(), self._stream()))

    def _stage_residual(self, h, g, f, prior, x, xn, r, conv, test, tol):
        from . import _native
        _native.check(_native.lib().gf_batch_stage_residual(self.R, self.ns, h.data_ptr(), float(g), f.data_ptr(),
                                                            prior.data_ptr(), x.data_ptr(), xn.data_ptr(),
                                                            self.w.data_ptr(), float(tol), r.data_ptr(),
                                                            conv.data_ptr(), int(test), self._stream()))

    # ---------------- forward ----------------
    def integrate(self, t_end, first_time_step=1.e-6, max_time_step=1.e6, transient_tolerance=1.e-10,
                  nonlinear_solve_tolerance=1.e-12, max_nonlinear_iter=20, dt_schedule=None, record=True,
                  max_steps=100000, log_rate=0, use_graphs=True):
        """Integrate every reactor from its initial state to t_end. With dt_schedule (a [steps, R] array of
        step sizes, 0 for no step) the step sizes are replayed instead of controlled (the frozen-step map that
        the adjoint differentiates). Returns the final states [R, ns].

        With use_graphs, the stage Newton iterations run as a captured HIP graph of two iterations (block solve,
        masked update, chemistry right-hand side, stage residual and convergence test), replayed until every
        reactor has converged: one launch per two iterations instead of a dozen, over the same device buffers.
        The results are bitwise those of the eager path (use_graphs=False)."""
        import torch
        R, ns = self.R, self.ns
        A, b, bh, g = self.A, self.b, self.bh, self.gamma
        w = self.w
        x = self.x0.clone()
        t = torch.zeros(R, dtype=torch.float64, device=self.dev)
        dt = torch.full((R,), float(first_time_step), dtype=torch.float64, device=self.dev)
        f = torch.empty_like(x)
        J = torch.empty((R, ns * ns), dtype=torch.float64, device=self.dev)
        eye = torch.eye(ns, dtype=torch.float64, device=self.dev)
        st = self._stream()
        recs = []
        nsteps = 0
        self.stats = {'steps': 0, 'newton_iterations': 0, 'max_newton_iterations_hit': 0}
        active_its = torch.zeros((), dtype=torch.int64, device=self.dev)
        # the Newton iteration's operands live in fixed buffers (a captured graph reads fixed addresses)
        M = torch.empty((R, ns * ns), dtype=torch.float64, device=self.dev)
        hb = torch.empty(R, dtype=torch.float64, device=self.dev)
        xn = torch.empty_like(x)
        xi = torch.empty_like(x)
        fi = torch.empty_like(x)
        prior = torch.empty_like(x)
        r = torch.empty_like(x)
        d = torch.empty_like(x)
        conv = torch.zeros(R, dtype=torch.bool, device=self.dev)
        tol = nonlinear_solve_tolerance

        its = torch.zeros(R, dtype=torch.int64, device=self.dev)  # iterations each reactor started unconverged

        def newton_iteration():
            # elementwise only (no reductions, which can record memory fills into a graph)
            its.add_(~conv)
            self.lu.solve(r, d, self._stream())
            self._masked_update(xi, d, conv)
            self._rhs(xi, fi)
            self._stage_residual(hb, g, fi, prior, xi, xn, r, conv, True, tol)

        use_graphs = use_graphs and self.dev.type == 'cuda'
        pair = None  # the captured graph of two Newton iterations, made after the first eager pair
        while True:
            if dt_schedule is not None:
                if nsteps >= len(dt_schedule):
                    break
                h = torch.tensor(np.asarray(dt_schedule[nsteps], dtype=np.float64), device=self.dev)
            else:
                active = t < t_end
                if not bool(active.any()) or nsteps >= max_steps:
                    break
                dt = torch.where(t + dt > t_end, t_end - t, dt)
                h = torch.where(active, dt, torch.zeros_like(dt))
            hb.copy_(h)
            xn.copy_(x)
            self._jac(xn, f, J)
            torch.sub(hb[:, None, None] * g * J.view(R, ns, ns), eye, out=M.view(R, ns, ns))
            self.lu.factor(M, st)
            ks = [f]  # f is not written again within the step
            X = [xn.clone()]
            for s in range(1, 6):
                prior.copy_(sum(A[s, j] * ks[j] for j in range(s)))
                if s == 1:  # later stages start from the previous stage's solution, which xi and fi still hold
                    xi.copy_(xn)
                    fi.copy_(f)
                conv.zero_()
                # stage residual r = h (gamma f + prior) - (x - xn), then Newton: x -= solve(r) for the
                # reactors not yet converged, f(x), residual and convergence test (fused kernels)
                self._stage_residual(hb, g, fi, prior, xi, xn, r, conv, False, tol)
                # whether all reactors have converged is read back (a host synchronization) every other
                # iteration: an iteration after all have converged changes nothing, every update being
                # masked, so the results are those of checking every iteration
                it = 0
                while it < max_nonlinear_iter:
                    if max_nonlinear_iter - it >= 2:
                        if pair is not None:
                            pair.replay()
                        else:
                            newton_iteration()
                            newton_iteration()
                            if use_graphs:
                                pair = self._capture(lambda: (newton_iteration(), newton_iteration()))
                        it += 2
                    else:
                        newton_iteration()
                        it += 1
                    if bool(conv.all()):
                        break
                else:
                    self.stats['max_newton_iterations_hit'] += 1
                # an iteration counts when some reactor was still unconverged at its start; a converged reactor
                # stays converged, so that count is the largest per-reactor count (device-side, no synchronization)
                active_its.add_(its.max())
                its.zero_()
                ks.append(fi.clone())
                X.append(xi.clone())
            dx = h[:, None] * sum(b[j] * ks[j] for j in range(6))
            err = torch.max(torch.abs(h[:, None] * sum((b[j] - bh[j]) * ks[j] for j in range(6)) * w), dim=1).values
            ok = torch.isfinite(dx).all(dim=1)
            hacc = torch.where(ok, h, torch.zeros_like(h))
            xnew = xn + torch.where(ok[:, None], dx, torch.zeros_like(dx))
            clip = xnew < 0.
            x = torch.where(clip, torch.zeros_like(xnew), xnew)
            t = t + hacc
            if dt_schedule is None:
                # PI control (stepcontrol.py:84-101; the step count is at least 1 here, so the history ratio is 1)
                ramp = torch.clamp((transient_tolerance / err) ** 0.1333333333, max=1.1)
                dnew = torch.where(err < 1.e-16, dt * 1.1, dt * ramp)
                dnew = torch.clamp(dnew, max=max_time_step)
                dt = torch.where(ok, dnew, dt * 0.8)  # non-finite update: reject, shrink (integrator.py)
            if record:
                Xs = torch.stack([torch.where(ok[:, None], Xi, xn) for Xi in X])
                recs.append((hacc.clone(), Xs, clip))
            nsteps += 1
            if log_rate and nsteps % log_rate == 0:
                import sys
                print('reactor batch step %d: min t %.3e' % (nsteps, float(t.min())), file=sys.stderr, flush=True)
        self.stats['steps'] = nsteps
        self.records = recs if record else None
        self.stats['newton_iterations'] = int(active_its.item())
        self.final_state = x
        self.final_time = t
        return x

    @staticmethod
    def _capture(fn):
        """fn's device work captured as a graph on a side stream (it is not run by the capture)"""
        import torch
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                fn()
        torch.cuda.current_stream().wait_stream(side)
        return graph

    def step_sizes(self):
        """[steps, R] accepted step sizes of the last recorded integration (0: no step)"""
        return np.stack([r[0].cpu().numpy() for r in self.records])

    # ---------------- adjoint ----------------
    def _terms(self, x, rhs, jac, aux, cpi, h, qa):
        self._g.batch_reactor_jac_isobaric_terms(x, self.p, rhs, jac, aux, cpi, h, qa)

    def heat_release(self):
        """L [R] of the recorded integration (the same quadrature the adjoint differentiates)"""
        import torch
        R, ns, nr, dev = self.R, self.ns, self.nr, self.dev
        f = torch.empty((R, ns), dtype=torch.float64, device=dev)
        J = torch.empty((R, ns * ns), dtype=torch.float64, device=dev)
        aux = torch.empty((R, 4), dtype=torch.float64, device=dev)
        cpi = torch.empty((R, ns), dtype=torch.float64, device=dev)
        hs = torch.empty((R, ns), dtype=torch.float64, device=dev)
        qa = torch.empty((R, nr), dtype=torch.float64, device=dev)
        L = torch.zeros(R, dtype=torch.float64, device=dev)
        for h, Xs, clip in self.records:
            for i in range(6):
                self._terms(Xs[i].contiguous(), f, J, aux, cpi, hs, qa)
                L += h * self.b[i] * aux[:, 0] * aux[:, 1] * f[:, 0]
        return L

    def heat_release_gradient(self, log_rate=0):
        """Discrete adjoint over the recorded integration. Returns (L [R], dL/dlnA [R, nr]) with reactions in
        the mechanism's declared (file) order."""
        import torch
        if not self.records:
            raise RuntimeError('integrate(record=True) first')
        R, ns, nr = self.R, self.ns, self.nr
        A, b = self.A, self.b
        dev = self.dev
        f = torch.empty((R, ns), dtype=torch.float64, device=dev)
        J = torch.empty((R, ns * ns), dtype=torch.float64, device=dev)
        aux = torch.empty((R, 4), dtype=torch.float64, device=dev)
        cpi = torch.empty((R, ns), dtype=torch.float64, device=dev)
        hs = torch.empty((R, ns), dtype=torch.float64, device=dev)
        qa = torch.empty((R, nr), dtype=torch.float64, device=dev)
        dinvm = (1. / self.mw[:-1] - 1. / self.mw[-1])
        lam = torch.zeros((R, ns), dtype=torch.float64, device=dev)
        grad = torch.zeros((R, nr), dtype=torch.float64, device=dev)
        L = torch.zeros(R, dtype=torch.float64, device=dev)
        st = self._stream()
        thb = torch.empty((6, R, ns), dtype=torch.float64, device=dev)  # each stage's theta, one step at a time
        Mt = torch.empty((R, ns * ns), dtype=torch.float64, device=dev)
        rhs = torch.empty((R, ns), dtype=torch.float64, device=dev)
        cm = torch.empty((R, ns), dtype=torch.float64, device=dev)
        for n_back, (h, Xs, clip) in enumerate(reversed(self.records)):
            if log_rate and n_back % log_rate == 0:
                import sys
                print('adjoint step %d of %d' % (n_back, len(self.records)), file=sys.stderr, flush=True)
            lam_p = torch.where(clip, torch.zeros_like(lam), lam)
            thetas = [None] * 6
            for i in range(5, -1, -1):
                x = Xs[i].contiguous()
                self._terms(x, f, J, aux, cpi, hs, qa)
                v = b[i] * lam_p
                for j in range(i + 1, 6):
                    if A[j, i] != 0.:
                        v = v + A[j, i] * thetas[j]
                v = v.contiguous()
                # rhs = h (J^T v + b_i grad qdot) and, for an implicit stage, I - h a_ii J^T; the explicit first
                # stage's rhs is its theta
                theta = thb[i]
                self._adjoint_prep(x, f, J, aux, cpi, dinvm, v, h, b[i], A[i, i], rhs if i > 0 else theta,
                                   Mt if i > 0 else None)
                if i > 0:
                    self.lu.factor(Mt, st)
                    self.lu.solve(rhs, theta, st)
                thetas[i] = theta
                self._adjoint_finish(v, theta, aux, hs, f, h, b[i], A[i, i], cm, L)
                grad += h[:, None] * qa * (cm @ self.N)
            lam = lam_p + sum(thetas)
        out = torch.empty_like(grad)
        out[:, torch.as_tensor(self.file_order, device=dev)] = grad
        self.lambda0 = lam
        return L, out


def _pre_exponential_key(rx):
    return 'A' if 'A' in rx else 'fwd-A'  # falloff reactions: the high-pressure (forward) A


def log_pre_exponentials(mech_spec):
    """ln A of every reaction in the mechanism's declared order, in the units of its mechanism data: the A of
    the forward rate, the high-pressure A for falloff reactions (the parameters heat_release_gradient
    differentiates against)."""
    return np.log([float(rx[_pre_exponential_key(rx)]) for rx in mech_spec.mech_data['reactions']])


def with_log_pre_exponentials(mech_spec, log_A):
    """A copy of the mechanism whose reactions (declared order) have the pre-exponential factors exp(log_A)"""
    import copy
    from .mechanism import ChemicalMechanismSpec
    log_A = np.asarray(log_A, dtype=np.float64)
    d = copy.deepcopy(mech_spec.mech_data)
    if log_A.shape != (len(d['reactions']),):
        raise ValueError('log_A needs one value per reaction (%d)' % len(d['reactions']))
    for rx, la in zip(d['reactions'], log_A):
        rx[_pre_exponential_key(rx)] = float(np.exp(la))
    return ChemicalMechanismSpec(d)


def _heat_release_function():
    import torch

    class HeatRelease(torch.autograd.Function):
        """L = integrated heat release of every reactor of a batch, as a differentiable function of ln A.

        forward(log_A, mech_spec, T0, Y0, pressure, t_end, integrate_kwargs) integrates the reactors with the
        mechanism's pre-exponential factors replaced by exp(log_A) (declared order; the mechanism is rebuilt
        only when log_A differs from its own values), recording every stage state, and returns L [R]. backward
        runs the discrete adjoint over the recorded stages (ReactorBatch.heat_release_gradient) and returns
        sum_r dLoss/dL_r dL_r/dlog_A."""

        @staticmethod
        def forward(ctx, log_A, mech_spec, T0, Y0, pressure, t_end, integrate_kwargs):
            la = log_A.detach().cpu().numpy().astype(np.float64)
            own = log_pre_exponentials(mech_spec)
            spec = mech_spec if np.array_equal(la, own) else with_log_pre_exponentials(mech_spec, la)
            rb = ReactorBatch(spec, T0, Y0, pressure)
            rb.integrate(t_end, record=True, **(integrate_kwargs or {}))
            ctx.batch = rb
            ctx.device = log_A.device
            return rb.heat_release().to(log_A.device)

        @staticmethod
        def backward(ctx, grad_L):
            _, G = ctx.batch.heat_release_gradient()
            g = (grad_L.to(G.device, torch.float64)[:, None] * G).sum(dim=0)
            return g.to(ctx.device), None, None, None, None, None, None

    return HeatRelease


_HEAT_RELEASE = None


def heat_release(log_A, mech_spec, T0, Y0, pressure=101325., t_end=2.e-3, **integrate_kwargs):
    """Differentiable integrated heat release [R] of reactors started at (T0 [R], Y0 [R, ns]) with
    pre-exponential factors exp(log_A) (a torch tensor, one per reaction in declared order): the forward
    integration of ReactorBatch, with the discrete adjoint as the backward pass (torch.autograd.Function)."""
    global _HEAT_RELEASE
    if _HEAT_RELEASE is None:
        _HEAT_RELEASE = _heat_release_function()
    return _HEAT_RELEASE.apply(log_A, mech_spec, T0, Y0, pressure, t_end, integrate_kwargs)
